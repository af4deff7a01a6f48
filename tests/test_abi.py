"""The C-ABI library loads without a GPU and exports every declared entry point."""
import os
import re

from beamz_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            text = re.sub(r"/\*.*?\*/", "", open(os.path.join(inc, f)).read(), flags=re.S)
            names |= set(re.findall(r"\b(bz_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = N.lib()
    names = declared()
    assert len(names) >= 10
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_every_declared_symbol_has_ctypes_signature():
    assert declared() <= set(N.SIGNATURES)


def test_errors_are_reported_without_gpu():
    lib = N.lib()
    assert lib.bz_version() >= 100
    import pytest
    with pytest.raises(ValueError, match="bz_raster_2d"):
        N.call("bz_raster_2d", None, 0, None, None, 0, 1.0, 1.0, 0.1, 0, 0, None, None, None)


def test_xcd_order_is_a_padded_permutation():
    """The XCD tile order keeps every tile exactly once, pads with skipped (-1) entries only, and gives each
    of the 8 XCDs (workgroup w -> XCD w % 8, 4 tiles per workgroup) one contiguous run of the input."""
    import numpy as np
    from beamz_amd.simulation.engine import xcd_order
    for n in (1, 3, 4, 31, 32, 33, 1000, 66347):
        tiles = np.stack([np.arange(n) // 7, np.arange(n) % 7, np.arange(n)], 1).astype(np.int32)
        out = xcd_order(tiles)
        assert len(out) % 4 == 0 and len(out) - n < 4
        real = out[out[:, 0] >= 0]
        assert len(real) == n
        assert np.array_equal(np.sort(real[:, 2]), np.arange(n))
        assert np.all(out[out[:, 0] < 0] == [-1, 0, 0])
        per_xcd = {}
        for w in range(len(out) // 4):
            for e in out[4 * w: 4 * w + 4]:
                if e[0] >= 0:
                    per_xcd.setdefault(w % 8, []).append(int(e[2]))
        for ids in per_xcd.values():
            assert ids == list(range(ids[0], ids[0] + len(ids)))      # contiguous, in order


def test_descriptor_struct_layouts_match_the_library():
    """The ctypes descriptors must have the sizes the C structs were compiled with (a field added on one side
    only would shift every later field the kernels read)."""
    import ctypes as C
    out = (C.c_int * 3)()
    N.call("bz_abi_sizes", out, 3)
    assert list(out) == [C.sizeof(N.Sim2D), C.sizeof(N.Adj2D), C.sizeof(N.BoxSource)]
