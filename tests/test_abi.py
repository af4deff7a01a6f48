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
