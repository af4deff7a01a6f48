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


def test_kstep_owner_tables_give_each_point_one_owner():
    """Host tables of the K-step kernel: every monitor point is listed once, by the tile whose inner cells
    hold it, and its packed (region row, region column, component) decodes back to its grid cell."""
    import numpy as np
    from beamz_amd.simulation.engine import kstep_owner_tables, xcd_order
    rng = np.random.default_rng(5)
    dims = np.array([(150, 90, 192), (37, 260, 64), (301, 33, 320)], np.int64)   # (nx, ny, ld)
    for K, rh in ((8, 64), (1, 32), (15, 32), (31, 64)):
        ix, iy = 64 - 2 * K, rh - 2 * K
        rows = []
        for s, (nx, ny, _) in enumerate(dims):
            gx, gy = -(-nx // ix), -(-ny // iy)
            t = np.stack(np.meshgrid(np.arange(gx), np.arange(gy), indexing="xy"), -1).reshape(-1, 2)
            rows.append(np.concatenate([np.full((len(t), 1), s), t], axis=1))
        tiles = xcd_order(np.concatenate(rows), per_wg=1)
        sim = rng.integers(0, 3, 400)
        jj = rng.integers(0, dims[sim, 1])
        ii = rng.integers(0, dims[sim, 0])
        comp = rng.integers(0, 3, 400)
        pts = np.stack([sim, comp, jj * dims[sim, 2] + ii], axis=1)
        start, ents = kstep_owner_tables(tiles, pts, dims, K, rh)
        assert len(start) == len(tiles) + 1 and start[-1] == len(pts)
        assert sorted(ents[:, 0].tolist()) == list(range(len(pts)))
        for t in range(len(tiles)):
            e = ents[start[t]:start[t + 1]]
            assert np.all(np.diff(e[:, 1] & 3) >= 0)              # sorted by component within a tile
            for p, code in e:
                s, tx, ty = tiles[t]
                row, col, c = code >> 10, (code >> 2) & 255, code & 3
                assert K <= row < rh - K and K <= col < 64 - K     # an inner cell of the region
                assert (sim[p], comp[p]) == (s, c)
                assert (ty * iy - K + row, tx * ix - K + col) == (jj[p], ii[p])


def test_kstep_shapes_are_valid():
    """Every automatic and tunable K-step shape satisfies the kernel's limits (a region of 64 x rows*waves
    cells with inner cells left after K steps on each side), and one-tile-per-workgroup ordering is a
    permutation (no padding is needed with one entry per workgroup)."""
    import numpy as np
    from beamz_amd.simulation.engine import Fused2D, xcd_order
    for K, rows, waves in Fused2D.KCANDIDATES + ((8, 8, 8), (8, 4, 8)):
        assert (rows, waves) in Fused2D.KSHAPES and 2 * K < min(64, rows * waves)
        assert waves * 64 <= 1024
    tiles = np.stack([np.zeros(37, np.int32), np.arange(37), np.arange(37) % 5], axis=1)
    out = xcd_order(tiles, per_wg=1)
    real = out[out[:, 0] >= 0]
    assert len(out) == len(tiles) and sorted(map(tuple, real)) == sorted(map(tuple, tiles))


def test_library_shares_torch_hip_runtime():
    """Loading the library first (a mesh rasterised before any GPU work) must not map a second HIP runtime
    next to torch's: the library's launches would then fail with "no ROCm-capable device"."""
    import subprocess
    import sys
    code = ("import os, sys; sys.path.insert(0, os.getcwd())\n"
            "from beamz_amd import _native\n_native.lib()\nimport torch\n"
            "print(sorted({l.split()[-1] for l in open('/proc/self/maps') if 'amdhip64' in l}))\n")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=root, timeout=300)
    assert out.returncode == 0, out.stderr
    libs = eval(out.stdout.strip().splitlines()[-1])
    assert len(libs) == 1, libs
