"""The allocation-free two-step 3D pass (bz_yee3d_step2_noalloc) against bz_yee3d_step2 with a caller-owned
staging buffer followed by bz_gather2: identical fields of step n+2, monitor series of both steps and FP64 DFT
sums (bit for bit), on consecutive calls, each later one on a smaller grid and staging box, so the library's
staging array keeps the previous call's larger contents. An oversized box is refused before any launch."""
import ctypes as C

import numpy as np
import pytest

from beamz_amd import _native as N

pytestmark = pytest.mark.gpu

MU0, EPS0 = 1.25663706212e-06, 8.8541878128e-12


def _case(torch, nx, ny, nz, seed, box, nfreq):
    """Random fields and materials on the padded [nz][ny][ld] layout, one box source, and a monitor table over
    every component of every cell of the staging box."""
    dev = torch.device("cuda:0")
    ld = -(-nx // 64) * 64
    g = torch.Generator().manual_seed(seed)
    shape = (nz, ny, ld)

    def field():
        a = torch.randn(shape, generator=g)
        a[..., nx:] = 0
        return a.to(dev)

    src = [field() for _ in range(6)]
    eps = (1 + 11 * torch.rand(shape, generator=g)).to(dev)
    sig = (5e4 * torch.rand(shape, generator=g) * (torch.rand(shape, generator=g) < 0.3)).to(dev)
    bx, by, bz_ = 3, 4, 2
    add = torch.rand((bz_, by, bx), generator=g).to(dev)
    sb = N.BoxSource(2, nx // 3, ny // 2, nz // 2, bx, by, bz_, add.data_ptr(), None, 0.75)
    x0, y0, z0, sx, sy, sz = box
    cc, b, a = np.meshgrid(np.arange(sz), np.arange(sy), np.arange(sx), indexing="ij")
    cc, b, a = cc.ravel(), b.ravel(), a.ravel()
    comp = np.repeat(np.arange(6), len(a)).astype(np.int32)
    stage_off = np.tile((cc * sy + b) * sx + a, 6).astype(np.int64)
    field_off = np.tile(((z0 + cc) * ny + (y0 + b)) * ld + (x0 + a), 6).astype(np.int64)
    npts = len(comp)
    rng = np.random.default_rng(seed)
    ph = [torch.from_numpy(rng.standard_normal((nfreq, 2))).to(dev) for _ in range(2)]
    dt = 0.99 * 25e-9 / (2.99792458e8 * np.sqrt(3))
    return dict(src=src, eps=eps, sig=sig, add=add, box=sb, amp1=0.5, stage_box=box, ld=ld, nx=nx, ny=ny, nz=nz,
                comp=torch.from_numpy(comp).to(dev), stage_off=torch.from_numpy(stage_off).to(dev),
                field_off=torch.from_numpy(field_off).to(dev), npts=npts, nfreq=nfreq, ph=ph, dt=dt)


def _outputs(torch, c):
    dev = c["src"][0].device
    return dict(dst=[torch.full_like(c["src"][0], 7.0) for _ in range(6)],
                s0=torch.zeros(c["npts"], device=dev), s1=torch.zeros(c["npts"], device=dev),
                dft=torch.zeros((c["nfreq"], c["npts"], 2), dtype=torch.float64, device=dev))


def _common(c, o, kseg):
    boxes = (N.BoxSource * 1)(c["box"])
    amp1 = (C.c_float * 1)(c["amp1"])
    return [N.ptrs6([t.data_ptr() for t in c["src"]]), N.ptrs6([t.data_ptr() for t in o["dst"]]),
            c["eps"].data_ptr(), c["sig"].data_ptr(), c["nx"], c["ny"], c["nz"], c["ld"], 25e-9, 25e-9, 25e-9,
            c["dt"], MU0, EPS0, boxes, amp1, 1]


def _with_buffer(torch, c, o, kseg, rows, waves, stream):
    x0, y0, z0, sx, sy, sz = c["stage_box"]
    stage = torch.full((6, sz, sy, sx), -3.0, device=c["src"][0].device)
    sbox = (C.c_int * 6)(*c["stage_box"])
    N.call("bz_yee3d_step2", *_common(c, o, kseg), stage.data_ptr(), sbox, kseg, rows, waves, stream)
    base0 = N.ptrs6([stage[q].data_ptr() for q in range(6)])
    base1 = N.ptrs6([t.data_ptr() for t in o["dst"]])
    N.call("bz_gather2", base0, c["stage_off"].data_ptr(), base1, c["comp"].data_ptr(), c["field_off"].data_ptr(),
           c["npts"], 1, 1, o["s0"].data_ptr(), o["s1"].data_ptr(), c["nfreq"], c["ph"][0].data_ptr(),
           c["ph"][1].data_ptr(), o["dft"].data_ptr(), stream)


def _without_alloc(torch, c, o, kseg, rows, waves, stream):
    sbox = (C.c_int * 6)(*c["stage_box"])
    N.call("bz_yee3d_step2_noalloc", *_common(c, o, kseg), sbox, c["comp"].data_ptr(), c["stage_off"].data_ptr(),
           c["field_off"].data_ptr(), c["npts"], 1, 1, o["s0"].data_ptr(), o["s1"].data_ptr(), c["nfreq"],
           c["ph"][0].data_ptr(), c["ph"][1].data_ptr(), o["dft"].data_ptr(), kseg, rows, waves, stream)


def test_allocation_free_pass_equals_the_buffered_pass(hip):
    torch = hip
    stream = torch.cuda.current_stream().cuda_stream
    cases = [((130, 70, 40), (20, 6, 5, 100, 58, 30), 7, 10, 2, 8),    # 1,044,000 staged floats, near the limit
             ((70, 36, 20), (10, 3, 2, 41, 29, 15), 3, 0, 4, 8),       # smaller grid and box: 107,010
             ((40, 20, 12), (30, 2, 2, 1, 15, 7), 5, 4, 2, 16)]     # 630
    for seed, ((nx, ny, nz), box, nfreq, kseg, rows, waves) in enumerate(cases):
        print(f"no-alloc case {seed}: grid {nx}x{ny}x{nz}, staging box {box}, kseg {kseg}, rows {rows}, "
              f"waves {waves}", flush=True)
        c = _case(torch, nx, ny, nz, 100 + seed, box, nfreq)
        a, b = _outputs(torch, c), _outputs(torch, c)
        _with_buffer(torch, c, a, kseg, rows, waves, stream)
        _without_alloc(torch, c, b, kseg, rows, waves, stream)
        torch.cuda.synchronize()
        for q in range(6):
            assert torch.equal(a["dst"][q], b["dst"][q]), (seed, q)
        for k in ("s0", "s1", "dft"):
            assert torch.equal(a[k], b[k]), (seed, k)
        assert torch.count_nonzero(b["dft"]) > 0 and torch.count_nonzero(b["s0"]) > 0
        assert not torch.equal(b["s0"], b["s1"])


def test_oversized_staging_box_is_refused(hip):
    torch = hip
    c = _case(torch, 300, 200, 40, 9, (0, 0, 0, 300, 200, 3), 1)       # 6 * 180,000 floats > 1,048,576
    o = _outputs(torch, c)
    before = [t.clone() for t in o["dst"]]
    with pytest.raises(ValueError, match="exceeds"):
        _without_alloc(torch, c, o, 0, 2, 8, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert all(torch.equal(x, y) for x, y in zip(before, o["dst"]))
