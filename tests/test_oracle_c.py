"""The oracle's C restatement of the 2D and 3D updates (oracle/yee_ref.c) gives the bits of the
NumPy restatement (oracle/fdtd_ref.py, itself pinned in test_oracle_fdtd.py), on random fields,
materials with loss and anisotropic cells, step after step, and through RefSim with sources
and monitors."""
import numpy as np
import pytest

from oracle import fdtd_ref as R


@pytest.fixture(scope="module", autouse=True)
def _built():
    from oracle.build import build
    build()


def test_c_2d_step_equals_numpy_bit_for_bit():
    rng = np.random.default_rng(21)
    ny, nx = 37, 53
    dx, dy, dt = 5e-8, 6e-8, 1.1e-16
    eps = 1 + 10 * rng.random((ny, nx)); sig = rng.random((ny, nx)) * 5e4
    F = {k: rng.standard_normal(s) for k, s in R.field_shapes(nx, ny).items()}
    G = {k: v.copy() for k, v in F.items()}
    for _ in range(25):
        F["Hx"], F["Hy"] = R.update_h_fields_2d(F["Hx"], F["Hy"], F["Ez"], sig, dx, dy, dt)
        F["Ez"] = R.update_e_field_2d(F["Ez"], F["Hx"], F["Hy"], sig, eps, dx, dy, dt)
        R.step_2d_c(G, eps, sig, dx, dy, dt)
    for k in F:
        assert np.array_equal(F[k], G[k]), k


def test_c_3d_step_equals_numpy_bit_for_bit():
    rng = np.random.default_rng(22)
    nz, ny, nx = 9, 11, 13
    dx, dy, dz, dt = 1e-7, 1.2e-7, 0.9e-7, 1.5e-16
    eps = 1 + 10 * rng.random((nz, ny, nx)); sig = rng.random((nz, ny, nx)) * 5e4
    F = {k: rng.standard_normal(s) for k, s in R.field_shapes(nx, ny, nz).items()}
    G = {k: v.copy() for k, v in F.items()}
    for _ in range(15):
        R.update_fields_3d(F, eps, sig, dx, dy, dz, dt)
        R.step_3d_c(G, eps, sig, dx, dy, dz, dt)
    for k in F:
        assert np.array_equal(F[k], G[k]), k


def test_refsim_native_equals_numpy_with_sources_and_monitors():
    from types import SimpleNamespace
    rng = np.random.default_rng(23)
    d = 5e-8
    t = np.arange(60) * 1e-16
    sig_t = np.sin(2e15 * t)
    for dims in (2, 3):
        shape = (30, 28) if dims == 2 else (14, 16, 18)
        eps = 1 + 3 * rng.random(shape); sig = rng.random(shape) * 1e4
        pos = (9 * d, 8 * d) if dims == 2 else (9 * d, 8 * d, 7 * d)
        src = SimpleNamespace(position=pos + ((0.0,) if dims == 2 else ()), width=d, signal=sig_t)
        mon = SimpleNamespace(start=(3 * d, 2 * d), end=(3 * d, 20 * d)) if dims == 2 else \
            SimpleNamespace(plane_normal="z", plane_position=5 * d, start=(0.0, 0.0, 5 * d), size=(18 * d, 16 * d))
        kw = dict(sources=[src], monitors=[mon])
        if dims == 3:
            kw.update(dz=d, design_dims=(18 * d, 16 * d, 14 * d))
        a = R.RefSim(eps, sig, d, d, 1e-16, len(t), **kw).run()
        b = R.RefSim(eps, sig, d, d, 1e-16, len(t), native=True, **kw).run()
        for k in a.F:
            assert np.array_equal(a.F[k], b.F[k]), (dims, k)
        for k in ("Ez", "power"):
            assert np.array_equal(np.array(a.mon_series[0][k]), np.array(b.mon_series[0][k])), (dims, k)


def test_f32_storage_step_rounds_the_fp64_step_once():
    """step_3d_c_f32store: the FP64 expressions on float32-stored fields. From float32 fields, one step's H
    is float32(the FP64 step's H) exactly; E, which reads the stored (rounded) H, differs from the FP64 step's
    E by that rounding carried through the curl, under 1e-6 of the largest |E| (a wrong term or index would
    differ by order one); over 40 steps the two runs stay close but not equal."""
    rng = np.random.default_rng(24)
    nz, ny, nx = 9, 11, 13
    dx, dy, dz, dt = 1e-7, 1.2e-7, 0.9e-7, 1.5e-16
    eps = 1 + 10 * rng.random((nz, ny, nx)); sig = rng.random((nz, ny, nx)) * 5e4
    F32 = {k: rng.standard_normal(s).astype(np.float32) for k, s in R.field_shapes(nx, ny, nz).items()}
    F64 = {k: v.astype(np.float64) for k, v in F32.items()}
    R.step_3d_c_f32store(F32, eps, sig, dx, dy, dz, dt)
    R.step_3d_c(F64, eps, sig, dx, dy, dz, dt)
    for k in ("Hx", "Hy", "Hz"):
        assert F32[k].dtype == np.float32 and np.array_equal(F32[k], F64[k].astype(np.float32)), k
    for k in ("Ex", "Ey", "Ez"):
        d = np.abs(F32[k].astype(np.float64) - F64[k]).max() / np.abs(F64[k]).max()
        assert 0 < d < 1e-6, (k, d)
    for _ in range(39):
        R.step_3d_c_f32store(F32, eps, sig, dx, dy, dz, dt)
        R.step_3d_c(F64, eps, sig, dx, dy, dz, dt)
    for k in F64:
        e = np.linalg.norm(F32[k] - F64[k]) / np.linalg.norm(F64[k])
        assert 0 < e < 1e-5, (k, e)
