"""The FP64 adjoint oracle is pinned by the mathematics: the transpose identity of
the one-step map and central finite differences of the forward objective."""
import numpy as np

from oracle import adjoint_ref as A
from oracle.fdtd_ref import LIGHT_SPEED


def small_problem(seed=0, nx=30, ny=24, nsteps=160, sigma_edge=True):
    rng = np.random.default_rng(seed)
    dx = dy = 60e-9
    dt = 0.99 * dx / (LIGHT_SPEED * np.sqrt(2))
    eps = 2.0 + 8.0 * rng.random((ny, nx))
    sig = np.zeros((ny, nx))
    if sigma_edge:
        sig[:, :4] = sig[:, -4:] = 2e4 * rng.random((ny, 4))
        sig[:3, :] += 1e4
    t = np.arange(nsteps) * dt
    f0 = LIGHT_SPEED / 1.55e-6
    signal = np.exp(-((t - 25 * dt) / (8 * dt)) ** 2) * np.cos(2 * np.pi * f0 * t)
    src = (ny // 2 - 1, ny // 2 + 2, 6, 8, rng.random((3, 2)))
    pts = [(nx - 7, y) for y in range(5, ny - 5)]
    c = rng.standard_normal(len(pts)) + 1j * rng.standard_normal(len(pts))
    return A.Problem2D(eps, sig, dx, dy, dt, nsteps, src, signal, pts, c, 2 * np.pi * f0), rng


def test_step_transpose_identity():
    rng = np.random.default_rng(1)
    ny, nx = 17, 23
    eps = 1.0 + 10 * rng.random((ny, nx))
    sig = 3e4 * rng.random((ny, nx))
    dx, dy, dt = 50e-9, 40e-9, 7e-17
    u = (rng.standard_normal((ny, nx)), rng.standard_normal((ny, nx - 1)) / 377, rng.standard_normal((ny - 1, nx)) / 377)
    v = (rng.standard_normal((ny, nx)), rng.standard_normal((ny, nx - 1)) * 377, rng.standard_normal((ny - 1, nx)) * 377)
    Lu = A.step_2d(*u, eps, sig, dx, dy, dt)
    LTv = A.step_2d_transpose(*v, eps, sig, dx, dy, dt)
    lhs = sum(np.vdot(a, b) for a, b in zip(Lu, v))
    rhs = sum(np.vdot(a, b) for a, b in zip(u, LTv))
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)


def test_adjoint_gradient_matches_central_differences():
    P, rng = small_problem()
    g = P.gradient()
    ny, nx = P.eps_r.shape
    cells = [(int(rng.integers(1, ny - 1)), int(rng.integers(1, nx - 1))) for _ in range(10)]
    cells += [(ny // 2, 10), (ny // 2, nx - 8), (2, 2)]    # next to the source, the monitor, inside sigma
    for (j, i) in cells:
        h = 1e-4 * P.eps_r[j, i]
        ep = P.eps_r.copy(); ep[j, i] += h
        em = P.eps_r.copy(); em[j, i] -= h
        fd = (P.forward(ep) - P.forward(em)) / (2 * h)
        assert abs(fd - g[j, i]) <= 1e-6 * np.abs(g).max() + 1e-5 * abs(fd), (j, i, fd, g[j, i])
    # boundary cells are never updated, so eps there has no effect
    assert np.all(g[0, :] == 0) and np.all(g[:, 0] == 0) and np.all(g[-1, :] == 0) and np.all(g[:, -1] == 0)


def test_multi_frequency_gradient_matches_central_differences():
    P, rng = small_problem(seed=2, nsteps=140)
    f0 = P.omega[0]
    npts = len(P.pts)
    c = rng.standard_normal((3, npts)) + 1j * rng.standard_normal((3, npts))
    M = A.Problem2D(P.eps_r, P.sigma, P.dx, P.dy, P.dt, P.nsteps, P.src, P.signal, P.pts, c,
                    f0 * np.array([0.95, 1.0, 1.05]), fscale=[0.5, 1.0, 2.0])
    g = M.gradient()
    ny, nx = M.eps_r.shape
    for (j, i) in [(ny // 2, 12), (5, nx - 9), (ny - 4, 8)]:
        h = 1e-4 * M.eps_r[j, i]
        ep = M.eps_r.copy(); ep[j, i] += h
        em = M.eps_r.copy(); em[j, i] -= h
        fd = (M.forward(ep) - M.forward(em)) / (2 * h)
        assert abs(fd - g[j, i]) <= 1e-6 * np.abs(g).max() + 1e-5 * abs(fd), (j, i, fd, g[j, i])
