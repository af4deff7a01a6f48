"""FP64 discrete adjoint of the 2D step: the checker of the HIP reverse-time kernel.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Forward map: the reference 2D step on real FP64 arrays in the reference's own
shapes (numpy_backend.py:34-88 through ``fdtd_ref.update_h_fields_2d`` /
``update_e_field_2d``), followed by an additive Ez patch source
(helper.py:96-118). Objective (no reference counterpart; the reference's own
gradient is the overlap heuristic of examples/4_topo.py:282-285, restated in
``topo_ref.overlap_gradient``):

    a_f = sum_n sum_p c_fp Ez_{n+1}(p) exp(-i w_f t_n),    J = sum_f s_f |a_f|^2

i.e. the DFT at w of the Ez samples a monitor records after each step,
projected on complex weights c_p (a mode profile). ``Problem2D.gradient``
returns dJ/d eps_r for every cell by the exact transpose of the step, run
backwards in time over stored forward states:

    lamE'' = lamE_{n+1}                          (additive source)
    w      = cb * lamE''  on the interior, 0 elsewhere
    muX    = lamHx_{n+1} + (w[:, 1:] - w[:, :-1]) / dy
    muY    = lamHy_{n+1} + (w[:-1, :] - w[1:, :]) / dx
    lamE_n = (interior ? ca : 1) * lamE'' + transpose of the H update's
             Ez differences (fb / dy on Hx, fb / dx on Hy)
    lamHx_n = fa * muX, lamHy_n = fa * muY
    dJ/deps += lamE'' * (dca/deps * Ez_n + dcb/deps * curlH_{n+1})  (interior)
    lamE_n(p) += Re(2 conj(a) c_p exp(-i w t_{n-1}))

with ca = (1-s)/(1+s), cb = dt/(eps0 eps)/(1+s), s = sigma dt/(2 eps0 eps):
dca/deps = 2 s / (eps (1+s)^2), dcb/deps = -cb / (eps (1+s)).

Pinned by (tests/test_oracle_adjoint.py): central finite differences of the
FP64 forward J(eps) at sampled cells (the gradient of the discrete map, which
no retyping of these formulas can fake), and the transpose identity
<L u, v> = <u, L^T v> of the one-step linear map on random states.
"""
from __future__ import annotations

import numpy as np

from .fdtd_ref import EPS_0, MU_0, update_e_field_2d, update_h_fields_2d


def coefficients_2d(eps_r, sigma, dt, mu_0=MU_0, eps_0=EPS_0):
    """Per-cell update coefficients on the full grid (same formulas as numpy_backend.py:45-58, 69-83)."""
    sm = sigma * mu_0 / eps_0
    den_h = 1.0 + sm * dt / (2.0 * mu_0)
    fa = (1.0 - sm * dt / (2.0 * mu_0)) / den_h
    fb = (dt / mu_0) / den_h
    s = sigma * dt / (2.0 * eps_0 * eps_r)
    ca = (1.0 - s) / (1.0 + s)
    cb = (dt / (eps_0 * eps_r)) / (1.0 + s)
    dca = 2.0 * s / (eps_r * (1.0 + s) ** 2)
    dcb = -cb / (eps_r * (1.0 + s))
    return fa, fb, ca, cb, dca, dcb


def step_2d(E, Hx, Hy, eps_r, sigma, dx, dy, dt):
    """One source-free reference step on copies: returns (E', Hx', Hy')."""
    Hx, Hy = update_h_fields_2d(Hx.copy(), Hy.copy(), E, sigma, dx, dy, dt)
    E = update_e_field_2d(E.copy(), Hx, Hy, sigma, eps_r, dx, dy, dt)
    return E, Hx, Hy


def step_2d_transpose(lE, lHx, lHy, eps_r, sigma, dx, dy, dt):
    """Transpose of the one-step linear map (E, Hx, Hy) -> (E', Hx', Hy') applied to (lE, lHx, lHy)."""
    fa, fb, ca, cb, _, _ = coefficients_2d(eps_r, sigma, dt)
    w = np.zeros_like(lE)
    w[1:-1, 1:-1] = cb[1:-1, 1:-1] * lE[1:-1, 1:-1]
    mux = lHx + (w[:, 1:] - w[:, :-1]) / dy
    muy = lHy + (w[:-1, :] - w[1:, :]) / dx
    out = lE.copy()
    out[1:-1, 1:-1] = ca[1:-1, 1:-1] * lE[1:-1, 1:-1]
    fbx, fby = fb[:, :-1], fb[:-1, :]
    out[:, :-1] += fbx * mux / dy
    out[:, 1:] -= fbx * mux / dy
    out[:-1, :] -= fby * muy / dx
    out[1:, :] += fby * muy / dx
    return out, fa[:, :-1] * mux, fa[:-1, :] * muy


class Problem2D:
    """Forward run, objective and exact adjoint gradient on a small 2D grid (FP64, real fields).

    src: (y0, y1, x0, x1, weights) additive Ez patch; signal[n] its amplitude at step n.
    pts: list of (x, y) Ez sample cells; c: complex weights, (npts,) or (nf, npts); omega: angular
    frequency, scalar or (nf,); fscale: s_f of J = sum_f s_f |a_f|^2 (default 1).
    """

    def __init__(self, eps_r, sigma, dx, dy, dt, nsteps, src, signal, pts, c, omega, fscale=None):
        self.eps_r = np.asarray(eps_r, dtype=np.float64)
        self.sigma = np.asarray(sigma, dtype=np.float64)
        self.dx, self.dy, self.dt, self.nsteps = dx, dy, dt, int(nsteps)
        self.src, self.signal = src, np.asarray(signal, dtype=np.float64)
        self.pts = [(int(x), int(y)) for x, y in pts]
        self.c = np.atleast_2d(np.asarray(c, dtype=np.complex128))
        self.omega = np.atleast_1d(np.asarray(omega, dtype=np.float64))
        self.fscale = np.ones(len(self.omega)) if fscale is None else np.asarray(fscale, dtype=np.float64)
        t = 0.0
        ts = []
        for _ in range(self.nsteps):      # fdtd.py accumulates t += dt
            ts.append(t)
            t += dt
        self.phasor = np.exp(-1j * np.outer(np.asarray(ts), self.omega))     # (N, nf)

    def forward(self, eps_r=None, keep=False):
        eps = self.eps_r if eps_r is None else np.asarray(eps_r, dtype=np.float64)
        ny, nx = eps.shape
        E = np.zeros((ny, nx)); Hx = np.zeros((ny, nx - 1)); Hy = np.zeros((ny - 1, nx))
        y0, y1, x0, x1, wsrc = self.src
        a = np.zeros(len(self.omega), dtype=np.complex128)
        xs = np.array([p[0] for p in self.pts]); ys = np.array([p[1] for p in self.pts])
        states = [(E.copy(), Hx.copy(), Hy.copy())] if keep else None
        for n in range(self.nsteps):
            E, Hx, Hy = step_2d(E, Hx, Hy, eps, self.sigma, self.dx, self.dy, self.dt)
            E[y0:y1, x0:x1] += wsrc * self.signal[n]
            a += self.phasor[n] * (self.c @ E[ys, xs])
            if keep:
                states.append((E.copy(), Hx.copy(), Hy.copy()))
        self.a = a
        self.states = states
        return float((self.fscale * np.abs(a) ** 2).sum())

    def gradient(self):
        """dJ/d eps_r on the full grid (zero outside the interior)."""
        J = self.forward(keep=True)
        S = self.states
        fa, fb, ca, cb, dca, dcb = coefficients_2d(self.eps_r, self.sigma, self.dt)
        ny, nx = self.eps_r.shape
        wgt = 2.0 * (self.fscale * np.conj(self.a))[:, None] * self.c       # (nf, npts)
        lE = np.zeros((ny, nx)); lHx = np.zeros((ny, nx - 1)); lHy = np.zeros((ny - 1, nx))
        N = self.nsteps

        def inject(lE, n):          # the sample recorded after step n-1 is E_n
            v = np.real(self.phasor[n - 1] @ wgt)                           # (npts,)
            for q, (x, y) in enumerate(self.pts):
                lE[y, x] += v[q]

        inject(lE, N)
        g = np.zeros((ny, nx))
        for n in range(N - 1, -1, -1):
            En = S[n][0]
            _, Hx1, Hy1 = S[n + 1]
            curl = np.zeros((ny, nx))
            curl[1:-1, 1:-1] = (-(Hx1[1:-1, 1:] - Hx1[1:-1, :-1]) / self.dy
                                + (Hy1[1:, 1:-1] - Hy1[:-1, 1:-1]) / self.dx)
            g[1:-1, 1:-1] += lE[1:-1, 1:-1] * (dca[1:-1, 1:-1] * En[1:-1, 1:-1] + dcb[1:-1, 1:-1] * curl[1:-1, 1:-1])
            lE, lHx, lHy = step_2d_transpose(lE, lHx, lHy, self.eps_r, self.sigma, self.dx, self.dy, self.dt)
            if n >= 1:
                inject(lE, n)
        self.J = J
        return g
