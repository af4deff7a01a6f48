"""1D waveguide mode solvers (beamz/devices/mode.py:9-227).

``solve_modes``: finite-difference Helmholtz operator eps + (Dxf Dxb + Dxb Dxf)/(2 k0^2)
with a cubic stretched-coordinate PML, shift-inverted at n_max^2 (as the
reference). Unlike the reference, whose ARPACK start vector is random so each
mode carries an arbitrary global phase, this version passes a fixed start
vector and rotates every mode so that its largest-magnitude sample is real and
positive, which makes injected profiles reproducible.

``slab_mode_source``: analytic TE modes of a symmetric slab by root bracketing
over 2000 samples and bisection, the reference's method and tolerances.
"""
from __future__ import annotations

import warnings

import numpy as np
import scipy.sparse as sps
import scipy.sparse.linalg as spl

from ..const import LIGHT_SPEED


def compute_derivative_matrices(N, dL, npml=0):
    ones = np.ones(N)
    Dxf = sps.diags([-ones, ones[:-1]], [0, 1], shape=(N, N), format="csr") / dL
    Dxb = sps.diags([ones[:-1] * -1, ones], [-1, 0], shape=(N, N), format="csr") / dL
    if npml > 1:
        s = np.ones(N, dtype=complex)
        depth = np.arange(npml)[::-1] / (npml - 1)          # 1 at the outer edge, 0 at the interface
        s[:npml] = 1.0 / (1.0 + 3.0j * depth ** 3)
        s[N - npml:] = 1.0 / (1.0 + 3.0j * (np.arange(npml) / (npml - 1)) ** 3)
        S, Si = sps.diags(s, 0), sps.diags(1 / s, 0)
        Dxf = Si @ Dxf @ S
        Dxb = Si @ Dxb @ S
    return Dxf, Dxb


def solve_modes(eps, omega, dL, npml=0, m=2):
    k0 = omega / LIGHT_SPEED
    eps = np.asarray(eps, dtype=float).ravel()
    N = eps.size
    Dxf, Dxb = compute_derivative_matrices(N, dL, npml)
    A = sps.diags(eps, 0) + 0.5 * (Dxf @ Dxb + Dxb @ Dxf) * (1 / k0) ** 2
    nmax = np.sqrt(np.max(eps))
    v0 = np.exp(-((np.arange(N) - N / 2) / (N / 4 + 1)) ** 2).astype(complex)
    vals, vecs = spl.eigs(A, k=m, sigma=nmax ** 2, which="LM", v0=v0)
    vecs = vecs / np.sqrt(np.sum(np.abs(vecs) ** 2, axis=0))
    for q in range(vecs.shape[1]):
        a = vecs[np.argmax(np.abs(vecs[:, q])), q]
        vecs[:, q] *= np.conj(a) / abs(a)
    return np.sqrt(vals), vecs


def slab_mode_source(x, w, n_WG, n0, wavelength, ind_m=0, x0=0):
    """Normalised TE mode profile of a symmetric slab and its effective index."""
    k0 = 2 * np.pi / wavelength
    lo, hi = n0 * k0, n_WG * k0

    def split(beta):
        if beta < lo or beta > hi:
            return None
        a = n_WG ** 2 * k0 ** 2 - beta ** 2
        b = beta ** 2 - n0 ** 2 * k0 ** 2
        if a <= 0 or b <= 0:
            return None
        return np.sqrt(a), np.sqrt(b)

    def f_even(beta):
        r = split(beta)
        return None if r is None else r[0] * np.tan(r[0] * w / 2) - r[1]

    def f_odd(beta):
        r = split(beta)
        if r is None:
            return None
        s = np.sin(r[0] * w / 2)
        if abs(s) < 1e-12:
            return None
        return -r[0] * (np.cos(r[0] * w / 2) / s) - r[1]

    def branch(beta):
        a = n_WG ** 2 * k0 ** 2 - beta ** 2
        if a <= 0:
            return None
        th = np.sqrt(a) * w / 2
        return th, int(np.floor(2 * th / np.pi))

    def ok_even(beta):
        r = branch(beta)
        if r is None:
            return False
        th, q = r
        return q % 2 == 0 and not (q == 0 and th > (np.pi / 2 - 0.1))

    def ok_odd(beta):
        r = branch(beta)
        return r is not None and r[1] % 2 == 1

    def bisect(f, a, b):
        mid = 0.5 * (a + b)
        for _ in range(50):
            mid = 0.5 * (a + b)
            fm = f(mid)
            if fm is None:
                b = mid
                continue
            if abs(fm) < 1e-9:
                return mid
            fa = f(a)
            if fa is None or fa * fm > 0:
                a = mid
            else:
                b = mid
        return mid

    N = 2000
    scan = np.linspace(lo, hi, N)
    modes = []
    for f, ok, par in ((f_even, ok_even, "even"), (f_odd, ok_odd, "odd")):
        vals = [f(b) for b in scan]
        for q in range(N - 1):
            if vals[q] is not None and vals[q + 1] is not None and vals[q] * vals[q + 1] < 0:
                root = bisect(f, scan[q], scan[q + 1])
                if ok(root):
                    modes.append((par, root))
    modes.sort(key=lambda t: t[1], reverse=True)
    if not modes:
        raise ValueError("No guided slab modes found in [n0*k0, n_WG*k0].")
    if ind_m >= len(modes):
        warnings.warn(f"Requested mode index {ind_m} >= found modes ({len(modes)}). "
                      f"Using highest mode index {len(modes) - 1}.", UserWarning)
        ind_m = len(modes) - 1
    par, beta = modes[ind_m]
    kx = np.sqrt(n_WG ** 2 * k0 ** 2 - beta ** 2)
    kap = np.sqrt(beta ** 2 - n0 ** 2 * k0 ** 2)
    xp = np.asarray(x, dtype=float) - x0
    inside = np.abs(xp) <= w / 2
    if par == "even":
        E = np.where(inside, np.cos(kx * xp), np.cos(kx * (w / 2)) * np.exp(-kap * (np.abs(xp) - w / 2)))
    else:
        E = np.where(inside, np.sin(kx * xp), np.sign(xp) * np.sin(kx * (w / 2)) * np.exp(-kap * (np.abs(xp) - w / 2)))
    E = E.astype(np.complex128)
    E /= np.sqrt(np.trapezoid(np.abs(E) ** 2, x))
    return E, beta / k0
