"""FP64 restatement of the reference FDTD time step (2D and 3D).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows, op for op and in the same order:
  * 2D H/E updates ........ beamz/simulation/backends/numpy_backend.py:34-88
  * 3D H/E updates ........ beamz/simulation/fdtd.py:200-310
  * step ordering ......... beamz/simulation/fdtd.py:411-429 (H, E, sources,
                            monitors, power, t += dt)
  * field shapes .......... beamz/simulation/fdtd.py:151-186
  * source injection ...... beamz/simulation/helper.py:5-118
  * monitor sampling ...... beamz/simulation/helper.py:231-256,
                            beamz/devices/monitors.py:204-273, 279-397
  * Poynting power ........ beamz/devices/monitors.py:407-439,
                            beamz/simulation/helper.py:120-187
  * (Hz, Ex, Ey) 2D family  no reference 2D counterpart; fdtd.py:200-310 with
                            d/dz = 0 (pinned against the 3D restatement's
                            middle plane of a z-invariant run)
  * DFT accumulation ...... no reference counterpart; defined as the FP64 DFT
                            F_k = sum_n f_n exp(-i w_k t_n) of the very samples
                            the reference's monitor records (same t_n).

Pinned by (tests/test_oracle_*.py):
  * 2D/3D updates: brute-force scalar loops on tiny grids (independent index
    arithmetic), sigma=0 discrete-energy invariant, PEC cavity resonance of the
    Yee scheme (closed form), discrete divergence preservation, linearity, x/y
    mirror symmetry.
  * sources: closed-form Gaussian weights on hand-computed patches.
  * monitors: brute-force indexing; DFT: closed-form DFT of a synthetic
    sinusoid and numpy.fft at bin frequencies.
"""
from __future__ import annotations

import numpy as np

# beamz/const.py:1-5 (non-CODATA values are part of parity)
LIGHT_SPEED = 299792458.0
EPS_0 = 8.85e-12
MU_0 = 1.256e-6


# --------------------------------------------------------------------------
# 2D updates: numpy_backend.py:34-88
# --------------------------------------------------------------------------
def update_h_fields_2d(Hx, Hy, Ez, sigma, dx, dy, dt, mu_0=MU_0, eps_0=EPS_0):
    """numpy_backend.py:34-61 (returns new Hx, Hy)."""
    if np.iscomplexobj(Ez):
        if not np.iscomplexobj(Hx):
            Hx = Hx.astype(np.complex128)
        if not np.iscomplexobj(Hy):
            Hy = Hy.astype(np.complex128)
    sm_x = sigma[:, :-1] * mu_0 / eps_0
    sm_y = sigma[:-1, :] * mu_0 / eps_0
    ce_x = (Ez[:, 1:] - Ez[:, :-1]) / dy          # x-difference over dy (quirk Q2)
    ce_y = (Ez[1:, :] - Ez[:-1, :]) / dx
    den_x = 1.0 + sm_x * dt / (2.0 * mu_0)
    fa_x = (1.0 - sm_x * dt / (2.0 * mu_0)) / den_x
    fb_x = (dt / mu_0) / den_x
    Hx = fa_x * Hx - fb_x * ce_x
    den_y = 1.0 + sm_y * dt / (2.0 * mu_0)
    fa_y = (1.0 - sm_y * dt / (2.0 * mu_0)) / den_y
    fb_y = (dt / mu_0) / den_y
    Hy = fa_y * Hy + fb_y * ce_y
    return Hx, Hy


def update_e_field_2d(Ez, Hx, Hy, sigma, eps_r, dx, dy, dt, eps_0=EPS_0):
    """numpy_backend.py:63-88 (mutates Ez's interior in place and returns it)."""
    cplx = np.iscomplexobj(Ez) or np.iscomplexobj(Hx) or np.iscomplexobj(Hy)
    dtype = np.complex128 if cplx else Ez.dtype
    chx = np.zeros_like(Ez, dtype=dtype)
    chy = np.zeros_like(Ez, dtype=dtype)
    chx[1:-1, 1:-1] = (Hx[1:-1, 1:] - Hx[1:-1, :-1]) / dy
    chy[1:-1, 1:-1] = (Hy[1:, 1:-1] - Hy[:-1, 1:-1]) / dx
    s = sigma[1:-1, 1:-1]
    e = eps_r[1:-1, 1:-1]
    den = 1.0 + s * dt / (2.0 * eps_0 * e)
    f1 = (1.0 - s * dt / (2.0 * eps_0 * e)) / den
    f2 = (dt / (eps_0 * e)) / den
    Ez[1:-1, 1:-1] = f1 * Ez[1:-1, 1:-1] + f2 * (-chx[1:-1, 1:-1] + chy[1:-1, 1:-1])
    return Ez


# --------------------------------------------------------------------------
# 3D update: fdtd.py:200-310 (in place on all six arrays)
# --------------------------------------------------------------------------
def update_fields_3d(F, eps_r, sigma, dx, dy, dz, dt, mu0=MU_0, eps0=EPS_0):
    Ex, Ey, Ez, Hx, Hy, Hz = (F[k] for k in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"))
    sm_hx = sigma[:-1, :-1, :] * mu0 / eps0
    sm_hy = sigma[:-1, :, :-1] * mu0 / eps0
    sm_hz = sigma[:, :-1, :-1] * mu0 / eps0
    cex = (Ez[:, 1:, :] - Ez[:, :-1, :]) / dy - (Ey[1:, :, :] - Ey[:-1, :, :]) / dz
    cey = (Ex[1:, :, :] - Ex[:-1, :, :]) / dz - (Ez[:, :, 1:] - Ez[:, :, :-1]) / dx
    cez = (Ey[:, :, 1:] - Ey[:, :, :-1]) / dx - (Ex[:, 1:, :] - Ex[:, :-1, :]) / dy
    for H, sm, ce in ((Hx, sm_hx, cex), (Hy, sm_hy, cey), (Hz, sm_hz, cez)):
        den = 1.0 + sm * dt / (2.0 * mu0)
        fa = (1.0 - sm * dt / (2.0 * mu0)) / den
        fb = (dt / mu0) / den
        H[:] = fa * H - fb * ce
    e_x, s_x = eps_r[1:-1, 1:-1, :-1], sigma[1:-1, 1:-1, :-1]
    e_y, s_y = eps_r[1:-1, :-1, 1:-1], sigma[1:-1, :-1, 1:-1]
    e_z, s_z = eps_r[:-1, 1:-1, 1:-1], sigma[:-1, 1:-1, 1:-1]
    dhz_dy = (Hz[:, 1:, :] - Hz[:, :-1, :]) / dy
    dhy_dz = (Hy[1:, :, :] - Hy[:-1, :, :]) / dz
    chx = dhz_dy[1:-1, :, :] - dhy_dz[:, 1:-1, :]
    dhx_dz = (Hx[1:, :, :] - Hx[:-1, :, :]) / dz
    dhz_dx = (Hz[:, :, 1:] - Hz[:, :, :-1]) / dx
    chy = dhx_dz[:, :, 1:-1] - dhz_dx[1:-1, :, :]
    dhy_dx = (Hy[:, :, 1:] - Hy[:, :, :-1]) / dx
    dhx_dy = (Hx[:, 1:, :] - Hx[:, :-1, :]) / dy
    chz = dhy_dx[:, 1:-1, :] - dhx_dy[:, :, 1:-1]
    coef = []
    for e, s in ((e_x, s_x), (e_y, s_y), (e_z, s_z)):
        den = 1.0 + s * dt / (2.0 * eps0 * e)
        coef.append(((1.0 - s * dt / (2.0 * eps0 * e)) / den, (dt / (eps0 * e)) / den))
    Ex[1:-1, 1:-1, :] = coef[0][0] * Ex[1:-1, 1:-1, :] + coef[0][1] * chx
    Ey[1:-1, :, 1:-1] = coef[1][0] * Ey[1:-1, :, 1:-1] + coef[1][1] * chy
    Ez[:, 1:-1, 1:-1] = coef[2][0] * Ez[:, 1:-1, 1:-1] + coef[2][1] * chz


# --------------------------------------------------------------------------
# 2D (Hz, Ex, Ey) family: no reference 2D counterpart (the reference's 2D path
# is the (Ez, Hx, Hy) family only). Defined as fdtd.py:200-310 with every
# d/dz = 0, kept to Hz, Ex, Ey: the same expressions with the z-differences
# (exactly 0 for z-invariant fields) left out, which changes no rounding
# (x - 0 = x, 0 - x = -x).
# --------------------------------------------------------------------------
def update_fields_2d_hz(F, eps_r, sigma, dx, dy, dt, mu0=MU_0, eps0=EPS_0):
    Ex, Ey, Hz = F["Ex"], F["Ey"], F["Hz"]
    sm = sigma[:-1, :-1] * mu0 / eps0                          # fdtd.py: sigma[:, :-1, :-1]
    cez = (Ey[:, 1:] - Ey[:, :-1]) / dx - (Ex[1:, :] - Ex[:-1, :]) / dy
    den = 1.0 + sm * dt / (2.0 * mu0)
    fa = (1.0 - sm * dt / (2.0 * mu0)) / den
    fb = (dt / mu0) / den
    Hz[:] = fa * Hz - fb * cez
    e_x, s_x = eps_r[1:-1, :-1], sigma[1:-1, :-1]
    e_y, s_y = eps_r[:-1, 1:-1], sigma[:-1, 1:-1]
    chx = (Hz[1:, :] - Hz[:-1, :]) / dy                      # dHz/dy on Ex's interior rows
    chy = -((Hz[:, 1:] - Hz[:, :-1]) / dx)                   # dHx/dz - dHz/dx with dHx/dz = 0
    coef = []
    for e, s in ((e_x, s_x), (e_y, s_y)):
        den = 1.0 + s * dt / (2.0 * eps0 * e)
        coef.append(((1.0 - s * dt / (2.0 * eps0 * e)) / den, (dt / (eps0 * e)) / den))
    Ex[1:-1, :] = coef[0][0] * Ex[1:-1, :] + coef[0][1] * chx
    Ey[:, 1:-1] = coef[1][0] * Ey[:, 1:-1] + coef[1][1] * chy


# --------------------------------------------------------------------------
# The same 2D/3D updates in C (oracle/yee_ref.c, built by oracle/build.py), for
# long runs: the same expressions in the same order, so the same bits
# (tests/test_oracle_c.py). Real FP64 fields only.
# --------------------------------------------------------------------------
_CLIB = None


def _clib():
    global _CLIB
    if _CLIB is None:
        import ctypes as C
        import os
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "liboref.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run python -m oracle.build")
        lib = C.CDLL(path)
        dp, i32, f64 = C.c_void_p, C.c_int, C.c_double
        lib.oref_step2d.argtypes = [dp] * 5 + [i32] * 2 + [f64] * 5
        lib.oref_step2d.restype = None
        lib.oref_step3d.argtypes = [dp] * 8 + [i32] * 3 + [f64] * 6
        lib.oref_step3d.restype = None
        lib.oref_step3d_f32store.argtypes = [dp] * 8 + [i32] * 3 + [f64] * 6
        lib.oref_step3d_f32store.restype = None
        _CLIB = lib
    return _CLIB


def _ptr(a, shape=None, dtype=np.float64):
    if a.dtype != dtype or not a.flags.c_contiguous or (shape is not None and a.shape != tuple(shape)):
        raise ValueError(f"the C oracle takes C-contiguous {np.dtype(dtype).name} arrays of the reference shapes")
    return a.ctypes.data


def step_2d_c(F, eps_r, sigma, dx, dy, dt, mu_0=MU_0, eps_0=EPS_0):
    """update_h_fields_2d then update_e_field_2d, in place on F (real fields)."""
    ny, nx = eps_r.shape
    shp = field_shapes(nx, ny)
    _clib().oref_step2d(*(_ptr(F[k], shp[k]) for k in ("Ez", "Hx", "Hy")), _ptr(sigma, (ny, nx)),
                        _ptr(eps_r, (ny, nx)), nx, ny, dx, dy, dt, mu_0, eps_0)


def step_3d_c(F, eps_r, sigma, dx, dy, dz, dt, mu0=MU_0, eps0=EPS_0):
    """update_fields_3d, in place on F (real fields)."""
    nz, ny, nx = eps_r.shape
    shp = field_shapes(nx, ny, nz)
    _clib().oref_step3d(*(_ptr(F[k], shp[k]) for k in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")),
                        _ptr(eps_r, (nz, ny, nx)), _ptr(sigma, (nz, ny, nx)), nx, ny, nz, dx, dy, dz, dt, mu0, eps0)


def step_3d_c_f32store(F, eps_r, sigma, dx, dy, dz, dt, mu0=MU_0, eps0=EPS_0):
    """update_fields_3d's FP64 expressions on fields stored in FP32 (float32 arrays in F, each result rounded
    once to float when stored). Not a reference path: it measures the accuracy floor that FP32 storage alone
    sets on a long run, whatever the arithmetic."""
    nz, ny, nx = eps_r.shape
    shp = field_shapes(nx, ny, nz)
    _clib().oref_step3d_f32store(*(_ptr(F[k], shp[k], np.float32) for k in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")),
                                 _ptr(eps_r, (nz, ny, nx)), _ptr(sigma, (nz, ny, nx)), nx, ny, nz, dx, dy, dz, dt,
                                 mu0, eps0)


def field_shapes(nx, ny, nz=None, polarization="Ez"):
    """fdtd.py:151-186: reference array shapes, [z,y,x] / [y,x] order; the 2D (Hz, Ex, Ey) family takes the
    shapes of those components in the 3D layout."""
    if nz is None:
        if polarization == "Hz":
            return {"Hz": (ny - 1, nx - 1), "Ex": (ny, nx - 1), "Ey": (ny - 1, nx)}
        return {"Ez": (ny, nx), "Hx": (ny, nx - 1), "Hy": (ny - 1, nx)}
    return {"Ex": (nz, ny, nx - 1), "Ey": (nz, ny - 1, nx), "Ez": (nz - 1, ny, nx),
            "Hx": (nz - 1, ny - 1, nx), "Hy": (nz - 1, ny, nx - 1), "Hz": (nz, ny - 1, nx - 1)}


# --------------------------------------------------------------------------
# Sources: helper.py:5-118
# --------------------------------------------------------------------------
def gaussian_patch_2d(position, width, dx, dy, nx, ny):
    """helper.py:96-116: (y0, y1, x0, x1, weights) the source adds, before x signal."""
    cxg = position[0] / dx
    cyg = position[1] / dy
    wxg = width / dx
    wyg = width / dy
    wxc = max(1, int(round(3 * wxg)))
    wyc = max(1, int(round(3 * wyg)))
    xc = int(round(cxg))
    yc = int(round(cyg))
    x0 = max(0, xc - wxc); x1 = min(nx, xc + wxc + 1)
    y0 = max(0, yc - wyc); y1 = min(ny, yc + wyc + 1)
    yg, xg = np.meshgrid(np.arange(y0, y1), np.arange(x0, x1), indexing="ij")
    dxs = (xg - cxg) ** 2
    dys = (yg - cyg) ** 2
    eps = 1e-9
    sx = wxg ** 2 + eps
    sy = wyg ** 2 + eps
    w = np.exp(-(dxs / (2 * sx) + dys / (2 * sy))) / 4
    return y0, y1, x0, x1, w


def gaussian_patch_3d(position, width, dx, dy, dz, nx, ny, nz, ez_nz):
    """helper.py:62-95: (z0, z1, y0, y1, x0, x1, weights) added into Ez."""
    cxg, cyg, czg = position[0] / dx, position[1] / dy, position[2] / dz
    wxg, wyg, wzg = width / dx, width / dy, width / dz
    wxc = max(1, int(round(3 * wxg)))
    wyc = max(1, int(round(3 * wyg)))
    wzc = max(1, int(round(3 * wzg)))
    xc, yc, zc = int(round(cxg)), int(round(cyg)), int(round(czg))
    x0 = max(0, xc - wxc); x1 = min(nx, xc + wxc + 1)
    y0 = max(0, yc - wyc); y1 = min(ny, yc + wyc + 1)
    z0 = max(0, zc - wzc); z1 = min(nz, zc + wzc + 1)
    z1 = min(z1, ez_nz)
    zg, yg, xg = np.meshgrid(np.arange(z0, z1), np.arange(y0, y1), np.arange(x0, x1), indexing="ij")
    eps = 1e-9
    ex = -((xg - cxg) ** 2 / (2 * (wxg ** 2 + eps)) + (yg - cyg) ** 2 / (2 * (wyg ** 2 + eps))
           + (zg - czg) ** 2 / (2 * (wzg ** 2 + eps)))
    w = np.exp(ex)
    zoff = max(0, min(ez_nz - 1, z0))
    return zoff, zoff + (z1 - z0), y0, y1, x0, x1, w[: (z1 - z0)]


def mode_points(source, dx, dy, dz, nx, ny, nz, ez_nz, is_3d):
    """helper.py:8-36: list of (z|None, y, x, amplitude) for a ModeSource profile."""
    pts = []
    for point in source.mode_profiles[0]:
        if len(point) == 4:
            amp, xr, yr, zr = point
        else:
            amp, xr, yr = point
            zr = 0
        x = int(round(xr / dx))
        y = int(round(yr / dy))
        if is_3d:
            z = int(round(zr / dz))
            if x < 0 or x >= nx or y < 0 or y >= ny or z < 0 or z >= nz:
                continue
            zt = min(z, ez_nz - 1) if z < ez_nz else ez_nz // 2
        else:
            if x < 0 or x >= nx or y < 0 or y >= ny:
                continue
            zt = None
        pts.append((zt, y, x, amp))
    return pts


# --------------------------------------------------------------------------
# Monitors: monitors.py:204-273
# --------------------------------------------------------------------------
def monitor_points_2d(start, end, dx, dy):
    """monitors.py:204-220."""
    sx = int(round(start[0] / dx)); sy = int(round(start[1] / dy))
    ex = int(round(end[0] / dx)); ey = int(round(end[1] / dy))
    if abs(ex - sx) > abs(ey - sy):
        n = abs(ex - sx) + 1
    else:
        n = abs(ey - sy) + 1
    xi = np.linspace(sx, ex, n, dtype=int)
    yi = np.linspace(sy, ey, n, dtype=int)
    return list(zip(xi, yi))


def monitor_slice_3d(mon, design_w, design_h, design_d, dx, dy, dz):
    """monitors.py:222-273."""
    bnx = max(1, int(round(design_w / dx)))
    bny = max(1, int(round(design_h / dy)))
    bnz = max(1, int(round((design_d or 0.0) / dz)))
    st, sz = mon.start, mon.size
    if mon.plane_normal == "z":
        zi = max(0, min(int(round(mon.plane_position / dz)), bnz - 1))
        xs = max(0, min(int(round(st[0] / dx)), bnx - 1)); xe = max(0, min(int(round((st[0] + sz[0]) / dx)), bnx))
        ys = max(0, min(int(round(st[1] / dy)), bny - 1)); ye = max(0, min(int(round((st[1] + sz[1]) / dy)), bny))
        return slice(ys, ye), slice(xs, xe), zi
    if mon.plane_normal == "y":
        yi = max(0, min(int(round(mon.plane_position / dy)), bny - 1))
        xs = max(0, min(int(round(st[0] / dx)), bnx - 1)); xe = max(0, min(int(round((st[0] + sz[0]) / dx)), bnx))
        zs = max(0, min(int(round(st[2] / dz)), bnz - 1)); ze = max(0, min(int(round((st[2] + sz[1]) / dz)), bnz))
        return yi, slice(xs, xe), slice(zs, ze)
    xi = max(0, min(int(round(mon.plane_position / dx)), bnx - 1))
    ys = max(0, min(int(round(st[1] / dy)), bny - 1)); ye = max(0, min(int(round((st[1] + sz[0]) / dy)), bny))
    zs = max(0, min(int(round(st[2] / dz)), bnz - 1)); ze = max(0, min(int(round((st[2] + sz[1]) / dz)), bnz))
    return slice(ys, ye), xi, slice(zs, ze)


def slice_field_3d(arr, sl, plane_normal):
    """monitors.py:324-360 (x/y-normal planes return the whole volume, quirk Q12)."""
    if isinstance(sl[2], (int, np.integer)):
        zi = min(max(0, sl[2]), arr.shape[0] - 1)
        ys = sl[0].start or 0; ye = sl[0].stop or arr.shape[1]
        xs = sl[1].start or 0; xe = sl[1].stop or arr.shape[2]
        ys = max(0, min(ys, arr.shape[1] - 1)); ye = max(ys, min(ye, arr.shape[1]))
        xs = max(0, min(xs, arr.shape[2] - 1)); xe = max(xs, min(xe, arr.shape[2]))
        return arr[zi, ys:ye, xs:xe].copy()
    return arr.copy()


def true_plane_3d(arr, sl, plane_normal):
    """x- or y-normal plane of one component as a plane (the reference's fallback copies the whole volume,
    monitors.py:356-358; this is the documented fix used for flux ports). sl is get_grid_slice_3d's
    (y_slice, x_index, z_slice) or (y_index, x_slice, z_slice); the fixed index and the ranges are clamped
    to the component's own shape as the z-plane branch does (monitors.py:326-337). Returns [z][y] (x-normal)
    or [z][x] (y-normal)."""
    a0, a1, a2 = arr.shape
    if plane_normal == "x":
        fixed = min(max(0, sl[1]), a2 - 1)
        u_rng, n_u = sl[0], a1
    else:
        fixed = min(max(0, sl[0]), a1 - 1)
        u_rng, n_u = sl[1], a2
    us = u_rng.start or 0; ue = u_rng.stop or n_u
    zs = sl[2].start or 0; ze = sl[2].stop or a0
    us = max(0, min(us, n_u - 1)); ue = max(us, min(ue, n_u))
    zs = max(0, min(zs, a0 - 1)); ze = max(zs, min(ze, a0))
    if plane_normal == "x":
        return arr[zs:ze, us:ue, fixed].copy()
    return arr[zs:ze, fixed, us:ue].copy()


def plane_flux(dft, plane_normal):
    """Net spectral flux through a plane from its DFT fields (no reference counterpart):
    sum over the plane of Re(E x conj(H)) . n, per frequency; dft[c] has shape (nfreq, ...)."""
    a, b = {"x": ("y", "z"), "y": ("z", "x"), "z": ("x", "y")}[plane_normal]
    s = dft["E" + a] * np.conj(dft["H" + b]) - dft["E" + b] * np.conj(dft["H" + a])
    return np.real(s).reshape(s.shape[0], -1).sum(axis=1)


def power_2d(Ez_v, Hx_v, Hy_v):
    """monitors.py:407-417: per-point |S| and the step total."""
    ez, hx, hy = np.array(Ez_v), np.array(Hx_v), np.array(Hy_v)
    sx = -ez * hy
    sy = ez * hx
    pm = np.sqrt(sx ** 2 + sy ** 2)
    return pm, np.sum(pm)


def power_2d_hz(Hz_v, Ex_v, Ey_v):
    """The monitor power of monitors.py:407-417 for the (Hz, Ex, Ey) family: S = E x H = (Ey Hz, -Ex Hz)."""
    hz, ex, ey = np.array(Hz_v), np.array(Ex_v), np.array(Ey_v)
    sx = ey * hz
    sy = -ex * hz
    pm = np.sqrt(sx ** 2 + sy ** 2)
    return pm, np.sum(pm)


def grid_power_2d_hz(Hz, Ex, Ey):
    """helper.py:152-179's Sx^2 + Sy^2 for the (Hz, Ex, Ey) family (real fields), every component
    zero-padded to the (ny, nx) box."""
    ny, nx = Ex.shape[0], Ey.shape[1]
    hz = np.zeros((ny, nx)); ex = np.zeros((ny, nx)); ey = np.zeros((ny, nx))
    hz[:-1, :-1] = np.real(Hz)
    ex[:, :-1] = np.real(Ex)
    ey[:-1, :] = np.real(Ey)
    sx = ey * hz
    sy = -ex * hz
    return sx ** 2 + sy ** 2


def power_3d(Ex, Ey, Ez, Hx, Hy, Hz):
    """monitors.py:426-434."""
    sx = Ey * Hz - Ez * Hy
    sy = Ez * Hx - Ex * Hz
    sz = Ex * Hy - Ey * Hx
    pm = np.sqrt(sx ** 2 + sy ** 2 + sz ** 2)
    return pm, np.sum(pm)


def grid_power_2d(Ez, Hx, Hy):
    """helper.py:152-179 (real fields): Sx^2 + Sy^2 with zero-padded H."""
    ezr = np.real(Ez)
    hxf = np.zeros_like(ezr); hyf = np.zeros_like(ezr)
    hxf[:, :-1] = np.real(Hx)
    hyf[:-1, :] = np.real(Hy)
    if np.iscomplexobj(Ez) or np.iscomplexobj(Hx) or np.iscomplexobj(Hy):
        ezi = np.imag(Ez)
        hxi = np.zeros_like(ezr); hyi = np.zeros_like(ezr)
        hxi[:, :-1] = np.imag(Hx); hyi[:-1, :] = np.imag(Hy)
        sx = -ezr * hyf - ezi * hyi
        sy = ezr * hxf + ezi * hxi
    else:
        sx = -ezr * hyf
        sy = ezr * hxf
    return sx ** 2 + sy ** 2


def grid_power_3d(F):
    """helper.py:124-143: |Re(E x conj H)| on the common corner-aligned box."""
    arrs = [F[k] for k in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")]
    mz = min(a.shape[0] for a in arrs); my = min(a.shape[1] for a in arrs); mx = min(a.shape[2] for a in arrs)
    Ex, Ey, Ez, Hx, Hy, Hz = (a[:mz, :my, :mx] for a in arrs)
    sx = np.real(Ey * np.conj(Hz) - Ez * np.conj(Hy))
    sy = np.real(Ez * np.conj(Hx) - Ex * np.conj(Hz))
    sz = np.real(Ex * np.conj(Hy) - Ey * np.conj(Hx))
    return np.sqrt(sx ** 2 + sy ** 2 + sz ** 2)


def dft_of_series(series, t, omegas):
    """New feature (no reference counterpart): F[k, ...] = sum_n f_n exp(-i w_k t_n), FP64."""
    series = np.asarray(series)
    t = np.asarray(t, dtype=np.float64)
    ph = np.exp(-1j * np.outer(np.asarray(omegas, dtype=np.float64), t))   # (K, N)
    flat = series.reshape(series.shape[0], -1)
    return (ph @ flat).reshape((len(omegas),) + series.shape[1:])


# --------------------------------------------------------------------------
# The driver: fdtd.py:411-429 ordering
# --------------------------------------------------------------------------
class RefSim:
    """Reference-order FDTD on FP64 (or complex128) NumPy arrays.

    eps_r, sigma: rasterised maps [y,x] or [z,y,x]. sources: objects with the
    reference's GaussianSource/ModeSource attributes. monitors: objects with
    the reference's Monitor attributes (2D line or 3D z-plane).
    """

    def __init__(self, eps_r, sigma, dx, dy, dt, num_steps, dz=None, sources=(), monitors=(),
                 dtype=np.float64, accumulate_power=False, design_dims=None, polarization="Ez", native=False,
                 true_planes=False, f32_storage=False):
        """native=True advances the (Ez, Hx, Hy) and 3D updates with the C restatement (same bits, real
        fields only). true_planes=True samples x- and y-normal 3D monitors as planes (true_plane_3d) instead
        of the reference's whole-volume copy. f32_storage=True (3D, native) keeps the fields in float32 and
        advances them with the FP64 expressions (step_3d_c_f32store): the FP32-storage accuracy floor."""
        self.native = bool(native)
        self.true_planes = bool(true_planes)
        self.f32_storage = bool(f32_storage)
        if self.native and dtype != np.float64:
            raise ValueError("the C updates take real FP64 fields")
        if self.f32_storage and not (self.native and eps_r.ndim == 3):
            raise ValueError("f32_storage needs native=True and a 3D grid")
        self.is_3d = eps_r.ndim == 3
        if polarization not in ("Ez", "Hz") or (self.is_3d and polarization != "Ez"):
            raise ValueError("polarization is 'Ez' or (2D only) 'Hz'")
        self.pol = polarization
        self.eps_r = np.ascontiguousarray(eps_r, dtype=np.float64)
        self.sigma = np.ascontiguousarray(sigma, dtype=np.float64)
        self.dx, self.dy, self.dz, self.dt = dx, dy, dz, dt
        self.num_steps = num_steps
        self.sources = list(sources)
        self.monitors = list(monitors)
        self.dtype = dtype
        self.accumulate_power = accumulate_power
        self.power_accumulated = None
        self.power_count = 0
        self.design_dims = design_dims
        if self.is_3d:
            self.nz, self.ny, self.nx = eps_r.shape
            shp = field_shapes(self.nx, self.ny, self.nz)
        else:
            self.ny, self.nx = eps_r.shape
            self.nz = None
            shp = field_shapes(self.nx, self.ny, polarization=polarization)
        # fdtd.py:178,185-186: 2D Ez complex, H real until first upcast (the Hz family: all of dtype)
        self.F = {}
        for k, s in shp.items():
            dt_k = dtype
            if (not self.is_3d) and self.pol == "Ez" and k != "Ez" and dtype == np.complex128:
                dt_k = np.float64
            self.F[k] = np.zeros(s, dtype=np.float32 if self.f32_storage else dt_k)
        self.names = tuple(shp)
        self.t = 0
        self.current_step = 0
        self.mon_series = [dict(t=[], Ez=[], Hx=[], Hy=[], Ex=[], Ey=[], Hz=[], power=[]) for _ in self.monitors]
        self._src_cache = {}

    # fdtd.py:188-198 / 200-310
    def simulate_step(self):
        if self.is_3d and self.f32_storage:
            step_3d_c_f32store(self.F, self.eps_r, self.sigma, self.dx, self.dy, self.dz, self.dt)
        elif self.is_3d and self.native:
            step_3d_c(self.F, self.eps_r, self.sigma, self.dx, self.dy, self.dz, self.dt)
        elif self.is_3d:
            update_fields_3d(self.F, self.eps_r, self.sigma, self.dx, self.dy, self.dz, self.dt)
        elif self.native and self.pol == "Ez":
            step_2d_c(self.F, self.eps_r, self.sigma, self.dx, self.dy, self.dt)
        elif self.pol == "Hz":
            update_fields_2d_hz(self.F, self.eps_r, self.sigma, self.dx, self.dy, self.dt)
        else:
            F = self.F
            F["Hx"], F["Hy"] = update_h_fields_2d(F["Hx"], F["Hy"], F["Ez"], self.sigma, self.dx, self.dy, self.dt)
            F["Ez"] = update_e_field_2d(F["Ez"], F["Hx"], F["Hy"], self.sigma, self.eps_r, self.dx, self.dy, self.dt)

    # helper.py:5-118
    def apply_sources(self):
        """helper.py:5-118 into Ez; the (Hz, Ex, Ey) family adds the same patches and profiles into Hz,
        clipped to Hz's (ny-1, nx-1) cells."""
        if self.pol == "Hz":
            return self._apply_sources_hz()
        Ez = self.F["Ez"]
        n = self.current_step
        for si, src in enumerate(self.sources):
            if hasattr(src, "mode_profiles"):
                mod = src.signal[n]
                key = ("m", si)
                if key not in self._src_cache:
                    self._src_cache[key] = mode_points(src, self.dx, self.dy, self.dz, self.nx, self.ny,
                                                       self.nz, Ez.shape[0] if self.is_3d else None, self.is_3d)
                enforce = getattr(src, "enforce_direction", True)
                for zt, y, x, amp in self._src_cache[key]:
                    val = amp * mod
                    if not np.iscomplexobj(Ez):
                        val = np.real(val)
                    if self.is_3d:
                        Ez[zt, y, x] += val
                        if enforce:
                            d = src.direction
                            if d == "+x" and x > 0: Ez[zt, y, x - 1] = 0
                            elif d == "-x" and x < self.nx - 1: Ez[zt, y, x + 1] = 0
                            elif d == "+y" and y > 0: Ez[zt, y - 1, x] = 0
                            elif d == "-y" and y < self.ny - 1: Ez[zt, y + 1, x] = 0
                            elif d == "+z" and zt > 0: Ez[zt - 1, y, x] = 0
                            elif d == "-z" and zt < Ez.shape[0] - 1: Ez[zt + 1, y, x] = 0
                    else:
                        Ez[y, x] += val
                        if enforce:
                            d = src.direction
                            if d == "+x" and x > 0: Ez[y, x - 1] = 0
                            elif d == "-x" and x < self.nx - 1: Ez[y, x + 1] = 0
                            elif d == "+y" and y > 0: Ez[y - 1, x] = 0
                            elif d == "-y" and y < self.ny - 1: Ez[y + 1, x] = 0
            else:
                mod = src.signal[n]
                key = ("g", si)
                if key not in self._src_cache:
                    if self.is_3d:
                        self._src_cache[key] = gaussian_patch_3d(src.position, src.width, self.dx, self.dy, self.dz,
                                                                 self.nx, self.ny, self.nz, Ez.shape[0])
                    else:
                        self._src_cache[key] = gaussian_patch_2d(src.position, src.width, self.dx, self.dy,
                                                                 self.nx, self.ny)
                if self.is_3d:
                    z0, z1, y0, y1, x0, x1, w = self._src_cache[key]
                    Ez[z0:z1, y0:y1, x0:x1] += w * mod
                else:
                    y0, y1, x0, x1, w = self._src_cache[key]
                    Ez[y0:y1, x0:x1] += w * mod

    def _apply_sources_hz(self):
        Hz = self.F["Hz"]
        hy, hx = Hz.shape
        n = self.current_step
        for si, src in enumerate(self.sources):
            mod = src.signal[n]
            if hasattr(src, "mode_profiles"):
                key = ("m", si)
                if key not in self._src_cache:
                    self._src_cache[key] = mode_points(src, self.dx, self.dy, None, hx, hy, None, None, False)
                enforce = getattr(src, "enforce_direction", True)
                for _, y, x, amp in self._src_cache[key]:
                    val = amp * mod
                    if not np.iscomplexobj(Hz):
                        val = np.real(val)
                    Hz[y, x] += val
                    if enforce:
                        d = src.direction
                        if d == "+x" and x > 0: Hz[y, x - 1] = 0
                        elif d == "-x" and x < hx - 1: Hz[y, x + 1] = 0
                        elif d == "+y" and y > 0: Hz[y - 1, x] = 0
                        elif d == "-y" and y < hy - 1: Hz[y + 1, x] = 0
            else:
                key = ("g", si)
                if key not in self._src_cache:
                    y0, y1, x0, x1, w = gaussian_patch_2d(src.position, src.width, self.dx, self.dy, self.nx, self.ny)
                    y1c, x1c = min(y1, hy), min(x1, hx)
                    self._src_cache[key] = (y0, y1c, x0, x1c, w[: max(0, y1c - y0), : max(0, x1c - x0)])
                y0, y1, x0, x1, w = self._src_cache[key]
                Hz[y0:y1, x0:x1] += w * mod

    # helper.py:231-256 -> monitors.py:279-397
    def record_monitors(self):
        F = self.F
        for mi, mon in enumerate(self.monitors):
            rec = self.mon_series[mi]
            if not self.is_3d:
                pts = monitor_points_2d(mon.start, mon.end, self.dx, self.dy)
                vals = {}
                for name in self.names:
                    a = F[name]
                    row = []
                    for xi, yi in pts:
                        if 0 <= yi < a.shape[0] and 0 <= xi < a.shape[1]:
                            v = a[yi, xi]
                            row.append(complex(v) if np.iscomplexobj(v) else float(v))
                        else:
                            row.append(0.0)
                    vals[name] = row
                    rec[name].append(row)
                rec["t"].append(self.t)
                if self.pol == "Hz":
                    _, tot = power_2d_hz(vals["Hz"], vals["Ex"], vals["Ey"])
                else:
                    _, tot = power_2d(vals["Ez"], vals["Hx"], vals["Hy"])
                rec["power"].append(tot)
            else:
                W, H, D = self.design_dims
                sl = monitor_slice_3d(mon, W, H, D, self.dx, self.dy, self.dz)
                cut = true_plane_3d if (self.true_planes and mon.plane_normal in ("x", "y")) else slice_field_3d
                sls = {k: cut(F[k], sl, mon.plane_normal) for k in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")}
                my = min(s.shape[0] for s in sls.values()); mx = min(s.shape[1] for s in sls.values())
                for k in sls:
                    sls[k] = sls[k][:my, :mx]
                    rec[k].append(sls[k])
                rec["t"].append(self.t)
                _, tot = power_3d(sls["Ex"], sls["Ey"], sls["Ez"], sls["Hx"], sls["Hy"], sls["Hz"])
                rec["power"].append(tot)

    def accumulate(self):
        """helper.py:120-187."""
        if not self.accumulate_power:
            return
        if self.is_3d:
            pm = grid_power_3d(self.F)
        elif self.pol == "Hz":
            pm = grid_power_2d_hz(self.F["Hz"], self.F["Ex"], self.F["Ey"])
        else:
            pm = grid_power_2d(self.F["Ez"], self.F["Hx"], self.F["Hy"])
        if self.power_accumulated is None:
            self.power_accumulated = pm.copy()
        else:
            self.power_accumulated += pm
        self.power_count += 1

    def step(self):
        """fdtd.py:411-429."""
        if self.current_step >= self.num_steps:
            return False
        self.simulate_step()
        self.apply_sources()
        self.record_monitors()
        self.accumulate()
        self.t += self.dt
        self.current_step += 1
        return True

    def run(self, steps=None):
        n = self.num_steps if steps is None else steps
        for _ in range(n):
            if not self.step():
                break
        return self
