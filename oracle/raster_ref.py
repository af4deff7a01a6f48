"""FP64 restatement of the reference rasterizer and PML conductivity.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows, with the same floating-point operation order:
  * RegularGrid.__rasterize__ ........ beamz/simulation/meshing.py:92-536
  * RegularGrid3D.__rasterize_3d__ ... beamz/simulation/meshing.py:616-755
  * RegularGrid3D._process_3d_pml .... beamz/simulation/meshing.py:757-781
  * PML.get_profile/get_conductivity . beamz/design/pml.py:21-62
  * Polygon point-in-polygon ......... beamz/design/structures.py:178-207

It reads duck-typed structures exposing the reference attribute names
(``vertices``, ``interiors``, ``position``, ``width``, ``height``, ``radius``,
``inner_radius``, ``outer_radius``, ``material``, ``is_pml``) and boundaries
exposing the reference PML attributes.

Pinned by (tests/test_oracle_raster.py): closed-form maps of axis-aligned
rectangles whose edges fall on sample lines (blend = k/9 exactly), the
closed-form PML profile at hand-computed distances, brute-force 9-sample
coverage of circles/rings/polygons, and numpy.linspace for cell centres.
"""
from __future__ import annotations

import math

import numpy as np

from .fdtd_ref import EPS_0, MU_0


# structures.py:178-196 (even-odd crossing test, x/y only)
def point_in_path(x, y, path):
    if not path:
        return False
    pts = [(v[0], v[1]) for v in path]
    n = len(pts)
    inside = False
    p1x, p1y = pts[0]
    for i in range(n + 1):
        p2x, p2y = pts[i % n]
        if y > min(p1y, p2y):
            if y <= max(p1y, p2y):
                if x <= max(p1x, p2x):
                    if p1y != p2y:
                        xin = (y - p1y) * (p2x - p1x) / (p2y - p1y) + p1x
                    else:
                        xin = p1x
                    if p1x == p2x or x <= xin:
                        inside = not inside
        p1x, p1y = p2x, p2y
    return inside


# structures.py:198-207
def point_in_polygon(s, x, y):
    if not s.vertices:
        return False
    if not point_in_path(x, y, s.vertices):
        return False
    for hole in s.interiors:
        if hole and point_in_path(x, y, hole):
            return False
    return True


# pml.py:21-27
def pml_profile(b, nd):
    d = min(max(nd, 0.0), 1.0)
    if d < 0.05:
        sig = 0.01 * (d / 0.05) ** 2
    else:
        sig = ((d - 0.05) / 0.95) ** b.polynomial_order
    alpha = b.alpha_max * d ** 2
    return sig, alpha


# pml.py:29-62
def pml_conductivity(b, x, y, dx=None, dt=None, eps_avg=None):
    if dx is not None and eps_avg is not None:
        eta = np.sqrt(MU_0 / (EPS_0 * eps_avg))
        smax = 1.2 / (eta * dx)
        smax *= b.sigma_factor
    else:
        smax = 1.0
    if b.region_type == "rect":
        if not (b.position[0] <= x <= b.position[0] + b.width and b.position[1] <= y <= b.position[1] + b.height):
            return 0.0
        o = b.orientation
        if o == "left":
            dist = 1.0 - (x - b.position[0]) / b.width
        elif o == "right":
            dist = (x - b.position[0]) / b.width
        elif o == "top":
            dist = (y - b.position[1]) / b.height
        elif o == "bottom":
            dist = 1.0 - (y - b.position[1]) / b.height
        else:
            return 0.0
    else:
        r = np.hypot(x - b.position[0], y - b.position[1])
        if r > b.radius:
            return 0.0
        ddx = x - b.position[0]
        ddy = y - b.position[1]
        o = b.orientation
        if o == "top-left" and (ddx > 0 or ddy < 0): return 0.0
        elif o == "top-right" and (ddx < 0 or ddy < 0): return 0.0
        elif o == "bottom-left" and (ddx > 0 or ddy > 0): return 0.0
        elif o == "bottom-right" and (ddx < 0 or ddy > 0): return 0.0
        dist = r / b.radius
    sp, ap = pml_profile(b, dist)
    c = smax * sp
    if dt is not None:
        c *= 1.0 / (1.0 + ap)
    return c


def _props(m, x=0, y=0, z=0):
    """meshing.py:21-51 (_get_material_properties_safe): a material with getters is evaluated at (x, y, z),
    .item() taken where present; any failure gives the defaults."""
    if m is None:
        return 1.0, 1.0, 0.0
    if hasattr(m, "get_permittivity"):
        try:
            p = m.get_permittivity(x, y, z)
            b = m.get_permeability(x, y, z)
            c = m.get_conductivity(x, y, z)
            if hasattr(p, "item"):
                p = p.item()
            if hasattr(b, "item"):
                b = b.item()
            if hasattr(c, "item"):
                c = c.item()
            return p, b, c
        except Exception:
            return (getattr(m, "default_permittivity", 1.0), getattr(m, "default_permeability", 1.0),
                    getattr(m, "default_conductivity", 0.0))
    return m.permittivity, m.permeability, m.conductivity


def _mat(m):
    """Background (meshing.py:131-136) and 3D (:674) reads: _props at (0, 0, 0)."""
    return _props(m)


def _is_rectangle(s):
    return any(c.__name__ == "Rectangle" for c in type(s).__mro__)


def _is_device(s):
    return any(c.__name__ in ("ModeSource", "GaussianSource", "Monitor") for c in type(s).__mro__)


def rasterize_2d(design, resolution):
    """meshing.py:72-75, 92-536. Returns (permittivity, permeability, conductivity)."""
    W, H = design.width, design.height
    gw, gh = int(W / resolution), int(H / resolution)
    cs = resolution
    xc = np.linspace(0.5 * cs, W - 0.5 * cs, gw)
    yc = np.linspace(0.5 * cs, H - 0.5 * cs, gh)
    offs = np.array([-0.25, 0, 0.25]) * cs
    odx, ody = np.meshgrid(offs, offs)
    odx = odx.flatten(); ody = ody.flatten()
    ns = len(odx)
    c = 3e8
    dt_est = 0.5 * resolution / (c * np.sqrt(2))
    P = np.ones((gh, gw)); M = np.ones((gh, gw)); C = np.zeros((gh, gw))
    if design.structures:
        bg = design.structures[0]
        if getattr(bg, "material", None) is not None:
            a, b, d = _mat(bg.material)
            P.fill(a); M.fill(b); C.fill(d)

    def blend(i, j, cnt, mp, mb, mc):
        bf = cnt / ns
        P[i, j] = P[i, j] * (1 - bf) + mp * bf
        M[i, j] = M[i, j] * (1 - bf) + mb * bf
        C[i, j] = C[i, j] * (1 - bf) + mc * bf

    for s in design.structures[1:]:
        if getattr(s, "is_pml", False):
            continue
        if _is_device(s):
            continue
        if getattr(s, "material", None) is None:
            continue
        custom = hasattr(s.material, "get_permittivity")      # meshing.py:162-169
        mp, mb, mc = (None, None, None) if custom else _mat(s.material)
        bbox = s.get_bounding_box()
        if len(bbox) == 6:
            minx, miny, _, maxx, maxy, _ = bbox
        else:
            minx, miny, maxx, maxy = bbox
        mi = max(0, int(miny / cs) - 1)
        mj = max(0, int(minx / cs) - 1)
        Mi = min(gh, int(np.ceil(maxy / cs)) + 1)
        Mj = min(gw, int(np.ceil(maxx / cs)) + 1)
        if mi >= gh or mj >= gw or Mi <= 0 or Mj <= 0:
            continue
        if _is_rectangle(s) and all(v == 0 for v in [s.vertices[0][0] - s.position[0], s.vertices[0][1] - s.position[1]]):
            px, py = s.position[0], s.position[1]
            r0j = max(0, int(px / cs)); r0i = max(0, int(py / cs))
            r1j = min(gw, int(np.ceil((px + s.width) / cs))); r1i = min(gh, int(np.ceil((py + s.height) / cs)))
            n0j = max(0, int((px + 0.25 * cs) / cs)); n0i = max(0, int((py + 0.25 * cs) / cs))
            n1j = min(gw, int(np.floor((px + s.width - 0.25 * cs) / cs)))
            n1i = min(gh, int(np.floor((py + s.height - 0.25 * cs) / cs)))
            if n1i > n0i and n1j > n0j:
                if custom:                                       # meshing.py:209-217
                    for i in range(n0i, n1i):
                        for j in range(n0j, n1j):
                            P[i, j], M[i, j], C[i, j] = _props(s.material, xc[j], yc[i])
                else:
                    P[n0i:n1i, n0j:n1j] = mp; M[n0i:n1i, n0j:n1j] = mb; C[n0i:n1i, n0j:n1j] = mc
            bm = np.zeros((r1i - r0i, r1j - r0j), dtype=bool)
            if r0i < n0i: bm[:n0i - r0i, :] = True
            if n1i < r1i: bm[n1i - r0i:, :] = True
            if r0j < n0j: bm[:, :n0j - r0j] = True
            if n1j < r1j: bm[:, n1j - r0j:] = True
            bi, bj = np.where(bm)
            for q in range(len(bi)):
                i, j = bi[q] + r0i, bj[q] + r0j
                cx, cy = xc[j], yc[i]
                cnt = 0
                for k in range(ns):
                    xs = cx + odx[k]; ys = cy + ody[k]
                    if px <= xs < px + s.width and py <= ys < py + s.height:
                        cnt += 1
                if cnt > 0:
                    if custom:                                   # meshing.py:243-247
                        mp, mb, mc = _props(s.material, xc[j], yc[i])
                    blend(i, j, cnt, mp, mb, mc)
        elif custom:
            # the reference writes None (NaN) into covered cells here and stops at its first blend
            raise NotImplementedError("a spatially varying material is restated on axis-aligned rectangles only")
        elif hasattr(s, "radius"):
            cx0, cy0 = s.position[0], s.position[1]
            R = s.radius
            X, Y = np.meshgrid(xc[np.arange(mj, Mj)], yc[np.arange(mi, Mi)])
            D = np.sqrt((X - cx0) ** 2 + (Y - cy0) ** 2)
            full = D + 0.3536 * cs <= R
            bnd = (D - 0.3536 * cs <= R) & ~full
            li, lj = np.where(full)
            if len(li) > 0:
                P[li + mi, lj + mj] = mp; M[li + mi, lj + mj] = mb; C[li + mi, lj + mj] = mc
            bi, bj = np.where(bnd)
            for q in range(len(bi)):
                i, j = bi[q] + mi, bj[q] + mj
                cnt = 0
                for k in range(ns):
                    if np.hypot(xc[j] + odx[k] - cx0, yc[i] + ody[k] - cy0) <= R:
                        cnt += 1
                if cnt > 0:
                    blend(i, j, cnt, mp, mb, mc)
        elif hasattr(s, "inner_radius") and hasattr(s, "outer_radius"):
            cx0, cy0 = s.position[0], s.position[1]
            ri, ro = s.inner_radius, s.outer_radius
            X, Y = np.meshgrid(xc[np.arange(mj, Mj)], yc[np.arange(mi, Mi)])
            D = np.sqrt((X - cx0) ** 2 + (Y - cy0) ** 2)
            full = (D - 0.3536 * cs >= ri) & (D + 0.3536 * cs <= ro)
            ib = (D - 0.3536 * cs <= ri) & (D + 0.3536 * cs >= ri)
            ob = (D - 0.3536 * cs <= ro) & (D + 0.3536 * cs >= ro)
            bnd = ib | ob
            li, lj = np.where(full)
            if len(li) > 0:
                P[li + mi, lj + mj] = mp; M[li + mi, lj + mj] = mb; C[li + mi, lj + mj] = mc
            bi, bj = np.where(bnd)
            for q in range(len(bi)):
                i, j = bi[q] + mi, bj[q] + mj
                cnt = 0
                for k in range(ns):
                    d = np.hypot(xc[j] + odx[k] - cx0, yc[i] + ody[k] - cy0)
                    if ri <= d <= ro:
                        cnt += 1
                if cnt > 0:
                    blend(i, j, cnt, mp, mb, mc)
        else:
            if not hasattr(s, "vertices") or len(s.vertices) == 0:
                raise NotImplementedError("only polygon-like structures are restated")
            inside = np.zeros((Mi - mi, Mj - mj), dtype=bool)
            bmask = np.zeros((Mi - mi, Mj - mj), dtype=bool)
            corners = [(-0.4, -0.4), (-0.4, 0.4), (0.4, -0.4), (0.4, 0.4)]
            for ir in range(Mi - mi):
                for jr in range(Mj - mj):
                    cx, cy = xc[jr + mj], yc[ir + mi]
                    npin = 0
                    cin = False
                    if point_in_polygon(s, cx, cy):
                        cin = True
                        npin += 1
                    for ox, oy in corners:
                        if point_in_polygon(s, cx + ox * cs, cy + oy * cs):
                            npin += 1
                    if cin and npin == 5:
                        inside[ir, jr] = True
                    elif npin > 0:
                        bmask[ir, jr] = True
            li, lj = np.where(inside)
            for q in range(len(li)):
                P[li[q] + mi, lj[q] + mj] = mp; M[li[q] + mi, lj[q] + mj] = mb; C[li[q] + mi, lj[q] + mj] = mc
            for mask in (bmask, ~inside & ~bmask):
                bi, bj = np.where(mask)
                for q in range(len(bi)):
                    i, j = bi[q] + mi, bj[q] + mj
                    cnt = 0
                    for k in range(ns):
                        if point_in_polygon(s, xc[j] + odx[k], yc[i] + ody[k]):
                            cnt += 1
                    if cnt > 0:
                        blend(i, j, cnt, mp, mb, mc)
    # PML pass, meshing.py:484-531
    for b in design.boundaries:
        if not hasattr(b, "position"):
            continue
        if hasattr(b, "width") and hasattr(b, "height"):
            px, py = b.position
            mi = max(0, int(py / cs)); mj = max(0, int(px / cs))
            Mi = min(gh, int(np.ceil((py + b.height) / cs))); Mj = min(gw, int(np.ceil((px + b.width) / cs)))
        elif hasattr(b, "radius"):
            cx0, cy0 = b.position
            mi = max(0, int((cy0 - b.radius) / cs)); mj = max(0, int((cx0 - b.radius) / cs))
            Mi = min(gh, int(np.ceil((cy0 + b.radius) / cs))); Mj = min(gw, int(np.ceil((cx0 + b.radius) / cs)))
        else:
            continue
        for i in range(mi, Mi):
            for j in range(mj, Mj):
                pc = pml_conductivity(b, xc[j], yc[i], dx=resolution, dt=dt_est, eps_avg=P[i, j])
                if pc > 0:
                    C[i, j] += pc
    return P, M, C


def rasterize_3d(design, res_xy, res_z=None):
    """meshing.py:572-614, 616-781. Returns (permittivity, permeability, conductivity) [z,y,x]."""
    if res_z is None:
        res_z = res_xy
    W, H, D = design.width, design.height, design.depth
    gw = int(W / res_xy); gh = int(H / res_xy)
    gd = int(D / res_z) if D > 0 else 1
    cxy, cz = res_xy, res_z
    xc = np.linspace(0.5 * cxy, W - 0.5 * cxy, gw)
    yc = np.linspace(0.5 * cxy, H - 0.5 * cxy, gh)
    zc = np.linspace(0.5 * cz, D - 0.5 * cz, gd) if D > 0 else [0]
    oxy = np.array([-0.25, 0, 0.25]) * cxy
    oz = np.array([-0.25, 0, 0.25]) * cz if D > 0 else [0]
    odx, ody, odz = np.meshgrid(oxy, oxy, oz)
    odx, ody, odz = odx.flatten(), ody.flatten(), odz.flatten()
    ns = len(odx)
    c = 3e8
    dt_est = 0.5 * res_xy / (c * np.sqrt(2))
    P = np.ones((gd, gh, gw)); M = np.ones((gd, gh, gw)); C = np.zeros((gd, gh, gw))
    if design.structures:
        bg = design.structures[0]
        if getattr(bg, "material", None) is not None:
            P.fill(bg.material.permittivity); M.fill(bg.material.permeability); C.fill(bg.material.conductivity)
    for s in design.structures[1:]:
        if getattr(s, "is_pml", False):
            continue
        if getattr(s, "material", None) is None:
            continue
        mp, mb, mc = _mat(s.material)
        bbox = s.get_bounding_box()
        if len(bbox) == 6:
            minx, miny, minz, maxx, maxy, maxz = bbox
        else:
            minx, miny, maxx, maxy = bbox
            minz, maxz = 0, 0
        mi = max(0, int(miny / cxy) - 1); mj = max(0, int(minx / cxy) - 1)
        mk = max(0, int(minz / cz) - 1) if D > 0 else 0
        Mi = min(gh, int(np.ceil(maxy / cxy)) + 1); Mj = min(gw, int(np.ceil(maxx / cxy)) + 1)
        Mk = min(gd, int(np.ceil(maxz / cz)) + 1) if D > 0 else 1
        if mi >= gh or mj >= gw or mk >= gd or Mi <= 0 or Mj <= 0 or Mk <= 0:
            continue
        if not hasattr(s, "vertices"):
            raise NotImplementedError("only polygon-like structures are restated")
        for k in range(mk, Mk):
            for i in range(mi, Mi):
                for j in range(mj, Mj):
                    cnt = 0
                    for q in range(ns):
                        if point_in_polygon(s, xc[j] + odx[q], yc[i] + ody[q]):
                            cnt += 1
                    if cnt > 0:
                        bf = cnt / ns
                        P[k, i, j] = P[k, i, j] * (1 - bf) + mp * bf
                        M[k, i, j] = M[k, i, j] * (1 - bf) + mb * bf
                        C[k, i, j] = C[k, i, j] * (1 - bf) + mc * bf
    for b in design.boundaries:
        for k in range(len(zc)):
            for i, y in enumerate(yc):
                for j, x in enumerate(xc):
                    pc = pml_conductivity(b, x, y, dx=res_xy, dt=dt_est, eps_avg=P[k, i, j])
                    if pc > 0:
                        C[k, i, j] += pc
    return P, M, C
