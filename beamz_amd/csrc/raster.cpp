// Host rasterizer: permittivity / permeability / conductivity maps in FP64.
//
// Bit-exact with the reference rasterizer: every floating-point expression
// below keeps the reference's operation order (beamz/simulation/meshing.py:
// 92-536 for 2D, 616-781 for 3D; PML profile beamz/design/pml.py:21-62;
// even-odd containment beamz/design/structures.py:178-207), and the file is
// compiled with -ffp-contract=off so no multiply-add is fused. Cell centres
// follow numpy.linspace: c[j] = j*step + start, with the last set to stop.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/beamz_amd.h"

namespace bz {
void set_error(const std::string& msg);
}

namespace {

std::vector<double> linspace(double start, double stop, int n) {
    std::vector<double> y(n > 0 ? n : 0);
    if (n <= 0) return y;
    if (n == 1) {
        y[0] = 0.0 * (stop - start) + start;
        return y;
    }
    const double delta = stop - start;
    const double step = delta / (double)(n - 1);
    for (int j = 0; j < n; ++j) y[j] = (double)j * step + start;
    y[n - 1] = stop;
    return y;
}

inline long ipart(double v) { return (long)v; }  // Python int(): truncation toward zero

bool in_path(double x, double y, const double* p, int n) {
    if (n <= 0) return false;
    bool inside = false;
    double p1x = p[0], p1y = p[1];
    for (int i = 1; i <= n; ++i) {
        const double p2x = p[2 * (i % n)], p2y = p[2 * (i % n) + 1];
        const double lo = p1y < p2y ? p1y : p2y;
        const double hi = p1y > p2y ? p1y : p2y;
        const double mx = p1x > p2x ? p1x : p2x;
        if (y > lo && y <= hi && x <= mx) {
            double xin;
            if (p1y != p2y) xin = (y - p1y) * (p2x - p1x) / (p2y - p1y) + p1x;
            else xin = p1x;
            if (p1x == p2x || x <= xin) inside = !inside;
        }
        p1x = p2x;
        p1y = p2y;
    }
    return inside;
}

bool in_shape(const bz_shape& s, double x, double y) {
    if (s.nv <= 0 || !in_path(x, y, s.v, s.nv)) return false;
    const double* h = s.hv;
    for (int k = 0; k < s.nholes; ++k) {
        if (s.hole_nv[k] > 0 && in_path(x, y, h, s.hole_nv[k])) return false;
        h += 2 * s.hole_nv[k];
    }
    return true;
}

// pml.py:21-27
void pml_profile(const bz_pml& b, double nd, double* sig, double* alpha) {
    double d = nd > 0.0 ? nd : 0.0;
    d = d < 1.0 ? d : 1.0;
    if (d < 0.05) *sig = 0.01 * std::pow(d / 0.05, 2.0);
    else *sig = std::pow((d - 0.05) / 0.95, b.order);
    *alpha = b.alpha_max * std::pow(d, 2.0);
}

// pml.py:29-62 with dx, dt and eps_avg all given (the rasterizer's call)
double pml_conductivity(const bz_pml& b, double x, double y, double dx, double eps_avg) {
    const double eta = std::sqrt(1.256e-6 / (8.85e-12 * eps_avg));
    double smax = 1.2 / (eta * dx);
    smax *= b.sigma_factor;
    double dist;
    if (b.kind == 0) {
        if (!(b.px <= x && x <= b.px + b.w && b.py <= y && y <= b.py + b.h)) return 0.0;
        switch (b.orient) {
            case 0: dist = 1.0 - (x - b.px) / b.w; break;
            case 1: dist = (x - b.px) / b.w; break;
            case 2: dist = (y - b.py) / b.h; break;
            case 3: dist = 1.0 - (y - b.py) / b.h; break;
            default: return 0.0;
        }
    } else {
        const double rr = std::hypot(x - b.px, y - b.py);
        if (rr > b.r) return 0.0;
        const double ox = x - b.px, oy = y - b.py;
        if (b.orient == 0 && (ox > 0 || oy < 0)) return 0.0;
        if (b.orient == 1 && (ox < 0 || oy < 0)) return 0.0;
        if (b.orient == 2 && (ox > 0 || oy > 0)) return 0.0;
        if (b.orient == 3 && (ox < 0 || oy > 0)) return 0.0;
        dist = rr / b.r;
    }
    double sp, ap;
    pml_profile(b, dist, &sp, &ap);
    double c = smax * sp;
    c *= 1.0 / (1.0 + ap);
    return c;
}

struct Maps {
    double *P, *M, *C;
    int gw;
    inline void blend(long idx, int cnt, int ns, const bz_shape& s) {
        const double bf = (double)cnt / (double)ns;
        P[idx] = P[idx] * (1 - bf) + s.perm * bf;
        M[idx] = M[idx] * (1 - bf) + s.permb * bf;
        C[idx] = C[idx] * (1 - bf) + s.cond * bf;
    }
    inline void fill(long idx, const bz_shape& s) {
        P[idx] = s.perm;
        M[idx] = s.permb;
        C[idx] = s.cond;
    }
    // the same with the material of cell (i, j) of a per-cell table (a spatially varying material)
    inline void blend_cell(long idx, int cnt, int ns, const bz_shape& s, long i, long j) {
        const double bf = (double)cnt / (double)ns;
        const long n = (long)s.cell_ni * s.cell_nj, c = (i - s.cell_i0) * s.cell_nj + (j - s.cell_j0);
        P[idx] = P[idx] * (1 - bf) + s.cell[c] * bf;
        M[idx] = M[idx] * (1 - bf) + s.cell[n + c] * bf;
        C[idx] = C[idx] * (1 - bf) + s.cell[2 * n + c] * bf;
    }
    inline void fill_cell(long idx, const bz_shape& s, long i, long j) {
        const long n = (long)s.cell_ni * s.cell_nj, c = (i - s.cell_i0) * s.cell_nj + (j - s.cell_j0);
        P[idx] = s.cell[c];
        M[idx] = s.cell[n + c];
        C[idx] = s.cell[2 * n + c];
    }
};

inline bool in_cells(const bz_shape& s, long i0, long i1, long j0, long j1) {
    return i0 >= s.cell_i0 && i1 <= (long)s.cell_i0 + s.cell_ni && j0 >= s.cell_j0 && j1 <= (long)s.cell_j0 + s.cell_nj;
}

bool check_shape(const bz_shape& s) {
    if (s.kind < 0 || s.kind > 3) return false;
    if (s.cell && (s.kind != 0 || s.cell_ni <= 0 || s.cell_nj <= 0)) return false;
    if (s.kind == 3 && (s.nv < 0 || (s.nv > 0 && !s.v))) return false;
    if (s.nholes > 0 && (!s.hole_nv || !s.hv)) return false;
    return true;
}

}  // namespace

extern "C" int bz_raster_2d(const bz_shape* shapes, int nshapes, const double* bg, const bz_pml* pmls, int npml,
                            double width, double height, double res, int gw, int gh, double* perm, double* permb,
                            double* cond) {
    if (gw <= 0 || gh <= 0 || !perm || !permb || !cond || res <= 0 || (nshapes > 0 && !shapes) ||
        (npml > 0 && !pmls)) {
        bz::set_error("bz_raster_2d: invalid arguments");
        return BZ_EINVAL;
    }
    for (int q = 0; q < nshapes; ++q)
        if (!check_shape(shapes[q])) {
            bz::set_error("bz_raster_2d: malformed shape descriptor");
            return BZ_EINVAL;
        }
    const double cs = res;
    const std::vector<double> xc = linspace(0.5 * cs, width - 0.5 * cs, gw);
    const std::vector<double> yc = linspace(0.5 * cs, height - 0.5 * cs, gh);
    const double off[3] = {-0.25 * cs, 0.0 * cs, 0.25 * cs};
    const int ns = 9;
    const long N = (long)gw * gh;
    for (long q = 0; q < N; ++q) {
        perm[q] = bg ? bg[0] : 1.0;
        permb[q] = bg ? bg[1] : 1.0;
        cond[q] = bg ? bg[2] : 0.0;
    }
    Maps mp{perm, permb, cond, gw};
    for (int q = 0; q < nshapes; ++q) {
        const bz_shape& s = shapes[q];
        const double minx = s.bbox[0], miny = s.bbox[1], maxx = s.bbox[3], maxy = s.bbox[4];
        long mi = ipart(miny / cs) - 1; if (mi < 0) mi = 0;
        long mj = ipart(minx / cs) - 1; if (mj < 0) mj = 0;
        long Mi = ipart(std::ceil(maxy / cs)) + 1; if (Mi > gh) Mi = gh;
        long Mj = ipart(std::ceil(maxx / cs)) + 1; if (Mj > gw) Mj = gw;
        if (mi >= gh || mj >= gw || Mi <= 0 || Mj <= 0) continue;
        if (s.kind == 0) {
            // meshing.py:193-257 fast path
            const double px = s.px, py = s.py;
            long r0j = ipart(px / cs); if (r0j < 0) r0j = 0;
            long r0i = ipart(py / cs); if (r0i < 0) r0i = 0;
            long r1j = ipart(std::ceil((px + s.w) / cs)); if (r1j > gw) r1j = gw;
            long r1i = ipart(std::ceil((py + s.h) / cs)); if (r1i > gh) r1i = gh;
            long n0j = ipart((px + 0.25 * cs) / cs); if (n0j < 0) n0j = 0;
            long n0i = ipart((py + 0.25 * cs) / cs); if (n0i < 0) n0i = 0;
            long n1j = ipart(std::floor((px + s.w - 0.25 * cs) / cs)); if (n1j > gw) n1j = gw;
            long n1i = ipart(std::floor((py + s.h - 0.25 * cs) / cs)); if (n1i > gh) n1i = gh;
            if (s.cell && ((n1i > n0i && n1j > n0j && !in_cells(s, n0i, n1i, n0j, n1j)) ||
                           (r1i > r0i && r1j > r0j && !in_cells(s, r0i, r1i, r0j, r1j)))) {
                bz::set_error("bz_raster_2d: a rectangle's per-cell material table does not cover its cells");
                return BZ_EINVAL;
            }
            if (n1i > n0i && n1j > n0j)
                for (long i = n0i; i < n1i; ++i)
                    for (long j = n0j; j < n1j; ++j) {
                        if (s.cell) mp.fill_cell(i * gw + j, s, i, j);
                        else mp.fill(i * gw + j, s);
                    }
            const long hi = r1i - r0i, wj = r1j - r0j;
            if (hi <= 0 || wj <= 0) continue;
            for (long a = 0; a < hi; ++a)
                for (long b = 0; b < wj; ++b) {
                    bool edge = false;
                    if (r0i < n0i && a < n0i - r0i) edge = true;
                    // numpy slice [n1i-r0i:] with a possibly negative start counts from the end
                    if (n1i < r1i) {
                        long st = n1i - r0i;
                        if (st < 0) st += hi;
                        if (st < 0) st = 0;
                        if (a >= st) edge = true;
                    }
                    if (r0j < n0j && b < n0j - r0j) edge = true;
                    if (n1j < r1j) {
                        long st = n1j - r0j;
                        if (st < 0) st += wj;
                        if (st < 0) st = 0;
                        if (b >= st) edge = true;
                    }
                    if (!edge) continue;
                    const long i = a + r0i, j = b + r0j;
                    int cnt = 0;
                    for (int ky = 0; ky < 3; ++ky)
                        for (int kx = 0; kx < 3; ++kx) {
                            const double xs = xc[j] + off[kx], ys = yc[i] + off[ky];
                            if (px <= xs && xs < px + s.w && py <= ys && ys < py + s.h) ++cnt;
                        }
                    if (cnt > 0) {
                        if (s.cell) mp.blend_cell(i * gw + j, cnt, ns, s, i, j);
                        else mp.blend(i * gw + j, cnt, ns, s);
                    }
                }
        } else if (s.kind == 1 || s.kind == 2) {
            // meshing.py:259-361 circle / ring paths: full cells first, then boundary blends
            const long hi = Mi - mi, wj = Mj - mj;
            std::vector<double> D(hi * wj);
            std::vector<char> full(hi * wj), bnd(hi * wj);
            for (long a = 0; a < hi; ++a)
                for (long b = 0; b < wj; ++b) {
                    const double ddx = xc[b + mj] - s.cx, ddy = yc[a + mi] - s.cy;
                    const double d = std::sqrt(ddx * ddx + ddy * ddy);
                    D[a * wj + b] = d;
                    if (s.kind == 1) {
                        const bool f = d + 0.3536 * cs <= s.r;
                        full[a * wj + b] = f;
                        bnd[a * wj + b] = (d - 0.3536 * cs <= s.r) && !f;
                    } else {
                        full[a * wj + b] = (d - 0.3536 * cs >= s.r_in) && (d + 0.3536 * cs <= s.r_out);
                        const bool ib = (d - 0.3536 * cs <= s.r_in) && (d + 0.3536 * cs >= s.r_in);
                        const bool ob = (d - 0.3536 * cs <= s.r_out) && (d + 0.3536 * cs >= s.r_out);
                        bnd[a * wj + b] = ib || ob;
                    }
                }
            for (long a = 0; a < hi; ++a)
                for (long b = 0; b < wj; ++b)
                    if (full[a * wj + b]) mp.fill((a + mi) * gw + (b + mj), s);
            for (long a = 0; a < hi; ++a)
                for (long b = 0; b < wj; ++b) {
                    if (!bnd[a * wj + b]) continue;
                    const long i = a + mi, j = b + mj;
                    int cnt = 0;
                    for (int ky = 0; ky < 3; ++ky)
                        for (int kx = 0; kx < 3; ++kx) {
                            const double d = std::hypot(xc[j] + off[kx] - s.cx, yc[i] + off[ky] - s.cy);
                            if (s.kind == 1) {
                                if (d <= s.r) ++cnt;
                            } else if (s.r_in <= d && d <= s.r_out) {
                                ++cnt;
                            }
                        }
                    if (cnt > 0) mp.blend(i * gw + j, cnt, ns, s);
                }
        } else {
            // meshing.py:362-455 general polygon path (5-point pre-classification, then 9 samples)
            const long hi = Mi - mi, wj = Mj - mj;
            std::vector<char> cls(hi * wj, 0);  // 1 inside, 2 boundary, 0 remaining
            const double cpt[4][2] = {{-0.4, -0.4}, {-0.4, 0.4}, {0.4, -0.4}, {0.4, 0.4}};
            for (long a = 0; a < hi; ++a)
                for (long b = 0; b < wj; ++b) {
                    const double cx = xc[b + mj], cy = yc[a + mi];
                    int n_in = 0;
                    const bool cin = in_shape(s, cx, cy);
                    if (cin) ++n_in;
                    for (int c = 0; c < 4; ++c)
                        if (in_shape(s, cx + cpt[c][0] * cs, cy + cpt[c][1] * cs)) ++n_in;
                    if (cin && n_in == 5) cls[a * wj + b] = 1;
                    else if (n_in > 0) cls[a * wj + b] = 2;
                }
            for (long a = 0; a < hi; ++a)
                for (long b = 0; b < wj; ++b)
                    if (cls[a * wj + b] == 1) mp.fill((a + mi) * gw + (b + mj), s);
            for (int pass = 2; pass >= 0; pass -= 2)
                for (long a = 0; a < hi; ++a)
                    for (long b = 0; b < wj; ++b) {
                        if (cls[a * wj + b] != pass) continue;
                        const long i = a + mi, j = b + mj;
                        int cnt = 0;
                        for (int ky = 0; ky < 3; ++ky)
                            for (int kx = 0; kx < 3; ++kx)
                                if (in_shape(s, xc[j] + off[kx], yc[i] + off[ky])) ++cnt;
                        if (cnt > 0) mp.blend(i * gw + j, cnt, ns, s);
                    }
        }
    }
    // meshing.py:484-531 PML pass over each boundary's index box
    for (int q = 0; q < npml; ++q) {
        const bz_pml& b = pmls[q];
        long mi, mj, Mi, Mj;
        if (b.kind == 0) {
            mi = ipart(b.py / cs); mj = ipart(b.px / cs);
            Mi = ipart(std::ceil((b.py + b.h) / cs)); Mj = ipart(std::ceil((b.px + b.w) / cs));
        } else {
            mi = ipart((b.py - b.r) / cs); mj = ipart((b.px - b.r) / cs);
            Mi = ipart(std::ceil((b.py + b.r) / cs)); Mj = ipart(std::ceil((b.px + b.r) / cs));
        }
        if (mi < 0) mi = 0;
        if (mj < 0) mj = 0;
        if (Mi > gh) Mi = gh;
        if (Mj > gw) Mj = gw;
        for (long i = mi; i < Mi; ++i)
            for (long j = mj; j < Mj; ++j) {
                const double pc = pml_conductivity(b, xc[j], yc[i], res, perm[i * gw + j]);
                if (pc > 0) cond[i * gw + j] += pc;
            }
    }
    return BZ_OK;
}

extern "C" int bz_raster_3d(const bz_shape* shapes, int nshapes, const double* bg, const bz_pml* pmls, int npml,
                            double width, double height, double depth, double res_xy, double res_z, int gw, int gh,
                            int gd, double* perm, double* permb, double* cond) {
    if (gw <= 0 || gh <= 0 || gd <= 0 || !perm || !permb || !cond || res_xy <= 0 || res_z <= 0 ||
        (nshapes > 0 && !shapes) || (npml > 0 && !pmls)) {
        bz::set_error("bz_raster_3d: invalid arguments");
        return BZ_EINVAL;
    }
    for (int q = 0; q < nshapes; ++q)
        if (!check_shape(shapes[q]) || shapes[q].cell) {
            bz::set_error("bz_raster_3d: malformed shape descriptor");
            return BZ_EINVAL;
        }
    const double cxy = res_xy, cz = res_z;
    const std::vector<double> xc = linspace(0.5 * cxy, width - 0.5 * cxy, gw);
    const std::vector<double> yc = linspace(0.5 * cxy, height - 0.5 * cxy, gh);
    const double off[3] = {-0.25 * cxy, 0.0 * cxy, 0.25 * cxy};
    const int nzs = depth > 0 ? 3 : 1;
    const int ns = 9 * nzs;
    const long plane = (long)gw * gh;
    const long N = plane * gd;
    for (long q = 0; q < N; ++q) {
        perm[q] = bg ? bg[0] : 1.0;
        permb[q] = bg ? bg[1] : 1.0;
        cond[q] = bg ? bg[2] : 0.0;
    }
    Maps mp{perm, permb, cond, gw};
    for (int q = 0; q < nshapes; ++q) {
        const bz_shape& s = shapes[q];
        // meshing.py:682-741: every structure takes the 27-sample polygon test (z ignored)
        const double minx = s.bbox[0], miny = s.bbox[1], minz = s.bbox[2];
        const double maxx = s.bbox[3], maxy = s.bbox[4], maxz = s.bbox[5];
        long mi = ipart(miny / cxy) - 1; if (mi < 0) mi = 0;
        long mj = ipart(minx / cxy) - 1; if (mj < 0) mj = 0;
        long mk = depth > 0 ? ipart(minz / cz) - 1 : 0; if (mk < 0) mk = 0;
        long Mi = ipart(std::ceil(maxy / cxy)) + 1; if (Mi > gh) Mi = gh;
        long Mj = ipart(std::ceil(maxx / cxy)) + 1; if (Mj > gw) Mj = gw;
        long Mk = depth > 0 ? ipart(std::ceil(maxz / cz)) + 1 : 1; if (Mk > gd) Mk = gd;
        if (mi >= gh || mj >= gw || mk >= gd || Mi <= 0 || Mj <= 0 || Mk <= 0) continue;
        for (long k = mk; k < Mk; ++k)
            for (long i = mi; i < Mi; ++i)
                for (long j = mj; j < Mj; ++j) {
                    int cnt = 0;
                    for (int ky = 0; ky < 3; ++ky)
                        for (int kx = 0; kx < 3; ++kx)
                            if (in_shape(s, xc[j] + off[kx], yc[i] + off[ky])) cnt += nzs;
                    if (cnt > 0) mp.blend(k * plane + i * gw + j, cnt, ns, s);
                }
    }
    // meshing.py:757-781: x/y PML profile on every z-layer, all cells
    for (int q = 0; q < npml; ++q) {
        const bz_pml& b = pmls[q];
        for (long k = 0; k < gd; ++k)
            for (long i = 0; i < gh; ++i)
                for (long j = 0; j < gw; ++j) {
                    const long idx = k * plane + i * gw + j;
                    const double pc = pml_conductivity(b, xc[j], yc[i], res_xy, perm[idx]);
                    if (pc > 0) cond[idx] += pc;
                }
    }
    return BZ_OK;
}
