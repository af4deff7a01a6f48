// Fused single-pass 3D Yee step for gfx950: H and E updates of one time step
// plus soft-source injection in one sweep over HBM.
//
// Step semantics are those of beamz/simulation/fdtd.py:200-310 followed by
// helper.py:5-118 (H on all valid entries, E on the interior slices, sources
// added to the updated E). Fields are ping-ponged: every cell of the step is
// read from `src` and written to `dst`, so neighbouring waves can read halos
// from `src` without any ordering between them.
//
// Work decomposition (wave64, no LDS, no barriers):
//   * a wave owns x-columns i0-1 .. i0+62 (lane l -> i = i0-1+l) and rows
//     j0-1 .. j0+R-1 (register row r -> j = j0-1+r) and marches along z over a
//     segment [k0, k1). Lanes 0 and 63 and row 0 are halo owners: lane 0 and
//     row 0 supply H at i-1 / j-1, lane 63 supplies E at i+1; they store
//     nothing, so a wave outputs 62 x R columns. Row R+1 supplies E at j+1.
//   * per plane every column reads its eight values (Ex/Ey of the next plane,
//     Ez, Hx, Hy, Hz, sigma, eps) once; Ex/Ey of the current plane and Hx/Hy
//     of the previous plane stay in registers; x neighbours move between
//     lanes by wave shuffles, y neighbours between register rows.
//   * both planes of a loop iteration are requested before any store of the
//     iteration (gfx9 counts loads and stores on one in-order counter).
//   * addressing is 32-bit: each lane keeps one clamped in-grid offset per
//     register row, and a plane adds one wave-uniform offset; halo lanes/rows
//     outside the grid read a clamped in-grid cell whose value only reaches
//     entries that the validity selects discard or that are not stored.
//   * a segment starting at k0 > 0 first recomputes H of plane k0-1 for its
//     own columns (bit-identical to what the previous segment computes).
//   * workgroups are 4 waves stacked in y; hardware workgroup ids are remapped
//     so each XCD (8 per MI355X, round-robin dispatch) gets a contiguous block
//     of tiles and the halos it re-reads stay in its own L2.
// HBM traffic per cell-update is 8 reads + 6 writes of FP32 = 56 B, plus the
// halo rows / lanes (mostly L2 hits) and one extra plane per segment.
#include "common.h"

// No multiply-add contraction in this file or in yee3d_tb2.hip: both kernels then evaluate every
// update with the same rounded operations, which is what makes them bit-identical.
#pragma clang fp contract(off)

namespace {

constexpr int LANES = 64;
constexpr int OX = LANES - 2;
constexpr int WPG = 4;       // waves per workgroup
constexpr int NXCD = 8;

struct Fields6 {
    float* p[6];
};

struct BoxSources {
    int n;
    bz_box_source b[BZ_MAX_BOX];
};

// Per-plane register set of one wave column block: row r = 0..R+1.
template <int R>
struct Plane {
    float ex1[R + 2], ey1[R + 1];   // Ex, Ey of plane k+1
    float ez[R + 2];                // Ez of plane k
    float hx[R + 1], hy[R + 1], hz[R + 1], sg[R + 1], ee[R + 1];
};

// Values carried from plane k-1 to plane k: Ex/Ey of plane k, Hx/Hy (new) of plane k-1.
template <int R>
struct Carry {
    float ex0[R + 2], ey0[R + 1], hxp[R + 1], hyp[R + 1];
};

template <int R>
struct Geo {
    int nx, ny, nz, jb, i, k0;
    uint32_t ps;
    uint32_t ro[R + 2];      // clamped in-grid offset of each register row within a plane
    bool lane_out, xin, hyok;
};

__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Element at a 32-bit byte offset from a wave-uniform base: lets the compiler use the scalar-base +
// 32-bit vector-offset addressing form instead of per-lane 64-bit address arithmetic.
__device__ __forceinline__ float at(const float* __restrict__ base, uint32_t byte_off) {
    return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + byte_off);
}
__device__ __forceinline__ void put(float* __restrict__ base, uint32_t byte_off, float v) {
    *reinterpret_cast<float*>(reinterpret_cast<char*>(base) + byte_off) = v;
}

__device__ __forceinline__ void h_coef_fast(float sig, const Coef& c, float& fa, float& fb) {
    const float s = sig * c.kh;
    const float r = fast_rcp(1.0f + s);
    fa = (1.0f - s) * r;
    fb = c.kbh * r;
}

__device__ __forceinline__ void e_coef_fast(float sig, float eps, const Coef& c, float& ca, float& cb) {
    const float ie = fast_rcp(eps);
    const float s = sig * c.ke * ie;
    const float r = fast_rcp(1.0f + s);
    ca = (1.0f - s) * r;
    cb = c.kbe * ie * r;
}

template <int R>
__device__ __forceinline__ void load_plane(Plane<R>& P, const Fields6& S, const float* __restrict__ eps,
                                           const float* __restrict__ sig, int k, const Geo<R>& g) {
    // planes past the grid are clamped to the last one (their values feed only discarded entries)
    const uint32_t kp = (uint32_t)min(k, g.nz - 1) * g.ps;
    const uint32_t kq = (uint32_t)min(k + 1, g.nz - 1) * g.ps;
#pragma unroll
    for (int r = 0; r < R + 2; ++r) {
        const uint32_t o = kp + g.ro[r], q = kq + g.ro[r];
        P.ex1[r] = at(S.p[0], q);
        P.ez[r] = at(S.p[2], o);
        if (r < R + 1) {
            P.ey1[r] = at(S.p[1], q);
            P.hx[r] = at(S.p[3], o);
            P.hz[r] = at(S.p[5], o);
            P.sg[r] = at(sig, o);
            P.hy[r] = (r > 0) ? at(S.p[4], o) : 0.0f;
            P.ee[r] = (r > 0) ? at(eps, o) : 1.0f;
        }
    }
}

// One z-plane: H of plane k for rows 0..R, then (if k >= k0) E of rows 1..R, sources, stores.
template <int R>
__device__ __forceinline__ void march_plane(const Plane<R>& P, Carry<R>& C, int k, const Geo<R>& g,
                                            const bool* col, const Fields6& D, const Coef& c, float rdx, float rdy,
                                            float rdz, const BoxSources& src) {
    float hxn[R + 1], hyn[R + 1], hzn[R + 1];
    const bool kh = k < g.nz - 1;
#pragma unroll
    for (int r = 0; r < R + 1; ++r) {
        const int j = g.jb + r;
        const float ezR = __shfl_down(P.ez[r], 1, LANES);
        const float eyR = __shfl_down(C.ey0[r], 1, LANES);
        float fa, fb;
        h_coef_fast(P.sg[r], c, fa, fb);
        const bool jh = j < g.ny - 1;
        hxn[r] = (kh && jh) ? fa * P.hx[r] - fb * ((P.ez[r + 1] - P.ez[r]) * rdy - (P.ey1[r] - C.ey0[r]) * rdz)
                            : P.hx[r];
        hyn[r] = (kh && g.hyok) ? fa * P.hy[r] - fb * ((P.ex1[r] - C.ex0[r]) * rdz - (ezR - P.ez[r]) * rdx)
                                : P.hy[r];
        hzn[r] = (jh && g.hyok) ? fa * P.hz[r] - fb * ((eyR - C.ey0[r]) * rdx - (C.ex0[r + 1] - C.ex0[r]) * rdy)
                                : P.hz[r];
    }
    if (k >= g.k0) {
        const bool kin = k >= 1 && k <= g.nz - 2;
        const int i = g.i;
        const uint32_t kp = (uint32_t)k * g.ps;
#pragma unroll
        for (int r = 1; r < R + 1; ++r) {
            const float hzL = __shfl_up(hzn[r], 1, LANES);
            const float hyL = __shfl_up(hyn[r], 1, LANES);
            const int j = g.jb + r;
            if (g.lane_out && col[r]) {
                const uint32_t o = kp + g.ro[r];
                const bool yin = j >= 1 && j <= g.ny - 2;
                float ca, cb;
                e_coef_fast(P.sg[r], P.ee[r], c, ca, cb);
                float exn = (kin && yin && i <= g.nx - 2)
                                ? ca * C.ex0[r] + cb * ((hzn[r] - hzn[r - 1]) * rdy - (hyn[r] - C.hyp[r]) * rdz)
                                : C.ex0[r];
                float eyn = (kin && j <= g.ny - 2 && g.xin)
                                ? ca * C.ey0[r] + cb * ((hxn[r] - C.hxp[r]) * rdz - (hzn[r] - hzL) * rdx)
                                : C.ey0[r];
                float ezn = (k <= g.nz - 2 && yin && g.xin)
                                ? ca * P.ez[r] + cb * ((hyn[r] - hyL) * rdx - (hxn[r] - hxn[r - 1]) * rdy)
                                : P.ez[r];
                for (int q = 0; q < src.n; ++q) {
                    const bz_box_source& b = src.b[q];
                    const int a = i - b.x0, bb = j - b.y0, cc = k - b.z0;
                    if (a >= 0 && a < b.sx && bb >= 0 && bb < b.sy && cc >= 0 && cc < b.sz) {
                        const long rr = ((long)cc * b.sy + bb) * b.sx + a;
                        const float m = b.mul ? b.mul[rr] : 1.0f;
                        if (b.comp == 0) exn = m * exn + b.add[rr] * b.amp;
                        else if (b.comp == 1) eyn = m * eyn + b.add[rr] * b.amp;
                        else ezn = m * ezn + b.add[rr] * b.amp;
                    }
                }
                put(D.p[0], o, exn);
                put(D.p[1], o, eyn);
                put(D.p[2], o, ezn);
                put(D.p[3], o, hxn[r]);
                put(D.p[4], o, hyn[r]);
                put(D.p[5], o, hzn[r]);
            }
        }
    }
#pragma unroll
    for (int r = 0; r < R + 2; ++r) {
        C.ex0[r] = P.ex1[r];
        if (r < R + 1) {
            C.ey0[r] = P.ey1[r];
            C.hxp[r] = hxn[r];
            C.hyp[r] = hyn[r];
        }
    }
}

template <int R>
__global__ __launch_bounds__(LANES* WPG) void k_yee3d_fused(Fields6 S, Fields6 D, const float* __restrict__ eps,
                                                           const float* __restrict__ sig, int nx, int ny, int nz,
                                                           int ld, int kseg, int gx, int gy, int gz, int per_xcd,
                                                           Coef c, float rdx, float rdy, float rdz, BoxSources src) {
    // XCD-aware tile id: hardware id h runs on XCD h % 8; give XCD q the tiles [q*per_xcd, (q+1)*per_xcd).
    const int h = blockIdx.x;
    const int tile = (h % NXCD) * per_xcd + h / NXCD;
    if (tile >= gx * gy * gz) return;
    const int tx = tile % gx;
    const int ty = (tile / gx) % gy;
    const int tz = tile / (gx * gy);

    Geo<R> g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    const int l = threadIdx.x;
    g.i = tx * OX - 1 + l;
    g.jb = (ty * WPG + threadIdx.y) * R - 1;   // row 0 of this wave (halo row)
    g.k0 = tz * kseg;
    g.ps = (uint32_t)ny * (uint32_t)ld * 4u;     // plane stride in bytes
    const int k1 = min(nz, g.k0 + kseg);
    const bool ci = g.i >= 0 && g.i < nx;
    g.lane_out = l > 0 && l < LANES - 1 && ci;
    g.xin = g.i >= 1 && g.i <= nx - 2;
    g.hyok = g.i < nx - 1;
    const int ic = min(max(g.i, 0), nx - 1);
    bool col[R + 2];
#pragma unroll
    for (int r = 0; r < R + 2; ++r) {
        const int j = g.jb + r;
        col[r] = ci && j >= 0 && j < ny;
        g.ro[r] = ((uint32_t)min(max(j, 0), ny - 1) * (uint32_t)ld + (uint32_t)ic) * 4u;
    }

    const int kstart = g.k0 > 0 ? g.k0 - 1 : g.k0;
    Carry<R> C;
    {
        const uint32_t kp = (uint32_t)kstart * g.ps;
#pragma unroll
        for (int r = 0; r < R + 2; ++r) {
            C.ex0[r] = at(S.p[0], kp + g.ro[r]);
            if (r < R + 1) {
                C.ey0[r] = at(S.p[1], kp + g.ro[r]);
                C.hxp[r] = 0.0f;
                C.hyp[r] = 0.0f;
            }
        }
    }
    // Both planes of an iteration are requested before any store of the iteration: the second plane's
    // latency hides behind the first plane's compute, and no loaded plane crosses the loop back-edge.
    Plane<R> A, B;
    for (int k = kstart; k < k1; k += 2) {
        load_plane<R>(A, S, eps, sig, k, g);
        load_plane<R>(B, S, eps, sig, k + 1, g);
        march_plane<R>(A, C, k, g, col, D, c, rdx, rdy, rdz, src);
        if (k + 1 >= k1) break;
        march_plane<R>(B, C, k + 1, g, col, D, c, rdx, rdy, rdz, src);
    }
}

template <int R>
int launch(const Fields6& S, const Fields6& D, const float* eps, const float* sig, int nx, int ny, int nz, int ld,
           int kseg, const Coef& c, float rdx, float rdy, float rdz, const BoxSources& bs, hipStream_t st) {
    const int gx = (nx + OX - 1) / OX;
    const int gy = (ny + WPG * R - 1) / (WPG * R);
    const int gz = (nz + kseg - 1) / kseg;
    const int n = gx * gy * gz;
    const int per_xcd = (n + NXCD - 1) / NXCD;
    auto kern = k_yee3d_fused<R>;
    hipLaunchKernelGGL(kern, dim3(per_xcd * NXCD), dim3(LANES, WPG), 0, st, S, D, eps, sig, nx, ny, nz, ld, kseg, gx,
                       gy, gz, per_xcd, c, rdx, rdy, rdz, bs);
    BZ_CHECK_LAUNCH("bz_yee3d_step");
    return BZ_OK;
}

}  // namespace

extern "C" int bz_yee3d_step(float* const src[6], float* const dst[6], const float* eps_r, const float* sigma, int nx,
                             int ny, int nz, int ld, double dx, double dy, double dz, double dt, double mu0,
                             double eps0, const bz_box_source* boxes, int nbox, int kseg, int rows, void* stream) {
    BZ_REQUIRE(src && dst && eps_r && sigma, "bz_yee3d_step: null pointer");
    for (int q = 0; q < 6; ++q) BZ_REQUIRE(src[q] && dst[q] && src[q] != dst[q], "bz_yee3d_step: bad field buffers");
    BZ_REQUIRE(nx >= 2 && ny >= 2 && nz >= 2 && ld >= nx, "bz_yee3d_step: bad shape");
    BZ_REQUIRE(4.0 * nz * ny * ld < 4294967295.0, "bz_yee3d_step: a field array exceeds 4 GiB (32-bit offsets)");
    BZ_REQUIRE(nbox >= 0 && nbox <= BZ_MAX_BOX && (nbox == 0 || boxes), "bz_yee3d_step: bad source list");
    BoxSources bs;
    bs.n = nbox;
    for (int q = 0; q < nbox; ++q) {
        const bz_box_source& b = boxes[q];
        BZ_REQUIRE(b.comp >= 0 && b.comp <= 2 && b.add && b.sx > 0 && b.sy > 0 && b.sz > 0 && b.x0 >= 0 &&
                       b.y0 >= 0 && b.z0 >= 0 && b.x0 + b.sx <= nx && b.y0 + b.sy <= ny && b.z0 + b.sz <= nz,
                   "bz_yee3d_step: source box outside the grid");
        bs.b[q] = b;
    }
    if (kseg <= 0) kseg = 16;
    Fields6 S, D;
    for (int q = 0; q < 6; ++q) {
        S.p[q] = src[q];
        D.p[q] = dst[q];
    }
    const Coef c = make_coef(dt, mu0, eps0);
    const float rx = (float)(1.0 / dx), ry = (float)(1.0 / dy), rz = (float)(1.0 / dz);
    const hipStream_t st = (hipStream_t)stream;
    switch (rows) {
        case 2: return launch<2>(S, D, eps_r, sigma, nx, ny, nz, ld, kseg, c, rx, ry, rz, bs, st);
        case 4: return launch<4>(S, D, eps_r, sigma, nx, ny, nz, ld, kseg, c, rx, ry, rz, bs, st);
        default: bz::set_error("bz_yee3d_step: rows must be 2 or 4"); return BZ_EINVAL;
    }
}
