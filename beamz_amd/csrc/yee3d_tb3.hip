// Three-step temporally blocked 3D Yee kernel for gfx950: steps n -> n+3 of
// FDTD._update_3d_fields (beamz/simulation/fdtd.py:200-310) with the soft
// sources of helper.py:62-118 after each E update, in one pass over HBM.
//
// Every cell update is the arithmetic of the one-step fused kernel
// (yee3d_fused.hip) applied three times, so three launches of that kernel and
// one launch of this one give identical bits. The intermediate states (steps
// n+1 and n+2) never go to HBM; monitor samples of them are written to two
// staging boxes.
//
// Decomposition:
//   * a workgroup of W waves owns a tile of 58 x-columns, T = W*R y-rows and a
//     z-segment [k0, k1); lane l of a wave covers column i0-3+l (lanes 0-2 and
//     61-63 are halo lanes), wave w covers rows J0 + w*R .. J0 + w*R + R - 1,
//     with J0 = tile_y*(T-6) - 3, so the first three and last three rows of
//     the workgroup are halo rows and T-6 rows are stored;
//   * the workgroup marches z with a three-level pipeline: at iteration k it
//     computes H and E of step n+1 on plane k, H of step n+2 on plane k-1, E
//     of step n+2 on plane k-2, H of step n+3 on plane k-3 and E of step n+3
//     on plane k-4 (which it stores with H of step n+3 on that plane). Each
//     level lags so that the rows it needs from the neighbour waves were
//     computed in an earlier iteration and cross in the first exchange, which
//     keeps two workgroup barriers per plane;
//   * x neighbours move between lanes by wave shuffles, y neighbours between
//     register rows and, at wave edges, through LDS; z neighbours are carried
//     in registers;
//   * a segment starting at k0 > 0 begins three planes early (recomputed halo).
// Validity: every value outside the cells the reference updates keeps its old
// value, exactly as in the one-step kernel; halo lanes, rows and planes
// (clamped in-grid reads) only feed values that are not stored.
// HBM traffic per three cell-updates: 8 reads + 6 writes of FP32 (56 B) plus
// the halo columns, rows and planes.
#include "common.h"

// No multiply-add contraction, as in yee3d_fused.hip and yee3d_tb2.hip: every update is then the same
// sequence of rounded operations in all three kernels.
#pragma clang fp contract(off)

namespace {

constexpr int LANES = 64;
constexpr int HALO = 3;
constexpr int OXT = LANES - 2 * HALO;   // stored columns per wave
constexpr int NXCD = 8;

struct Fields6 {
    float* p[6];
};

struct Boxes3 {
    int n;
    bz_box_source b[BZ_MAX_BOX];   // b[q].amp: amplitude of step n
    float amp1[BZ_MAX_BOX];        // amplitude of step n + 1
    float amp2[BZ_MAX_BOX];        // amplitude of step n + 2
};

struct Stage3 {
    float* p1;   // [6][sz][sy][sx] (Ex, Ey, Ez, Hx, Hy, Hz of step n+1), or NULL
    float* p2;   // the same for step n+2, or NULL
    int x0, y0, z0, sx, sy, sz;
};

// The update is evaluated on every lane and then selected (a lane select, not a branch around the update):
// for this kernel that measured 2 % faster than the conditional expressions of yee3d_tb2.hip.
__device__ __forceinline__ float sel(bool c, float a, float b) { return c ? a : b; }

__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

__device__ __forceinline__ float at(const float* __restrict__ base, uint32_t byte_off) {
    return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + byte_off);
}
__device__ __forceinline__ void put(float* __restrict__ base, uint32_t byte_off, float v) {
    *reinterpret_cast<float*>(reinterpret_cast<char*>(base) + byte_off) = v;
}

// the coefficient roundings of yee3d_fused.hip
__device__ __forceinline__ void hcoef(float sig, const Coef& c, float& fa, float& fb) {
    const float s = sig * c.kh;
    const float r = fast_rcp(1.0f + s);
    fa = (1.0f - s) * r;
    fb = c.kbh * r;
}

__device__ __forceinline__ void ecoef(float sig, float eps, const Coef& c, float& ca, float& cb) {
    const float ie = fast_rcp(eps);
    const float s = sig * c.ke * ie;
    const float r = fast_rcp(1.0f + s);
    ca = (1.0f - s) * r;
    cb = c.kbe * ie * r;
}

// soft sources of step n + STEP at cell (i, j, k), after its E update
template <int STEP>
__device__ __forceinline__ void add_sources(float& ex, float& ey, float& ez, int i, int j, int k, const Boxes3& src) {
#pragma unroll
    for (int q = 0; q < BZ_MAX_BOX; ++q) {
        if (q >= src.n) break;
        const bz_box_source& b = src.b[q];
        const int a = i - b.x0, bb = j - b.y0, cc = k - b.z0;
        if (a >= 0 && a < b.sx && bb >= 0 && bb < b.sy && cc >= 0 && cc < b.sz) {
            const long rr = ((long)cc * b.sy + bb) * b.sx + a;
            const float m = b.mul ? b.mul[rr] : 1.0f;
            const float amp = STEP == 0 ? b.amp : (STEP == 1 ? src.amp1[q] : src.amp2[q]);
            const float v = b.add[rr] * amp;
            ex = b.comp == 0 ? m * ex + v : ex;
            ey = b.comp == 1 ? m * ey + v : ey;
            ez = b.comp == 2 ? m * ez + v : ez;
        }
    }
}

// Register state of one wave's R rows carried from plane to plane (k: the plane loaded this iteration).
template <int R>
struct Carry3 {
    float ex0[R], ey0[R];                   // E(n) x, y at k
    float h1x[R], h1y[R], h1z[R];           // H(n+1) at k-1
    float e1x[R], e1y[R], e1z[R];           // E(n+1) at k-1
    float f1x[R], f1y[R], f1z[R];           // E(n+1) at k-2
    float h2x[R], h2y[R], h2z[R];           // H(n+2) at k-2
    float g2x[R], g2y[R], g2z[R];           // H(n+2) at k-3
    float p2x[R], p2y[R], p2z[R];           // E(n+2) at k-3
    float q2x[R], q2y[R], q2z[R];           // E(n+2) at k-4
    float h3x[R], h3y[R], h3z[R];           // H(n+3) at k-4
    float g3x[R], g3y[R];                   // H(n+3) x, y at k-5
    float fa1[R], fb1[R], ca1[R], cb1[R];   // update coefficients at k-1
    float fa2[R], fb2[R], ca2[R], cb2[R];   // at k-2
    float fa3[R], fb3[R], ca3[R], cb3[R];   // at k-3
    float ca4[R], cb4[R];                   // E coefficients at k-4
};

template <int R, int W>
__global__ __launch_bounds__(LANES* W) void k_yee3d_tb3(Fields6 S, Fields6 D, const float* __restrict__ eps,
                                                       const float* __restrict__ sig, int nx, int ny, int nz, int ld,
                                                       int kseg, int gx, int gy, int gz, int per_xcd, Coef c,
                                                       float rdx, float rdy, float rdz, Boxes3 src, Stage3 stg) {
    constexpr int T = W * R;
    // first row of each wave: Ez(n), Ex(n) at k; Ez(n+1), Ex(n+1) at k-1; Ez(n+2), Ex(n+2) at k-3
    __shared__ float xa[W][6][LANES];
    // last row: Hz(n+2), Hx(n+2) at k-2; Hz(n+3), Hx(n+3) at k-4
    __shared__ float xh[W][4][LANES];
    // last row: Hz(n+1), Hx(n+1) at k
    __shared__ float xb[W][2][LANES];

    const int h = blockIdx.x;
    const int tile = (h % NXCD) * per_xcd + h / NXCD;
    if (tile >= gx * gy * gz) return;     // uniform over the workgroup
    const int tx = tile % gx;
    const int ty = (tile / gx) % gy;
    const int tz = tile / (gx * gy);
    const int l = threadIdx.x;
    const int w = threadIdx.y;
    const int i = tx * OXT - HALO + l;
    const int J0 = ty * (T - 2 * HALO) - HALO;
    const int jb = J0 + w * R;            // first row of this wave
    const int k0 = tz * kseg;
    const int k1 = min(nz, k0 + kseg);
    const uint32_t ps = (uint32_t)ny * (uint32_t)ld * 4u;
    const int ic = min(max(i, 0), nx - 1);
    uint32_t ro[R];
#pragma unroll
    for (int r = 0; r < R; ++r)
        ro[r] = ((uint32_t)min(max(jb + r, 0), ny - 1) * (uint32_t)ld + (uint32_t)ic) * 4u;
    const bool hyok = i < nx - 1;
    const bool xin = i >= 1 && i <= nx - 2;
    const bool xe = i <= nx - 2;
    const bool lane_out = l >= HALO && l < LANES - HALO && i >= 0 && i < nx;
    const int wd = min(w + 1, W - 1), wu = max(w - 1, 0);   // neighbour waves (edge waves read their own slot)

    Carry3<R> C;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        C.h1x[r] = C.h1y[r] = C.h1z[r] = 0.0f;
        C.e1x[r] = C.e1y[r] = C.e1z[r] = 0.0f;
        C.f1x[r] = C.f1y[r] = C.f1z[r] = 0.0f;
        C.h2x[r] = C.h2y[r] = C.h2z[r] = 0.0f;
        C.g2x[r] = C.g2y[r] = C.g2z[r] = 0.0f;
        C.p2x[r] = C.p2y[r] = C.p2z[r] = 0.0f;
        C.q2x[r] = C.q2y[r] = C.q2z[r] = 0.0f;
        C.h3x[r] = C.h3y[r] = C.h3z[r] = 0.0f;
        C.g3x[r] = C.g3y[r] = 0.0f;
        // the coefficients of sigma 0, eps 1 (planes before the march has loaded any)
        hcoef(0.0f, c, C.fa1[r], C.fb1[r]);
        ecoef(0.0f, 1.0f, c, C.ca1[r], C.cb1[r]);
        C.fa2[r] = C.fa3[r] = C.fa1[r];
        C.fb2[r] = C.fb3[r] = C.fb1[r];
        C.ca2[r] = C.ca3[r] = C.ca4[r] = C.ca1[r];
        C.cb2[r] = C.cb3[r] = C.cb4[r] = C.cb1[r];
    }
    const int kstart = k0 > 0 ? k0 - HALO : 0;
    {
        const uint32_t kp = (uint32_t)min(max(kstart, 0), nz - 1) * ps;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            C.ex0[r] = at(S.p[0], kp + ro[r]);
            C.ey0[r] = at(S.p[1], kp + ro[r]);
        }
    }
    const long splane = (long)stg.sz * stg.sy * stg.sx;

    for (int k = kstart; k <= k1 + HALO; ++k) {
        // ---- loads of plane k (E x, y of plane k+1) ----
        const uint32_t kp = (uint32_t)min(max(k, 0), nz - 1) * ps;
        const uint32_t kq = (uint32_t)min(max(k + 1, 0), nz - 1) * ps;
        float ex1[R], ey1[R], ez[R], hx[R], hy[R], hz[R], sg[R], ee[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            ex1[r] = at(S.p[0], kq + ro[r]);
            ey1[r] = at(S.p[1], kq + ro[r]);
            ez[r] = at(S.p[2], kp + ro[r]);
            hx[r] = at(S.p[3], kp + ro[r]);
            hy[r] = at(S.p[4], kp + ro[r]);
            hz[r] = at(S.p[5], kp + ro[r]);
            sg[r] = at(sig, kp + ro[r]);
            ee[r] = at(eps, kp + ro[r]);
        }
        // the update coefficients of plane k, once: step n+1 uses them here, the later steps further on
        float fa0[R], fb0[R], ca0[R], cb0[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            hcoef(sg[r], c, fa0[r], fb0[r]);
            ecoef(sg[r], ee[r], c, ca0[r], cb0[r]);
        }
        // ---- exchange A: first rows go up a wave, last rows go down ----
        xa[w][0][l] = ez[0];
        xa[w][1][l] = C.ex0[0];
        xa[w][2][l] = C.e1z[0];
        xa[w][3][l] = C.e1x[0];
        xa[w][4][l] = C.p2z[0];
        xa[w][5][l] = C.p2x[0];
        xh[w][0][l] = C.h2z[R - 1];
        xh[w][1][l] = C.h2x[R - 1];
        xh[w][2][l] = C.h3z[R - 1];
        xh[w][3][l] = C.h3x[R - 1];
        __syncthreads();
        const float dEz0 = xa[wd][0][l], dEx0 = xa[wd][1][l], dE1z = xa[wd][2][l], dE1x = xa[wd][3][l];
        const float dE2z = xa[wd][4][l], dE2x = xa[wd][5][l];
        const float uH2z = xh[wu][0][l], uH2x = xh[wu][1][l], uH3z = xh[wu][2][l], uH3x = xh[wu][3][l];

        // ---- H of step n+1 at plane k ----
        float h1x[R], h1y[R], h1z[R];
        {
            const bool kh = k < nz - 1;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int j = jb + r;
                const float ezN = r + 1 < R ? ez[r + 1] : dEz0;
                const float exN = r + 1 < R ? C.ex0[r + 1] : dEx0;
                const float ezR = __shfl_down(ez[r], 1, LANES);
                const float eyR = __shfl_down(C.ey0[r], 1, LANES);
                const float fa = fa0[r], fb = fb0[r];
                const bool jh = j < ny - 1;
                h1x[r] = sel((kh && jh), fa * hx[r] - fb * ((ezN - ez[r]) * rdy - (ey1[r] - C.ey0[r]) * rdz), hx[r]);
                h1y[r] = sel((kh && hyok), fa * hy[r] - fb * ((ex1[r] - C.ex0[r]) * rdz - (ezR - ez[r]) * rdx), hy[r]);
                h1z[r] = sel((jh && hyok), fa * hz[r] - fb * ((eyR - C.ey0[r]) * rdx - (exN - C.ex0[r]) * rdy), hz[r]);
            }
        }
        // ---- exchange B: last row of H(n+1) goes down a wave ----
        xb[w][0][l] = h1z[R - 1];
        xb[w][1][l] = h1x[R - 1];
        __syncthreads();
        const float uH1z = xb[wu][0][l], uH1x = xb[wu][1][l];

        // ---- E of step n+1 at plane k ----
        float e1x[R], e1y[R], e1z[R];
        {
            const bool kin = k >= 1 && k <= nz - 2;
            const bool kz = k <= nz - 2;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int j = jb + r;
                const float hzU = r > 0 ? h1z[r - 1] : uH1z;
                const float hxU = r > 0 ? h1x[r - 1] : uH1x;
                const float hzL = __shfl_up(h1z[r], 1, LANES);
                const float hyL = __shfl_up(h1y[r], 1, LANES);
                const bool yin = j >= 1 && j <= ny - 2;
                const float ca = ca0[r], cb = cb0[r];
                float ex = sel((kin && yin && xe), ca * C.ex0[r] + cb * ((h1z[r] - hzU) * rdy - (h1y[r] - C.h1y[r]) * rdz), C.ex0[r]);
                float ey = sel((kin && j <= ny - 2 && xin), ca * C.ey0[r] + cb * ((h1x[r] - C.h1x[r]) * rdz - (h1z[r] - hzL) * rdx), C.ey0[r]);
                float ezz = sel((kz && yin && xin), ca * ez[r] + cb * ((h1y[r] - hyL) * rdx - (h1x[r] - hxU) * rdy), ez[r]);
                if (src.n) add_sources<0>(ex, ey, ezz, i, j, k, src);
                e1x[r] = ex;
                e1y[r] = ey;
                e1z[r] = ezz;
            }
        }
        // ---- monitor samples of step n+1 (owned cells inside the staging box) ----
        if (stg.p1 && k >= k0 && k < k1 && lane_out) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int j = jb + r;
                const int a = i - stg.x0, b = j - stg.y0, cc = k - stg.z0;
                if (j >= J0 + HALO && j < J0 + T - HALO && j < ny && a >= 0 && a < stg.sx && b >= 0 && b < stg.sy &&
                    cc >= 0 && cc < stg.sz) {
                    float* q = stg.p1 + ((long)cc * stg.sy + b) * stg.sx + a;
                    q[0] = e1x[r];
                    q[splane] = e1y[r];
                    q[2 * splane] = e1z[r];
                    q[3 * splane] = h1x[r];
                    q[4 * splane] = h1y[r];
                    q[5 * splane] = h1z[r];
                }
            }
        }

        // ---- H of step n+2 at plane k-1 ----
        float h2x[R], h2y[R], h2z[R];
        {
            const int m = k - 1;
            const bool kh = m < nz - 1;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int j = jb + r;
                const float ezN = r + 1 < R ? C.e1z[r + 1] : dE1z;
                const float exN = r + 1 < R ? C.e1x[r + 1] : dE1x;
                const float ezR = __shfl_down(C.e1z[r], 1, LANES);
                const float eyR = __shfl_down(C.e1y[r], 1, LANES);
                const float fa = C.fa1[r], fb = C.fb1[r];
                const bool jh = j < ny - 1;
                h2x[r] = sel((kh && jh), fa * C.h1x[r] - fb * ((ezN - C.e1z[r]) * rdy - (e1y[r] - C.e1y[r]) * rdz), C.h1x[r]);
                h2y[r] = sel((kh && hyok), fa * C.h1y[r] - fb * ((e1x[r] - C.e1x[r]) * rdz - (ezR - C.e1z[r]) * rdx), C.h1y[r]);
                h2z[r] = sel((jh && hyok), fa * C.h1z[r] - fb * ((eyR - C.e1y[r]) * rdx - (exN - C.e1x[r]) * rdy), C.h1z[r]);
            }
        }
        // ---- E of step n+2 at plane k-2 (every cell: later steps read it) ----
        float e2x[R], e2y[R], e2z[R];
        {
            const int m = k - 2;
            const bool kin = m >= 1 && m <= nz - 2;
            const bool kz = m <= nz - 2;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int j = jb + r;
                const float hzU = r > 0 ? C.h2z[r - 1] : uH2z;
                const float hxU = r > 0 ? C.h2x[r - 1] : uH2x;
                const float hzL = __shfl_up(C.h2z[r], 1, LANES);
                const float hyL = __shfl_up(C.h2y[r], 1, LANES);
                const bool yin = j >= 1 && j <= ny - 2;
                const float ca = C.ca2[r], cb = C.cb2[r];
                float ex = sel((kin && yin && xe), ca * C.f1x[r] + cb * ((C.h2z[r] - hzU) * rdy - (C.h2y[r] - C.g2y[r]) * rdz), C.f1x[r]);
                float ey = sel((kin && j <= ny - 2 && xin), ca * C.f1y[r] + cb * ((C.h2x[r] - C.g2x[r]) * rdz - (C.h2z[r] - hzL) * rdx), C.f1y[r]);
                float ezz = sel((kz && yin && xin), ca * C.f1z[r] + cb * ((C.h2y[r] - hyL) * rdx - (C.h2x[r] - hxU) * rdy), C.f1z[r]);
                if (src.n) add_sources<1>(ex, ey, ezz, i, j, m, src);
                e2x[r] = ex;
                e2y[r] = ey;
                e2z[r] = ezz;
            }
        }
        // ---- monitor samples of step n+2 at plane k-2 ----
        if (stg.p2 && k - 2 >= k0 && k - 2 < k1 && lane_out) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int j = jb + r;
                const int a = i - stg.x0, b = j - stg.y0, cc = k - 2 - stg.z0;
                if (j >= J0 + HALO && j < J0 + T - HALO && j < ny && a >= 0 && a < stg.sx && b >= 0 && b < stg.sy &&
                    cc >= 0 && cc < stg.sz) {
                    float* q = stg.p2 + ((long)cc * stg.sy + b) * stg.sx + a;
                    q[0] = e2x[r];
                    q[splane] = e2y[r];
                    q[2 * splane] = e2z[r];
                    q[3 * splane] = C.h2x[r];
                    q[4 * splane] = C.h2y[r];
                    q[5 * splane] = C.h2z[r];
                }
            }
        }

        // ---- H of step n+3 at plane k-3 ----
        float h3x[R], h3y[R], h3z[R];
        {
            const int m = k - 3;
            const bool kh = m < nz - 1;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int j = jb + r;
                const float ezN = r + 1 < R ? C.p2z[r + 1] : dE2z;
                const float exN = r + 1 < R ? C.p2x[r + 1] : dE2x;
                const float ezR = __shfl_down(C.p2z[r], 1, LANES);
                const float eyR = __shfl_down(C.p2y[r], 1, LANES);
                const float fa = C.fa3[r], fb = C.fb3[r];
                const bool jh = j < ny - 1;
                h3x[r] = sel((kh && jh), fa * C.g2x[r] - fb * ((ezN - C.p2z[r]) * rdy - (e2y[r] - C.p2y[r]) * rdz), C.g2x[r]);
                h3y[r] = sel((kh && hyok), fa * C.g2y[r] - fb * ((e2x[r] - C.p2x[r]) * rdz - (ezR - C.p2z[r]) * rdx), C.g2y[r]);
                h3z[r] = sel((jh && hyok), fa * C.g2z[r] - fb * ((eyR - C.p2y[r]) * rdx - (exN - C.p2x[r]) * rdy), C.g2z[r]);
            }
        }
        // ---- E of step n+3 at plane m = k-4, stores of plane m ----
        {
            const int m = k - 4;
            const bool kin = m >= 1 && m <= nz - 2;
            const bool kz = m <= nz - 2;
            const bool kst = m >= k0 && m < k1;
            const uint32_t mp = (uint32_t)max(m, 0) * ps;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int j = jb + r;
                const float hzU = r > 0 ? C.h3z[r - 1] : uH3z;
                const float hxU = r > 0 ? C.h3x[r - 1] : uH3x;
                const float hzL = __shfl_up(C.h3z[r], 1, LANES);
                const float hyL = __shfl_up(C.h3y[r], 1, LANES);
                const bool yin = j >= 1 && j <= ny - 2;
                const float ca = C.ca4[r], cb = C.cb4[r];
                float ex = sel((kin && yin && xe), ca * C.q2x[r] + cb * ((C.h3z[r] - hzU) * rdy - (C.h3y[r] - C.g3y[r]) * rdz), C.q2x[r]);
                float ey = sel((kin && j <= ny - 2 && xin), ca * C.q2y[r] + cb * ((C.h3x[r] - C.g3x[r]) * rdz - (C.h3z[r] - hzL) * rdx), C.q2y[r]);
                float ezz = sel((kz && yin && xin), ca * C.q2z[r] + cb * ((C.h3y[r] - hyL) * rdx - (C.h3x[r] - hxU) * rdy), C.q2z[r]);
                if (kst && lane_out && j >= J0 + HALO && j < J0 + T - HALO && j < ny) {
                    if (src.n) add_sources<2>(ex, ey, ezz, i, j, m, src);
                    const uint32_t o = mp + ro[r];
                    put(D.p[0], o, ex);
                    put(D.p[1], o, ey);
                    put(D.p[2], o, ezz);
                    put(D.p[3], o, C.h3x[r]);
                    put(D.p[4], o, C.h3y[r]);
                    put(D.p[5], o, C.h3z[r]);
                }
            }
        }
        // ---- rotate the carried planes ----
#pragma unroll
        for (int r = 0; r < R; ++r) {
            C.g3x[r] = C.h3x[r];
            C.g3y[r] = C.h3y[r];
            C.h3x[r] = h3x[r];
            C.h3y[r] = h3y[r];
            C.h3z[r] = h3z[r];
            C.q2x[r] = C.p2x[r];
            C.q2y[r] = C.p2y[r];
            C.q2z[r] = C.p2z[r];
            C.p2x[r] = e2x[r];
            C.p2y[r] = e2y[r];
            C.p2z[r] = e2z[r];
            C.g2x[r] = C.h2x[r];
            C.g2y[r] = C.h2y[r];
            C.g2z[r] = C.h2z[r];
            C.h2x[r] = h2x[r];
            C.h2y[r] = h2y[r];
            C.h2z[r] = h2z[r];
            C.f1x[r] = C.e1x[r];
            C.f1y[r] = C.e1y[r];
            C.f1z[r] = C.e1z[r];
            C.e1x[r] = e1x[r];
            C.e1y[r] = e1y[r];
            C.e1z[r] = e1z[r];
            C.h1x[r] = h1x[r];
            C.h1y[r] = h1y[r];
            C.h1z[r] = h1z[r];
            C.ca4[r] = C.ca3[r];
            C.cb4[r] = C.cb3[r];
            C.fa3[r] = C.fa2[r];
            C.fb3[r] = C.fb2[r];
            C.ca3[r] = C.ca2[r];
            C.cb3[r] = C.cb2[r];
            C.fa2[r] = C.fa1[r];
            C.fb2[r] = C.fb1[r];
            C.ca2[r] = C.ca1[r];
            C.cb2[r] = C.cb1[r];
            C.fa1[r] = fa0[r];
            C.fb1[r] = fb0[r];
            C.ca1[r] = ca0[r];
            C.cb1[r] = cb0[r];
            C.ex0[r] = ex1[r];
            C.ey0[r] = ey1[r];
        }
        // No barrier here: every wave read the exchange-A slots before barrier B and the exchange-B slots
        // before it can reach the next barrier A, so the next iteration's writes cannot race those reads.
    }
}

template <int R, int W>
int launch(const Fields6& S, const Fields6& D, const float* eps, const float* sig, int nx, int ny, int nz, int ld,
           int kseg, const Coef& c, float rdx, float rdy, float rdz, const Boxes3& bs, const Stage3& stg,
           hipStream_t st) {
    constexpr int T = W * R;
    const int gx = (nx + OXT - 1) / OXT;
    const int gy = (ny + (T - 2 * HALO) - 1) / (T - 2 * HALO);
    const int gz = (nz + kseg - 1) / kseg;
    const int n = gx * gy * gz;
    const int per_xcd = (n + NXCD - 1) / NXCD;
    auto kern = k_yee3d_tb3<R, W>;
    hipLaunchKernelGGL(kern, dim3(per_xcd * NXCD), dim3(LANES, W), 0, st, S, D, eps, sig, nx, ny, nz, ld, kseg, gx,
                       gy, gz, per_xcd, c, rdx, rdy, rdz, bs, stg);
    BZ_CHECK_LAUNCH("bz_yee3d_step3");
    return BZ_OK;
}

}  // namespace

extern "C" int bz_yee3d_step3(float* const src[6], float* const dst[6], const float* eps_r, const float* sigma,
                              int nx, int ny, int nz, int ld, double dx, double dy, double dz, double dt, double mu0,
                              double eps0, const bz_box_source* boxes, const float* amp1, const float* amp2, int nbox,
                              float* stage1, float* stage2, const int* stage_box, int kseg, int rows, int waves,
                              void* stream) {
    BZ_REQUIRE(src && dst && eps_r && sigma, "bz_yee3d_step3: null pointer");
    for (int q = 0; q < 6; ++q)
        BZ_REQUIRE(src[q] && dst[q] && src[q] != dst[q], "bz_yee3d_step3: bad field buffers");
    BZ_REQUIRE(nx >= 2 && ny >= 2 && nz >= 2 && ld >= nx, "bz_yee3d_step3: bad shape");
    BZ_REQUIRE(4.0 * nz * ny * ld < 4294967295.0, "bz_yee3d_step3: a field array exceeds 4 GiB (32-bit offsets)");
    BZ_REQUIRE(nbox >= 0 && nbox <= BZ_MAX_BOX && (nbox == 0 || (boxes && amp1 && amp2)),
               "bz_yee3d_step3: bad source list");
    Boxes3 bs;
    bs.n = nbox;
    for (int q = 0; q < nbox; ++q) {
        const bz_box_source& b = boxes[q];
        BZ_REQUIRE(b.comp >= 0 && b.comp <= 2 && b.add && b.sx > 0 && b.sy > 0 && b.sz > 0 && b.x0 >= 0 &&
                       b.y0 >= 0 && b.z0 >= 0 && b.x0 + b.sx <= nx && b.y0 + b.sy <= ny && b.z0 + b.sz <= nz,
                   "bz_yee3d_step3: source box outside the grid");
        bs.b[q] = b;
        bs.amp1[q] = amp1[q];
        bs.amp2[q] = amp2[q];
    }
    Stage3 stg;
    stg.p1 = stage1;
    stg.p2 = stage2;
    stg.x0 = stg.y0 = stg.z0 = stg.sx = stg.sy = stg.sz = 0;
    if (stage1 || stage2) {
        BZ_REQUIRE(stage_box, "bz_yee3d_step3: staging buffer without its box");
        stg.x0 = stage_box[0];
        stg.y0 = stage_box[1];
        stg.z0 = stage_box[2];
        stg.sx = stage_box[3];
        stg.sy = stage_box[4];
        stg.sz = stage_box[5];
        BZ_REQUIRE(stg.sx > 0 && stg.sy > 0 && stg.sz > 0 && stg.x0 >= 0 && stg.y0 >= 0 && stg.z0 >= 0 &&
                       stg.x0 + stg.sx <= nx && stg.y0 + stg.sy <= ny && stg.z0 + stg.sz <= nz,
                   "bz_yee3d_step3: staging box outside the grid");
    }
    if (kseg <= 0) kseg = 32;
    BZ_REQUIRE(kseg >= 2, "bz_yee3d_step3: kseg must be at least 2");
    Fields6 S, D;
    for (int q = 0; q < 6; ++q) {
        S.p[q] = src[q];
        D.p[q] = dst[q];
    }
    const Coef c = make_coef(dt, mu0, eps0);
    const float rx = (float)(1.0 / dx), ry = (float)(1.0 / dy), rz = (float)(1.0 / dz);
    const hipStream_t st = (hipStream_t)stream;
    if (waves == 8 && rows == 2) return launch<2, 8>(S, D, eps_r, sigma, nx, ny, nz, ld, kseg, c, rx, ry, rz, bs, stg, st);
    if (waves == 8 && rows == 3) return launch<3, 8>(S, D, eps_r, sigma, nx, ny, nz, ld, kseg, c, rx, ry, rz, bs, stg, st);
    if (waves == 16 && rows == 1) return launch<1, 16>(S, D, eps_r, sigma, nx, ny, nz, ld, kseg, c, rx, ry, rz, bs, stg, st);
    if (waves == 16 && rows == 2) return launch<2, 16>(S, D, eps_r, sigma, nx, ny, nz, ld, kseg, c, rx, ry, rz, bs, stg, st);
    bz::set_error("bz_yee3d_step3: (rows, waves) must be (2, 8), (3, 8), (1, 16) or (2, 16)");
    return BZ_EINVAL;
}
