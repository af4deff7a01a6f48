// Reverse-time (adjoint) 2D Yee step for gfx950: the exact transpose of the
// forward step of yee2d_fused.hip, batched over many simulations.
//
// No reference counterpart: the reference's topology example estimates the
// gradient with a time-reversed overlap of two forward runs
// (examples/4_topo.py:282-285). This kernel computes the gradient of the
// discrete forward map itself. For forward step n (numpy_backend.py:34-88 and
// an additive or multiplicative Ez source box), given the adjoint state
// lam_{n+1} = (lE, lHx, lHy) it forms lam_n:
//
//   lE''  = M lE                    (M: the source box's multiplier, 1 if none)
//   w     = cb lE''  on the interior, 0 elsewhere
//   muX   = lHx + (w(j,i+1) - w(j,i)) / dy        (Hx entries i <= nx-2)
//   muY   = lHy + (w(j,i) - w(j+1,i)) / dx        (Hy entries j <= ny-2)
//   lE_n  = (interior ? ca : 1) lE'' + (fb muX(j,i) - fb muX(j,i-1)) / dy
//                                    + (fb muY(j-1,i) - fb muY(j,i)) / dx
//   lHx_n = fa muX,  lHy_n = fa muY
//   lE_n(p) += sum_f Re(w_fp phasor[n-1][f])  (objective: DFTs of the Ez sample after step n-1)
//
// and accumulates, inside the design box, dJ/deps += lE'' (dca/deps Ez_n +
// dcb/deps curlH_{n+1}) from the forward snapshot that bz_yee2d_run stored
// while recomputing the segment (the "x dy" quirk of the reference, an
// x-difference divided by dy, is transposed as it is). The dependency pattern
// is the forward one mirrored (w like Ez, mu like H), so the tiling is the
// forward engine's: waves of 64 lanes (lanes 0 and 63 halo), SEG rows per
// tile, all loads of a tile issued at once, ping-pong adjoint buffers
// (reverse step n reads parity (n+1)&1 and writes n&1).
// HBM traffic per cell: lE, lHx, lHy, eps, sigma read + 3 written = 32 B, as
// the forward step.
#include "common.h"

// No multiply-add contraction: the one-step and the two-step kernels round every operation alike, so a
// pair of reverse steps gives the bits of two single steps.
#pragma clang fp contract(off)

namespace {

constexpr int LANES = 64;
constexpr int OX = LANES - 2;
constexpr int OX2 = LANES - 4;   // stored columns of a two-step tile (two halo lanes on each side)
constexpr int WPG = 4;

__device__ __forceinline__ float at(const float* __restrict__ base, uint32_t byte_off) {
    return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + byte_off);
}
__device__ __forceinline__ void put(float* __restrict__ base, uint32_t byte_off, float v) {
    *reinterpret_cast<float*>(reinterpret_cast<char*>(base) + byte_off) = v;
}

// dJ/dEz of the sample after step n-1 at injection cell p: sum_f Re(inj_w[f][p] * phasor[n-1][f]).
__device__ __forceinline__ double injection(const bz_adj2d& A, int p, int n) {
    const long cells = (long)A.inj_sx * A.inj_sy;
    const double* ph = A.phasor + 2 * (long)(n - 1) * A.nf;
    double s = 0.0;
    for (int f = 0; f < A.nf; ++f) {
        const double* wp = A.inj_w + 2 * (f * cells + p);
        s += wp[0] * ph[2 * f] - wp[1] * ph[2 * f + 1];
    }
    return s;
}

// Source-box multiplier of the forward step at cell (j, i) (1 outside the box or without one).
__device__ __forceinline__ float src_mul_at(const bz_adj2d& A, int i, int j) {
    if (A.src_mul) {
        const int a = i - A.src_x0, b = j - A.src_y0;
        if (a >= 0 && a < A.src_sx && b >= 0 && b < A.src_sy) return A.src_mul[(long)b * A.src_sx + a];
    }
    return 1.0f;
}

// Injection of the objective into lE at cell (j, i) before reverse step n's output is stored.
__device__ __forceinline__ void inject_at(const bz_adj2d& A, float& lnew, int i, int j, int n) {
    if (A.inj_w && n >= 1) {
        const int a = i - A.inj_x0, b = j - A.inj_y0;
        if (a >= 0 && a < A.inj_sx && b >= 0 && b < A.inj_sy) lnew += (float)injection(A, b * A.inj_sx + a, n);
    }
}

// dJ/deps of reverse step n at an interior cell (j, i) of the design box: l2 (dca/deps Ez_n + dcb/deps curlH),
// from the forward snapshot of step n (slot n - snap_step0; the snapshot box is the design box grown by one
// cell on the low x and y sides).
// With l2hist set, only l2 is stored (slot n - snap_step0) and k_adj2d_grad adds the term later.
__device__ __forceinline__ void grad_at(const bz_adj2d& A, const Coef& c, int i, int j, int n, float l2, float ee,
                                        float sg, float cb, float rdx, float rdy) {
    const int a = i - A.box_x0, b = j - A.box_y0;
    if (a >= 0 && a < A.box_sx && b >= 0 && b < A.box_sy && A.l2hist) {
        A.l2hist[(long)(n - A.snap_step0) * A.box_sx * A.box_sy + (long)b * A.box_sx + a] = l2;
        return;
    }
    if (a >= 0 && a < A.box_sx && b >= 0 && b < A.box_sy) {
        const int sx = A.box_sx + 1, sy = A.box_sy + 1;
        const long plane = (long)sx * sy;
        const float* p = A.snap + (long)(n - A.snap_step0) * 3 * plane + (long)(b + 1) * sx + (a + 1);
        const float en = p[0];
        const float* hx = p + plane;
        const float* hy = p + 2 * plane;
        const float curl = -((hx[0] - hx[-1]) * rdy) + (hy[0] - hy[-sx]) * rdx;
        const float e = ee;
        const float s = sg * c.ke / e;
        const float dca = 2.0f * s / (e * (1.0f + s) * (1.0f + s));
        const float dcb = -cb / (e * (1.0f + s));
        A.grad[(long)b * A.box_sx + a] += (double)l2 * ((double)dca * en + (double)dcb * curl);
    }
}

template <int SEG>
__global__ __launch_bounds__(LANES* WPG) void k_adj2d(const bz_adj2d* __restrict__ adjs, const int* __restrict__ tiles,
                                                     int ntiles, int n, int flip) {
    const int t = __builtin_amdgcn_readfirstlane(blockIdx.x * WPG + threadIdx.y);
    if (t >= ntiles) return;
    const int rd = (n + 1 + flip) & 1, wr = (n + flip) & 1;
    const int s = tiles[3 * t], tx = tiles[3 * t + 1], ty = tiles[3 * t + 2];
    if (s < 0) return;   // padding entry of a reordered tile list
    const bz_adj2d& A = adjs[s];
    const int nx = A.nx, ny = A.ny, ld = A.ld;
    const int l = threadIdx.x;
    const int i = tx * OX - 1 + l;
    const int j0 = ty * SEG;
    Coef c;
    c.kh = A.kh;
    c.kbh = A.kbh;
    c.ke = A.ke;
    c.kbe = A.kbe;
    const float rdx = A.rdx, rdy = A.rdy;
    const uint32_t ic = (uint32_t)min(max(i, 0), nx - 1);
    const bool xin = i >= 1 && i <= nx - 2;

    uint32_t ro[SEG + 2];
#pragma unroll
    for (int r = 0; r < SEG + 2; ++r) {
        const int j = min(max(j0 - 1 + r, 0), ny - 1);
        ro[r] = ((uint32_t)j * (uint32_t)ld + ic) * 4u;
    }
    const float* __restrict__ gle = A.le[rd];
    const float* __restrict__ glx = A.lhx[rd];
    const float* __restrict__ gly = A.lhy[rd];
    float le[SEG + 2], ee[SEG + 2], sg[SEG + 2], lx[SEG + 1], ly[SEG + 1];
#pragma unroll
    for (int r = 0; r < SEG + 2; ++r) {
        le[r] = at(gle, ro[r]);
        ee[r] = at(A.eps, ro[r]);
        sg[r] = at(A.sig, ro[r]);
        if (r < SEG + 1) {
            ly[r] = at(gly, ro[r]);
            lx[r] = r > 0 ? at(glx, ro[r]) : 0.0f;
        }
    }
    // lE'' and w on rows j0-1 .. j0+SEG (the E coefficients of rows 1..SEG are kept for the update below)
    float l2[SEG + 2], w[SEG + 2], ca_[SEG + 2], cb_[SEG + 2];
#pragma unroll
    for (int r = 0; r < SEG + 2; ++r) {
        const int j = j0 - 1 + r;
        l2[r] = src_mul_at(A, i, j) * le[r];
        e_coef(sg[r], ee[r], c, ca_[r], cb_[r]);
        w[r] = (xin && j >= 1 && j <= ny - 2) ? cb_[r] * l2[r] : 0.0f;
    }
    // fb muY on rows j0-1 .. j0+SEG-1, fb muX on rows j0 .. j0+SEG-1 (zero where the H entry does not exist)
    float fmy[SEG + 1], fmx[SEG + 1], fmxl[SEG + 1], my[SEG + 1], mx[SEG + 1], fa_[SEG + 1];
    const bool hxok = i <= nx - 2;
#pragma unroll
    for (int r = 0; r < SEG + 1; ++r) {
        const int j = j0 - 1 + r;
        float fa, fb;
        h_coef(sg[r], c, fa, fb);
        fa_[r] = fa;
        my[r] = (j >= 0 && j <= ny - 2) ? ly[r] + (w[r] - w[r + 1]) * rdx : 0.0f;
        fmy[r] = fb * my[r];
        const float wR = __shfl_down(w[r], 1, LANES);
        mx[r] = hxok ? lx[r] + (wR - w[r]) * rdy : 0.0f;
        fmx[r] = fb * mx[r];
        const float left = __shfl_up(fmx[r], 1, LANES);
        fmxl[r] = i >= 1 ? left : 0.0f;
    }

    const bool out = l > 0 && l < LANES - 1 && i < nx;
    float* __restrict__ ole = A.le[wr];
    float* __restrict__ olx = A.lhx[wr];
    float* __restrict__ oly = A.lhy[wr];
#pragma unroll
    for (int r = 1; r < SEG + 1; ++r) {
        const int j = j0 - 1 + r;
        if (out && j < ny) {
            const float ca = ca_[r], cb = cb_[r];
            const bool in = xin && j >= 1 && j <= ny - 2;
            float lnew = (in ? ca * l2[r] : l2[r]) + (fmx[r] - fmxl[r]) * rdy + (fmy[r - 1] - fmy[r]) * rdx;
            inject_at(A, lnew, i, j, n);
            put(ole, ro[r], lnew);
            put(olx, ro[r], fa_[r] * mx[r]);
            put(oly, ro[r], fa_[r] * my[r]);
            if (A.grad && in) grad_at(A, c, i, j, n, l2[r], ee[r], sg[r], cb, rdx, rdy);
        }
    }
}

// Reverse steps n and n-1 in one pass over a tile of 60 stored columns x seg rows, marching down the rows.
// Row r of the march computes reverse step n on row r (it needs lam_{n+1} on rows r and r+1: the lE,
// sigma and eps of row r+1 are loaded one iteration ahead) and reverse step n-1 on row r-1 (it needs
// lam_n on rows r-2 .. r, i.e. this row's step-n output and the previous rows' ones carried in registers).
// The march starts two rows above the tile. Step n is valid on lanes 1..62 and step n-1 on lanes 2..61,
// the stored ones. The buffers: lam_{n+1} is read from buffer (n + 1 + flip) & 1 and lam_{n-1} written to
// the other one (a pair moves the state across buffers, so the caller toggles flip after each pair). Same
// operations in the same order as two k_adj2d steps, gradient contributions of step n before those of n-1.
__global__ __launch_bounds__(LANES* WPG) void k_adj2d_tb2(const bz_adj2d* __restrict__ adjs,
                                                         const int* __restrict__ tiles, int ntiles, int seg, int n,
                                                         int flip) {
    const int t = __builtin_amdgcn_readfirstlane(blockIdx.x * WPG + threadIdx.y);
    if (t >= ntiles) return;
    const int s = tiles[3 * t];
    if (s < 0) return;   // padding entry of a reordered tile list
    const bz_adj2d& A = adjs[s];
    const int tx = tiles[3 * t + 1], ty = tiles[3 * t + 2];
    const int nx = A.nx, ny = A.ny, ld = A.ld;
    const int l = threadIdx.x;
    const int i = tx * OX2 - 2 + l;
    const uint32_t ic = (uint32_t)min(max(i, 0), nx - 1);
    const int j0 = ty * seg, j1 = min(ny, j0 + seg);
    const int rb = (n + 1 + flip) & 1, wb = rb ^ 1;
    const float* __restrict__ gle = A.le[rb];
    const float* __restrict__ glx = A.lhx[rb];
    const float* __restrict__ gly = A.lhy[rb];
    float* __restrict__ ole = A.le[wb];
    float* __restrict__ olx = A.lhx[wb];
    float* __restrict__ oly = A.lhy[wb];
    Coef c;
    c.kh = A.kh;
    c.kbh = A.kbh;
    c.ke = A.ke;
    c.kbe = A.kbe;
    const float rdx = A.rdx, rdy = A.rdy;
    const bool xin = i >= 1 && i <= nx - 2;
    const bool hxok = i <= nx - 2;
    const bool out = l >= 2 && l <= LANES - 3 && i < nx;
    const int rstart = j0 > 0 ? j0 - 2 : 0;
    auto off = [&](int r) { return ((uint32_t)min(max(r, 0), ny - 1) * (uint32_t)ld + ic) * 4u; };
    auto interior = [&](int j) { return xin && j >= 1 && j <= ny - 2; };

    // lam_{n+1}-side quantities of the current row r: l2 = M lE, E coefficients, w = cb l2 on the interior
    float sg0 = at(A.sig, off(rstart)), ee0 = at(A.eps, off(rstart));
    float l2_0 = src_mul_at(A, i, rstart) * at(gle, off(rstart));
    float ca0, cb0;
    e_coef(sg0, ee0, c, ca0, cb0);
    float w0 = interior(rstart) ? cb0 * l2_0 : 0.0f;
    // loads one row ahead: lE, sigma, eps of row r+1 and lHx, lHy of row r
    float nle = at(gle, off(rstart + 1)), nsg = at(A.sig, off(rstart + 1)), nee = at(A.eps, off(rstart + 1));
    float nlx = at(glx, off(rstart)), nly = at(gly, off(rstart));
    // carried from row r-1: fb muY of step n; step n's outputs there and its cell's coefficients for step n-1
    float fmy_p = 0.0f;
    float l2b_p = 0.0f, wb_p = 0.0f, lxn_p = 0.0f, lyn_p = 0.0f, fmyb_p = 0.0f;
    float fa_p = 0.0f, fb_p = 0.0f, ca_p = 1.0f, cb_p = 0.0f, sg_p = 0.0f, ee_p = 1.0f;
    for (int r = rstart; r <= j1; ++r) {
        const float le1 = nle, sg1 = nsg, ee1 = nee, lx0 = nlx, ly0 = nly;
        {
            const uint32_t o2 = off(r + 2), o1 = off(r + 1);
            nle = at(gle, o2);
            nsg = at(A.sig, o2);
            nee = at(A.eps, o2);
            nlx = at(glx, o1);
            nly = at(gly, o1);
        }
        // lam_{n+1}-side quantities of row r+1
        const float l2_1 = src_mul_at(A, i, r + 1) * le1;
        float ca1, cb1;
        e_coef(sg1, ee1, c, ca1, cb1);
        const float w1 = interior(r + 1) ? cb1 * l2_1 : 0.0f;

        // ---- reverse step n on row r ----
        float fa0, fb0;
        h_coef(sg0, c, fa0, fb0);
        const float wR = __shfl_down(w0, 1, LANES);
        const float mx = hxok ? lx0 + (wR - w0) * rdy : 0.0f;
        const float my = (r >= 0 && r <= ny - 2) ? ly0 + (w0 - w1) * rdx : 0.0f;
        const float fmx = fb0 * mx, fmy = fb0 * my;
        const float left = __shfl_up(fmx, 1, LANES);
        const float fmxl = i >= 1 ? left : 0.0f;
        const bool in0 = interior(r);
        float lnew = (in0 ? ca0 * l2_0 : l2_0) + (fmx - fmxl) * rdy + (fmy_p - fmy) * rdx;
        inject_at(A, lnew, i, r, n);
        const float lxn = fa0 * mx, lyn = fa0 * my;
        if (A.grad && out && in0 && r >= j0 && r < j1) grad_at(A, c, i, r, n, l2_0, ee0, sg0, cb0, rdx, rdy);
        // lam_n-side quantities of row r for step n-1 (same cell coefficients)
        const float l2b = src_mul_at(A, i, r) * lnew;
        const float wbv = in0 ? cb0 * l2b : 0.0f;

        // ---- reverse step n-1 on row m = r-1 ----
        const int m = r - 1;
        const float wbR = __shfl_down(wb_p, 1, LANES);
        const float mx2 = hxok ? lxn_p + (wbR - wb_p) * rdy : 0.0f;
        const float my2 = (m >= 0 && m <= ny - 2) ? lyn_p + (wb_p - wbv) * rdx : 0.0f;
        const float fmx2 = fb_p * mx2, fmy2 = fb_p * my2;
        const float left2 = __shfl_up(fmx2, 1, LANES);
        const float fmxl2 = i >= 1 ? left2 : 0.0f;
        const bool inm = interior(m);
        float lnew2 = (inm ? ca_p * l2b_p : l2b_p) + (fmx2 - fmxl2) * rdy + (fmyb_p - fmy2) * rdx;
        if (out && m >= j0 && m < j1) {
            inject_at(A, lnew2, i, m, n - 1);
            const uint32_t om = off(m);
            put(ole, om, lnew2);
            put(olx, om, fa_p * mx2);
            put(oly, om, fa_p * my2);
            if (A.grad && inm) grad_at(A, c, i, m, n - 1, l2b_p, ee_p, sg_p, cb_p, rdx, rdy);
        }

        // carry: row r's cell quantities and step-n outputs become the previous row's, row r+1's the current
        ca_p = ca0;
        cb_p = cb0;
        fa_p = fa0;
        fb_p = fb0;
        sg_p = sg0;
        ee_p = ee0;
        l2b_p = l2b;
        wb_p = wbv;
        lxn_p = lxn;
        lyn_p = lyn;
        fmyb_p = fmy2;
        fmy_p = fmy;
        l2_0 = l2_1;
        w0 = w1;
        ca0 = ca1;
        cb0 = cb1;
        sg0 = sg1;
        ee0 = ee1;
    }
}


// ---------------- K reverse steps per pass ----------------
// The forward K-step scheme (yee2d_fused.hip) mirrored: a workgroup of KW waves holds a region of 64 columns
// x KW*B rows, wave w rows w*B .. w*B+B-1 in registers (lane l = column), and runs reverse steps n_top ..
// n_top-K+1 on the whole region; the cells within K of its edge go stale one ring per step, the inner ones
// are stored. w = cb lE'' plays Ez (muY needs w of the next row, muX w of the next column) and fb mu plays H
// (lE needs fb muY of the previous row, fb muX of the previous column): right and left neighbours by DPP
// wave shifts, band edges through LDS. Each inner cell is owned by one tile, which injects the objective
// and accumulates its gradient, step by step in the one-step order. Reads buffer rb, writes rb ^ 1.
__device__ __forceinline__ float from_right(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x130, 0xf, 0xf, false));
}
__device__ __forceinline__ float from_left(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x138, 0xf, 0xf, false));
}

template <int B, int KW>
__global__ __launch_bounds__(LANES* KW) void k_adj2d_tbk(const bz_adj2d* __restrict__ adjs,
                                                        const int* __restrict__ tiles, int ntiles, int K, int n_top,
                                                        int rb) {
    constexpr int RH = KW * B;
    __shared__ float w_top[KW][LANES];
    __shared__ float fmy_bot[KW][LANES];
    __shared__ int ibox[4];     // the injection box (x0, y0, sx, sy), read after the first barrier
    const int t = blockIdx.x;
    if (t >= ntiles) return;
    const int s = tiles[3 * t];
    if (s < 0) return;   // padding entry of a reordered tile list (uniform over the workgroup)
    const bz_adj2d& A = adjs[s];
    const int tx = tiles[3 * t + 1], ty = tiles[3 * t + 2];
    const int nx = A.nx, ny = A.ny, ld = A.ld;
    const int l = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int rx0 = tx * (LANES - 2 * K) - K;
    const int i = rx0 + l;
    const int jb = ty * (RH - 2 * K) - K + w * B;
    const uint32_t ic = (uint32_t)min(max(i, 0), nx - 1);
    Coef c;
    c.kh = A.kh;
    c.kbh = A.kbh;
    c.ke = A.ke;
    c.kbe = A.kbe;
    const float rdx = A.rdx, rdy = A.rdy;
    const bool xin = i >= 1 && i <= nx - 2;
    const bool hxok = i <= nx - 2;
    const bool lane_inner = l >= K && l < LANES - K;
    // uniform over the wave: does it meet the source box, the injection box, the design box (inner cells)
    const bool src_band = A.src_mul && jb < A.src_y0 + A.src_sy && jb + B > A.src_y0 && rx0 < A.src_x0 + A.src_sx &&
                          rx0 + LANES > A.src_x0;
    const bool inj_band = A.inj_w && jb < A.inj_y0 + A.inj_sy && jb + B > A.inj_y0 &&
                          rx0 < A.inj_x0 + A.inj_sx && rx0 + LANES > A.inj_x0;
    // with the injection table, the next step's injection values are loaded one step ahead (their latency
    // hides behind a whole step); the box comes from LDS, so it adds no registers
    const float* __restrict__ itab = A.inj_tab;
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        ibox[0] = A.inj_x0;
        ibox[1] = A.inj_y0;
        ibox[2] = A.inj_sx;
        ibox[3] = A.inj_sy;
    }
    float injv[B];
    auto inj_fetch = [&](int nn) {       // values of step nn (0 outside the box or for nn < 1)
#pragma unroll
        for (int r = 0; r < B; ++r) {
            const int a = i - ibox[0], bb = jb + r - ibox[1];
            injv[r] = (nn >= 1 && a >= 0 && a < ibox[2] && bb >= 0 && bb < ibox[3])
                          ? itab[(long)nn * ibox[2] * ibox[3] + (long)bb * ibox[2] + a]
                          : 0.0f;
        }
    };
    const bool grad_band = A.grad && jb < A.box_y0 + A.box_sy && jb + B > A.box_y0 &&
                           rx0 + K < A.box_x0 + A.box_sx && rx0 + LANES - K > A.box_x0;
    const bool inner = jb >= 1 && jb + B - 1 <= ny - 2 && rx0 >= 1 && rx0 + LANES - 1 <= nx - 2;

    float le[B], lx[B], ly[B], fa[B], fb[B], ca[B], cb[B];
    {
        const float* __restrict__ gle = A.le[rb];
        const float* __restrict__ glx = A.lhx[rb];
        const float* __restrict__ gly = A.lhy[rb];
        const float* __restrict__ cf = A.coef;
        uint32_t o[B];
#pragma unroll
        for (int r = 0; r < B; ++r) {
            o[r] = ((uint32_t)min(max(jb + r, 0), ny - 1) * (uint32_t)ld + ic) * 4u;
            le[r] = at(gle, o[r]);
            lx[r] = at(glx, o[r]);
            ly[r] = at(gly, o[r]);
        }
        if (cf) {
#pragma unroll
            for (int r = 0; r < B; ++r) {
                const float4 q = *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(cf) + 4u * o[r]);
                fa[r] = q.x;
                fb[r] = q.y;
                ca[r] = q.z;
                cb[r] = q.w;
            }
        } else {
#pragma unroll
            for (int r = 0; r < B; ++r) {
                const float sg = at(A.sig, o[r]);
                h_coef(sg, c, fa[r], fb[r]);
                e_coef(sg, at(A.eps, o[r]), c, ca[r], cb[r]);
            }
        }
    }

    for (int k = 0; k < K; ++k) {
        const int n = n_top - k;
        // lE'' = M lE (kept in le) and w = cb lE'' on the interior
        if (src_band) {
#pragma unroll
            for (int r = 0; r < B; ++r) le[r] = src_mul_at(A, i, jb + r) * le[r];
        }
        float wv[B];
#pragma unroll
        for (int r = 0; r < B; ++r) {
            const int j = jb + r;
            wv[r] = (inner || (xin && j >= 1 && j <= ny - 2)) ? cb[r] * le[r] : 0.0f;
        }
        // mu: muY of the band's last row needs w of the next band's first row
        w_top[w][l] = wv[0];
        __syncthreads();
        if (itab && inj_band && k == 0) inj_fetch(n);     // the first step's values (ibox is written by now)
        const float wb = w < KW - 1 ? w_top[w + 1][l] : wv[B - 1];
        float fmx[B], fmy[B];
#pragma unroll
        for (int r = 0; r < B; ++r) {
            const int j = jb + r;
            const float wR = from_right(wv[r]);
            const float wd = r < B - 1 ? wv[r + 1] : wb;
            const float mx = (inner || hxok) ? lx[r] + (wR - wv[r]) * rdy : 0.0f;
            const float my = (inner || (j >= 0 && j <= ny - 2)) ? ly[r] + (wv[r] - wd) * rdx : 0.0f;
            fmx[r] = fb[r] * mx;
            fmy[r] = fb[r] * my;
            lx[r] = fa[r] * mx;
            ly[r] = fa[r] * my;
        }
        // lE: the band's first row needs fb muY of the previous band's last row
        fmy_bot[w][l] = fmy[B - 1];
        __syncthreads();
        const float fu = w > 0 ? fmy_bot[w - 1][l] : fmy[0];
#pragma unroll
        for (int r = 0; r < B; ++r) {
            const int j = jb + r;
            const float left = from_left(fmx[r]);
            const float fmxl = (inner || i >= 1) ? left : 0.0f;
            const float fmyu = r > 0 ? fmy[r - 1] : fu;
            const bool in = inner || (xin && j >= 1 && j <= ny - 2);
            const float l2 = le[r];
            float lnew = (in ? ca[r] * l2 : l2) + (fmx[r] - fmxl) * rdy + (fmyu - fmy[r]) * rdx;
            if (inj_band) {
                if (itab) {
                    const int a = i - ibox[0], bb = j - ibox[1];
                    if (n >= 1 && a >= 0 && a < ibox[2] && bb >= 0 && bb < ibox[3]) lnew += injv[r];
                } else {
                    inject_at(A, lnew, i, j, n);
                }
            }
            const int rr = w * B + r;
            if (grad_band && in && lane_inner && rr >= K && rr < RH - K) {
                if (A.l2hist) {
                    const int a = i - A.box_x0, bb = j - A.box_y0;
                    if (a >= 0 && a < A.box_sx && bb >= 0 && bb < A.box_sy)
                        A.l2hist[(long)(n - A.snap_step0) * A.box_sx * A.box_sy + (long)bb * A.box_sx + a] = l2;
                } else {
                    const uint32_t o = ((uint32_t)min(max(j, 0), ny - 1) * (uint32_t)ld + ic) * 4u;
                    grad_at(A, c, i, j, n, l2, at(A.eps, o), at(A.sig, o), cb[r], rdx, rdy);
                }
            }
            le[r] = lnew;
        }
        if (itab && inj_band && k + 1 < K) inj_fetch(n - 1);   // the next step's values load during this one
    }

    // the inner cells hold lam_{n_top-K+1}
    if (!lane_inner || i >= nx) return;
    float* __restrict__ ole = A.le[rb ^ 1];
    float* __restrict__ olx = A.lhx[rb ^ 1];
    float* __restrict__ oly = A.lhy[rb ^ 1];
#pragma unroll
    for (int r = 0; r < B; ++r) {
        const int rr = w * B + r, j = jb + r;
        if (rr >= K && rr < RH - K && j < ny) {
            const uint32_t o = ((uint32_t)j * (uint32_t)ld + (uint32_t)i) * 4u;
            put(ole, o, le[r]);
            put(olx, o, lx[r]);
            put(oly, o, ly[r]);
        }
    }
}

// lE_n += Re(w_p phasor[n-1]) on the injection boxes: the terminal condition lam_N = dJ/du_N.
__global__ void k_adj_inject(const bz_adj2d* __restrict__ adjs, int n) {
    const bz_adj2d& A = adjs[blockIdx.y];
    if (!A.inj_w || n < 1) return;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= A.inj_sx * A.inj_sy) return;
    const int a = p % A.inj_sx, b = p / A.inj_sx;
    float* le = A.le[n & 1] + (long)(A.inj_y0 + b) * A.ld + (A.inj_x0 + a);
    *le += (float)injection(A, p, n);
}

template <int SEG>
int run(const bz_adj2d* adjs, const int* tiles, int ntiles, int n_hi, int nsteps, int flip, hipStream_t st) {
    const dim3 g((ntiles + WPG - 1) / WPG), b(LANES, WPG);
    auto kern = k_adj2d<SEG>;
    for (int q = 0; q < nsteps; ++q) {
        hipLaunchKernelGGL(kern, g, b, 0, st, adjs, tiles, ntiles, n_hi - 1 - q, flip);
        BZ_CHECK_LAUNCH("bz_adj2d_run");
    }
    return BZ_OK;
}

int run_seg(const bz_adj2d* adjs, const int* tiles, int ntiles, int seg, int n_hi, int nsteps, int flip,
            hipStream_t st) {
    switch (seg) {
        case 2: return run<2>(adjs, tiles, ntiles, n_hi, nsteps, flip, st);
        case 4: return run<4>(adjs, tiles, ntiles, n_hi, nsteps, flip, st);
        case 8: return run<8>(adjs, tiles, ntiles, n_hi, nsteps, flip, st);
        default: bz::set_error("bz_adj2d_run: seg must be 2, 4 or 8"); return BZ_EINVAL;
    }
}

}  // namespace

extern "C" int bz_adj2d_run(const bz_adj2d* adjs, int nadj, const int* tiles, int ntiles, int seg, int n_hi,
                            int nsteps, void* stream) {
    BZ_REQUIRE(adjs && nadj > 0 && tiles && ntiles > 0 && nsteps >= 0 && n_hi - nsteps >= 0,
               "bz_adj2d_run: bad arguments");
    return run_seg(adjs, tiles, ntiles, seg, n_hi, nsteps, 0, (hipStream_t)stream);
}

extern "C" int bz_adj2d_run2(const bz_adj2d* adjs, int nadj, const int* tiles2, int ntiles2, int seg2,
                             const int* tiles, int ntiles, int seg, int n_hi, int nsteps, int flip, void* stream) {
    BZ_REQUIRE(adjs && nadj > 0 && tiles2 && ntiles2 > 0 && seg2 >= 2 && nsteps >= 0 && n_hi - nsteps >= 0 &&
                   (flip == 0 || flip == 1),
               "bz_adj2d_run2: bad arguments");
    BZ_REQUIRE(nsteps % 2 == 0 || (tiles && ntiles > 0), "bz_adj2d_run2: an odd step count needs one-step tiles");
    const hipStream_t st = (hipStream_t)stream;
    const dim3 g((ntiles2 + WPG - 1) / WPG), b(LANES, WPG);
    const int npairs = nsteps / 2;
    for (int q = 0; q < npairs; ++q) {
        hipLaunchKernelGGL(k_adj2d_tb2, g, b, 0, st, adjs, tiles2, ntiles2, seg2, n_hi - 1 - 2 * q, flip);
        BZ_CHECK_LAUNCH("bz_adj2d_run2");
        flip ^= 1;
    }
    if (nsteps % 2) return run_seg(adjs, tiles, ntiles, seg, n_hi - 2 * npairs, 1, flip, st);
    return BZ_OK;
}

// The gradient terms of steps n_hi-1 .. n_lo of one design cell (x: cell of the box, y: adjoint), added in
// that order to grad: grad_at's expressions, with the coefficient derivatives computed once per cell.
__global__ __launch_bounds__(256) void k_adj2d_grad(const bz_adj2d* __restrict__ adjs, int n_lo, int n_hi) {
    const bz_adj2d& A = adjs[blockIdx.y];
    const int nbox = A.box_sx * A.box_sy;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nbox || !A.grad || !A.l2hist) return;
    const int b = q / A.box_sx, a = q - b * A.box_sx;
    const long o = (long)(A.box_y0 + b) * A.ld + (A.box_x0 + a);
    const float ee = A.eps[o], sg = A.sig[o];
    Coef c;
    c.kh = A.kh;
    c.kbh = A.kbh;
    c.ke = A.ke;
    c.kbe = A.kbe;
    float ca, cb;
    e_coef(sg, ee, c, ca, cb);
    const float e = ee;
    const float s = sg * c.ke / e;
    const float dca = 2.0f * s / (e * (1.0f + s) * (1.0f + s));
    const float dcb = -cb / (e * (1.0f + s));
    const float rdx = A.rdx, rdy = A.rdy;
    const int sx = A.box_sx + 1;
    const long plane = (long)sx * (A.box_sy + 1);
    const float* __restrict__ snap = A.snap + (long)(b + 1) * sx + (a + 1);
    const float* __restrict__ l2h = A.l2hist + q;
    double acc = A.grad[q];
#pragma unroll 8
    for (int n = n_hi - 1; n >= n_lo; --n) {
        const long slot = n - A.snap_step0;
        const float* p = snap + slot * 3 * plane;
        const float en = p[0];
        const float* hx = p + plane;
        const float* hy = p + 2 * plane;
        const float curl = -((hx[0] - hx[-1]) * rdy) + (hy[0] - hy[-sx]) * rdx;
        const float l2 = l2h[slot * nbox];
        acc += (double)l2 * ((double)dca * en + (double)dcb * curl);
    }
    A.grad[q] = acc;
}

extern "C" int bz_adj2d_grad(const bz_adj2d* adjs, int nadj, int max_box, int n_lo, int n_hi, void* stream) {
    BZ_REQUIRE(adjs && nadj > 0 && nadj <= 65535 && max_box >= 0 && n_lo >= 0 && n_hi >= n_lo,
               "bz_adj2d_grad: bad arguments");
    if (max_box == 0 || n_hi == n_lo) return BZ_OK;
    hipLaunchKernelGGL(k_adj2d_grad, dim3((max_box + 255) / 256, nadj), dim3(256), 0, (hipStream_t)stream, adjs,
                       n_lo, n_hi);
    BZ_CHECK_LAUNCH("bz_adj2d_grad");
    return BZ_OK;
}

// inj_tab[n][cell] for steps n_lo .. n_hi-1 (n >= 1) of one adjoint (y) and injection cell (x).
// one thread per (step, cell) of the steps n_lo .. n_hi - 1 (n_lo >= 1), flat index q = (n - n_lo) * cells + p
__global__ __launch_bounds__(256) void k_adj2d_injtab(const bz_adj2d* __restrict__ adjs, int n_lo, int n_hi) {
    const bz_adj2d& A = adjs[blockIdx.y];
    const long cells = (long)A.inj_sx * A.inj_sy;
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (!A.inj_w || !A.inj_tab || q >= cells * (n_hi - n_lo)) return;
    const int n = n_lo + (int)(q / cells);
    const int p = (int)(q - (long)(n - n_lo) * cells);
    A.inj_tab[(long)n * cells + p] = (float)injection(A, p, n);
}

extern "C" int bz_adj2d_injtab(const bz_adj2d* adjs, int nadj, int max_inj, int n_lo, int n_hi, void* stream) {
    BZ_REQUIRE(adjs && nadj > 0 && nadj <= 65535 && max_inj >= 0 && n_lo >= 0 && n_hi >= n_lo,
               "bz_adj2d_injtab: bad arguments");
    n_lo = n_lo < 1 ? 1 : n_lo;
    if (max_inj == 0 || n_hi <= n_lo) return BZ_OK;
    const long total = (long)max_inj * (n_hi - n_lo);
    BZ_REQUIRE((total + 255) / 256 <= 0x7fffffffL, "bz_adj2d_injtab: table too large");
    hipLaunchKernelGGL(k_adj2d_injtab, dim3((unsigned)((total + 255) / 256), nadj), dim3(256), 0,
                       (hipStream_t)stream, adjs, n_lo, n_hi);
    BZ_CHECK_LAUNCH("bz_adj2d_injtab");
    return BZ_OK;
}

extern "C" int bz_adj2d_inject(const bz_adj2d* adjs, int nadj, int max_inj, int n, void* stream) {
    BZ_REQUIRE(adjs && nadj > 0 && max_inj >= 0 && n >= 0, "bz_adj2d_inject: bad arguments");
    if (max_inj == 0 || n < 1) return BZ_OK;
    hipLaunchKernelGGL(k_adj_inject, dim3((max_inj + 255) / 256, nadj), dim3(256), 0, (hipStream_t)stream, adjs, n);
    BZ_CHECK_LAUNCH("bz_adj2d_inject");
    return BZ_OK;
}

extern "C" int bz_adj2d_runk(const bz_adj2d* adjs, int nadj, const int* tilesk, int ntilesk, int rows, int waves,
                             int ksteps, const int* tiles, int ntiles, int seg, int n_hi, int nsteps, int flip,
                             void* stream) {
    BZ_REQUIRE(adjs && nadj > 0 && tilesk && ntilesk > 0 && nsteps >= 0 && n_hi - nsteps >= 0 &&
                   (flip == 0 || flip == 1),
               "bz_adj2d_runk: bad arguments");
    const int shape = rows * 100 + waves;
    BZ_REQUIRE(shape == 804 || shape == 808 || shape == 816 || shape == 408 || shape == 416 || shape == 1604,
               "bz_adj2d_runk: (rows, waves) must be (8, 4), (8, 8), (8, 16), (4, 8), (4, 16) or (16, 4)");
    BZ_REQUIRE(ksteps >= 1 && 2 * ksteps < LANES && 2 * ksteps < waves * rows,
               "bz_adj2d_runk: ksteps must leave inner cells (2 * ksteps < 64 and < rows * waves)");
    BZ_REQUIRE(nsteps % ksteps == 0 || (tiles && ntiles > 0),
               "bz_adj2d_runk: a step count that is not a multiple of ksteps needs one-step tiles");
    const hipStream_t st = (hipStream_t)stream;
    const int nl = nsteps / ksteps;
    for (int q = 0; q < nl; ++q) {
        const int n_top = n_hi - 1 - q * ksteps;
        const int rb = (n_top + 1 + flip) & 1;
        const dim3 b(LANES, waves);
#define BZ_ADJK(R, W) hipLaunchKernelGGL((k_adj2d_tbk<R, W>), dim3(ntilesk), b, 0, st, adjs, tilesk, ntilesk, ksteps, \
                                         n_top, rb)
        switch (shape) {
            case 804: BZ_ADJK(8, 4); break;
            case 808: BZ_ADJK(8, 8); break;
            case 816: BZ_ADJK(8, 16); break;
            case 408: BZ_ADJK(4, 8); break;
            case 416: BZ_ADJK(4, 16); break;
            default: BZ_ADJK(16, 4); break;
        }
#undef BZ_ADJK
        BZ_CHECK_LAUNCH("bz_adj2d_runk");
        flip = (flip + ksteps - 1) & 1;
    }
    if (nsteps % ksteps) return run_seg(adjs, tiles, ntiles, seg, n_hi - nl * ksteps, nsteps % ksteps, flip, st);
    return BZ_OK;
}
