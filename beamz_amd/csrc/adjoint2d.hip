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

namespace {

constexpr int LANES = 64;
constexpr int OX = LANES - 2;
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

template <int SEG>
__global__ __launch_bounds__(LANES* WPG) void k_adj2d(const bz_adj2d* __restrict__ adjs, const int* __restrict__ tiles,
                                                     int ntiles, int n) {
    const int t = __builtin_amdgcn_readfirstlane(blockIdx.x * WPG + threadIdx.y);
    if (t >= ntiles) return;
    const int rd = (n + 1) & 1, wr = n & 1;
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
        float m = 1.0f;
        if (A.src_mul) {
            const int a = i - A.src_x0, b = j - A.src_y0;
            if (a >= 0 && a < A.src_sx && b >= 0 && b < A.src_sy) m = A.src_mul[(long)b * A.src_sx + a];
        }
        l2[r] = m * le[r];
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
            if (A.inj_w && n >= 1) {
                const int a = i - A.inj_x0, b = j - A.inj_y0;
                if (a >= 0 && a < A.inj_sx && b >= 0 && b < A.inj_sy) lnew += (float)injection(A, b * A.inj_sx + a, n);
            }
            put(ole, ro[r], lnew);
            put(olx, ro[r], fa_[r] * mx[r]);
            put(oly, ro[r], fa_[r] * my[r]);
            if (A.grad && in) {
                const int a = i - A.box_x0, b = j - A.box_y0;
                if (a >= 0 && a < A.box_sx && b >= 0 && b < A.box_sy) {
                    // snapshot box = design box grown by one cell on the low x and y sides
                    const int sx = A.box_sx + 1, sy = A.box_sy + 1;
                    const long plane = (long)sx * sy;
                    const float* p = A.snap + (long)(n - A.snap_step0) * 3 * plane + (long)(b + 1) * sx + (a + 1);
                    const float en = p[0];
                    const float* hx = p + plane;
                    const float* hy = p + 2 * plane;
                    const float curl = -((hx[0] - hx[-1]) * rdy) + (hy[0] - hy[-sx]) * rdx;
                    const float e = ee[r];
                    const float s = sg[r] * c.ke / e;
                    const float dca = 2.0f * s / (e * (1.0f + s) * (1.0f + s));
                    const float dcb = -cb / (e * (1.0f + s));
                    A.grad[(long)b * A.box_sx + a] += (double)l2[r] * ((double)dca * en + (double)dcb * curl);
                }
            }
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
int run(const bz_adj2d* adjs, const int* tiles, int ntiles, int n_hi, int nsteps, hipStream_t st) {
    const dim3 g((ntiles + WPG - 1) / WPG), b(LANES, WPG);
    auto kern = k_adj2d<SEG>;
    for (int q = 0; q < nsteps; ++q) {
        hipLaunchKernelGGL(kern, g, b, 0, st, adjs, tiles, ntiles, n_hi - 1 - q);
        BZ_CHECK_LAUNCH("bz_adj2d_run");
    }
    return BZ_OK;
}

}  // namespace

extern "C" int bz_adj2d_run(const bz_adj2d* adjs, int nadj, const int* tiles, int ntiles, int seg, int n_hi,
                            int nsteps, void* stream) {
    BZ_REQUIRE(adjs && nadj > 0 && tiles && ntiles > 0 && nsteps >= 0 && n_hi - nsteps >= 0,
               "bz_adj2d_run: bad arguments");
    const hipStream_t st = (hipStream_t)stream;
    switch (seg) {
        case 2: return run<2>(adjs, tiles, ntiles, n_hi, nsteps, st);
        case 4: return run<4>(adjs, tiles, ntiles, n_hi, nsteps, st);
        case 8: return run<8>(adjs, tiles, ntiles, n_hi, nsteps, st);
        default: bz::set_error("bz_adj2d_run: seg must be 2, 4 or 8"); return BZ_EINVAL;
    }
}

extern "C" int bz_adj2d_inject(const bz_adj2d* adjs, int nadj, int max_inj, int n, void* stream) {
    BZ_REQUIRE(adjs && nadj > 0 && max_inj >= 0 && n >= 0, "bz_adj2d_inject: bad arguments");
    if (max_inj == 0 || n < 1) return BZ_OK;
    hipLaunchKernelGGL(k_adj_inject, dim3((max_inj + 255) / 256, nadj), dim3(256), 0, (hipStream_t)stream, adjs, n);
    BZ_CHECK_LAUNCH("bz_adj2d_inject");
    return BZ_OK;
}
