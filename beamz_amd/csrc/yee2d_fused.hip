// Fused single-pass 2D (Ez, Hx, Hy) Yee step for gfx950, batched over many
// independent simulations of different sizes in one launch.
//
// One step = numpy_backend.py:34-88 (H on all valid entries, Ez on the
// interior), then helper.py:96-118 / 8-54 (source box added to Ez), then
// helper.py:152-187 (optional grid-power accumulation), for every simulation
// of the batch. Fields are ping-ponged by step parity (read buffer n & 1,
// write buffer (n + 1) & 1), so any tile can read its halo from the read
// buffer without ordering between tiles.
//
// Work decomposition: the batch is cut into tiles of 63 columns x `seg` rows,
// listed on the host (the tile work queue); each wave takes one tile: lane l
// owns column i0-1+l (lane 0 is a halo lane that computes H but stores
// nothing) and marches along y, keeping Ez of the next row and Hy of the
// previous row in registers; the x neighbours come from adjacent lanes by
// shuffles. A tile starting at j0 > 0 first recomputes Hy of row j0-1.
// HBM traffic per cell-update: Ez, Hx, Hy, eps, sigma read + Ez, Hx, Hy
// written = 32 B (+ the halo lane and one row per tile).
//
// The step index comes from device memory (`step`), so a sequence of launches
// can be captured once in a hipGraph and replayed; bz_yee2d_tick advances it.
#include "common.h"

namespace {

constexpr int LANES = 64;
constexpr int OX = LANES - 1;
constexpr int WPG = 4;

__device__ __forceinline__ float ldg(const float* __restrict__ p, long o, bool ok) { return ok ? p[o] : 0.0f; }

__global__ __launch_bounds__(LANES* WPG) void k_yee2d_fused(const bz_sim2d* __restrict__ sims,
                                                           const int* __restrict__ tiles, int ntiles, int seg,
                                                           const int* __restrict__ step) {
    const int t = blockIdx.x * WPG + threadIdx.y;
    if (t >= ntiles) return;
    const int n = *step;
    const int rd = n & 1, wr = rd ^ 1;
    const int s = tiles[3 * t], tx = tiles[3 * t + 1], ty = tiles[3 * t + 2];
    const bz_sim2d& S = sims[s];
    const int nx = S.nx, ny = S.ny, ld = S.ld;
    const int l = threadIdx.x;
    const int i = tx * OX - 1 + l;
    const bool ci = i >= 0 && i < nx;
    const int j0 = ty * seg;
    const int j1 = min(ny, j0 + seg);
    const float* __restrict__ gez = S.ez[rd];
    const float* __restrict__ ghx = S.hx[rd];
    const float* __restrict__ ghy = S.hy[rd];
    float* __restrict__ oez = S.ez[wr];
    float* __restrict__ ohx = S.hx[wr];
    float* __restrict__ ohy = S.hy[wr];
    Coef c;
    c.kh = S.kh;
    c.kbh = S.kbh;
    c.ke = S.ke;
    c.kbe = S.kbe;
    const float rdx = S.rdx, rdy = S.rdy;
    const bool xin = i >= 1 && i <= nx - 2;
    const bool hxok = i < nx - 1;
    const float amp = S.signal ? S.signal[n] : 0.0f;
    const bool out = ci && l > 0;
    const bool rx = i + 1 < nx;

    const int jstart = j0 > 0 ? j0 - 1 : j0;
    float ez = ldg(gez, (long)jstart * ld + i, ci);
    float hyp = 0.0f;
    for (int j = jstart; j < j1; ++j) {
        const long o = (long)j * ld + i;
        const bool nj = j + 1 < ny;
        const float eznx = ldg(gez, o + ld, ci && nj);
        const float hx = ldg(ghx, o, ci);
        const float hy = ldg(ghy, o, ci);
        const float sg = ldg(S.sig, o, ci);
        float ezR = __shfl_down(ez, 1, LANES);
        if (l == LANES - 1) ezR = ldg(gez, o + 1, ci && rx);
        float fa, fb;
        h_coef(sg, c, fa, fb);
        const float hxn = hxok ? fa * hx - fb * ((ezR - ez) * rdy) : hx;
        const float hyn = nj ? fa * hy + fb * ((eznx - ez) * rdx) : hy;
        if (j >= j0) {
            const float hxL = __shfl_up(hxn, 1, LANES);
            if (out) {
                const float ee = S.eps[o];
                float ca, cb;
                e_coef(sg, ee, c, ca, cb);
                float ezn = (j >= 1 && j <= ny - 2 && xin)
                                ? ca * ez + cb * (-((hxn - hxL) * rdy) + (hyn - hyp) * rdx)
                                : ez;
                const int a = i - S.src_x0, b = j - S.src_y0;
                if (a >= 0 && a < S.src_sx && b >= 0 && b < S.src_sy) {
                    const long r = (long)b * S.src_sx + a;
                    ezn = (S.src_mul ? S.src_mul[r] : 1.0f) * ezn + S.src_add[r] * amp;
                }
                oez[o] = ezn;
                ohx[o] = hxn;
                ohy[o] = hyn;
                if (S.power) {
                    const double e = ezn;
                    const double sx = -e * (double)hyn, sy = e * (double)hxn;
                    S.power[o] += sx * sx + sy * sy;
                }
            }
        }
        hyp = hyn;
        ez = eznx;
    }
}

// Monitor samples of the batch after a step: v = field[comp][off] of simulation sim, read from the
// buffer the step wrote. Recorded steps have slot[n] >= 0.
__global__ void k_gather2d(const bz_sim2d* __restrict__ sims, const int* __restrict__ pts, int npts,
                           const int* __restrict__ step, const int* __restrict__ slot, float* __restrict__ series,
                           int nfreq, const double* __restrict__ phasor, double* __restrict__ dft) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npts) return;
    const int n = *step;
    const int q = slot[n];
    if (q < 0) return;
    const int par = (n + 1) & 1;
    const bz_sim2d& S = sims[pts[3 * p]];
    const int comp = pts[3 * p + 1];
    const float* f = comp == 0 ? S.ez[par] : (comp == 1 ? S.hx[par] : S.hy[par]);
    const float v = f[pts[3 * p + 2]];
    if (series) series[(long)q * npts + p] = v;
    const double dv = v;
    const double* ph = phasor + 2 * (long)q * nfreq;
    for (int k = 0; k < nfreq; ++k) {
        double* a = dft + 2 * ((long)k * npts + p);
        a[0] += dv * ph[2 * k];
        a[1] += dv * ph[2 * k + 1];
    }
}

__global__ void k_tick(int* step) {
    if (threadIdx.x == 0) step[0] += 1;
}

}  // namespace

extern "C" int bz_yee2d_run(const bz_sim2d* sims, int nsims, const int* tiles, int ntiles, int seg, int* step,
                            int nsteps, const int* pts, int npts, const int* slot, float* series, int nfreq,
                            const double* phasor, double* dft, void* stream) {
    BZ_REQUIRE(sims && nsims > 0 && tiles && ntiles > 0 && seg > 0 && step && nsteps >= 0,
               "bz_yee2d_run: bad arguments");
    BZ_REQUIRE(npts == 0 || (pts && slot), "bz_yee2d_run: monitor points without a slot table");
    BZ_REQUIRE(nfreq == 0 || (phasor && dft), "bz_yee2d_run: DFT requested without buffers");
    const hipStream_t st = (hipStream_t)stream;
    const dim3 g((ntiles + WPG - 1) / WPG), b(LANES, WPG);
    for (int q = 0; q < nsteps; ++q) {
        hipLaunchKernelGGL(k_yee2d_fused, g, b, 0, st, sims, tiles, ntiles, seg, step);
        BZ_CHECK_LAUNCH("bz_yee2d_run(step)");
        if (npts > 0) {
            hipLaunchKernelGGL(k_gather2d, dim3((npts + 255) / 256), dim3(256), 0, st, sims, pts, npts, step, slot,
                               series, nfreq, phasor, dft);
            BZ_CHECK_LAUNCH("bz_yee2d_run(gather)");
        }
        hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, st, step);
        BZ_CHECK_LAUNCH("bz_yee2d_run(tick)");
    }
    return BZ_OK;
}
