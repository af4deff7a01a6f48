// Fused single-pass 2D (Ez, Hx, Hy) Yee step for gfx950, batched over many
// independent simulations of different sizes in one launch.
//
// One step = numpy_backend.py:34-88 (H on all valid entries, Ez on the
// interior), then helper.py:96-118 / 8-54 (source box added to Ez), then
// helper.py:152-187 (optional grid-power accumulation), for every simulation
// of the batch, then the monitor samples of helper.py:231-256 (+ FP64 DFT).
// Fields are ping-ponged by step parity (read buffer n & 1, write buffer
// (n + 1) & 1), so any tile can read its halo from the read buffer without
// ordering between tiles.
//
// Work decomposition: the batch is cut into tiles of 62 columns x SEG rows,
// listed on the host (the tile work queue); each wave takes one tile: lane l
// owns column i0-1+l, lanes 0 and 63 are halo lanes that compute but store
// nothing (lane 0 supplies Hx at i-1, lane 63 Ez at i+1, by wave shuffles).
// A wave issues every load of its tile at once (rows j0-1 .. j0+SEG of Ez,
// j0-1 .. j0+SEG-1 of Hy and sigma, j0 .. j0+SEG-1 of Hx and eps), so a tile
// costs one memory round trip; halo lanes and rows outside the grid read a
// clamped in-grid cell whose value only reaches entries that are not stored.
// Addressing is 32-bit byte offsets from the wave-uniform array bases.
// HBM traffic per cell-update: Ez, Hx, Hy, eps, sigma read + Ez, Hx, Hy
// written = 32 B (+ the two halo lanes and three halo rows per tile).
//
// One launch per step: step n is `*step + q` (base in device memory, q a
// launch argument), and the monitor samples of step n-1 are taken by extra
// workgroups of step n's launch, which read step n-1's output while it is
// step n's unmodified input. A run of nsteps launches ends with one gather
// launch for its last step and one launch that advances the base by nsteps,
// so the whole run can be captured once in a hipGraph and replayed.
#include "common.h"

namespace {

constexpr int LANES = 64;
constexpr int OX = LANES - 2;
constexpr int WPG = 4;
constexpr int GPB = LANES * WPG;   // monitor points per gather workgroup

struct Gather2D {
    const int* pts;      // (sim, comp, offset) per point
    int npts;
    const int* slot;     // slot[n] = recorded row of step n, or -1
    float* series;       // [rows][npts] or NULL
    int nfreq;
    const double* phasor;   // [rows][nfreq] complex
    double* dft;            // [nfreq][npts] complex
};

__device__ __forceinline__ float at(const float* __restrict__ base, uint32_t byte_off) {
    return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + byte_off);
}
__device__ __forceinline__ void put(float* __restrict__ base, uint32_t byte_off, float v) {
    *reinterpret_cast<float*>(reinterpret_cast<char*>(base) + byte_off) = v;
}

// Monitor sample p of step n, read from the buffer step n wrote.
__device__ __forceinline__ void gather_point(const bz_sim2d* __restrict__ sims, const Gather2D& G, int p, int n) {
    if (p >= G.npts) return;
    const int q = G.slot[n];
    if (q < 0) return;
    const int par = (n + 1) & 1;
    const bz_sim2d& S = sims[G.pts[3 * p]];
    const int comp = G.pts[3 * p + 1];
    const float* f = comp == 0 ? S.ez[par] : (comp == 1 ? S.hx[par] : S.hy[par]);
    const float v = f[G.pts[3 * p + 2]];
    if (G.series) G.series[(long)q * G.npts + p] = v;
    const double dv = v;
    const double* ph = G.phasor + 2 * (long)q * G.nfreq;
    for (int k = 0; k < G.nfreq; ++k) {
        double* a = G.dft + 2 * ((long)k * G.npts + p);
        a[0] += dv * ph[2 * k];
        a[1] += dv * ph[2 * k + 1];
    }
}

template <int SEG>
__global__ __launch_bounds__(LANES* WPG) void k_yee2d_fused(const bz_sim2d* __restrict__ sims,
                                                           const int* __restrict__ tiles, int ntiles,
                                                           const int* __restrict__ step, int q, Gather2D G) {
    const int n = *step + q;
    const int tile_blocks = (ntiles + WPG - 1) / WPG;
    if ((int)blockIdx.x >= tile_blocks) {
        // monitor samples of the previous step (launched only when q > 0)
        gather_point(sims, G, ((int)blockIdx.x - tile_blocks) * GPB + threadIdx.y * LANES + threadIdx.x, n - 1);
        return;
    }
    // the tile index is uniform over the wave: say so, so the tile and simulation descriptors are read
    // with scalar loads and the array bases stay in scalar registers
    const int t = __builtin_amdgcn_readfirstlane(blockIdx.x * WPG + threadIdx.y);
    if (t >= ntiles) return;
    const int rd = n & 1, wr = rd ^ 1;
    const int s = tiles[3 * t], tx = tiles[3 * t + 1], ty = tiles[3 * t + 2];
    const bz_sim2d& S = sims[s];
    const int nx = S.nx, ny = S.ny, ld = S.ld;
    const int l = threadIdx.x;
    const int i = tx * OX - 1 + l;
    const int j0 = ty * SEG;
    const float* __restrict__ gez = S.ez[rd];
    const float* __restrict__ ghx = S.hx[rd];
    const float* __restrict__ ghy = S.hy[rd];
    const float* __restrict__ gsg = S.sig;
    const float* __restrict__ gee = S.eps;
    Coef c;
    c.kh = S.kh;
    c.kbh = S.kbh;
    c.ke = S.ke;
    c.kbe = S.kbe;
    const float rdx = S.rdx, rdy = S.rdy;
    const uint32_t ic = (uint32_t)min(max(i, 0), nx - 1);

    // byte offset of register row r = row j0-1+r, clamped into the grid
    uint32_t ro[SEG + 2];
#pragma unroll
    for (int r = 0; r < SEG + 2; ++r) {
        const int j = min(max(j0 - 1 + r, 0), ny - 1);
        ro[r] = ((uint32_t)j * (uint32_t)ld + ic) * 4u;
    }
    float ez[SEG + 2], hx[SEG + 1], hy[SEG + 1], sg[SEG + 1], ee[SEG + 1];
#pragma unroll
    for (int r = 0; r < SEG + 2; ++r) {
        ez[r] = at(gez, ro[r]);
        if (r < SEG + 1) {
            hy[r] = at(ghy, ro[r]);
            sg[r] = at(gsg, ro[r]);
            hx[r] = r > 0 ? at(ghx, ro[r]) : 0.0f;
            ee[r] = r > 0 ? at(gee, ro[r]) : 1.0f;
        }
    }

    // H of rows j0-1 .. j0+SEG-1 (Hy of row j0-1 feeds Ez of row j0; its Hx is not needed)
    float hxn[SEG + 1], hyn[SEG + 1], hxl[SEG + 1];
    const bool hxok = i < nx - 1;
#pragma unroll
    for (int r = 0; r < SEG + 1; ++r) {
        const int j = j0 - 1 + r;
        float fa, fb;
        h_coef(sg[r], c, fa, fb);
        hyn[r] = (j < ny - 1) ? fa * hy[r] + fb * ((ez[r + 1] - ez[r]) * rdx) : hy[r];
        const float ezR = __shfl_down(ez[r], 1, LANES);
        hxn[r] = hxok ? fa * hx[r] - fb * ((ezR - ez[r]) * rdy) : hx[r];
        hxl[r] = __shfl_up(hxn[r], 1, LANES);
    }

    // E of rows j0 .. j0+SEG-1, sources, stores
    const bool out = l > 0 && l < LANES - 1 && i < nx;
    const bool xin = i >= 1 && i <= nx - 2;
    const float amp = S.signal ? S.signal[n] : 0.0f;
    float* __restrict__ oez = S.ez[wr];
    float* __restrict__ ohx = S.hx[wr];
    float* __restrict__ ohy = S.hy[wr];
#pragma unroll
    for (int r = 1; r < SEG + 1; ++r) {
        const int j = j0 - 1 + r;
        if (out && j < ny) {
            float ca, cb;
            e_coef(sg[r], ee[r], c, ca, cb);
            float ezn = (j >= 1 && j <= ny - 2 && xin)
                            ? ca * ez[r] + cb * (-((hxn[r] - hxl[r]) * rdy) + (hyn[r] - hyn[r - 1]) * rdx)
                            : ez[r];
            const int a = i - S.src_x0, b = j - S.src_y0;
            if (a >= 0 && a < S.src_sx && b >= 0 && b < S.src_sy) {
                const long rr = (long)b * S.src_sx + a;
                ezn = (S.src_mul ? S.src_mul[rr] : 1.0f) * ezn + S.src_add[rr] * amp;
            }
            put(oez, ro[r], ezn);
            put(ohx, ro[r], hxn[r]);
            put(ohy, ro[r], hyn[r]);
            if (S.power) {
                const double e = ezn;
                const double sx = -e * (double)hyn[r], sy = e * (double)hxn[r];
                S.power[(long)j * ld + i] += sx * sx + sy * sy;
            }
            if (S.snap) {
                const int sa = i - S.snap_x0, sb = j - S.snap_y0;
                if (sa >= 0 && sa < S.snap_sx && sb >= 0 && sb < S.snap_sy) {
                    const long plane = (long)S.snap_sy * S.snap_sx;
                    float* p = S.snap + (long)q * 3 * plane + (long)sb * S.snap_sx + sa;
                    p[0] = ez[r];
                    p[plane] = hxn[r];
                    p[2 * plane] = hyn[r];
                }
            }
        }
    }
}

__global__ void k_gather2d(const bz_sim2d* __restrict__ sims, Gather2D G, const int* __restrict__ step, int q) {
    gather_point(sims, G, blockIdx.x * blockDim.x + threadIdx.x, *step + q);
}

__global__ void k_tick(int* step, int nsteps) {
    if (threadIdx.x == 0) step[0] += nsteps;
}

template <int SEG>
int run(const bz_sim2d* sims, const int* tiles, int ntiles, int* step, int nsteps, const Gather2D& G,
        hipStream_t st) {
    const int tile_blocks = (ntiles + WPG - 1) / WPG;
    const int gather_blocks = (G.npts + GPB - 1) / GPB;
    const dim3 b(LANES, WPG);
    auto kern = k_yee2d_fused<SEG>;
    for (int q = 0; q < nsteps; ++q) {
        const dim3 g(tile_blocks + (q > 0 ? gather_blocks : 0));
        hipLaunchKernelGGL(kern, g, b, 0, st, sims, tiles, ntiles, step, q, G);
        BZ_CHECK_LAUNCH("bz_yee2d_run(step)");
    }
    if (G.npts > 0) {
        hipLaunchKernelGGL(k_gather2d, dim3(gather_blocks), dim3(GPB), 0, st, sims, G, step, nsteps - 1);
        BZ_CHECK_LAUNCH("bz_yee2d_run(gather)");
    }
    hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, st, step, nsteps);
    BZ_CHECK_LAUNCH("bz_yee2d_run(tick)");
    return BZ_OK;
}

}  // namespace

extern "C" int bz_yee2d_run(const bz_sim2d* sims, int nsims, const int* tiles, int ntiles, int seg, int* step,
                            int nsteps, const int* pts, int npts, const int* slot, float* series, int nfreq,
                            const double* phasor, double* dft, void* stream) {
    BZ_REQUIRE(sims && nsims > 0 && tiles && ntiles > 0 && step && nsteps >= 0, "bz_yee2d_run: bad arguments");
    BZ_REQUIRE(npts >= 0 && (npts == 0 || (pts && slot)), "bz_yee2d_run: monitor points without a slot table");
    BZ_REQUIRE(nfreq >= 0 && (nfreq == 0 || (phasor && dft)), "bz_yee2d_run: DFT requested without buffers");
    if (nsteps == 0) return BZ_OK;
    Gather2D G;
    G.pts = pts;
    G.npts = npts;
    G.slot = slot;
    G.series = series;
    G.nfreq = nfreq;
    G.phasor = phasor;
    G.dft = dft;
    const hipStream_t st = (hipStream_t)stream;
    switch (seg) {
        case 2: return run<2>(sims, tiles, ntiles, step, nsteps, G, st);
        case 4: return run<4>(sims, tiles, ntiles, step, nsteps, G, st);
        case 8: return run<8>(sims, tiles, ntiles, step, nsteps, G, st);
        case 16: return run<16>(sims, tiles, ntiles, step, nsteps, G, st);
        default: bz::set_error("bz_yee2d_run: seg must be 2, 4, 8 or 16"); return BZ_EINVAL;
    }
}
