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

#include <algorithm>
#include <type_traits>

// No multiply-add contraction: the one-step and two-step kernels then round identically, so a run gives
// the same bits whichever mix of them it uses.
#pragma clang fp contract(off)

namespace {

constexpr int LANES = 64;
constexpr int OX = LANES - 2;
constexpr int WPG = 4;
constexpr int GPB = LANES * WPG;   // monitor points per gather workgroup

struct Gather2D {
    const int* pts;      // (sim, comp, offset) per point
    int npts;
    const int* slot;     // slot[g][n] = recorded row of step n for point group g, or -1
    const int* grp;      // group of each point, or NULL (one group)
    int nslot;           // steps per group row of slot (groups > 0 only)
    int nrow;            // rows per group block of phasor (groups > 0 only)
    float* series;       // [rows][npts] or NULL (each group numbers its rows from 0)
    int nfreq;
    const double* phasor;   // [group][rows][nfreq] complex (frequencies a group lacks hold 0)
    double* dft;            // [nfreq][npts] complex
};

// Recorded row of step n for point p (-1: not recorded), and the phasor of that row at frequency k.
__device__ __forceinline__ int gslot(const Gather2D& G, int p, int n) {
    return G.grp ? G.slot[(long)G.grp[p] * G.nslot + n] : G.slot[n];
}
__device__ __forceinline__ const double* gphasor(const Gather2D& G, int p, int row, int k) {
    const long g = G.grp ? G.grp[p] : 0;
    return G.phasor + 2 * ((g * G.nrow + row) * G.nfreq + k);
}

// Every buffer a descriptor points at is device (global) memory. The pointers are read from the descriptor, so
// the compiler cannot tell their address space and would issue flat instructions, which also count against
// the LDS wait counter; these accessors issue global loads and stores (with the scalar base and a 32-bit
// offset).
#define BZ_GLOBAL __attribute__((address_space(1)))
__device__ __forceinline__ float at(const float* __restrict__ base, uint32_t byte_off) {
    return *(const BZ_GLOBAL float*)((const char*)base + byte_off);
}
__device__ __forceinline__ void put(float* __restrict__ base, uint32_t byte_off, float v) {
    *(BZ_GLOBAL float*)((char*)base + byte_off) = v;
}
typedef float f32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 at4(const float* __restrict__ base, uint32_t byte_off) {
    const f32x4 v = *(const BZ_GLOBAL f32x4*)((const char*)base + byte_off);
    return make_float4(v.x, v.y, v.z, v.w);
}
template <typename T>
__device__ __forceinline__ T gload(const T* p, long i) {
    return ((const BZ_GLOBAL T*)p)[i];
}
template <typename T>
__device__ __forceinline__ void gstore(T* p, long i, T v) {
    ((BZ_GLOBAL T*)p)[i] = v;
}

// Amplitude of step n for cell rr of the source box: that cell's own source when the box holds several
// (disjoint) sources, else the single signal's value amp.
__device__ __forceinline__ float box_amp(const bz_sim2d& S, long rr, int n, float amp) {
    return S.src_sel ? gload(S.signal, (long)gload(S.src_sel, rr) * S.src_sig_len + n) : amp;
}

// The (Ez, Hx, Hy) updates of every kernel below, in one arithmetic: the y-spacing folded into the cell's
// coefficients (gb = fb * rdy, gc = cb * rdy, each rounded once), r = rdx / rdy (exactly 1 on square cells,
// where its products are exact), explicit fused multiply-adds. 11 to 12 vector operations per cell instead
// of 18, and the same bits in every kernel.
//   Hx' = fa Hx - gb (Ez[i+1] - Ez)            (= fa Hx - fb dEz/dy)
//   Hy' = fa Hy + gb r (Ez[j+1] - Ez)          (= fa Hy + fb dEz/dx)
//   Ez' = ca Ez + gc (r (Hy - Hy[j-1]) - (Hx - Hx[i-1]))
// The field's own term is the multiply of the last fused multiply-add: a separate product (fa Hx) needs no
// neighbour, so the compiler hoists it above the step's branches and holds it in registers.
__device__ __forceinline__ float upd_hx(float fa, float gb, float hx, float ezR, float ez) {
    return fmaf(fa, hx, -(gb * (ezR - ez)));
}
__device__ __forceinline__ float upd_hy(float fa, float gb, float r, float hy, float ed, float ez) {
    return fmaf(fa, hy, gb * (r * (ed - ez)));
}
__device__ __forceinline__ float upd_ez(float ca, float gc, float r, float ez, float hx, float hxl, float hy,
                                        float hu) {
    return fmaf(ca, ez, gc * fmaf(r, hy - hu, -(hx - hxl)));
}

// Monitor sample p of step n, read from the buffer step n wrote.
__device__ __forceinline__ void gather_point(const bz_sim2d* __restrict__ sims, const Gather2D& G, int p, int n) {
    if (p >= G.npts) return;
    const int q = gslot(G, p, n);
    if (q < 0) return;
    const int par = (n + 1) & 1;
    const bz_sim2d& S = sims[G.pts[3 * p]];
    const int comp = G.pts[3 * p + 1];
    const float* f = comp == 0 ? S.ez[par] : (comp == 1 ? S.hx[par] : S.hy[par]);
    const float v = f[G.pts[3 * p + 2]];
    if (G.series) G.series[(long)q * G.npts + p] = v;
    const double dv = v;
    const double* ph = gphasor(G, p, q, 0);
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
    if (s < 0) return;   // padding entry of a reordered tile list
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
    const float rxy = rdx / rdy;   // 1 on square cells
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
        const float gb = fb * rdy;
        hyn[r] = (j < ny - 1) ? upd_hy(fa, gb, rxy, hy[r], ez[r + 1], ez[r]) : hy[r];
        const float ezR = __shfl_down(ez[r], 1, LANES);
        hxn[r] = hxok ? upd_hx(fa, gb, hx[r], ezR, ez[r]) : hx[r];
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
                            ? upd_ez(ca, cb * rdy, rxy, ez[r], hxn[r], hxl[r], hyn[r], hyn[r - 1])
                            : ez[r];
            const int a = i - S.src_x0, b = j - S.src_y0;
            if (a >= 0 && a < S.src_sx && b >= 0 && b < S.src_sy) {
                const long rr = (long)b * S.src_sx + a;
                ezn = (S.src_mul ? S.src_mul[rr] : 1.0f) * ezn + S.src_add[rr] * box_amp(S, rr, n, amp);
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

// ---------------- the (Hz, Ex, Ey) polarisation ----------------
// The z-invariant restriction of the reference's 3D step (fdtd.py:200-310 with every d/dz = 0) to the
// components Hz (ny-1, nx-1), Ex (ny, nx-1), Ey (ny-1, nx); the descriptor's ez / hx / hy slots hold Hz / Ex / Ey.
//   Hz[j,i] <- fa Hz - fb ((Ey[j,i+1] - Ey[j,i]) / dx - (Ex[j+1,i] - Ex[j,i]) / dy),  fa, fb of sigma[j,i]
//   Ex[j,i] <- ca Ex + cb (Hz[j,i] - Hz[j-1,i]) / dy           (1 <= j <= ny-2; every i <= nx-2)
//   Ey[j,i] <- ca Ey - cb (Hz[j,i] - Hz[j,i-1]) / dx           (1 <= i <= nx-2; every j <= ny-2)
// with ca, cb of eps_r[j,i], sigma[j,i]; the source box is added to Hz after the E update (the reference's
// order: H, E, sources), and grid power accumulates Sx^2 + Sy^2 with S = E x H = (Ey Hz, -Ex Hz).
// Same tiling as k_yee2d_fused: lane l owns column i0-1+l, lane 63 supplies Ey at i+1 and lane 0 Hz at i-1.
template <int SEG>
__global__ __launch_bounds__(LANES* WPG) void k_yee2d_hz(const bz_sim2d* __restrict__ sims,
                                                        const int* __restrict__ tiles, int ntiles,
                                                        const int* __restrict__ step, int q, Gather2D G) {
    const int n = *step + q;
    const int tile_blocks = (ntiles + WPG - 1) / WPG;
    if ((int)blockIdx.x >= tile_blocks) {
        gather_point(sims, G, ((int)blockIdx.x - tile_blocks) * GPB + threadIdx.y * LANES + threadIdx.x, n - 1);
        return;
    }
    const int t = __builtin_amdgcn_readfirstlane(blockIdx.x * WPG + threadIdx.y);
    if (t >= ntiles) return;
    const int rd = n & 1, wr = rd ^ 1;
    const int s = tiles[3 * t], tx = tiles[3 * t + 1], ty = tiles[3 * t + 2];
    if (s < 0) return;
    const bz_sim2d& S = sims[s];
    const int nx = S.nx, ny = S.ny, ld = S.ld;
    const int l = threadIdx.x;
    const int i = tx * OX - 1 + l;
    const int j0 = ty * SEG;
    const float* __restrict__ ghz = S.ez[rd];
    const float* __restrict__ gex = S.hx[rd];
    const float* __restrict__ gey = S.hy[rd];
    Coef c;
    c.kh = S.kh;
    c.kbh = S.kbh;
    c.ke = S.ke;
    c.kbe = S.kbe;
    const float rdx = S.rdx, rdy = S.rdy;
    const uint32_t ic = (uint32_t)min(max(i, 0), nx - 1);
    // register row r = row j0-1+r, clamped into the grid
    uint32_t ro[SEG + 2];
#pragma unroll
    for (int r = 0; r < SEG + 2; ++r) {
        const int j = min(max(j0 - 1 + r, 0), ny - 1);
        ro[r] = ((uint32_t)j * (uint32_t)ld + ic) * 4u;
    }
    float ex[SEG + 2], hz[SEG + 1], ey[SEG + 1], sg[SEG + 1], ee[SEG + 1];
#pragma unroll
    for (int r = 0; r < SEG + 2; ++r) {
        ex[r] = at(gex, ro[r]);
        if (r < SEG + 1) {
            hz[r] = at(ghz, ro[r]);
            ey[r] = at(gey, ro[r]);
            sg[r] = at(S.sig, ro[r]);
            ee[r] = r > 0 ? at(S.eps, ro[r]) : 1.0f;
        }
    }
    // Hz of rows j0-1 .. j0+SEG-1 (row j0-1 feeds Ex of row j0)
    float hzn[SEG + 1], hzl[SEG + 1];
    const bool hzx = i < nx - 1;
#pragma unroll
    for (int r = 0; r < SEG + 1; ++r) {
        const int j = j0 - 1 + r;
        float fa, fb;
        h_coef(sg[r], c, fa, fb);
        const float eyR = __shfl_down(ey[r], 1, LANES);
        hzn[r] = (hzx && j < ny - 1) ? fa * hz[r] - fb * ((eyR - ey[r]) * rdx - (ex[r + 1] - ex[r]) * rdy) : hz[r];
        hzl[r] = __shfl_up(hzn[r], 1, LANES);
    }
    // Ex, Ey of rows j0 .. j0+SEG-1, the source on Hz, stores
    const bool out = l > 0 && l < LANES - 1 && i < nx;
    const float amp = S.signal ? S.signal[n] : 0.0f;
    float* __restrict__ ohz = S.ez[wr];
    float* __restrict__ oex = S.hx[wr];
    float* __restrict__ oey = S.hy[wr];
#pragma unroll
    for (int r = 1; r < SEG + 1; ++r) {
        const int j = j0 - 1 + r;
        if (out && j < ny) {
            float ca, cb;
            e_coef(sg[r], ee[r], c, ca, cb);
            const float exn = (j >= 1 && j <= ny - 2 && hzx) ? ca * ex[r] + cb * ((hzn[r] - hzn[r - 1]) * rdy) : ex[r];
            const float eyn = (j <= ny - 2 && i >= 1 && i <= nx - 2) ? ca * ey[r] + cb * (-((hzn[r] - hzl[r]) * rdx))
                                                                     : ey[r];
            float hzo = hzn[r];
            const int a = i - S.src_x0, b = j - S.src_y0;
            if (a >= 0 && a < S.src_sx && b >= 0 && b < S.src_sy) {
                const long rr = (long)b * S.src_sx + a;
                hzo = (S.src_mul ? S.src_mul[rr] : 1.0f) * hzo + S.src_add[rr] * box_amp(S, rr, n, amp);
            }
            put(ohz, ro[r], hzo);
            put(oex, ro[r], exn);
            put(oey, ro[r], eyn);
            if (S.power) {
                const double h = hzo;
                const double sx = (double)eyn * h, sy = -(double)exn * h;
                S.power[(long)j * ld + i] += sx * sx + sy * sy;
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

// ---------------- two steps per pass ----------------
// Steps n, n+1 of a tile of 60 columns x seg rows: the wave marches y computing step n+1 (H then E) on
// row r and step n+2 on row r-1 (its E needs only the H of step n+2 on rows r-1 and r-2, carried), so
// one pass reads the five arrays and writes the three fields once per two steps. Lanes 0, 1, 62, 63 are
// halo lanes; a segment starting at j0 > 0 begins two rows early. Reads buffer rd, writes rd ^ 1.
// Monitor samples of both steps of pair q: step n from staging slot q & 1, step n+1 from buffer wr.
// Thread g handles (point p, frequency k), p fastest, so the FP64 accumulators dft[k][p] are read and
// written coalesced and the latency of the read-modify-write is spread over npts * nfreq threads. Each
// accumulator still receives step n's term and then step n+1's, as two rounded adds in that order.
__device__ __forceinline__ void gather_pair(const bz_sim2d* __restrict__ sims, const Gather2D& G,
                                            const int* __restrict__ spts, const int* __restrict__ step, int q, int wr,
                                            long g) {
    const int p = (int)(g % G.npts);
    const int k = (int)(g / G.npts);
    if (k >= max(G.nfreq, 1)) return;
    const int n = *step + 2 * q;
    const bz_sim2d& S = sims[G.pts[3 * p]];
    const int comp = G.pts[3 * p + 1];
    int row[2];
    float v[2];
    for (int s = 0; s < 2; ++s) {
        row[s] = gslot(G, p, n + s);
        v[s] = 0.0f;
        if (row[s] < 0) continue;
        if (s == 0) {
            const int so = spts[p];
            const long plane = (long)S.stg_sx * S.stg_sy;
            v[s] = so >= 0 ? S.stg[((long)(q & 1) * 3 + comp) * plane + so] : 0.0f;
        } else {
            const float* f = comp == 0 ? S.ez[wr] : (comp == 1 ? S.hx[wr] : S.hy[wr]);
            v[s] = f[G.pts[3 * p + 2]];
        }
        if (k == 0 && G.series) G.series[(long)row[s] * G.npts + p] = v[s];
    }
    if (k >= G.nfreq || (row[0] < 0 && row[1] < 0)) return;
    double* a = G.dft + 2 * ((long)k * G.npts + p);
    double re = a[0], im = a[1];
    for (int s = 0; s < 2; ++s) {
        if (row[s] < 0) continue;
        const double dv = v[s];
        const double* ph = gphasor(G, p, row[s], k);
        re += dv * ph[0];
        im += dv * ph[1];
    }
    a[0] = re;
    a[1] = im;
}

// Snapshot of the step with call-relative index slot at cell (j, i), if inside the box: the Ez the step read
// and the Hx, Hy it wrote (the layout of bz_yee2d_run's snapshots).
// The snapshot box, read from the descriptor once per kernel: read through the descriptor after the snapshot
// stores of the previous row, the fields would be loaded again (a store through a float pointer may alias the
// descriptor), and each such load waits until those stores have completed.
struct SnapBox {
    float* p;
    int x0, y0, sx, sy;
};
__device__ __forceinline__ SnapBox snap_box(const bz_sim2d& S) {
    return SnapBox{S.snap, S.snap_x0, S.snap_y0, S.snap_sx, S.snap_sy};
}
__device__ __forceinline__ void snap_put(const SnapBox& Q, int slot, int i, int j, float ez, float hx, float hy) {
    const int sa = i - Q.x0, sb = j - Q.y0;
    if (sa >= 0 && sa < Q.sx && sb >= 0 && sb < Q.sy) {
        const long plane = (long)Q.sy * Q.sx;
        float* p = Q.p + (long)slot * 3 * plane + (long)sb * Q.sx + sa;
        gstore(p, 0, ez);
        gstore(p, plane, hx);
        gstore(p, 2 * plane, hy);
    }
}

// Same rounded operations as k_yee2d_fused (no contraction in this file), so the results are identical.
constexpr int OX2 = LANES - 4;

__device__ __forceinline__ void box_src(float& e, const bz_sim2d& S, int i, int j, int n, float amp) {
    const int a = i - S.src_x0, b = j - S.src_y0;
    if (a >= 0 && a < S.src_sx && b >= 0 && b < S.src_sy) {
        const long r = (long)b * S.src_sx + a;
        e = (S.src_mul ? S.src_mul[r] : 1.0f) * e + S.src_add[r] * box_amp(S, r, n, amp);
    }
}

__device__ __forceinline__ void add_power(const bz_sim2d& S, long o, float e, float hx, float hy) {
    const double ed = e;
    const double sx = -ed * (double)hy, sy = ed * (double)hx;
    S.power[o] += sx * sx + sy * sy;
}

__global__ __launch_bounds__(LANES* WPG) void k_yee2d_tb2(const bz_sim2d* __restrict__ sims,
                                                         const int* __restrict__ tiles, int ntiles, int seg,
                                                         const int* __restrict__ step, int q, int rd, Gather2D G,
                                                         const int* __restrict__ spts) {
    const int tile_blocks = (ntiles + WPG - 1) / WPG;
    if ((int)blockIdx.x >= tile_blocks) {
        // monitor samples of the previous pair (launched only for q > 0): its first step sits in staging slot
        // (q - 1) & 1, which this launch does not write, and its second step in buffer rd, which it only reads
        gather_pair(sims, G, spts, step, q - 1, rd,
                    (long)((int)blockIdx.x - tile_blocks) * GPB + threadIdx.y * LANES + threadIdx.x);
        return;
    }
    const int t = __builtin_amdgcn_readfirstlane(blockIdx.x * WPG + threadIdx.y);
    if (t >= ntiles) return;
    const int n = *step + 2 * q;
    const int wr = rd ^ 1;
    const int s = tiles[3 * t];
    if (s < 0) return;   // padding entry of a reordered tile list
    const bz_sim2d& S = sims[s];
    const SnapBox snq = snap_box(S);
    const int tx = tiles[3 * t + 1], ty = tiles[3 * t + 2];
    const int nx = S.nx, ny = S.ny, ld = S.ld;
    const int l = threadIdx.x;
    const int i = tx * OX2 - 2 + l;
    const uint32_t ic = (uint32_t)min(max(i, 0), nx - 1);
    const int j0 = ty * seg, j1 = min(ny, j0 + seg);
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
    const float rxy = rdx / rdy;   // 1 on square cells
    const bool hxok = i < nx - 1;
    const bool xin = i >= 1 && i <= nx - 2;
    const bool out = l >= 2 && l <= LANES - 3 && i < nx;
    const float amp0 = S.signal ? S.signal[n] : 0.0f;
    const float amp1 = S.signal ? S.signal[n + 1] : 0.0f;
    const int rstart = j0 > 0 ? j0 - 2 : 0;
    auto off = [&](int r) { return ((uint32_t)min(max(r, 0), ny - 1) * (uint32_t)ld + ic) * 4u; };

    float ez0 = at(gez, off(rstart));
    float e1p = 0.0f, h1xp = 0.0f, h1yp = 0.0f, sgp = 0.0f, eep = 1.0f, h2yp = 0.0f;
    // software pipeline: the loads of row r+1 are in flight while row r is computed and stored
    float nez = at(gez, off(rstart + 1)), nhx = at(ghx, off(rstart)), nhy = at(ghy, off(rstart));
    float nsg = at(S.sig, off(rstart)), nee = at(S.eps, off(rstart));
    for (int r = rstart; r <= j1; ++r) {
        const float ez1 = nez, hx0 = nhx, hy0 = nhy, sg = nsg, ee = nee;
        {
            const uint32_t o1 = off(r + 1);
            nez = at(gez, off(r + 2));
            nhx = at(ghx, o1);
            nhy = at(ghy, o1);
            nsg = at(S.sig, o1);
            nee = at(S.eps, o1);
        }
        // step n+1 on row r
        const float ezR = __shfl_down(ez0, 1, LANES);
        float fa, fb;
        h_coef(sg, c, fa, fb);
        const float gb = fb * rdy;
        const float h1x = hxok ? upd_hx(fa, gb, hx0, ezR, ez0) : hx0;
        const float h1y = (r < ny - 1) ? upd_hy(fa, gb, rxy, hy0, ez1, ez0) : hy0;
        const float h1xL = __shfl_up(h1x, 1, LANES);
        float ca, cb;
        e_coef(sg, ee, c, ca, cb);
        float e1 = (r >= 1 && r <= ny - 2 && xin) ? upd_ez(ca, cb * rdy, rxy, ez0, h1x, h1xL, h1y, h1yp) : ez0;
        if (S.src_sx) box_src(e1, S, i, r, n, amp0);
        // step n+2 on row m = r - 1 (step n+1 of that row is e1p, h1xp, h1yp)
        const int m = r - 1;
        const float e1pR = __shfl_down(e1p, 1, LANES);
        float fa2, fb2;
        h_coef(sgp, c, fa2, fb2);
        const float gb2 = fb2 * rdy;
        const float h2x = hxok ? upd_hx(fa2, gb2, h1xp, e1pR, e1p) : h1xp;
        const float h2y = (m < ny - 1) ? upd_hy(fa2, gb2, rxy, h1yp, e1, e1p) : h1yp;
        const float h2xL = __shfl_up(h2x, 1, LANES);
        float ca2, cb2;
        e_coef(sgp, eep, c, ca2, cb2);
        float e2 = (m >= 1 && m <= ny - 2 && xin) ? upd_ez(ca2, cb2 * rdy, rxy, e1p, h2x, h2xL, h2y, h2yp) : e1p;
        if (out && m >= j0 && m < j1) {
            if (S.src_sx) box_src(e2, S, i, m, n + 1, amp1);
            const uint32_t om = off(m);
            put(oez, om, e2);
            put(ohx, om, h2x);
            put(ohy, om, h2y);
            if (S.power) add_power(S, (long)m * ld + i, e2, h2x, h2y);
            if (snq.p) snap_put(snq, 2 * q + 1, i, m, e1p, h2x, h2y);
        }
        if (out && r >= j0 && r < j1) {
            if (S.power) add_power(S, (long)r * ld + i, e1, h1x, h1y);
            if (snq.p) snap_put(snq, 2 * q, i, r, ez0, h1x, h1y);
            if (S.stg) {
                const int a = i - S.stg_x0, b = r - S.stg_y0;
                if (a >= 0 && a < S.stg_sx && b >= 0 && b < S.stg_sy) {
                    const long plane = (long)S.stg_sx * S.stg_sy;
                    float* p = S.stg + (long)(q & 1) * 3 * plane + (long)b * S.stg_sx + a;
                    p[0] = e1;
                    p[plane] = h1x;
                    p[2 * plane] = h1y;
                }
            }
        }
        ez0 = ez1;
        e1p = e1;
        h1xp = h1x;
        h1yp = h1y;
        sgp = sg;
        eep = ee;
        h2yp = h2y;
    }
}

// Gather of pair q as a launch of its own (the last pair of a call).
__global__ void k_gather2d_pair(const bz_sim2d* __restrict__ sims, Gather2D G, const int* __restrict__ spts,
                                const int* __restrict__ step, int q, int wr) {
    gather_pair(sims, G, spts, step, q, wr, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

template <int SEG, bool HZ = false>
int run(const bz_sim2d* sims, const int* tiles, int ntiles, int* step, int nsteps, const Gather2D& G,
        hipStream_t st) {
    const int tile_blocks = (ntiles + WPG - 1) / WPG;
    const int gather_blocks = (G.npts + GPB - 1) / GPB;
    const dim3 b(LANES, WPG);
    auto kern = HZ ? k_yee2d_hz<SEG> : k_yee2d_fused<SEG>;
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


// ---------------- K steps per pass ----------------
// A workgroup of KW waves takes one tile: a region of 64 columns x KW*B rows, wave w holding rows
// w*B .. w*B+B-1 of it in registers (lane l = column), with the H and E coefficients of its cells computed
// once. It advances the whole region K steps; the cells within K of the region's edge go stale one ring
// per step (they lack neighbours outside the region), so after K steps the inner (64-2K) x (KW*B-2K) cells
// hold step n+K and are stored. Tiles overlap by 2K cells, so one pass over HBM serves K steps.
// Neighbours along a row come from wave shuffles, neighbours across band edges through LDS (two barriers
// per step). Cells outside the grid compute garbage that never reaches a cell inside it: the update rules
// read only in-grid neighbours for every updated in-grid entry. Monitor samples of each step are copied
// from the owning tile to staging slot q & 1 of launch q, stage[q & 1][k][point], and accumulated into the
// FP64 DFT in step order by extra workgroups of the next launch (by k_gather2d_k after the last one).
// Reads buffer rd, writes rd ^ 1.
// Update coefficients of every cell: blockIdx.y is the simulation, the x blocks stride over its cells.
__global__ __launch_bounds__(256) void k_coef2d(const bz_sim2d* __restrict__ sims) {
    const bz_sim2d& S = sims[blockIdx.y];
    float* __restrict__ cf = S.coef;
    if (!cf) return;
    Coef c;
    c.kh = S.kh;
    c.kbh = S.kbh;
    c.ke = S.ke;
    c.kbe = S.kbe;
    const long plane = (long)S.ny * S.ld;
    for (long o = (long)blockIdx.x * blockDim.x + threadIdx.x; o < plane; o += (long)gridDim.x * blockDim.x) {
        float fa, fb, ca, cb;
        h_coef(S.sig[o], c, fa, fb);
        e_coef(S.sig[o], S.eps[o], c, ca, cb);
        reinterpret_cast<float4*>(cf)[o] = make_float4(fa, fb, ca, cb);
    }
}

// Row neighbours by DPP wave shifts: lane l receives lane l+1 (shl) or l-1 (shr); the end lanes read 0,
// which reaches only cells within K of the region's edge.
__device__ __forceinline__ float from_right(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x130, 0xf, 0xf, false));
}
__device__ __forceinline__ float from_left(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x138, 0xf, 0xf, false));
}
// Keeps a value computed on every lane, so that a per-lane choice below is a select, not a branch.
__device__ __forceinline__ float computed(float x) {
    asm volatile("" : "+v"(x));
    return x;
}

// Monitor samples of the K steps of launch q from stage[k][p]: thread g handles (point p, frequency f), and
// each accumulator receives the K terms in step order, one rounded add each, as with one step per launch.
__device__ __forceinline__ void gather_k(const Gather2D& G, const float* __restrict__ stage, int n0, int K, long g) {
    const int p = (int)(g % G.npts);
    const int f = (int)(g / G.npts);
    if (f >= max(G.nfreq, 1)) return;
    double* a = f < G.nfreq ? G.dft + 2 * ((long)f * G.npts + p) : nullptr;
    double re = a ? a[0] : 0.0, im = a ? a[1] : 0.0;
    bool any = false;
    for (int k = 0; k < K; ++k) {
        const int row = gslot(G, p, n0 + k);
        if (row < 0) continue;
        any = true;
        const float v = stage[(long)k * G.npts + p];
        if (f == 0 && G.series) G.series[(long)row * G.npts + p] = v;
        if (a) {
            const double dv = v;
            const double* ph = gphasor(G, p, row, f);
            re += dv * ph[0];
            im += dv * ph[1];
        }
    }
    if (a && any) {
        a[0] = re;
        a[1] = im;
    }
}

// The gather of the last launch of a call (the others run inside the next launch).
__global__ void k_gather2d_k(Gather2D G, const float* __restrict__ stage, const int* __restrict__ step, int q, int K) {
    gather_k(G, stage, *step + K * q, K, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// SEL: some simulation's source box holds several sources (bz_sim2d.src_sel); without it the per-cell
// source lookup is compiled out, which keeps the kernel at the register count of four waves per SIMD.
template <int B, int KW, bool SEL>
__global__ __launch_bounds__(LANES* KW) void k_yee2d_tbk(const bz_sim2d* __restrict__ sims,
                                                        const int* __restrict__ tiles, int ntiles, int K,
                                                        const int* __restrict__ step, int q, int rd,
                                                        const int* __restrict__ ent_start,
                                                        const int* __restrict__ ents, float* __restrict__ stage_all,
                                                        Gather2D G) {
    constexpr int RH = KW * B;
    __shared__ float ez_top[KW][LANES];
    __shared__ float hy_bot[KW][LANES];
    __shared__ float amps[32];          // the signal's K <= 31 samples of this launch
    extern __shared__ float smp[];      // [3][RH][LANES] when there are monitor points (launch-sized)
    const int npts = G.npts;
    const int t = blockIdx.x;
    if (t >= ntiles) {
        // monitor samples of the previous launch (q > 0): staging slot (q - 1) & 1, which this launch does not
        // write
        gather_k(G, stage_all + (long)((q - 1) & 1) * K * npts, *step + K * (q - 1), K,
                 (long)(t - ntiles) * (LANES * KW) + threadIdx.y * LANES + threadIdx.x);
        return;
    }
    float* __restrict__ stage = stage_all + (long)(q & 1) * K * npts;
    const int s = tiles[3 * t];
    if (s < 0) return;   // padding entry of a reordered tile list (uniform over the workgroup)
    const bz_sim2d& S = sims[s];
    const int tx = tiles[3 * t + 1], ty = tiles[3 * t + 2];
    const int nx = S.nx, ny = S.ny, ld = S.ld;
    const int l = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int i = tx * (LANES - 2 * K) - K + l;
    const int jb = ty * (RH - 2 * K) - K + w * B;   // grid row of the wave's first register row
    const uint32_t ic = (uint32_t)min(max(i, 0), nx - 1);
    const int wr = rd ^ 1;
    const int n0 = *step + K * q;
    Coef c;
    c.kh = S.kh;
    c.kbh = S.kbh;
    c.ke = S.ke;
    c.kbe = S.kbe;
    const float rdx = S.rdx, rdy = S.rdy;
    // 1 on square cells; uniform, so kept in a scalar register
    const float rxy = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(rdx / rdy)));
    const bool hxok = i < nx - 1;
    const bool xin = i >= 1 && i <= nx - 2;
    // the source box, held in registers (the step loop stores to memory, so fields read through S would be
    // reloaded on every row)
    const int sx0 = S.src_x0, sy0 = S.src_y0, ssx = S.src_sx, ssy = S.src_sy;
    const float* __restrict__ sadd = S.src_add;
    const float* __restrict__ smul = S.src_mul;
    const float* __restrict__ sgl = S.signal;
    const bool src_col = ssx > 0 && i >= sx0 && i < sx0 + ssx;
    // does the source box meet this wave's rows and this region's columns (uniform over the wave)
    const int rx0 = tx * (LANES - 2 * K) - K;
    const bool src_band = ssx > 0 && jb < sy0 + ssy && jb + B > sy0 && rx0 < sx0 + ssx && rx0 + LANES > sx0;
    // every row of the wave and every column of the region is interior to the grid (the usual case): no edge
    // rules per row or per lane
    const bool inner = jb >= 1 && jb + B - 1 <= ny - 2 && rx0 >= 1 && rx0 + LANES - 1 <= nx - 2;
    // does the snapshot box meet this wave's rows and this region's inner columns (uniform over the wave)
    // the snapshot box in LDS (written before the first barrier, read after it): in registers it would add
    // scalar spills to every wave
    __shared__ SnapBox snq;
    if (threadIdx.x == 0 && threadIdx.y == 0) snq = snap_box(S);
    const bool snap_band = S.snap && jb < S.snap_y0 + S.snap_sy && jb + B > S.snap_y0 &&
                           rx0 + K < S.snap_x0 + S.snap_sx && rx0 + LANES - K > S.snap_x0;
    const bool lane_inner = l >= K && l < LANES - K;

    // fb and cb hold fb * rdy and cb * rdy (upd_hx, upd_hy, upd_ez)
    float ez[B], hx[B], hy[B], fa[B], fb[B], ca[B], cb[B];
    {
        const float* __restrict__ gez = S.ez[rd];
        const float* __restrict__ ghx = S.hx[rd];
        const float* __restrict__ ghy = S.hy[rd];
        const float* __restrict__ cf = S.coef;
        uint32_t o[B];
#pragma unroll
        for (int r = 0; r < B; ++r) {
            o[r] = ((uint32_t)min(max(jb + r, 0), ny - 1) * (uint32_t)ld + ic) * 4u;
            ez[r] = at(gez, o[r]);
            hx[r] = at(ghx, o[r]);
            hy[r] = at(ghy, o[r]);
        }
        if (cf) {
            // precomputed by k_coef2d (the same roundings), one 16-byte (fa, fb, ca, cb) per cell
#pragma unroll
            for (int r = 0; r < B; ++r) {
                const float4 q = at4(cf, 4u * o[r]);
                fa[r] = q.x;
                fb[r] = q.y * rdy;
                ca[r] = q.z;
                cb[r] = q.w * rdy;
            }
        } else {
#pragma unroll
            for (int r = 0; r < B; ++r) {
                const float sg = at(S.sig, o[r]);
                h_coef(sg, c, fa[r], fb[r]);
                e_coef(sg, at(S.eps, o[r]), c, ca[r], cb[r]);
                fb[r] *= rdy;
                cb[r] *= rdy;
            }
        }
    }
    const int e0 = ent_start ? ent_start[t] : 0, e1 = ent_start ? ent_start[t + 1] : 0;
    const int tid = w * LANES + l;
    // the tile's monitor entries: thread tid stages entries e0 + tid, e0 + tid + LANES*KW, ...; the first one
    // is read here, once for all K steps
    const int ef = e0 + tid;
    const int epnt = ef < e1 ? ents[2 * ef] : -1;
    const int ecode = ef < e1 ? ents[2 * ef + 1] : 0;

    // the launch's signal samples, staged in LDS and read after the step's barriers: a vector load inside the
    // step loop would wait behind the snapshot and staging stores of the previous step (the vector memory
    // counter retires in order)
    if (sgl && tid < K) amps[tid] = gload(sgl, n0 + tid);
    for (int k = 0; k < K; ++k) {
        // H of step n0+k+1: Hy of the band's last row needs Ez of the next band's first row
        ez_top[w][l] = ez[0];
        __syncthreads();
        const float ezb = w < KW - 1 ? ez_top[w + 1][l] : ez[B - 1];
        auto h_rows = [&](auto check) {
            constexpr bool C = decltype(check)::value;
#pragma unroll
            for (int r = 0; r < B; ++r) {
                const int j = jb + r;
                const float ezR = from_right(ez[r]);
                const float ed = r < B - 1 ? ez[r + 1] : ezb;
                if (C) {
                    const float hn = computed(upd_hx(fa[r], fb[r], hx[r], ezR, ez[r]));
                    hx[r] = hxok ? hn : hx[r];
                } else {
                    hx[r] = upd_hx(fa[r], fb[r], hx[r], ezR, ez[r]);
                }
                if (!C || j < ny - 1) hy[r] = upd_hy(fa[r], fb[r], rxy, hy[r], ed, ez[r]);
            }
        };
        if (inner) h_rows(std::false_type{});
        else h_rows(std::true_type{});
        // snapshot of call-relative step q*K+k at the inner cells: the Ez this step reads, the Hx, Hy it wrote
        if (snap_band) {
#pragma unroll
            for (int r = 0; r < B; ++r) {
                const int rr = w * B + r;
                if (rr >= K && rr < RH - K && lane_inner) snap_put(snq, q * K + k, i, jb + r, ez[r], hx[r], hy[r]);
            }
        }
        // E: Ez of the band's first row needs the new Hy of the previous band's last row
        hy_bot[w][l] = hy[B - 1];
        __syncthreads();
        const float hya = w > 0 ? hy_bot[w - 1][l] : hy[0];
        auto e_rows = [&](auto check) {
            constexpr bool C = decltype(check)::value;
#pragma unroll
            for (int r = 0; r < B; ++r) {
                const int j = jb + r;
                const float hxl = from_left(hx[r]);
                const float hu = r > 0 ? hy[r - 1] : hya;
                if (!C) {
                    ez[r] = upd_ez(ca[r], cb[r], rxy, ez[r], hx[r], hxl, hy[r], hu);
                } else if (j >= 1 && j <= ny - 2) {
                    const float en = computed(upd_ez(ca[r], cb[r], rxy, ez[r], hx[r], hxl, hy[r], hu));
                    ez[r] = xin ? en : ez[r];
                }
            }
        };
        if (inner) e_rows(std::false_type{});
        else e_rows(std::true_type{});
        // the source box, after E (it touches only the Ez of its own cells)
        if (src_band) {
#pragma unroll
            for (int r = 0; r < B; ++r) {
                const int j = jb + r;
                if (j >= sy0 && j < sy0 + ssy && src_col) {
                    const long rr = (long)(j - sy0) * ssx + (i - sx0);
                    const float amp = sgl ? amps[k] : 0.0f;        // written before the first barrier
                    const float a = SEL ? box_amp(S, rr, n0 + k, amp) : amp;
                    ez[r] = (smul ? gload(smul, rr) : 1.0f) * ez[r] + gload(sadd, rr) * a;
                }
            }
        }
        // monitor samples of step n0+k owned by this tile, all three components behind one barrier. The next
        // step's writes to smp come after its two barriers, which every thread reaches only after these reads.
        if (e1 > e0) {
#pragma unroll
            for (int r = 0; r < B; ++r) {
                const int o = (w * B + r) * LANES + l;
                smp[o] = ez[r];
                smp[RH * LANES + o] = hx[r];
                smp[2 * RH * LANES + o] = hy[r];
            }
            __syncthreads();
            // code = (region row << 10) | (region column << 2) | component
            auto sample = [&](int code) { return smp[((code & 3) * RH + (code >> 10)) * LANES + ((code >> 2) & 255)]; };
            if (epnt >= 0) gstore(stage, (long)k * npts + epnt, sample(ecode));
            for (int e = ef + LANES * KW; e < e1; e += LANES * KW)
                gstore(stage, (long)k * npts + ents[2 * e], sample(ents[2 * e + 1]));
        }
    }

    // the inner cells of the region hold step n0+K
    if (l < K || l >= LANES - K || i >= nx) return;
    float* __restrict__ oez = S.ez[wr];
    float* __restrict__ ohx = S.hx[wr];
    float* __restrict__ ohy = S.hy[wr];
#pragma unroll
    for (int r = 0; r < B; ++r) {
        const int rr = w * B + r, j = jb + r;
        if (rr >= K && rr < RH - K && j < ny) {
            const uint32_t o = ((uint32_t)j * (uint32_t)ld + (uint32_t)i) * 4u;
            put(oez, o, ez[r]);
            put(ohx, o, hx[r]);
            put(ohy, o, hy[r]);
        }
    }
}

template <int B, int KW>
int runk(const bz_sim2d* sims, const int* tiles, int ntiles, int K, int* step, int rd0, int nlaunch,
         const int* ent_start, const int* ents, float* stage, const Gather2D& G, bool sel, hipStream_t st) {
    const dim3 b(LANES, KW);
    // sample staging in LDS, [3][KW*B][LANES] floats, only when there are monitor points
    const size_t lds = G.npts > 0 ? 3u * KW * B * LANES * sizeof(float) : 0;
    const long gather_threads = (long)G.npts * (G.nfreq > 0 ? G.nfreq : 1);
    const int gather_blocks = (int)((gather_threads + LANES * KW - 1) / (LANES * KW));
    // the gather of launch q-1 runs as extra workgroups of launch q; the last launch's gets a launch of its own
    for (int q = 0; q < nlaunch; ++q) {
        const dim3 g(ntiles + (q > 0 && G.npts > 0 ? gather_blocks : 0));
        if (sel)
            hipLaunchKernelGGL((k_yee2d_tbk<B, KW, true>), g, b, lds, st, sims, tiles, ntiles, K, step, q, (rd0 + q) & 1,
                               ent_start, ents, stage, G);
        else
            hipLaunchKernelGGL((k_yee2d_tbk<B, KW, false>), g, b, lds, st, sims, tiles, ntiles, K, step, q,
                               (rd0 + q) & 1, ent_start, ents, stage, G);
        BZ_CHECK_LAUNCH("bz_yee2d_runk(steps)");
    }
    if (G.npts > 0) {
        const int q = nlaunch - 1;
        hipLaunchKernelGGL(k_gather2d_k, dim3((int)((gather_threads + 255) / 256)), dim3(256), 0, st, G,
                           stage + (long)(q & 1) * K * G.npts, step, q, K);
        BZ_CHECK_LAUNCH("bz_yee2d_runk(gather)");
    }
    hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, st, step, K * nlaunch);
    BZ_CHECK_LAUNCH("bz_yee2d_runk(tick)");
    return BZ_OK;
}

}  // namespace

extern "C" int bz_yee2d_run(const bz_sim2d* sims, int nsims, const int* tiles, int ntiles, int seg, int* step,
                            int nsteps, const int* pts, int npts, const int* slot, const int* grp, int nslot,
                            int nrow, float* series, int nfreq, const double* phasor, double* dft, void* stream) {
    BZ_REQUIRE(sims && nsims > 0 && tiles && ntiles > 0 && step && nsteps >= 0, "bz_yee2d_run: bad arguments");
    BZ_REQUIRE(npts >= 0 && (npts == 0 || (pts && slot)), "bz_yee2d_run: monitor points without a slot table");
    BZ_REQUIRE(nfreq >= 0 && (nfreq == 0 || (phasor && dft)), "bz_yee2d_run: DFT requested without buffers");
    BZ_REQUIRE(!grp || (nslot > 0 && nrow > 0), "bz_yee2d_run: point groups need nslot > 0 and nrow > 0");
    if (nsteps == 0) return BZ_OK;
    Gather2D G;
    G.pts = pts;
    G.npts = npts;
    G.slot = slot;
    G.grp = grp;
    G.nslot = nslot;
    G.nrow = nrow;
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

extern "C" int bz_yee2d_hz_run(const bz_sim2d* sims, int nsims, const int* tiles, int ntiles, int seg, int* step,
                               int nsteps, const int* pts, int npts, const int* slot, const int* grp, int nslot,
                               int nrow, float* series, int nfreq, const double* phasor, double* dft,
                               void* stream) {
    BZ_REQUIRE(sims && nsims > 0 && tiles && ntiles > 0 && step && nsteps >= 0, "bz_yee2d_hz_run: bad arguments");
    BZ_REQUIRE(npts >= 0 && (npts == 0 || (pts && slot)), "bz_yee2d_hz_run: monitor points without a slot table");
    BZ_REQUIRE(nfreq >= 0 && (nfreq == 0 || (phasor && dft)), "bz_yee2d_hz_run: DFT requested without buffers");
    BZ_REQUIRE(!grp || (nslot > 0 && nrow > 0), "bz_yee2d_hz_run: point groups need nslot > 0 and nrow > 0");
    if (nsteps == 0) return BZ_OK;
    Gather2D G;
    G.pts = pts;
    G.npts = npts;
    G.slot = slot;
    G.grp = grp;
    G.nslot = nslot;
    G.nrow = nrow;
    G.series = series;
    G.nfreq = nfreq;
    G.phasor = phasor;
    G.dft = dft;
    const hipStream_t st = (hipStream_t)stream;
    switch (seg) {
        case 2: return run<2, true>(sims, tiles, ntiles, step, nsteps, G, st);
        case 4: return run<4, true>(sims, tiles, ntiles, step, nsteps, G, st);
        case 8: return run<8, true>(sims, tiles, ntiles, step, nsteps, G, st);
        case 16: return run<16, true>(sims, tiles, ntiles, step, nsteps, G, st);
        default: bz::set_error("bz_yee2d_hz_run: seg must be 2, 4, 8 or 16"); return BZ_EINVAL;
    }
}

extern "C" int bz_yee2d_run2(const bz_sim2d* sims, int nsims, const int* tiles, int ntiles, int seg, int* step,
                             int rd0, int npairs, const int* pts, const int* spts, int npts, const int* slot,
                             const int* grp, int nslot, int nrow, float* series, int nfreq, const double* phasor,
                             double* dft, void* stream) {
    BZ_REQUIRE(sims && nsims > 0 && tiles && ntiles > 0 && step && npairs >= 0 && seg >= 2 && (rd0 == 0 || rd0 == 1),
               "bz_yee2d_run2: bad arguments");
    BZ_REQUIRE(npts >= 0 && (npts == 0 || (pts && spts && slot)), "bz_yee2d_run2: monitor points without tables");
    BZ_REQUIRE(nfreq >= 0 && (nfreq == 0 || (phasor && dft)), "bz_yee2d_run2: DFT requested without buffers");
    BZ_REQUIRE(!grp || (nslot > 0 && nrow > 0), "bz_yee2d_run2: point groups need nslot > 0 and nrow > 0");
    if (npairs == 0) return BZ_OK;
    Gather2D G;
    G.pts = pts;
    G.npts = npts;
    G.slot = slot;
    G.grp = grp;
    G.nslot = nslot;
    G.nrow = nrow;
    G.series = series;
    G.nfreq = nfreq;
    G.phasor = phasor;
    G.dft = dft;
    const hipStream_t st = (hipStream_t)stream;
    const dim3 b(LANES, WPG);
    const long gather_threads = (long)npts * (nfreq > 0 ? nfreq : 1);
    BZ_REQUIRE(gather_threads / GPB < (1L << 31) - 1, "bz_yee2d_run2: too many monitor points x frequencies");
    const int gather_blocks = (int)((gather_threads + GPB - 1) / GPB);
    // the gather of pair q-1 runs as extra workgroups of pair q's launch; the last pair's gets its own launch
    const int tile_blocks = (ntiles + WPG - 1) / WPG;
    for (int q = 0; q < npairs; ++q) {
        const int rd = (rd0 + q) & 1;
        const dim3 gq(tile_blocks + (q > 0 && npts > 0 ? gather_blocks : 0));
        hipLaunchKernelGGL(k_yee2d_tb2, gq, b, 0, st, sims, tiles, ntiles, seg, step, q, rd, G, spts);
        BZ_CHECK_LAUNCH("bz_yee2d_run2(steps)");
    }
    if (npts > 0) {
        hipLaunchKernelGGL(k_gather2d_pair, dim3(gather_blocks), dim3(GPB), 0, st, sims, G, spts, step, npairs - 1,
                           ((rd0 + npairs - 1) & 1) ^ 1);
        BZ_CHECK_LAUNCH("bz_yee2d_run2(gather)");
    }
    hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, st, step, 2 * npairs);
    BZ_CHECK_LAUNCH("bz_yee2d_run2(tick)");
    return BZ_OK;
}

extern "C" int bz_yee2d_runk(const bz_sim2d* sims, int nsims, const int* tiles, int ntiles, int rows, int waves,
                             int ksteps, int* step, int rd0, int nlaunch, const int* ent_start, const int* ents,
                             float* stage, int npts, const int* slot, const int* grp, int nslot, int nrow,
                             float* series, int nfreq, const double* phasor, double* dft, int src_sel,
                             void* stream) {
    BZ_REQUIRE(sims && nsims > 0 && tiles && ntiles > 0 && step && nlaunch >= 0 && (rd0 == 0 || rd0 == 1),
               "bz_yee2d_runk: bad arguments");
    const int shape = rows * 100 + waves;
    BZ_REQUIRE(shape == 804 || shape == 808 || shape == 816 || shape == 408 || shape == 416 || shape == 1604,
               "bz_yee2d_runk: (rows, waves) must be (8, 4), (8, 8), (8, 16), (4, 8), (4, 16) or (16, 4)");
    BZ_REQUIRE(ksteps >= 1 && 2 * ksteps < LANES && 2 * ksteps < waves * rows,
               "bz_yee2d_runk: ksteps must leave inner cells (2 * ksteps < 64 and < rows * waves)");
    BZ_REQUIRE(npts >= 0 && (npts == 0 || (ent_start && ents && stage && slot)),
               "bz_yee2d_runk: monitor points without tables");
    BZ_REQUIRE(nfreq >= 0 && (nfreq == 0 || (phasor && dft)), "bz_yee2d_runk: DFT requested without buffers");
    BZ_REQUIRE(!grp || (nslot > 0 && nrow > 0), "bz_yee2d_runk: point groups need nslot > 0 and nrow > 0");
    BZ_REQUIRE((long)npts * (nfreq > 0 ? nfreq : 1) / 256 < (1L << 31) - 1,
               "bz_yee2d_runk: too many monitor points x frequencies");
    if (nlaunch == 0) return BZ_OK;
    Gather2D G;
    G.pts = nullptr;
    G.npts = npts;
    G.slot = slot;
    G.grp = grp;
    G.nslot = nslot;
    G.nrow = nrow;
    G.series = series;
    G.nfreq = nfreq;
    G.phasor = phasor;
    G.dft = dft;
    const hipStream_t st = (hipStream_t)stream;
    if (npts == 0) ent_start = ents = nullptr;
#define BZ_RUNK(R, W) \
    runk<R, W>(sims, tiles, ntiles, ksteps, step, rd0, nlaunch, ent_start, ents, stage, G, src_sel != 0, st)
    switch (shape) {
        case 804: return BZ_RUNK(8, 4);
        case 808: return BZ_RUNK(8, 8);
        case 816: return BZ_RUNK(8, 16);
        case 408: return BZ_RUNK(4, 8);
        case 416: return BZ_RUNK(4, 16);
        default: return BZ_RUNK(16, 4);
    }
#undef BZ_RUNK
}

extern "C" int bz_yee2d_coef(const bz_sim2d* sims, int nsims, int max_cells, void* stream) {
    BZ_REQUIRE(sims && nsims > 0 && nsims <= 65535 && max_cells >= 0, "bz_yee2d_coef: bad arguments");
    const int bx = (int)std::min<long>(64, ((long)max_cells + 255) / 256 + 1);
    hipLaunchKernelGGL(k_coef2d, dim3(bx, nsims), dim3(256), 0, (hipStream_t)stream, sims);
    BZ_CHECK_LAUNCH("bz_yee2d_coef");
    return BZ_OK;
}
