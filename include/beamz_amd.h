/*
 * beamz_amd C ABI: MI355X (gfx950) FDTD engine behind the beamz Python API.
 *
 * Conventions (all entry points):
 *   - return 0 on success, a negative BZ_E* code on failure; the message of
 *     the last failure on the calling thread is returned by bz_last_error().
 *   - device pointers are plain float / double pointers into memory the
 *     caller owns (the Python layer allocates them as torch tensors); the
 *     library never frees them.
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream). All
 *     launches are asynchronous on that stream.
 *   - 2D arrays are row-major [ny][ld] with x fastest, ld >= nx; 3D arrays are
 *     [nz][ny][ld]. Every field component is stored on the full padded box;
 *     entries outside the reference's staggered shape stay zero.
 *
 * Reference interfaces replaced (paths relative to the beamz repository):
 *   bz_update_h_2d ........ Backend.update_h_fields, beamz/simulation/backends/base.py:36-39,
 *                           implemented by numpy_backend.py:34-61
 *   bz_update_e_2d ........ Backend.update_e_field, base.py:41-44; numpy_backend.py:63-88
 *   bz_update_3d .......... FDTD._update_3d_fields, beamz/simulation/fdtd.py:200-310
 *   bz_add_points / bz_affine_points
 *                           source injection, beamz/simulation/helper.py:5-118
 *   bz_gather / bz_gather2 / bz_gather3  monitor sampling of one step / of the steps of a two- or
 *                           three-step pass,
 *                           beamz/simulation/helper.py:231-256 with beamz/devices/monitors.py:279-397,
 *                           plus the on-the-fly DFT (no reference counterpart)
 *   bz_grid_power_2d/3d ... grid power accumulation, beamz/simulation/helper.py:120-187
 *   bz_grid_power_2d/3d_cplx  the same for complex fields (a real and an imaginary run)
 *   bz_raster_2d/3d ....... RegularGrid.__rasterize__ / RegularGrid3D.__rasterize_3d__ +
 *                           _process_3d_pml, beamz/simulation/meshing.py:92-536, 616-781,
 *                           with PML.get_conductivity, beamz/design/pml.py:21-62
 *   bz_yee3d_step ......... one fused 3D step: FDTD._update_3d_fields (fdtd.py:200-310) followed by
 *                           the soft-source adds of helper.py:62-118, i.e. the H, E, sources part of
 *                           the step order of fdtd.py:411-429, in one pass over HBM
 *   bz_yee2d_run .......... fused multi-step 2D engine for many independent simulations at once:
 *                           per step numpy_backend.py:34-88 (H, E), helper.py:96-118 (source),
 *                           helper.py:231-256 (monitor samples + DFT), helper.py:120-187 (power)
 *   bz_yee2d_hz_run ....... the (Hz, Ex, Ey) polarisation of bz_yee2d_run: the reference's 3D update
 *                           fdtd.py:200-310 restricted to d/dz = 0 (its 2D path numpy_backend.py:34-88 has
 *                           only the (Ez, Hx, Hy) family)
 *   bz_yee3d_step2 ........ two consecutive bz_yee3d_step steps in one pass (same bits)
 *   bz_yee3d_step3 ........ three consecutive bz_yee3d_step steps in one pass (same bits)
 *   bz_yee2d_run2 ......... two consecutive bz_yee2d_run steps per pass (same bits, no grid power)
 *   bz_yee2d_runk ......... K consecutive bz_yee2d_run steps per pass (same bits, no grid power), with
 *                           bz_yee2d_coef filling the per-cell update coefficients it reads
 *   bz_adj2d_run / _inject  reverse-time adjoint of the 2D step for the topology gradient of
 *                           examples/4_topo.py:255-321 (the reference differentiates with a second
 *                           forward run, 4_topo.py:282-285; this is the exact discrete adjoint);
 *                           bz_adj2d_run2 / _runk: two / K reverse steps per pass (same bits)
 *   bz_abi_sizes .......... descriptor struct sizes for binding checks; no reference counterpart
 *   bz_nn_* ............... CNN surrogate kernels; no reference counterpart
 */
#ifndef BEAMZ_AMD_H
#define BEAMZ_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BZ_OK 0
#define BZ_EINVAL (-1)   /* bad argument (shape, null pointer, out-of-range index) */
#define BZ_EHIP (-2)     /* a HIP runtime call failed */
#define BZ_ENOTSUP (-3)  /* unsupported configuration */

const char* bz_last_error(void);
int bz_version(void);
int bz_device_count(int* count);

/* ---------------- 2D (Ez, Hx, Hy) plugin updates ----------------
 * Arrays in the reference's shapes, Ez/sigma/eps_r [ny][nx], Hx [ny][nx-1], Hy [ny-1][nx], each with its
 * own row pitch ld_* (elements). Updated in place (the reference returns new H arrays). */
int bz_update_h_2d(float* Hx, int ld_hx, float* Hy, int ld_hy, const float* Ez, int ld_ez,
                   const float* sigma, int ld_sig, int nx, int ny, double dx, double dy, double dt,
                   double mu0, double eps0, void* stream);
int bz_update_e_2d(float* Ez, int ld_ez, const float* Hx, int ld_hx, const float* Hy, int ld_hy,
                   const float* sigma, int ld_sig, const float* eps_r, int ld_eps, int nx, int ny,
                   double dx, double dy, double dt, double eps0, void* stream);

/* ---------------- 3D six-component update (H then E) ---------------- */
/* f[0..5] = Ex, Ey, Ez, Hx, Hy, Hz, each [nz][ny][ld]. */
int bz_update_3d(float* const f[6], const float* eps_r, const float* sigma, int nx, int ny, int nz, int ld,
                 double dx, double dy, double dz, double dt, double mu0, double eps0, void* stream);

/* ---------------- fused 3D step ----------------
 * One time step (H, E, sources) of fdtd.py:200-310 + helper.py:5-118 in a single pass:
 * reads src[6], writes every cell of dst[6] (ping-pong buffers; src != dst).
 * Box sources: inside box b, after the E update, E_comp = mul * E_comp + add * amp
 * (mul NULL means 1). A GaussianSource patch is a box with mul NULL; a ModeSource
 * line/plane is a box whose mul zeroes the upstream cells (helper.py:37-54). */
#define BZ_MAX_BOX 4
typedef struct bz_box_source {
    int comp;               /* 3D: 0 Ex, 1 Ey, 2 Ez; 2D: 0 Ez */
    int x0, y0, z0, sx, sy, sz;
    const float* add;       /* device, sz*sy*sx, x fastest */
    const float* mul;       /* device or NULL */
    float amp;              /* signal value of this step */
} bz_box_source;

/* kseg: z-planes per wave segment (<= 0: default 16); rows: y-rows per wave (2 or 4). */
int bz_yee3d_step(float* const src[6], float* const dst[6], const float* eps_r, const float* sigma,
                  int nx, int ny, int nz, int ld, double dx, double dy, double dz, double dt,
                  double mu0, double eps0, const bz_box_source* boxes, int nbox, int kseg, int rows, void* stream);

/* ---------------- two-step 3D kernel ----------------
 * Steps n and n+1 of bz_yee3d_step in one pass: reads src[6] (step n), writes every cell of dst[6]
 * (step n+2); bit-identical to two bz_yee3d_step calls. boxes[q].amp is the amplitude of step n,
 * amp1[q] that of step n+1. If stage is not NULL, the fields of step n+1 inside the box
 * stage_box = {x0, y0, z0, sx, sy, sz} are written to stage[c][z][y][x] (c = Ex, Ey, Ez, Hx, Hy, Hz)
 * for monitor sampling. kseg: z-planes per segment (>= 2; <= 0: 32); rows: y-rows per wave (2, 3, 4);
 * waves: waves per workgroup stacked in y (8, or 16 with rows 2). */
int bz_yee3d_step2(float* const src[6], float* const dst[6], const float* eps_r, const float* sigma,
                   int nx, int ny, int nz, int ld, double dx, double dy, double dz, double dt,
                   double mu0, double eps0, const bz_box_source* boxes, const float* amp1, int nbox,
                   float* stage, const int* stage_box, int kseg, int rows, int waves, void* stream);

/* ---------------- three-step 3D kernel ----------------
 * Steps n, n+1 and n+2 of bz_yee3d_step (fdtd.py:200-310 with the soft sources of helper.py:62-118, the
 * H, E, sources part of the step order of fdtd.py:411-429) in one pass: reads src[6] (step n), writes every
 * cell of dst[6] (step n+3); bit-identical to three bz_yee3d_step calls. boxes[q].amp is the amplitude
 * of step n, amp1[q] and amp2[q] those of steps n+1 and n+2. stage1 / stage2 (each may be NULL) receive
 * the fields of steps n+1 / n+2 inside stage_box = {x0, y0, z0, sx, sy, sz}, laid out as for
 * bz_yee3d_step2. kseg: z-planes per segment (>= 2; <= 0: 32); (rows, waves): (2, 8), (3, 8), (1, 16)
 * or (2, 16). */
int bz_yee3d_step3(float* const src[6], float* const dst[6], const float* eps_r, const float* sigma,
                   int nx, int ny, int nz, int ld, double dx, double dy, double dz, double dt,
                   double mu0, double eps0, const bz_box_source* boxes, const float* amp1, const float* amp2,
                   int nbox, float* stage1, float* stage2, const int* stage_box, int kseg, int rows, int waves,
                   void* stream);

/* Monitor samples of the three steps of a bz_yee3d_step3 pass (helper.py:231-256 with
 * monitors.py:279-397, plus the on-the-fly DFT): steps n+1 and n+2 from base0 / base1 at off0 (the staging
 * boxes), step n+3 from base2 at off2 (the fields); act0..act2 select the steps sampled. Each DFT
 * accumulator receives its terms in step order, as from three bz_gather calls. */
int bz_gather3(const float* const base0[6], const float* const base1[6], const int64_t* off0,
               const float* const base2[6], const int* comp, const int64_t* off2, int n, int act0, int act1,
               int act2, float* series0, float* series1, float* series2, int nfreq, const double* phasor0,
               const double* phasor1, const double* phasor2, double* dft, void* stream);

/* ---------------- allocation-free two-step 3D pass with monitor sampling ----------------
 * No counterpart in the reference. For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): bz_yee3d_step2 with a staging box
 * followed by bz_gather2 of both steps, where the staging box lives in a device array the library owns
 * (BZ_NOALLOC_STAGE_FLOATS floats per GPU, held until the process ends) instead of a caller buffer.
 *   stage_box = {x0, y0, z0, sx, sy, sz} (host), 6*sx*sy*sz <= BZ_NOALLOC_STAGE_FLOATS, else BZ_EINVAL;
 *   comp[npts], stage_off[npts] (offset into a component's [sz][sy][sx] block), field_off[npts] (offset
 *   into a field array) and the series / phasor / dft arguments are those of bz_gather2 with base0 = the
 *   staged step n+1 and base1 = dst (step n+2). npts == 0 or act0 == act1 == 0 samples nothing.
 * Allocates nothing (no hipMalloc, no library allocation). Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (two streams, two threads) race on it. */
#define BZ_NOALLOC_STAGE_FLOATS (1 << 20)   /* 4 MiB per GPU */
int bz_yee3d_step2_noalloc(float* const src[6], float* const dst[6], const float* eps_r, const float* sigma,
                           int nx, int ny, int nz, int ld, double dx, double dy, double dz, double dt,
                           double mu0, double eps0, const bz_box_source* boxes, const float* amp1, int nbox,
                           const int* stage_box, const int* comp, const int64_t* stage_off, const int64_t* field_off,
                           int npts, int act0, int act1, float* series0, float* series1, int nfreq,
                           const double* phasor0, const double* phasor1, double* dft, int kseg, int rows, int waves,
                           void* stream);

/* ---------------- fused, batched 2D engine ----------------
 * Runs nsteps full steps (numpy_backend.py:34-88, then the source box, then optional grid power,
 * then monitor sampling) of every simulation described in `sims` (device array), starting at the
 * step held in device memory `step` (advanced by nsteps when the run's work completes; the call
 * only enqueues launches, so it can be captured in a hipGraph and replayed). Field buffers are
 * ping-ponged by step parity: step n reads buffer n & 1 and writes buffer (n + 1) & 1.
 * tiles: device int[3*ntiles] = (sim, x-tile of 62 columns, y-tile of `seg` rows), seg one of
 * 2, 4, 8, 16; an entry with sim < 0 is padding and is skipped (tiles are taken four to a workgroup,
 * in list order, so the order decides which XCD runs which tile); every array of a simulation must
 * hold fewer than 2^31 elements.
 * pts: device int[3*npts] = (sim, comp 0 Ez / 1 Hx / 2 Hy, offset); slot: device int[steps],
 * slot[n] = recorded row of step n or -1; series [rows][npts] FP32 (may be NULL);
 * phasor [rows][nfreq] complex FP64; dft [nfreq][npts] complex FP64 accumulators.
 * Point groups (monitors with their own record steps and DFT frequencies, in one batch): with grp (device
 * int[npts], the group of each point) not NULL, slot is [groups][nslot] (slot[g*nslot + n]: row of step n
 * in group g's own numbering, or -1) and phasor is [groups][nrow][nfreq] (a group with fewer frequencies
 * holds 0 in the others, whose accumulators then stay 0). grp NULL: one group, nslot and nrow unused.
 * Snapshot box (snap_sx > 0): the q-th step of the call (q = 0..nsteps-1) stores, for every cell
 * (j, i) of the box, the Ez it read and the Hx, Hy it wrote, at
 * snap[((q*3 + c)*snap_sy + j - snap_y0)*snap_sx + i - snap_x0], c = 0 Ez, 1 Hx, 2 Hy
 * (the forward states the reverse-time adjoint needs inside a design region). */
typedef struct bz_sim2d {
    int nx, ny, ld;
    float* ez[2];
    float* hx[2];
    float* hy[2];
    const float* eps;
    const float* sig;
    float kh, kbh, ke, kbe;   /* dt/(2 eps0), dt/mu0, dt/(2 eps0), dt/eps0 */
    float rdx, rdy;
    int src_x0, src_y0, src_sx, src_sy;   /* box source on Ez (src_sx = 0: none) */
    const float* src_add;
    const float* src_mul;                 /* NULL means 1 */
    const float* signal;                  /* device FP32 amplitude per step (NULL: no source) */
    double* power;                        /* NULL or [ny][ld] FP64 grid-power accumulator */
    int snap_x0, snap_y0, snap_sx, snap_sy;   /* snapshot box (snap_sx = 0: none) */
    float* snap;                          /* [nsteps][3][snap_sy][snap_sx] */
    float* stg;                           /* bz_yee2d_run2: [2][3][stg_sy][stg_sx] staging of the first step's */
    int stg_x0, stg_y0, stg_sx, stg_sy;   /* Ez, Hx, Hy for monitor samples, slot q & 1 for pair q (NULL: none) */
    float* coef;                          /* NULL, or [ny][ld][4] FP32 update coefficients (fa, fb, ca, cb) of every
                                           * cell as bz_yee2d_coef writes them (read by bz_yee2d_runk) */
    const unsigned char* src_sel;         /* NULL (one source), or per source-box cell the index of the source it
                                           * belongs to: several sources on disjoint cells in one box, whose
                                           * amplitude at step n is signal[src_sel[cell] * src_sig_len + n] */
    int src_sig_len;                      /* samples per source in signal when src_sel is set */
} bz_sim2d;

int bz_yee2d_run(const bz_sim2d* sims, int nsims, const int* tiles, int ntiles, int seg, int* step, int nsteps,
                 const int* pts, int npts, const int* slot, const int* grp, int nslot, int nrow, float* series,
                 int nfreq, const double* phasor, double* dft, void* stream);

/* The (Hz, Ex, Ey) polarisation: nsteps steps of the z-invariant restriction of fdtd.py:200-310 (the
 * reference's 3D update with every d/dz = 0, kept to the components Hz (ny-1, nx-1), Ex (ny, nx-1),
 * Ey (ny-1, nx)); the reference's 2D path has only the (Ez, Hx, Hy) family (numpy_backend.py:34-88), and
 * this is its E <-> H counterpart. Same arguments as bz_yee2d_run, with the descriptor's ez / hx / hy slots
 * holding Hz / Ex / Ey and monitor comp 0 / 1 / 2 meaning Hz / Ex / Ey. The source box is added to Hz after
 * the E update; grid power accumulates Sx^2 + Sy^2 with (Sx, Sy) = (Ey Hz, -Ex Hz) on the Hz cells' box.
 * No snapshot boxes. */
int bz_yee2d_hz_run(const bz_sim2d* sims, int nsims, const int* tiles, int ntiles, int seg, int* step,
                    int nsteps, const int* pts, int npts, const int* slot, const int* grp, int nslot, int nrow,
                    float* series, int nfreq, const double* phasor, double* dft, void* stream);

/* Two steps per pass: npairs pairs of the bz_yee2d_run step (same results, bit for bit), pair q reading
 * buffer (rd0 + q) & 1 and writing the other, so after the call the state sits in buffer
 * (rd0 + npairs) & 1 and the step counter has advanced by 2 * npairs. tiles: (sim, x-tile of 60
 * columns, y-tile of `seg` rows), seg >= 2. Monitor samples of the first step of a pair come from each
 * simulation's staging box: spts[p] = offset of point p inside its box (-1: not in it, reads 0).
 * Snapshot boxes as for bz_yee2d_run: pair q writes slots 2q and 2q+1. */
int bz_yee2d_run2(const bz_sim2d* sims, int nsims, const int* tiles, int ntiles, int seg, int* step, int rd0,
                  int npairs, const int* pts, const int* spts, int npts, const int* slot, const int* grp,
                  int nslot, int nrow, float* series, int nfreq, const double* phasor, double* dft, void* stream);

/* The update coefficients of every cell of every simulation whose coef is not NULL, from its eps and sig:
 * s = sig*kh, fa = (1-s)/(1+s), fb = kbh/(1+s) (H); s = sig*ke/eps, ca = (1-s)/(1+s), cb = kbe/eps/(1+s)
 * (E), rounded exactly as the step kernels round them. Must run again after eps or sig change. */
int bz_yee2d_coef(const bz_sim2d* sims, int nsims, int max_cells, void* stream);

/* Sizes of the descriptor structs as this library was compiled: out[0] = sizeof(bz_sim2d),
 * out[1] = sizeof(bz_adj2d), out[2] = sizeof(bz_box_source) (bindings check their own layouts). */
int bz_abi_sizes(int* out, int n);

/* K steps per pass: nlaunch launches of ksteps steps of the bz_yee2d_run step each (same results, bit for
 * bit), launch q reading buffer (rd0 + q) & 1 and writing the other; the step counter advances by
 * ksteps * nlaunch. A tile is a region of 64 columns x (waves * rows) rows, one workgroup of `waves` waves
 * each holding `rows` rows in registers, (rows, waves) one of (8, 4), (8, 8), (8, 16), (4, 8), (4, 16),
 * (16, 4); the region advances ksteps steps and its inner (64 - 2 ksteps) x (waves * rows - 2 ksteps)
 * cells are stored. tiles: (sim, x-tile, y-tile)
 * in units of the inner size, one tile per workgroup (sim < 0: padding). Monitor point p is sampled by the
 * tile whose inner cells hold it: ent_start[t] .. ent_start[t+1]-1 index ents[2e] = p,
 * ents[2e+1] = (region row << 10) | (region column << 2) | comp, sorted by comp within a tile;
 * src_sel: nonzero if any descriptor's src_sel is set (several sources in one box; the kernel variant that
 * reads the per-cell selector runs only then).
 * stage: device FP32 [2][ksteps][npts] (launch q writes slot q & 1). Snapshot boxes as for bz_yee2d_run:
 * step k of launch q writes slot q*ksteps + k. No grid power (use bz_yee2d_run / run2). */
int bz_yee2d_runk(const bz_sim2d* sims, int nsims, const int* tiles, int ntiles, int rows, int waves, int ksteps,
                  int* step, int rd0, int nlaunch, const int* ent_start, const int* ents, float* stage, int npts,
                  const int* slot, const int* grp, int nslot, int nrow, float* series, int nfreq,
                  const double* phasor, double* dft, int src_sel, void* stream);

/* ---------------- reverse-time 2D adjoint ----------------
 * No reference counterpart (the reference's gradient is the overlap heuristic of
 * examples/4_topo.py:282-285). Exact transpose of the bz_yee2d_run step for the objective
 * J = sum_f s_f |a_f|^2, a_f = sum_n sum_p c_fp Ez_{n+1}(p) phasor[n][f]: reverse step n maps the adjoint state
 * lam_{n+1} (buffer parity (n+1)&1) to lam_n (parity n&1) and, inside the design box, adds
 * dJ/d eps_r of step n to grad using the forward snapshot that bz_yee2d_run stored for the
 * segment (snapshot box = design box grown by one cell on the low x and y sides, slot
 * n - snap_step0). The design box must lie in the interior (1 <= x, x + sx <= nx - 1, same in y).
 * inj_w: complex dJ/dEz-sample weights 2 s_f conj(a_f) c_fp per frequency and cell of the injection box. */
typedef struct bz_adj2d {
    int nx, ny, ld;
    float* le[2];
    float* lhx[2];
    float* lhy[2];
    const float* eps;
    const float* sig;
    float kh, kbh, ke, kbe;
    float rdx, rdy;
    int src_x0, src_y0, src_sx, src_sy;
    const float* src_mul;                 /* the forward source box's multiplier (NULL: additive source) */
    int box_x0, box_y0, box_sx, box_sy;   /* design box */
    const float* snap;                    /* [K][3][box_sy+1][box_sx+1] forward snapshot of the segment */
    int snap_step0;
    double* grad;                         /* [box_sy][box_sx] FP64 accumulator (NULL: no gradient) */
    int inj_x0, inj_y0, inj_sx, inj_sy;
    const double* inj_w;                  /* [nf][inj_sy][inj_sx] complex, or NULL */
    const double* phasor;                 /* [num_steps][nf] complex phasors of the sample after step n */
    int nf;                               /* number of objective frequencies */
    const float* coef;                    /* NULL, or [ny][ld][4] update coefficients as bz_yee2d_coef writes
                                           * them for the same eps and sig (read by bz_adj2d_runk) */
    float* l2hist;                        /* NULL, or [K][box_sy][box_sx]: the reverse kernels store there the
                                           * lambda value each design cell's gradient term takes at step n (slot
                                           * n - snap_step0) instead of adding the term into grad;
                                           * bz_adj2d_grad then adds the terms, in the same order */
    float* inj_tab;                       /* NULL, or [num_steps][inj_sy][inj_sx]: the objective's injection
                                           * value of each cell at step n >= 1, as bz_adj2d_injtab computes it;
                                           * bz_adj2d_runk then reads it instead of summing inj_w * phasor */
} bz_adj2d;

/* Reverse steps n = n_hi-1 down to n_hi-nsteps of every adjoint described in `adjs` (device
 * array); tiles as for bz_yee2d_run (seg 2, 4 or 8; sim < 0 marks padding). */
int bz_adj2d_run(const bz_adj2d* adjs, int nadj, const int* tiles, int ntiles, int seg, int n_hi, int nsteps,
                 void* stream);
/* The same reverse steps two per launch (same bits), on tiles2 of 60 columns x seg2 rows (the layout of
 * bz_yee2d_run2's tiles), with a last single step on `tiles` (seg 2, 4 or 8) when nsteps is odd. A pair
 * reads lam_{n+1} from buffer (n + 1 + flip) & 1 and writes lam_{n-1} to the other buffer, so the caller
 * passes flip = 0 for the first call after bz_adj2d_inject and afterwards flip ^ ((nsteps / 2) & 1) of
 * the previous call; bz_adj2d_run is this with flip = 0 throughout. */
int bz_adj2d_run2(const bz_adj2d* adjs, int nadj, const int* tiles2, int ntiles2, int seg2, const int* tiles,
                  int ntiles, int seg, int n_hi, int nsteps, int flip, void* stream);
/* lE_n += sum_f Re(inj_w[f] * phasor[n-1][f]) over each injection box (max_inj >= the largest box's cell count):
 * the terminal condition lam_N = dJ/du_N before the first reverse step. */
/* K reverse steps per pass: nsteps / ksteps launches, each doing reverse steps n .. n-ksteps+1 over regions
 * of 64 columns x (rows * waves) rows ((rows, waves) as for bz_yee2d_runk) whose inner cells are stored,
 * then the nsteps % ksteps last steps with the one-step kernel (tiles, ntiles, seg as for bz_adj2d_run).
 * Same results, bit for bit, as bz_adj2d_run. The adjoint state of lam_m sits in buffer (m + flip) & 1;
 * a K-step launch moves it so that flip becomes (flip + ksteps - 1) & 1, which the caller tracks. */
int bz_adj2d_runk(const bz_adj2d* adjs, int nadj, const int* tilesk, int ntilesk, int rows, int waves, int ksteps,
                  const int* tiles, int ntiles, int seg, int n_hi, int nsteps, int flip, void* stream);

int bz_adj2d_inject(const bz_adj2d* adjs, int nadj, int max_inj, int n, void* stream);

/* grad[cell] += the gradient terms of reverse steps n = n_hi-1 down to n_lo, in that order, from the
 * snapshots and the l2hist values the reverse kernels stored (slot n - snap_step0), for every adjoint with
 * grad and l2hist set: the same bits as adding each term inside the reverse kernels, with the per-step
 * read-modify-write of grad taken off the reverse sweep. max_box >= every box_sx * box_sy. */
int bz_adj2d_grad(const bz_adj2d* adjs, int nadj, int max_box, int n_lo, int n_hi, void* stream);

/* inj_tab[n][cell] = the injection value sum_f Re(inj_w[f][cell] phasor[n-1][f]), rounded to FP32, for steps
 * n_lo <= n < n_hi (n >= 1) of every adjoint with inj_w and inj_tab set: the value the reverse kernels add,
 * so results do not change; max_inj >= every inj_sx * inj_sy. */
int bz_adj2d_injtab(const bz_adj2d* adjs, int nadj, int max_inj, int n_lo, int n_hi, void* stream);

/* Density filter of the topology parametrisation (replaces the blur of examples/4_topo.py:120-129 in the
 * product's filter -> projection chain): for each of B maps of H x W FP32 values,
 *   out[y][x] = oscale[y][x] * sum_{|dy|,|dx| <= r} K[dy+r][dx+r] * iscale[y+dy][x+dx] * in[y+dy][x+dx]
 * with zero padding (iscale and oscale may be NULL; K is (2r+1)^2 row-major, r <= 8; in != out). The
 * normalised cone filter is oscale = 1 / (K * 1); its gradient is the flipped K with iscale = 1 / (K * 1). */
int bz_filter2d(const float* in, float* out, int B, int H, int W, const float* K, int r, const float* iscale,
                const float* oscale, void* stream);

/* ---------------- CNN surrogate (BF16 MFMA, FP32 accumulate) ----------------
 * No reference counterpart: the surrogate maps a density map to port transmission spectra and,
 * by back-propagation, to d(spectrum)/d(density). Activations are NHWC bfloat16 (uint16 bit
 * patterns), convolution weights [Co][3][3][Ci], gradients and master weights FP32.
 * Gradient entry points accumulate (+=) into their FP32 outputs; the caller zeroes them. */
int bz_nn_conv3x3(const uint16_t* x, const uint16_t* w16, const float* bias, uint16_t* y, int B, int H, int W,
                  int Ci, int Co, int relu, void* stream);          /* Ci % 32 == 0, Co % 32 == 0 */
int bz_nn_conv3x3_dgrad(const uint16_t* gz, const uint16_t* wt16, uint16_t* gx, int B, int H, int W, int Ci,
                        int Co, void* stream);                      /* wt16 = [Ci][3][3][Co] */
/* bz_nn_conv3x3_dgrad followed by bz_nn_relu_mask(gx, x): gx = 0 where x (the forward input, an activation after
 * ReLU) is not positive; the same bits as the two calls, in one pass. */
int bz_nn_conv3x3_dgrad_relu(const uint16_t* gz, const uint16_t* wt16, const uint16_t* x, uint16_t* gx, int B, int H,
                             int W, int Ci, int Co, void* stream);
/* Weight and bias gradients. With a workspace (work != NULL, nwork >= bz_nn_wgrad_work(...) floats) every
 * partial sum is written out and added in a fixed order: the result is deterministic. With work == NULL
 * the partial sums go through FP32 atomics, whose order varies from run to run. */
long bz_nn_wgrad_work(int B, int H, int W, int Ci, int Co, int splits);
int bz_nn_conv3x3_wgrad(const uint16_t* x, const uint16_t* gz, float* gw, float* gb, int B, int H, int W,
                        int Ci, int Co, int splits, float* work, long nwork, void* stream);
int bz_nn_conv_first(const uint16_t* x, const float* w, const float* b, uint16_t* y, int B, int H, int W, int Co,
                     void* stream);                                 /* one input channel, ReLU */
long bz_nn_first_work(int B, int H, int W, int Co);                 /* workspace floats, as bz_nn_wgrad_work */
int bz_nn_conv_first_bwd(const uint16_t* x, const uint16_t* gz, const float* w, float* gw, float* gb, float* gx,
                         int B, int H, int W, int Co, float* work, long nwork,
                         void* stream);                             /* gx (may be NULL) is written, not added */
int bz_nn_relu_mask(uint16_t* gz, const uint16_t* y, long n, void* stream);
int bz_nn_avgpool2(const uint16_t* x, uint16_t* y, int B, int H, int W, int C, void* stream);
int bz_nn_avgpool2_bwd(const uint16_t* gy, uint16_t* gx, int B, int H, int W, int C, void* stream);
/* bz_nn_avgpool2_bwd followed by bz_nn_relu_mask(gx, y) (C % 8 == 0, 16-byte aligned tensors), same bits. */
int bz_nn_avgpool2_bwd_relu(const uint16_t* gy, const uint16_t* y, uint16_t* gx, int B, int H, int W, int C,
                            void* stream);
int bz_nn_head(const uint16_t* a, const float* fw, const float* fb, float* gap, float* out, int B, int HW, int C,
               int nout, void* stream);                             /* global average pool + fully connected */
int bz_nn_head_bwd(const float* gout, const float* gap, const float* fw, float* gfw, float* gfb, uint16_t* ga,
                   int B, int HW, int C, int nout, void* stream);
int bz_nn_weights(const float* w, uint16_t* w16, uint16_t* wt16, int Co, int Ci, void* stream);
/* bz_nn_weights of n <= 16 layers in one launch (host arrays of n pointers and sizes; wt16[j] may be NULL) */
int bz_nn_weights_n(const float* const* w, uint16_t* const* w16, uint16_t* const* wt16, const int* Co, const int* Ci,
                    int n, void* stream);
int bz_nn_adam(float* p, const float* g, float* m, float* v, long n, float lr, float b1, float b2, float eps,
               int step, void* stream);
int bz_nn_to_bf16(const float* x, uint16_t* y, long n, void* stream);

/* ---------------- sources ---------------- */
/* field[off[p]] += w[p] * amp for p < n (offsets unique within one call). */
int bz_add_points(float* field, const int64_t* off, const float* w, int n, float amp, void* stream);
/* field[off[p]] = mul[p] * field[off[p]] + w[p] * amp for p < n (offsets unique).
 * A ModeSource's sequential "add, then zero the upstream cell" (helper.py:37-54) reduces to one
 * such affine map per touched cell, precomputed on the host. */
int bz_affine_points(float* field, const int64_t* off, const float* mul, const float* w, int n, float amp,
                     void* stream);

/* ---------------- monitors ---------------- */
/* Two consecutive steps' samples in one launch: step n from comp_base0/off0 (if act0), step n+1 from
 * comp_base1/off1 (if act1), same comp table; series rows series0/series1 (either may be NULL), phasor rows
 * phasor0/phasor1; each accumulator gets step n's term and then step n+1's (the bits of two bz_gather calls). */
int bz_gather2(const float* const comp_base0[6], const int64_t* off0, const float* const comp_base1[6],
               const int* comp, const int64_t* off1, int n, int act0, int act1, float* series0, float* series1,
               int nfreq, const double* phasor0, const double* phasor1, double* dft, void* stream);

/* For p < n: v = comp_base[comp[p]][off[p]].
 *   series (if not NULL): series[p] = v            (caller offsets it to the step's row)
 *   dft (if nfreq > 0):   dft[k*n + p] += v * phasor[k]   for k < nfreq, FP64 complex
 * comp_base holds up to 6 device pointers. */
int bz_gather(const float* const comp_base[6], const int* comp, const int64_t* off, int n,
              float* series, int nfreq, const double* phasor, double* dft, void* stream);

/* Grid power: acc[j*ld+i] += Sx^2 + Sy^2 with Sx = -Ez*Hy, Sy = Ez*Hx (zero-padded H). */
int bz_grid_power_2d(double* acc, const float* Ez, const float* Hx, const float* Hy, int nx, int ny, int ld,
                     void* stream);
/* Grid power 3D: acc (on the [nz-1][ny-1][nx-1] corner box, pitch ld) += |E x H|. */
int bz_grid_power_3d(double* acc, float* const f[6], int nx, int ny, int nz, int ld, void* stream);
/* The same with complex fields held as a real and an imaginary set of arrays (the reference's complex128
 * fields, helper.py:124-143 and :152-179): 2D Sx = -(Ez_re Hy_re + Ez_im Hy_im), Sy = Ez_re Hx_re + Ez_im Hx_im,
 * acc += Sx^2 + Sy^2; 3D acc += |Re(E x conj H)|. */
int bz_grid_power_2d_cplx(double* acc, const float* Ez, const float* Hx, const float* Hy, const float* Ez_im,
                          const float* Hx_im, const float* Hy_im, int nx, int ny, int ld, void* stream);
int bz_grid_power_3d_cplx(double* acc, float* const f[6], float* const f_im[6], int nx, int ny, int nz, int ld,
                          void* stream);

/* ---------------- numerical watchdog ----------------
 * count (device, caller-zeroed) += number of NaN or infinite entries of p[0 .. n). No reference
 * counterpart: the reference has no numerical-failure detection. */
int bz_nonfinite_count(const float* p, long n, unsigned int* count, void* stream);

/* ---------------- rasterizer (host, FP64, bit-exact with the reference) ---------------- */
typedef struct bz_shape {
    int kind;               /* 0 axis-aligned rectangle (fast path), 1 circle, 2 ring, 3 polygon */
    double perm, permb, cond;
    double bbox[6];         /* min_x, min_y, min_z, max_x, max_y, max_z */
    double px, py, w, h;    /* kind 0 */
    double cx, cy, r, r_in, r_out; /* kinds 1, 2 */
    int nv;                 /* exterior vertex count; v = x0,y0,x1,y1,... */
    const double* v;
    int nholes;             /* holes: hole_nv[h] vertices each, packed in hv */
    const int* hole_nv;
    const double* hv;
    /* kind 0 in 2D with a spatially varying material (the reference's CustomMaterial, meshing.py:162-169,
     * 209-217): cell = {perm, permb, cond} sampled at the cell centres of rows cell_i0 .. cell_i0+cell_ni-1
     * and columns cell_j0 .. cell_j0+cell_nj-1, as [3][cell_ni][cell_nj]; the interior fill and the edge
     * blends take the cell's own values. NULL: the constant perm, permb, cond. Every cell the rectangle
     * writes must lie in the box, else BZ_EINVAL. */
    const double* cell;
    int cell_i0, cell_j0, cell_ni, cell_nj;
} bz_shape;

typedef struct bz_pml {
    int kind;               /* 0 rect, 1 corner */
    int orient;             /* rect: 0 left 1 right 2 top 3 bottom; corner: 0 tl 1 tr 2 bl 3 br */
    double px, py, w, h, r;
    double order, sigma_factor, alpha_max;
} bz_pml;

/* bg = {permittivity, permeability, conductivity} of structures[0] (or NULL for vacuum).
 * Outputs [gh][gw] (2D) or [gd][gh][gw] (3D), FP64. */
int bz_raster_2d(const bz_shape* shapes, int nshapes, const double* bg, const bz_pml* pmls, int npml,
                 double width, double height, double res, int gw, int gh,
                 double* perm, double* permb, double* cond);
int bz_raster_3d(const bz_shape* shapes, int nshapes, const double* bg, const bz_pml* pmls, int npml,
                 double width, double height, double depth, double res_xy, double res_z,
                 int gw, int gh, int gd, double* perm, double* permb, double* cond);

#ifdef __cplusplus
}
#endif
#endif
