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
//     segment [k0, k1). Lane 0 and row 0 are halo owners: they compute H (the
//     E update of lane 1 / row 1 needs H at i-1 / j-1) and write nothing, so a
//     wave outputs 63 x R columns.
//   * per plane every column reads its eight values (Ex/Ey of the next plane,
//     Ez, Hx, Hy, Hz, sigma, eps) once; Ex/Ey of the current plane and Hx/Hy
//     of the previous plane stay in registers; the +x / -x neighbours come from
//     the adjacent lane by wave shuffles, the +y / -y neighbours from the
//     adjacent register row. One extra row (j0+R: Ez and the Ex queue) and,
//     for lane 63, the +x column are read for the halo.
//   * a segment starting at k0 > 0 first recomputes H of plane k0-1 for its
//     own columns (bit-identical to what the previous segment computes).
//   * workgroups are 4 waves stacked in y; hardware workgroup ids are remapped
//     so each XCD (8 per MI355X, round-robin dispatch) gets a contiguous block
//     of tiles and the halos it re-reads stay in its own L2.
// HBM traffic per cell-update is 8 reads + 6 writes of FP32 = 56 B, plus the
// halo rows / lanes (mostly L2 hits) and one extra plane per segment.
#include "common.h"

namespace {

constexpr int LANES = 64;
constexpr int OX = LANES - 1;
constexpr int WPG = 4;       // waves per workgroup
constexpr int NXCD = 8;

struct Fields6 {
    float* p[6];
};

struct BoxSources {
    int n;
    bz_box_source b[BZ_MAX_BOX];
};

__device__ __forceinline__ float ldg(const float* __restrict__ p, long o, bool ok) { return ok ? p[o] : 0.0f; }

__device__ __forceinline__ float shfl_dn(float v) { return __shfl_down(v, 1, LANES); }
__device__ __forceinline__ float shfl_up(float v) { return __shfl_up(v, 1, LANES); }

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

    const int l = threadIdx.x;
    const int i = tx * OX - 1 + l;
    const int jb = (ty * WPG + threadIdx.y) * R - 1;   // row 0 of this wave (halo row)
    const int k0 = tz * kseg;
    const int k1 = min(nz, k0 + kseg);
    const bool ci = i >= 0 && i < nx;
    const long ps = (long)ny * ld;

    const float* __restrict__ gEx = S.p[0];
    const float* __restrict__ gEy = S.p[1];
    const float* __restrict__ gEz = S.p[2];
    const float* __restrict__ gHx = S.p[3];
    const float* __restrict__ gHy = S.p[4];
    const float* __restrict__ gHz = S.p[5];

    bool col[R + 2];
#pragma unroll
    for (int r = 0; r < R + 2; ++r) col[r] = ci && (jb + r) >= 0 && (jb + r) < ny;
    const bool rx = i + 1 < nx;                  // +x neighbour exists (used by lane 63)
    const bool lane_out = l > 0 && ci;
    const bool xin = i >= 1 && i <= nx - 2;
    const bool hyok = i < nx - 1;

    const int kstart = k0 > 0 ? k0 - 1 : k0;
    float ex0[R + 2], ey0[R + 1], hxp[R + 1], hyp[R + 1];
#pragma unroll
    for (int r = 0; r < R + 2; ++r) {
        const long o = (long)kstart * ps + (long)(jb + r) * ld + i;
        ex0[r] = ldg(gEx, o, col[r]);
        if (r < R + 1) {
            ey0[r] = ldg(gEy, o, col[r]);
            hxp[r] = 0.0f;
            hyp[r] = 0.0f;
        }
    }

    for (int k = kstart; k < k1; ++k) {
        const long ok = (long)k * ps;
        const bool nk = k + 1 < nz;
        float ex1[R + 2], ey1[R + 1], ez[R + 2], hx[R + 1], hy[R + 1], hz[R + 1], sg[R + 1], ee[R + 1];
#pragma unroll
        for (int r = 0; r < R + 2; ++r) {
            const long o = ok + (long)(jb + r) * ld + i;
            ex1[r] = ldg(gEx, o + ps, col[r] && nk);
            ez[r] = ldg(gEz, o, col[r]);
            if (r < R + 1) {
                ey1[r] = ldg(gEy, o + ps, col[r] && nk);
                hx[r] = ldg(gHx, o, col[r]);
                hz[r] = ldg(gHz, o, col[r]);
                sg[r] = ldg(sig, o, col[r]);
                hy[r] = (r > 0) ? ldg(gHy, o, col[r]) : 0.0f;
                ee[r] = (r > 0) ? ldg(eps, o, col[r] && k >= k0) : 1.0f;
            }
        }
        // +x neighbours of Ez and Ey (lane 63 reads the next tile's first column itself)
        float ezR[R + 1], eyR[R + 1];
#pragma unroll
        for (int r = 0; r < R + 1; ++r) {
            ezR[r] = shfl_dn(ez[r]);
            eyR[r] = shfl_dn(ey0[r]);
        }
        if (l == LANES - 1) {
#pragma unroll
            for (int r = 0; r < R + 1; ++r) {
                const long o = ok + (long)(jb + r) * ld + i + 1;
                ezR[r] = ldg(gEz, o, col[r] && rx);
                eyR[r] = ldg(gEy, o, col[r] && rx);
            }
        }
        float hxn[R + 1], hyn[R + 1], hzn[R + 1];
        const bool kh = k < nz - 1;
#pragma unroll
        for (int r = 0; r < R + 1; ++r) {
            const int j = jb + r;
            float fa, fb;
            h_coef(sg[r], c, fa, fb);
            const bool jh = j < ny - 1;
            hxn[r] = (kh && jh) ? fa * hx[r] - fb * ((ez[r + 1] - ez[r]) * rdy - (ey1[r] - ey0[r]) * rdz) : hx[r];
            hyn[r] = (kh && hyok) ? fa * hy[r] - fb * ((ex1[r] - ex0[r]) * rdz - (ezR[r] - ez[r]) * rdx) : hy[r];
            hzn[r] = (jh && hyok) ? fa * hz[r] - fb * ((eyR[r] - ey0[r]) * rdx - (ex0[r + 1] - ex0[r]) * rdy) : hz[r];
        }
        if (k >= k0) {
            const bool kin = k >= 1 && k <= nz - 2;
#pragma unroll
            for (int r = 1; r < R + 1; ++r) {
                const float hzL = shfl_up(hzn[r]);
                const float hyL = shfl_up(hyn[r]);
                const int j = jb + r;
                if (lane_out && col[r]) {
                    const long o = ok + (long)j * ld + i;
                    const bool yin = j >= 1 && j <= ny - 2;
                    float ca, cb;
                    e_coef(sg[r], ee[r], c, ca, cb);
                    float exn = (kin && yin && i <= nx - 2)
                                    ? ca * ex0[r] + cb * ((hzn[r] - hzn[r - 1]) * rdy - (hyn[r] - hyp[r]) * rdz)
                                    : ex0[r];
                    float eyn = (kin && j <= ny - 2 && xin)
                                    ? ca * ey0[r] + cb * ((hxn[r] - hxp[r]) * rdz - (hzn[r] - hzL) * rdx)
                                    : ey0[r];
                    float ezn = (k <= nz - 2 && yin && xin)
                                    ? ca * ez[r] + cb * ((hyn[r] - hyL) * rdx - (hxn[r] - hxn[r - 1]) * rdy)
                                    : ez[r];
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
                    D.p[0][o] = exn;
                    D.p[1][o] = eyn;
                    D.p[2][o] = ezn;
                    D.p[3][o] = hxn[r];
                    D.p[4][o] = hyn[r];
                    D.p[5][o] = hzn[r];
                }
            }
        }
#pragma unroll
        for (int r = 0; r < R + 2; ++r) {
            ex0[r] = ex1[r];
            if (r < R + 1) {
                ey0[r] = ey1[r];
                hxp[r] = hxn[r];
                hyp[r] = hyn[r];
            }
        }
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
    hipLaunchKernelGGL(k_yee3d_fused<R>, dim3(per_xcd * NXCD), dim3(LANES, WPG), 0, st, S, D, eps, sig, nx, ny, nz,
                       ld, kseg, gx, gy, gz, per_xcd, c, rdx, rdy, rdz, bs);
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
        case 8: return launch<8>(S, D, eps_r, sigma, nx, ny, nz, ld, kseg, c, rx, ry, rz, bs, st);
        default: bz::set_error("bz_yee3d_step: rows must be 2, 4 or 8"); return BZ_EINVAL;
    }
}
