// Fused single-pass 3D Yee step for gfx950: H and E updates of one time step
// plus soft-source injection in one sweep over HBM.
//
// Step semantics are those of beamz/simulation/fdtd.py:200-310 followed by
// helper.py:5-118 (H on all valid entries, E on the interior slices, sources
// added to the updated E). Fields are ping-ponged: every cell of the step is
// read from `src` and written to `dst`, so neighbouring workgroups can read
// halos from `src` without ordering.
//
// Work decomposition (wave64):
//   * a workgroup is TY waves of 64 lanes; lane l of wave w owns column
//     (i, j) = (i0 - 1 + l, j0 - 1 + w). Lane 0 and wave 0 are halo owners:
//     they compute H (needed by the E update of their +x / +y neighbours) but
//     write nothing. Each workgroup therefore outputs 63 x (TY-1) columns.
//   * the workgroup marches along z over a segment [k0, k1): at each plane the
//     eight per-cell values (Ex/Ey of the next plane, Ez, Hx, Hy, Hz, eps,
//     sigma) are read once from HBM; Ex/Ey of the current plane and Hx/Hy of
//     the previous plane stay in registers; in-plane neighbours are exchanged
//     through LDS. A segment starting at k0 > 0 first recomputes H of plane
//     k0-1 (its own columns only).
// HBM traffic per cell-update is therefore 8 reads + 6 writes of FP32 = 56 B,
// plus the halo lanes/rows (served mostly from L2) and one extra plane per
// segment.
#include "common.h"

namespace {

constexpr int LANES = 64;
constexpr int TY = 8;        // waves per workgroup (rows incl. one halo row)
constexpr int OX = LANES - 1;
constexpr int OY = TY - 1;

struct Fields6 {
    float* p[6];
};

struct BoxSources {
    int n;
    bz_box_source b[BZ_MAX_BOX];
};

__device__ __forceinline__ float ldg(const float* p, long o, bool ok) { return ok ? p[o] : 0.0f; }

__global__ __launch_bounds__(LANES* TY) void k_yee3d_fused(Fields6 S, Fields6 D, const float* __restrict__ eps,
                                                          const float* __restrict__ sig, int nx, int ny, int nz,
                                                          int ld, int kseg, Coef c, float rdx, float rdy,
                                                          float rdz, BoxSources src) {
    __shared__ float sEz[TY][LANES], sEx[TY][LANES], sEy[TY][LANES];
    __shared__ float sHz[TY][LANES], sHx[TY][LANES], sHy[TY][LANES];

    const int l = threadIdx.x, w = threadIdx.y;
    const int i = blockIdx.x * OX - 1 + l;
    const int j = blockIdx.y * OY - 1 + w;
    const int k0 = blockIdx.z * kseg;
    const int k1 = min(nz, k0 + kseg);
    const bool col = (i >= 0 && i < nx && j >= 0 && j < ny);
    const bool out = col && l > 0 && w > 0;
    const long ps = (long)ny * ld;
    const long cb = (long)j * ld + i;           // column base (valid only if col)

    const float* __restrict__ sEx_g = S.p[0];
    const float* __restrict__ sEy_g = S.p[1];
    const float* __restrict__ sEz_g = S.p[2];
    const float* __restrict__ sHx_g = S.p[3];
    const float* __restrict__ sHy_g = S.p[4];
    const float* __restrict__ sHz_g = S.p[5];

    const bool xin = i >= 1 && i <= nx - 2, yin = j >= 1 && j <= ny - 2;
    const bool hxok = j < ny - 1, hyok = i < nx - 1, hzok = j < ny - 1 && i < nx - 1;
    const bool rcol = col && (i + 1 < nx);      // +x neighbour exists
    const bool ucol = col && (j + 1 < ny);      // +y neighbour exists

    const int kstart = k0 > 0 ? k0 - 1 : k0;
    float ex0 = ldg(sEx_g, (long)kstart * ps + cb, col);
    float ey0 = ldg(sEy_g, (long)kstart * ps + cb, col);
    float hxp = 0.0f, hyp = 0.0f;                // H_new of plane k-1 at this column

    for (int k = kstart; k < k1; ++k) {
        const long o = (long)k * ps + cb;
        const bool nk = k + 1 < nz;
        const float ex1 = ldg(sEx_g, o + ps, col && nk);
        const float ey1 = ldg(sEy_g, o + ps, col && nk);
        const float ez = ldg(sEz_g, o, col);
        const float hx = ldg(sHx_g, o, col);
        const float hy = ldg(sHy_g, o, col);
        const float hz = ldg(sHz_g, o, col);
        const float sg = ldg(sig, o, col);

        sEz[w][l] = ez;
        sEx[w][l] = ex0;
        sEy[w][l] = ey0;
        float ezR = 0.0f, eyR = 0.0f, ezU = 0.0f, exU = 0.0f;
        if (l == LANES - 1) {
            ezR = ldg(sEz_g, o + 1, rcol);
            eyR = ldg(sEy_g, o + 1, rcol);
        }
        if (w == TY - 1) {
            ezU = ldg(sEz_g, o + ld, ucol);
            exU = ldg(sEx_g, o + ld, ucol);
        }
        __syncthreads();
        if (l < LANES - 1) {
            ezR = sEz[w][l + 1];
            eyR = sEy[w][l + 1];
        }
        if (w < TY - 1) {
            ezU = sEz[w + 1][l];
            exU = sEx[w + 1][l];
        }
        float fa, fb;
        h_coef(sg, c, fa, fb);
        const bool kh = k < nz - 1;
        const float hxn = (kh && hxok) ? fa * hx - fb * ((ezU - ez) * rdy - (ey1 - ey0) * rdz) : hx;
        const float hyn = (kh && hyok) ? fa * hy - fb * ((ex1 - ex0) * rdz - (ezR - ez) * rdx) : hy;
        const float hzn = hzok ? fa * hz - fb * ((eyR - ey0) * rdx - (exU - ex0) * rdy) : hz;
        sHz[w][l] = hzn;
        sHx[w][l] = hxn;
        sHy[w][l] = hyn;
        __syncthreads();
        if (k >= k0 && out) {
            const float hzL = sHz[w][l - 1], hyL = sHy[w][l - 1];
            const float hzD = sHz[w - 1][l], hxD = sHx[w - 1][l];
            const float e = eps[o];
            float ca, cbv;
            e_coef(sg, e, c, ca, cbv);
            const bool kin = k >= 1 && k <= nz - 2;
            float exn = (kin && yin && i <= nx - 2) ? ca * ex0 + cbv * ((hzn - hzD) * rdy - (hyn - hyp) * rdz) : ex0;
            float eyn = (kin && j <= ny - 2 && xin) ? ca * ey0 + cbv * ((hxn - hxp) * rdz - (hzn - hzL) * rdx) : ey0;
            float ezn = (k <= nz - 2 && yin && xin) ? ca * ez + cbv * ((hyn - hyL) * rdx - (hxn - hxD) * rdy) : ez;
            for (int q = 0; q < src.n; ++q) {
                const bz_box_source& b = src.b[q];
                const int a = i - b.x0, bb = j - b.y0, cc = k - b.z0;
                if (a >= 0 && a < b.sx && bb >= 0 && bb < b.sy && cc >= 0 && cc < b.sz) {
                    const long r = ((long)cc * b.sy + bb) * b.sx + a;
                    const float m = b.mul ? b.mul[r] : 1.0f;
                    if (b.comp == 0) exn = m * exn + b.add[r] * b.amp;
                    else if (b.comp == 1) eyn = m * eyn + b.add[r] * b.amp;
                    else ezn = m * ezn + b.add[r] * b.amp;
                }
            }
            D.p[0][o] = exn;
            D.p[1][o] = eyn;
            D.p[2][o] = ezn;
            D.p[3][o] = hxn;
            D.p[4][o] = hyn;
            D.p[5][o] = hzn;
        }
        hxp = hxn;
        hyp = hyn;
        ex0 = ex1;
        ey0 = ey1;
    }
}

}  // namespace

extern "C" int bz_yee3d_step(float* const src[6], float* const dst[6], const float* eps_r, const float* sigma, int nx,
                             int ny, int nz, int ld, double dx, double dy, double dz, double dt, double mu0,
                             double eps0, const bz_box_source* boxes, int nbox, int kseg, void* stream) {
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
    const dim3 g((nx + OX - 1) / OX, (ny + OY - 1) / OY, (nz + kseg - 1) / kseg), b(LANES, TY);
    hipLaunchKernelGGL(k_yee3d_fused, g, b, 0, (hipStream_t)stream, S, D, eps_r, sigma, nx, ny, nz, ld, kseg,
                       make_coef(dt, mu0, eps0), (float)(1.0 / dx), (float)(1.0 / dy), (float)(1.0 / dz), bs);
    BZ_CHECK_LAUNCH("bz_yee3d_step");
    return BZ_OK;
}
