// One-cell-per-thread Yee kernels: the plugin-level 2D updates, the two-pass
// 3D update, point sources, monitor gathers with FP64 DFT, grid power.
//
// These are the direct MI355X counterparts of the reference's array-op groups
// (numpy_backend.py:34-88, fdtd.py:200-310, helper.py:5-256). The fused
// engine kernels (yee3d_fused.hip, yee2d_tiled.hip) are checked against them.
#include "common.h"

namespace {

constexpr int TX = 64, TY = 4;  // one wave per x-row segment, 4 rows per workgroup

// numpy_backend.py:34-61. Hx[j][i] (i < nx-1), Hy[j][i] (j < ny-1). Each array has its own row pitch.
struct Pitch2 {
    int hx, hy, ez, sig, eps;
};

__global__ __launch_bounds__(TX* TY) void k_h2d(float* __restrict__ hx, float* __restrict__ hy,
                                                 const float* __restrict__ ez, const float* __restrict__ sig,
                                                 int nx, int ny, Pitch2 ld, Coef c, float rdx, float rdy) {
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    if (i >= nx || j >= ny) return;
    const long oe = (long)j * ld.ez + i;
    float fa, fb;
    h_coef(sig[(long)j * ld.sig + i], c, fa, fb);
    const float e0 = ez[oe];
    if (i < nx - 1) {
        const long o = (long)j * ld.hx + i;
        hx[o] = fa * hx[o] - fb * ((ez[oe + 1] - e0) * rdy);   // x-difference over dy (reference)
    }
    if (j < ny - 1) {
        const long o = (long)j * ld.hy + i;
        hy[o] = fa * hy[o] + fb * ((ez[oe + ld.ez] - e0) * rdx);
    }
}

// numpy_backend.py:63-88. Ez interior 1..ny-2 x 1..nx-2.
__global__ __launch_bounds__(TX* TY) void k_e2d(float* __restrict__ ez, const float* __restrict__ hx,
                                                 const float* __restrict__ hy, const float* __restrict__ sig,
                                                 const float* __restrict__ eps, int nx, int ny, Pitch2 ld, Coef c,
                                                 float rdx, float rdy) {
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    if (i < 1 || i >= nx - 1 || j < 1 || j >= ny - 1) return;
    float ca, cb;
    e_coef(sig[(long)j * ld.sig + i], eps[(long)j * ld.eps + i], c, ca, cb);
    const long ox = (long)j * ld.hx + i, oy = (long)j * ld.hy + i, oe = (long)j * ld.ez + i;
    const float curl = -((hx[ox] - hx[ox - 1]) * rdy) + (hy[oy] - hy[oy - ld.hy]) * rdx;
    ez[oe] = ca * ez[oe] + cb * curl;
}

// fdtd.py:210-247. All H entries; components use sigma of the same index.
__global__ __launch_bounds__(TX* TY) void k_h3d(float* __restrict__ Ex, float* __restrict__ Ey,
                                                 float* __restrict__ Ez, float* __restrict__ Hx,
                                                 float* __restrict__ Hy, float* __restrict__ Hz,
                                                 const float* __restrict__ sig, int nx, int ny, int nz, int ld,
                                                 Coef c, float rdx, float rdy, float rdz) {
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= nx || j >= ny) return;
    const long ps = (long)ny * ld;
    const long o = (long)k * ps + (long)j * ld + i;
    float fa, fb;
    h_coef(sig[o], c, fa, fb);
    const bool kx = k < nz - 1, jy = j < ny - 1, ix = i < nx - 1;
    if (kx && jy) {
        const float cx = (Ez[o + ld] - Ez[o]) * rdy - (Ey[o + ps] - Ey[o]) * rdz;
        Hx[o] = fa * Hx[o] - fb * cx;
    }
    if (kx && ix) {
        const float cy = (Ex[o + ps] - Ex[o]) * rdz - (Ez[o + 1] - Ez[o]) * rdx;
        Hy[o] = fa * Hy[o] - fb * cy;
    }
    if (jy && ix) {
        const float cz = (Ey[o + 1] - Ey[o]) * rdx - (Ex[o + ld] - Ex[o]) * rdy;
        Hz[o] = fa * Hz[o] - fb * cz;
    }
}

// fdtd.py:249-310. Interior E slices; all three use eps/sigma of the same index.
__global__ __launch_bounds__(TX* TY) void k_e3d(float* __restrict__ Ex, float* __restrict__ Ey,
                                                 float* __restrict__ Ez, const float* __restrict__ Hx,
                                                 const float* __restrict__ Hy, const float* __restrict__ Hz,
                                                 const float* __restrict__ sig, const float* __restrict__ eps,
                                                 int nx, int ny, int nz, int ld, Coef c, float rdx, float rdy,
                                                 float rdz) {
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= nx || j >= ny) return;
    const long ps = (long)ny * ld;
    const long o = (long)k * ps + (long)j * ld + i;
    const bool kin = k >= 1 && k <= nz - 2, jin = j >= 1 && j <= ny - 2, iin = i >= 1 && i <= nx - 2;
    const bool ux = kin && jin && i <= nx - 2;
    const bool uy = kin && j <= ny - 2 && iin;
    const bool uz = k <= nz - 2 && jin && iin;
    if (!(ux || uy || uz)) return;
    float ca, cb;
    e_coef(sig[o], eps[o], c, ca, cb);
    if (ux) Ex[o] = ca * Ex[o] + cb * ((Hz[o] - Hz[o - ld]) * rdy - (Hy[o] - Hy[o - ps]) * rdz);
    if (uy) Ey[o] = ca * Ey[o] + cb * ((Hx[o] - Hx[o - ps]) * rdz - (Hz[o] - Hz[o - 1]) * rdx);
    if (uz) Ez[o] = ca * Ez[o] + cb * ((Hy[o] - Hy[o - 1]) * rdx - (Hx[o] - Hx[o - ld]) * rdy);
}

__global__ void k_add_points(float* __restrict__ f, const int64_t* __restrict__ off, const float* __restrict__ w,
                             int n, float amp) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) f[off[p]] += w[p] * amp;
}

__global__ void k_affine_points(float* __restrict__ f, const int64_t* __restrict__ off, const float* __restrict__ mul,
                                const float* __restrict__ w, int n, float amp) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) f[off[p]] = mul[p] * f[off[p]] + w[p] * amp;
}

struct CompBase {
    const float* p[6];
};

__global__ void k_gather(CompBase base, const int* __restrict__ comp, const int64_t* __restrict__ off, int n,
                         float* __restrict__ series, int nfreq, const double* __restrict__ phasor,
                         double* __restrict__ dft) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float v = base.p[comp[p]][off[p]];
    if (series) series[p] = v;
    const double dv = (double)v;
    for (int k = 0; k < nfreq; ++k) {
        const double re = phasor[2 * k], im = phasor[2 * k + 1];
        double* a = dft + 2 * ((long)k * n + p);
        a[0] += dv * re;
        a[1] += dv * im;
    }
}

// The samples of two consecutive steps (n from base0/off0, n+1 from base1/off1) in one launch, thread per
// (point, frequency), p fastest; each accumulator receives step n's term and then step n+1's, the two
// rounded adds of two k_gather launches in that order.
__global__ void k_gather2(CompBase base0, const int64_t* __restrict__ off0, CompBase base1,
                          const int* __restrict__ comp, const int64_t* __restrict__ off1, int n, int act0, int act1,
                          float* __restrict__ series0, float* __restrict__ series1, int nfreq,
                          const double* __restrict__ phasor0, const double* __restrict__ phasor1,
                          double* __restrict__ dft) {
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int p = (int)(g % n), k = (int)(g / n);
    if (k >= max(nfreq, 1)) return;
    const int c = comp[p];
    const float v0 = act0 ? base0.p[c][off0[p]] : 0.0f;
    const float v1 = act1 ? base1.p[c][off1[p]] : 0.0f;
    if (k == 0) {
        if (act0 && series0) series0[p] = v0;
        if (act1 && series1) series1[p] = v1;
    }
    if (k >= nfreq) return;
    double* a = dft + 2 * ((long)k * n + p);
    double re = a[0], im = a[1];
    if (act0) {
        const double dv = (double)v0;
        re += dv * phasor0[2 * k];
        im += dv * phasor0[2 * k + 1];
    }
    if (act1) {
        const double dv = (double)v1;
        re += dv * phasor1[2 * k];
        im += dv * phasor1[2 * k + 1];
    }
    a[0] = re;
    a[1] = im;
}

// The samples of three consecutive steps (n and n+1 from base0/base1 at off0, n+2 from base2 at off2) in one
// launch, thread per (point, frequency); each accumulator receives the three terms in step order, the rounded
// adds of three k_gather launches.
__global__ void k_gather3(CompBase base0, CompBase base1, const int64_t* __restrict__ off0, CompBase base2,
                          const int* __restrict__ comp, const int64_t* __restrict__ off2, int n, int act0, int act1,
                          int act2, float* __restrict__ series0, float* __restrict__ series1,
                          float* __restrict__ series2, int nfreq, const double* __restrict__ phasor0,
                          const double* __restrict__ phasor1, const double* __restrict__ phasor2,
                          double* __restrict__ dft) {
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int p = (int)(g % n), k = (int)(g / n);
    if (k >= max(nfreq, 1)) return;
    const int c = comp[p];
    const float v0 = act0 ? base0.p[c][off0[p]] : 0.0f;
    const float v1 = act1 ? base1.p[c][off0[p]] : 0.0f;
    const float v2 = act2 ? base2.p[c][off2[p]] : 0.0f;
    if (k == 0) {
        if (act0 && series0) series0[p] = v0;
        if (act1 && series1) series1[p] = v1;
        if (act2 && series2) series2[p] = v2;
    }
    if (k >= nfreq) return;
    double* a = dft + 2 * ((long)k * n + p);
    double re = a[0], im = a[1];
    if (act0) {
        const double dv = (double)v0;
        re += dv * phasor0[2 * k];
        im += dv * phasor0[2 * k + 1];
    }
    if (act1) {
        const double dv = (double)v1;
        re += dv * phasor1[2 * k];
        im += dv * phasor1[2 * k + 1];
    }
    if (act2) {
        const double dv = (double)v2;
        re += dv * phasor2[2 * k];
        im += dv * phasor2[2 * k + 1];
    }
    a[0] = re;
    a[1] = im;
}

// helper.py:152-187 (real fields)
__global__ __launch_bounds__(TX* TY) void k_pow2d(double* __restrict__ acc, const float* __restrict__ ez,
                                                   const float* __restrict__ hx, const float* __restrict__ hy,
                                                   int nx, int ny, int ld) {
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    if (i >= nx || j >= ny) return;
    const long o = (long)j * ld + i;
    const double e = ez[o];
    const double hxv = (i < nx - 1) ? (double)hx[o] : 0.0;
    const double hyv = (j < ny - 1) ? (double)hy[o] : 0.0;
    const double sx = -e * hyv, sy = e * hxv;
    acc[o] += sx * sx + sy * sy;
}

// helper.py:152-179 with complex fields held as a real and an imaginary set:
// Sx = -Ez_re Hy_re - Ez_im Hy_im, Sy = Ez_re Hx_re + Ez_im Hx_im, acc += Sx^2 + Sy^2
__global__ __launch_bounds__(TX* TY) void k_pow2d_cplx(double* __restrict__ acc, const float* __restrict__ ez,
                                                        const float* __restrict__ hx, const float* __restrict__ hy,
                                                        const float* __restrict__ ezi, const float* __restrict__ hxi,
                                                        const float* __restrict__ hyi, int nx, int ny, int ld) {
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    if (i >= nx || j >= ny) return;
    const long o = (long)j * ld + i;
    const double er = ez[o], ei = ezi[o];
    const bool xin = i < nx - 1, yin = j < ny - 1;
    const double hxr = xin ? (double)hx[o] : 0.0, hxm = xin ? (double)hxi[o] : 0.0;
    const double hyr = yin ? (double)hy[o] : 0.0, hym = yin ? (double)hyi[o] : 0.0;
    const double sx = -er * hyr - ei * hym, sy = er * hxr + ei * hxm;
    acc[o] += sx * sx + sy * sy;
}

// helper.py:124-151 (real fields) on the corner-aligned (nz-1, ny-1, nx-1) box
__global__ __launch_bounds__(TX* TY) void k_pow3d(double* __restrict__ acc, const float* __restrict__ Ex,
                                                   const float* __restrict__ Ey, const float* __restrict__ Ez,
                                                   const float* __restrict__ Hx, const float* __restrict__ Hy,
                                                   const float* __restrict__ Hz, int nx, int ny, int nz, int ld) {
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= nx - 1 || j >= ny - 1 || k >= nz - 1) return;
    const long o = (long)k * ny * ld + (long)j * ld + i;
    const double ex = Ex[o], ey = Ey[o], ez = Ez[o], hx = Hx[o], hy = Hy[o], hz = Hz[o];
    const double sx = ey * hz - ez * hy, sy = ez * hx - ex * hz, sz = ex * hy - ey * hx;
    acc[o] += sqrt(sx * sx + sy * sy + sz * sz);
}

// helper.py:124-143 with complex fields held as a real (f) and an imaginary (g) set:
// S = Re(E x conj(H)), e.g. Sx = (Ey_re Hz_re + Ey_im Hz_im) - (Ez_re Hy_re + Ez_im Hy_im); acc += |S|
struct Six {
    const float* p[6];
};
__global__ __launch_bounds__(TX* TY) void k_pow3d_cplx(double* __restrict__ acc, Six f, Six g, int nx, int ny,
                                                        int nz, int ld) {
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= nx - 1 || j >= ny - 1 || k >= nz - 1) return;
    const long o = (long)k * ny * ld + (long)j * ld + i;
    double r[6], m[6];
    for (int c = 0; c < 6; ++c) {
        r[c] = f.p[c][o];
        m[c] = g.p[c][o];
    }
    // re(a conj b) = a_re b_re + a_im b_im; components 0..5 = Ex, Ey, Ez, Hx, Hy, Hz
    auto rc = [&](int a, int b) { return r[a] * r[b] + m[a] * m[b]; };
    const double sx = rc(1, 5) - rc(2, 4), sy = rc(2, 3) - rc(0, 5), sz = rc(0, 4) - rc(1, 3);
    acc[o] += sqrt(sx * sx + sy * sy + sz * sz);
}

}  // namespace

extern "C" int bz_update_h_2d(float* Hx, int ld_hx, float* Hy, int ld_hy, const float* Ez, int ld_ez,
                              const float* sigma, int ld_sig, int nx, int ny, double dx, double dy, double dt,
                              double mu0, double eps0, void* stream) {
    BZ_REQUIRE(Hx && Hy && Ez && sigma, "bz_update_h_2d: null pointer");
    BZ_REQUIRE(nx >= 2 && ny >= 2 && ld_hx >= nx - 1 && ld_hy >= nx && ld_ez >= nx && ld_sig >= nx,
               "bz_update_h_2d: bad shape");
    const dim3 g((nx + TX - 1) / TX, (ny + TY - 1) / TY), b(TX, TY);
    const Pitch2 ld{ld_hx, ld_hy, ld_ez, ld_sig, 0};
    hipLaunchKernelGGL(k_h2d, g, b, 0, (hipStream_t)stream, Hx, Hy, Ez, sigma, nx, ny, ld, make_coef(dt, mu0, eps0),
                       (float)(1.0 / dx), (float)(1.0 / dy));
    BZ_CHECK_LAUNCH("bz_update_h_2d");
    return BZ_OK;
}

extern "C" int bz_update_e_2d(float* Ez, int ld_ez, const float* Hx, int ld_hx, const float* Hy, int ld_hy,
                              const float* sigma, int ld_sig, const float* eps_r, int ld_eps, int nx, int ny,
                              double dx, double dy, double dt, double eps0, void* stream) {
    BZ_REQUIRE(Ez && Hx && Hy && sigma && eps_r, "bz_update_e_2d: null pointer");
    BZ_REQUIRE(nx >= 2 && ny >= 2 && ld_hx >= nx - 1 && ld_hy >= nx && ld_ez >= nx && ld_sig >= nx && ld_eps >= nx,
               "bz_update_e_2d: bad shape");
    const dim3 g((nx + TX - 1) / TX, (ny + TY - 1) / TY), b(TX, TY);
    const Pitch2 ld{ld_hx, ld_hy, ld_ez, ld_sig, ld_eps};
    hipLaunchKernelGGL(k_e2d, g, b, 0, (hipStream_t)stream, Ez, Hx, Hy, sigma, eps_r, nx, ny, ld,
                       make_coef(dt, 1.0, eps0), (float)(1.0 / dx), (float)(1.0 / dy));
    BZ_CHECK_LAUNCH("bz_update_e_2d");
    return BZ_OK;
}

extern "C" int bz_update_3d(float* const f[6], const float* eps_r, const float* sigma, int nx, int ny, int nz, int ld,
                            double dx, double dy, double dz, double dt, double mu0, double eps0, void* stream) {
    BZ_REQUIRE(f && eps_r && sigma, "bz_update_3d: null pointer");
    for (int q = 0; q < 6; ++q) BZ_REQUIRE(f[q], "bz_update_3d: null field pointer");
    BZ_REQUIRE(nx >= 2 && ny >= 2 && nz >= 2 && ld >= nx, "bz_update_3d: bad shape");
    const Coef c = make_coef(dt, mu0, eps0);
    const dim3 g((nx + TX - 1) / TX, (ny + TY - 1) / TY, nz), b(TX, TY);
    hipLaunchKernelGGL(k_h3d, g, b, 0, (hipStream_t)stream, f[0], f[1], f[2], f[3], f[4], f[5], sigma, nx, ny, nz, ld,
                       c, (float)(1.0 / dx), (float)(1.0 / dy), (float)(1.0 / dz));
    BZ_CHECK_LAUNCH("bz_update_3d(H)");
    hipLaunchKernelGGL(k_e3d, g, b, 0, (hipStream_t)stream, f[0], f[1], f[2], f[3], f[4], f[5], sigma, eps_r, nx, ny,
                       nz, ld, c, (float)(1.0 / dx), (float)(1.0 / dy), (float)(1.0 / dz));
    BZ_CHECK_LAUNCH("bz_update_3d(E)");
    return BZ_OK;
}

extern "C" int bz_add_points(float* field, const int64_t* off, const float* w, int n, float amp, void* stream) {
    if (n == 0) return BZ_OK;
    BZ_REQUIRE(field && off && w && n > 0, "bz_add_points: bad arguments");
    hipLaunchKernelGGL(k_add_points, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, field, off, w, n, amp);
    BZ_CHECK_LAUNCH("bz_add_points");
    return BZ_OK;
}

extern "C" int bz_affine_points(float* field, const int64_t* off, const float* mul, const float* w, int n, float amp,
                                void* stream) {
    if (n == 0) return BZ_OK;
    BZ_REQUIRE(field && off && mul && w && n > 0, "bz_affine_points: bad arguments");
    hipLaunchKernelGGL(k_affine_points, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, field, off, mul, w,
                       n, amp);
    BZ_CHECK_LAUNCH("bz_affine_points");
    return BZ_OK;
}

extern "C" int bz_gather(const float* const comp_base[6], const int* comp, const int64_t* off, int n, float* series,
                         int nfreq, const double* phasor, double* dft, void* stream) {
    if (n == 0) return BZ_OK;
    BZ_REQUIRE(comp_base && comp && off && n > 0, "bz_gather: bad arguments");
    BZ_REQUIRE(nfreq == 0 || (phasor && dft), "bz_gather: DFT requested without buffers");
    CompBase cb;
    for (int q = 0; q < 6; ++q) cb.p[q] = comp_base[q];
    hipLaunchKernelGGL(k_gather, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, cb, comp, off, n, series,
                       nfreq, phasor, dft);
    BZ_CHECK_LAUNCH("bz_gather");
    return BZ_OK;
}

extern "C" int bz_grid_power_2d(double* acc, const float* Ez, const float* Hx, const float* Hy, int nx, int ny, int ld,
                                void* stream) {
    BZ_REQUIRE(acc && Ez && Hx && Hy && nx >= 2 && ny >= 2 && ld >= nx, "bz_grid_power_2d: bad arguments");
    const dim3 g((nx + TX - 1) / TX, (ny + TY - 1) / TY), b(TX, TY);
    hipLaunchKernelGGL(k_pow2d, g, b, 0, (hipStream_t)stream, acc, Ez, Hx, Hy, nx, ny, ld);
    BZ_CHECK_LAUNCH("bz_grid_power_2d");
    return BZ_OK;
}

extern "C" int bz_grid_power_2d_cplx(double* acc, const float* Ez, const float* Hx, const float* Hy,
                                     const float* Ez_im, const float* Hx_im, const float* Hy_im, int nx, int ny,
                                     int ld, void* stream) {
    BZ_REQUIRE(acc && Ez && Hx && Hy && Ez_im && Hx_im && Hy_im && nx >= 2 && ny >= 2 && ld >= nx,
               "bz_grid_power_2d_cplx: bad arguments");
    const dim3 g((nx + TX - 1) / TX, (ny + TY - 1) / TY), b(TX, TY);
    hipLaunchKernelGGL(k_pow2d_cplx, g, b, 0, (hipStream_t)stream, acc, Ez, Hx, Hy, Ez_im, Hx_im, Hy_im, nx, ny, ld);
    BZ_CHECK_LAUNCH("bz_grid_power_2d_cplx");
    return BZ_OK;
}

extern "C" int bz_grid_power_3d_cplx(double* acc, float* const f[6], float* const f_im[6], int nx, int ny, int nz,
                                     int ld, void* stream) {
    BZ_REQUIRE(acc && f && f_im && nx >= 2 && ny >= 2 && nz >= 2 && ld >= nx, "bz_grid_power_3d_cplx: bad arguments");
    Six a, c;
    for (int q = 0; q < 6; ++q) {
        BZ_REQUIRE(f[q] && f_im[q], "bz_grid_power_3d_cplx: null field");
        a.p[q] = f[q];
        c.p[q] = f_im[q];
    }
    const dim3 g((nx + TX - 1) / TX, (ny + TY - 1) / TY, nz), b(TX, TY);
    hipLaunchKernelGGL(k_pow3d_cplx, g, b, 0, (hipStream_t)stream, acc, a, c, nx, ny, nz, ld);
    BZ_CHECK_LAUNCH("bz_grid_power_3d_cplx");
    return BZ_OK;
}

extern "C" int bz_grid_power_3d(double* acc, float* const f[6], int nx, int ny, int nz, int ld, void* stream) {
    BZ_REQUIRE(acc && f && nx >= 2 && ny >= 2 && nz >= 2 && ld >= nx, "bz_grid_power_3d: bad arguments");
    const dim3 g((nx + TX - 1) / TX, (ny + TY - 1) / TY, nz), b(TX, TY);
    hipLaunchKernelGGL(k_pow3d, g, b, 0, (hipStream_t)stream, acc, f[0], f[1], f[2], f[3], f[4], f[5], nx, ny, nz, ld);
    BZ_CHECK_LAUNCH("bz_grid_power_3d");
    return BZ_OK;
}

namespace {
// Non-finite (NaN or infinite) entries of p[0 .. n): per-thread counts over a grid-stride loop, summed in
// LDS, one global atomic per workgroup.
__global__ void k_nonfinite(const float* __restrict__ p, long n, unsigned int* __restrict__ count) {
    __shared__ unsigned int c;
    if (threadIdx.x == 0) c = 0;
    __syncthreads();
    unsigned int mine = 0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        mine += (__float_as_uint(p[i]) & 0x7f800000u) == 0x7f800000u ? 1u : 0u;
    if (mine) atomicAdd(&c, mine);
    __syncthreads();
    if (threadIdx.x == 0 && c) atomicAdd(count, c);
}
}  // namespace

extern "C" int bz_nonfinite_count(const float* p, long n, unsigned int* count, void* stream) {
    BZ_REQUIRE(p && count && n >= 0, "bz_nonfinite_count: bad arguments");
    if (n == 0) return BZ_OK;
    const long blocks = (n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096;
    hipLaunchKernelGGL(k_nonfinite, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, p, n, count);
    BZ_CHECK_LAUNCH("bz_nonfinite_count");
    return BZ_OK;
}

extern "C" int bz_gather2(const float* const base0[6], const int64_t* off0, const float* const base1[6],
                          const int* comp, const int64_t* off1, int n, int act0, int act1, float* series0,
                          float* series1, int nfreq, const double* phasor0, const double* phasor1, double* dft,
                          void* stream) {
    if (n == 0 || (!act0 && !act1)) return BZ_OK;
    BZ_REQUIRE(base0 && off0 && base1 && comp && off1 && n > 0 && nfreq >= 0, "bz_gather2: bad arguments");
    BZ_REQUIRE(nfreq == 0 || (dft && (!act0 || phasor0) && (!act1 || phasor1)),
               "bz_gather2: DFT requested without buffers");
    CompBase c0, c1;
    for (int q = 0; q < 6; ++q) {
        c0.p[q] = base0[q];
        c1.p[q] = base1[q];
    }
    const long threads = (long)n * (nfreq > 0 ? nfreq : 1);
    BZ_REQUIRE(threads / 256 < (1L << 31) - 1, "bz_gather2: too many points x frequencies");
    hipLaunchKernelGGL(k_gather2, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, (hipStream_t)stream, c0, off0,
                       c1, comp, off1, n, act0, act1, series0, series1, nfreq, phasor0, phasor1, dft);
    BZ_CHECK_LAUNCH("bz_gather2");
    return BZ_OK;
}

extern "C" int bz_gather3(const float* const base0[6], const float* const base1[6], const int64_t* off0,
                          const float* const base2[6], const int* comp, const int64_t* off2, int n, int act0,
                          int act1, int act2, float* series0, float* series1, float* series2, int nfreq,
                          const double* phasor0, const double* phasor1, const double* phasor2, double* dft,
                          void* stream) {
    if (n == 0 || (!act0 && !act1 && !act2)) return BZ_OK;
    BZ_REQUIRE(base0 && base1 && off0 && base2 && comp && off2 && n > 0 && nfreq >= 0, "bz_gather3: bad arguments");
    BZ_REQUIRE(nfreq == 0 || (dft && (!act0 || phasor0) && (!act1 || phasor1) && (!act2 || phasor2)),
               "bz_gather3: DFT requested without buffers");
    CompBase c0, c1, c2;
    for (int q = 0; q < 6; ++q) {
        c0.p[q] = base0[q];
        c1.p[q] = base1[q];
        c2.p[q] = base2[q];
    }
    const long threads = (long)n * (nfreq > 0 ? nfreq : 1);
    BZ_REQUIRE(threads / 256 < (1L << 31) - 1, "bz_gather3: too many points x frequencies");
    hipLaunchKernelGGL(k_gather3, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, (hipStream_t)stream, c0, c1,
                       off0, c2, comp, off2, n, act0, act1, act2, series0, series1, series2, nfreq, phasor0, phasor1,
                       phasor2, dft);
    BZ_CHECK_LAUNCH("bz_gather3");
    return BZ_OK;
}
