/*
 * FP64 C restatement of the reference's Yee updates, for long oracle runs.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).
 *
 * Same expressions, in the same evaluation order, as the NumPy restatement in
 * oracle/fdtd_ref.py (update_h_fields_2d / update_e_field_2d and
 * update_fields_3d), which follow
 *   beamz/simulation/backends/numpy_backend.py:34-88 (2D, real fields) and
 *   beamz/simulation/fdtd.py:200-310 (3D, real fields).
 * Built with -O2 -ffp-contract=off and no fast-math, so every operation is one
 * IEEE double rounding as in NumPy: tests/test_oracle_c.py checks that both
 * give the same bits. Loops over cells are independent (each output reads only
 * the previous half-step), so the OpenMP split does not change any result.
 *
 * Arrays are C-order with the reference's shapes:
 *   2D  Ez (ny, nx), Hx (ny, nx-1), Hy (ny-1, nx); eps, sig (ny, nx)
 *   3D  Ex (nz, ny, nx-1), Ey (nz, ny-1, nx), Ez (nz-1, ny, nx),
 *       Hx (nz-1, ny-1, nx), Hy (nz-1, ny, nx-1), Hz (nz, ny-1, nx-1); eps, sig (nz, ny, nx)
 */
#include <stddef.h>

#define I2(j, i, w) ((size_t)(j) * (size_t)(w) + (size_t)(i))
#define I3(k, j, i, h, w) (((size_t)(k) * (size_t)(h) + (size_t)(j)) * (size_t)(w) + (size_t)(i))

/* One 2D step: numpy_backend.py:34-61 (H, new arrays there, in place here), then :63-88 (Ez interior). */
void oref_step2d(double* Ez, double* Hx, double* Hy, const double* sig, const double* eps, int nx, int ny,
                 double dx, double dy, double dt, double mu0, double eps0) {
    const double two_mu0 = 2.0 * mu0, dt_mu0 = dt / mu0, two_eps0 = 2.0 * eps0;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j) {
        for (int i = 0; i < nx - 1; ++i) {           /* Hx: sigma[:, :-1], x-difference over dy (quirk) */
            const double sm = sig[I2(j, i, nx)] * mu0 / eps0;
            const double ce = (Ez[I2(j, i + 1, nx)] - Ez[I2(j, i, nx)]) / dy;
            const double den = 1.0 + sm * dt / two_mu0;
            const double fa = (1.0 - sm * dt / two_mu0) / den;
            const double fb = dt_mu0 / den;
            Hx[I2(j, i, nx - 1)] = fa * Hx[I2(j, i, nx - 1)] - fb * ce;
        }
        if (j < ny - 1) {
            for (int i = 0; i < nx; ++i) {           /* Hy: sigma[:-1, :], y-difference over dx */
                const double sm = sig[I2(j, i, nx)] * mu0 / eps0;
                const double ce = (Ez[I2(j + 1, i, nx)] - Ez[I2(j, i, nx)]) / dx;
                const double den = 1.0 + sm * dt / two_mu0;
                const double fa = (1.0 - sm * dt / two_mu0) / den;
                const double fb = dt_mu0 / den;
                Hy[I2(j, i, nx)] = fa * Hy[I2(j, i, nx)] + fb * ce;
            }
        }
    }
#pragma omp parallel for schedule(static)
    for (int j = 1; j < ny - 1; ++j) {
        for (int i = 1; i < nx - 1; ++i) {
            const double chx = (Hx[I2(j, i, nx - 1)] - Hx[I2(j, i - 1, nx - 1)]) / dy;
            const double chy = (Hy[I2(j, i, nx)] - Hy[I2(j - 1, i, nx)]) / dx;
            const double s = sig[I2(j, i, nx)], e = eps[I2(j, i, nx)];
            const double den = 1.0 + s * dt / (two_eps0 * e);
            const double f1 = (1.0 - s * dt / (two_eps0 * e)) / den;
            const double f2 = (dt / (eps0 * e)) / den;
            Ez[I2(j, i, nx)] = f1 * Ez[I2(j, i, nx)] + f2 * (-chx + chy);
        }
    }
}

static inline void h_fac(double sg, double mu0, double eps0, double dt, double* fa, double* fb) {
    const double sm = sg * mu0 / eps0;
    const double den = 1.0 + sm * dt / (2.0 * mu0);
    *fa = (1.0 - sm * dt / (2.0 * mu0)) / den;
    *fb = (dt / mu0) / den;
}

static inline void e_fac(double s, double e, double eps0, double dt, double* f1, double* f2) {
    const double den = 1.0 + s * dt / (2.0 * eps0 * e);
    *f1 = (1.0 - s * dt / (2.0 * eps0 * e)) / den;
    *f2 = (dt / (eps0 * e)) / den;
}

/* One 3D step: fdtd.py:218-247 (all H from E), then :249-310 (E interiors from the new H). */
/* T = double: the FP64 oracle. T = float: the same FP64 expressions on fields stored in FP32 (each
   result rounded once to float when stored): the accuracy floor of any FP32-stored field, against which
   long GPU runs are judged where the fields have decayed (test infrastructure, not a reference path). */
#define DEFINE_STEP3D(NAME, T) \
void NAME(T* Ex, T* Ey, T* Ez, T* Hx, T* Hy, T* Hz, const double* eps, \
                 const double* sig, int nx, int ny, int nz, double dx, double dy, double dz, double dt, double mu0, \
                 double eps0) { \
_Pragma("omp parallel for schedule(static)") \
    for (int k = 0; k < nz; ++k) { \
        double fa, fb; \
        for (int j = 0; j < ny; ++j) { \
            for (int i = 0; i < nx; ++i) { \
                const double sg = sig[I3(k, j, i, ny, nx)]; \
                h_fac(sg, mu0, eps0, dt, &fa, &fb); \
                if (k < nz - 1 && j < ny - 1) {      /* Hx (nz-1, ny-1, nx) */ \
                    const double ce = ((double)Ez[I3(k, j + 1, i, ny, nx)] - (double)Ez[I3(k, j, i, ny, nx)]) / dy - \
                                      ((double)Ey[I3(k + 1, j, i, ny - 1, nx)] - (double)Ey[I3(k, j, i, ny - 1, nx)]) / dz; \
                    T* h = &Hx[I3(k, j, i, ny - 1, nx)]; \
                    *h = fa * *h - fb * ce; \
                } \
                if (k < nz - 1 && i < nx - 1) {      /* Hy (nz-1, ny, nx-1) */ \
                    const double ce = ((double)Ex[I3(k + 1, j, i, ny, nx - 1)] - (double)Ex[I3(k, j, i, ny, nx - 1)]) / dz - \
                                      ((double)Ez[I3(k, j, i + 1, ny, nx)] - (double)Ez[I3(k, j, i, ny, nx)]) / dx; \
                    T* h = &Hy[I3(k, j, i, ny, nx - 1)]; \
                    *h = fa * *h - fb * ce; \
                } \
                if (j < ny - 1 && i < nx - 1) {      /* Hz (nz, ny-1, nx-1) */ \
                    const double ce = ((double)Ey[I3(k, j, i + 1, ny - 1, nx)] - (double)Ey[I3(k, j, i, ny - 1, nx)]) / dx - \
                                      ((double)Ex[I3(k, j + 1, i, ny, nx - 1)] - (double)Ex[I3(k, j, i, ny, nx - 1)]) / dy; \
                    T* h = &Hz[I3(k, j, i, ny - 1, nx - 1)]; \
                    *h = fa * *h - fb * ce; \
                } \
            } \
        } \
    } \
_Pragma("omp parallel for schedule(static)") \
    for (int k = 0; k < nz; ++k) { \
        double f1, f2; \
        for (int j = 0; j < ny; ++j) { \
            for (int i = 0; i < nx; ++i) { \
                const double s = sig[I3(k, j, i, ny, nx)], e = eps[I3(k, j, i, ny, nx)]; \
                const int kin = k >= 1 && k <= nz - 2, jin = j >= 1 && j <= ny - 2, iin = i >= 1 && i <= nx - 2; \
                if (!((kin && jin && i <= nx - 2) || (kin && j <= ny - 2 && iin) || (k <= nz - 2 && jin && iin))) \
                    continue; \
                e_fac(s, e, eps0, dt, &f1, &f2); \
                if (kin && jin && i <= nx - 2) {     /* (double)Ex[1:-1, 1:-1, :] */ \
                    const double dhz_dy = ((double)Hz[I3(k, j, i, ny - 1, nx - 1)] - (double)Hz[I3(k, j - 1, i, ny - 1, nx - 1)]) / dy; \
                    const double dhy_dz = ((double)Hy[I3(k, j, i, ny, nx - 1)] - (double)Hy[I3(k - 1, j, i, ny, nx - 1)]) / dz; \
                    T* p = &Ex[I3(k, j, i, ny, nx - 1)]; \
                    *p = f1 * *p + f2 * (dhz_dy - dhy_dz); \
                } \
                if (kin && j <= ny - 2 && iin) {     /* (double)Ey[1:-1, :, 1:-1] */ \
                    const double dhx_dz = ((double)Hx[I3(k, j, i, ny - 1, nx)] - (double)Hx[I3(k - 1, j, i, ny - 1, nx)]) / dz; \
                    const double dhz_dx = ((double)Hz[I3(k, j, i, ny - 1, nx - 1)] - (double)Hz[I3(k, j, i - 1, ny - 1, nx - 1)]) / dx; \
                    T* p = &Ey[I3(k, j, i, ny - 1, nx)]; \
                    *p = f1 * *p + f2 * (dhx_dz - dhz_dx); \
                } \
                if (k <= nz - 2 && jin && iin) {     /* (double)Ez[:, 1:-1, 1:-1] */ \
                    const double dhy_dx = ((double)Hy[I3(k, j, i, ny, nx - 1)] - (double)Hy[I3(k, j, i - 1, ny, nx - 1)]) / dx; \
                    const double dhx_dy = ((double)Hx[I3(k, j, i, ny - 1, nx)] - (double)Hx[I3(k, j - 1, i, ny - 1, nx)]) / dy; \
                    T* p = &Ez[I3(k, j, i, ny, nx)]; \
                    *p = f1 * *p + f2 * (dhy_dx - dhx_dy); \
                } \
            } \
        } \
    } \
}

DEFINE_STEP3D(oref_step3d, double)
DEFINE_STEP3D(oref_step3d_f32store, float)
