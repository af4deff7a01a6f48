// Shared helpers for the beamz_amd HIP sources (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/beamz_amd.h"

namespace bz {
void set_error(const std::string& msg);
int hip_fail(const char* where, hipError_t e);
}  // namespace bz

#define BZ_CHECK_LAUNCH(name)                                  \
    do {                                                       \
        hipError_t _e = hipGetLastError();                     \
        if (_e != hipSuccess) return bz::hip_fail(name, _e);   \
    } while (0)

#define BZ_REQUIRE(cond, msg)             \
    do {                                  \
        if (!(cond)) {                    \
            bz::set_error(msg);           \
            return BZ_EINVAL;             \
        }                                 \
    } while (0)

// Update coefficients of the semi-implicit conductivity scheme, evaluated in
// FP32 from the per-cell FP32 relative permittivity and conductivity:
//   H:  s = sigma*dt/(2 eps0); fa = (1-s)/(1+s); fb = (dt/mu0)/(1+s)
//   E:  s = sigma*dt/(2 eps0 eps_r); ca = (1-s)/(1+s); cb = (dt/(eps0 eps_r))/(1+s)
struct Coef {
    float kh;   // dt/(2 eps0)      (= sigma_m * dt / (2 mu0) per unit sigma)
    float kbh;  // dt/mu0
    float ke;   // dt/(2 eps0)
    float kbe;  // dt/eps0
};

__device__ __forceinline__ void h_coef(float sig, const Coef& c, float& fa, float& fb) {
    const float s = sig * c.kh;
    const float r = 1.0f / (1.0f + s);
    fa = (1.0f - s) * r;
    fb = c.kbh * r;
}

__device__ __forceinline__ void e_coef(float sig, float eps, const Coef& c, float& ca, float& cb) {
    const float ie = 1.0f / eps;
    const float s = sig * c.ke * ie;
    const float r = 1.0f / (1.0f + s);
    ca = (1.0f - s) * r;
    cb = c.kbe * ie * r;
}

inline Coef make_coef(double dt, double mu0, double eps0) {
    Coef c;
    c.kh = (float)((mu0 / eps0) * dt / (2.0 * mu0));
    c.kbh = (float)(dt / mu0);
    c.ke = (float)(dt / (2.0 * eps0));
    c.kbe = (float)(dt / eps0);
    return c;
}
