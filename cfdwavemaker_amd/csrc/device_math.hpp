// Device helpers shared by the kernel files: the reference's frequency ramp,
// the component phase and its sine/cosine in each working precision.
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace cfdwm {
namespace dm {

constexpr double kInv2Pi = 0.15915494309189533577;

// Irregular::freq_ramp (Irregular.cpp:200-204)
__device__ __forceinline__ double freq_ramp(double t, double rt, double ts, double du) {
    const double a = 1. - ((t - ts - du + rt) / (rt + 1e-12));
    const double b = (t - ts) / (rt + 1e-12);
    const double down = a < 1. ? (a < 0. ? 0. : a) : 1.;
    const double up = b < 1. ? (b < 0. ? 0. : b) : 1.;
    return up * down;
}

__device__ __forceinline__ double phase_of(const PhaseRec& p, double dx, double dy, double t) {
    return p.k * (p.cth * dx + p.sth * dy) - p.omega * t + p.ph0;
}

__device__ __forceinline__ void cs_of(double ph, double& c, double& s) { sincos(ph, &s, &c); }
__device__ __forceinline__ void cs_of(double ph, float& c, float& s) {
    // FP64 range reduction to [-1/2, 1/2] revolutions, then the hardware
    // sin/cos (argument in revolutions).
    double r = ph * kInv2Pi;
    r -= rint(r);
    const float f = (float)r;
    s = __builtin_amdgcn_sinf(f);
    c = __builtin_amdgcn_cosf(f);
}

}  // namespace dm
}  // namespace cfdwm
