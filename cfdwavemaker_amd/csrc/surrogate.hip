// Linear-field features of the learned second-order surrogate and the
// assembly of its output (no counterpart in the reference).
//
// Per column (x, y) at time t, in two passes over the components:
//   pass 1: eta1, its slopes d eta1/dx, d eta1/dy and the first-order Taylor
//           gradients at z = 0 (Irregular.cpp:178-293 terms: A, -A k cos(theta_dir),
//           phi1_d*dz);
//   pass 2: u1, v1, w1 and the first-order dynamic pressure p1 at the nlev
//           levels z_l = eta1 + sigma_l (h + eta1), profiles at min(0, z_l)
//           and the Taylor term times max(0, z_l) (the s-grid's first-order
//           column semantics with the linear elevation).
// Outputs per column: features X = [eta1, eta1_x, eta1_y, (u1, v1, w1) at
// every fstride-th level, h, kp] and the linear fields
// Lin = [eta1, (u1, v1, w1, p1) at every level].
#include <hip/hip_runtime.h>

#include <cstring>
#include <memory>

#include "engine.hpp"
#include "../../include/cfdwm.h"

namespace cfdwm {

namespace {

constexpr double kInv2PiS = 0.15915494309189533577;

__device__ __forceinline__ double s_ramp(double t, double rt, double ts, double du) {
    const double a = 1. - ((t - ts - du + rt) / (rt + 1e-12));
    const double b = (t - ts) / (rt + 1e-12);
    const double down = a < 1. ? (a < 0. ? 0. : a) : 1.;
    const double up = b < 1. ? (b < 0. ? 0. : b) : 1.;
    return up * down;
}
__device__ __forceinline__ void s_cs(double ph, double& c, double& s) { sincos(ph, &s, &c); }
__device__ __forceinline__ void s_cs(double ph, float& c, float& s) {
    double r = ph * kInv2PiS;
    r -= rint(r);
    const float f = (float)r;
    s = __builtin_amdgcn_sinf(f);
    c = __builtin_amdgcn_cosf(f);
}
__device__ __forceinline__ void s_prof(double pk, double pg, double z, double& E, double& G) {
    E = exp(pk * z);
    G = exp(fma(-pk, z, pg));
}
__device__ __forceinline__ void s_prof(float pk, float pg, float z, float& E, float& G) {
    E = __builtin_amdgcn_exp2f(pk * z);
    G = __builtin_amdgcn_exp2f(fmaf(-pk, z, pg));
}

template <typename R>
__global__ __launch_bounds__(256) void k_surr_linear(SeaDev<R> S, SurrJob J) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= J.n) return;
    const double dx = J.x[q] - S.x0, dy = J.y[q] - S.y0, t = J.t;
    double e1 = 0, sx = 0, sy = 0, gx = 0, gy = 0, gz = 0, gp = 0;
    double eh = 0, hx = 0, hy = 0;                 // quadrature (Hilbert pair) of eta1 and its slopes
    for (int i = 0; i < S.n; ++i) {
        const PhaseRec ph = S.ph[i];
        const CompRec<R>& c = S.cr[i];
        R ci, si;
        s_cs(ph.k * (ph.cth * dx + ph.sth * dy) - ph.omega * t + ph.ph0, ci, si);
        const R ri = (R)s_ramp(t, ph.rt, ph.ts, ph.du);
        const R cr = ci * ri, sr = si * ri;
        const R ak = c.A * (R)ph.k;
        e1 += (double)(c.A * cr);
        sx -= (double)(ak * (R)ph.cth * sr);
        sy -= (double)(ak * (R)ph.sth * sr);
        gx += (double)(c.gxz * cr);
        gy += (double)(c.gyz * cr);
        gz += (double)(c.gzz * sr);
        gp += (double)(c.gpz * cr);
        eh += (double)(c.A * sr);
        hx += (double)(ak * (R)ph.cth * cr);
        hy += (double)(ak * (R)ph.sth * cr);
    }
    float* X = J.X + q * J.xstride;
    if (J.quad) {
        X[J.nx] = (float)eh; X[J.nx + 1] = (float)hx; X[J.nx + 2] = (float)hy;
    }
    float* L = J.Lin + q * (1 + 4 * J.nlev);
    X[0] = (float)e1; X[1] = (float)sx; X[2] = (float)sy;
    L[0] = (float)e1;
    // levels in chunks of 8 (register budget)
    for (int l0 = 0; l0 < J.nlev; l0 += 8) {
        double U[8], V[8], W[8], P[8];
        R zp[8];
        double zt[8];
#pragma unroll
        for (int l = 0; l < 8; ++l) {
            const int lv = l0 + l < J.nlev ? l0 + l : J.nlev - 1;
            const double z = e1 + J.sigma[lv] * (J.h + e1);
            zp[l] = (R)(z < 0. ? z : 0.);
            zt[l] = z > 0. ? z : 0.;
            U[l] = V[l] = W[l] = P[l] = 0.;
        }
        // FP32: the terms of 8 components meet in FP32 before each FP64 fold
        // (the outputs are FP32); FP64: every term folds into the totals
        constexpr int BLK = sizeof(R) == 4 ? 8 : 1;
        for (int i0 = 0; i0 < S.n; i0 += BLK) {
            R Ub[8], Vb[8], Wb[8], Pb[8];
#pragma unroll
            for (int l = 0; l < 8; ++l) Ub[l] = Vb[l] = Wb[l] = Pb[l] = 0;
            const int i1 = i0 + BLK < S.n ? i0 + BLK : S.n;
            for (int i = i0; i < i1; ++i) {
                const PhaseRec ph = S.ph[i];
                const CompRec<R>& c = S.cr[i];
                R ci, si;
                s_cs(ph.k * (ph.cth * dx + ph.sth * dy) - ph.omega * t + ph.ph0, ci, si);
                const R ri = (R)s_ramp(t, ph.rt, ph.ts, ph.du);
                const R cr = ci * ri, sr = si * ri;
#pragma unroll
                for (int l = 0; l < 8; ++l) {
                    R E, G;
                    s_prof(c.pk, c.pg, zp[l], E, G);
                    const R ch = E + G, sh = E - G;
                    Ub[l] += c.lu * cr * ch;
                    Vb[l] += c.lv * cr * ch;
                    Wb[l] += c.lw * sr * sh;
                    Pb[l] += c.lp * cr * ch;
                }
            }
#pragma unroll
            for (int l = 0; l < 8; ++l) {
                U[l] += (double)Ub[l];
                V[l] += (double)Vb[l];
                W[l] += (double)Wb[l];
                P[l] += (double)Pb[l];
            }
        }
#pragma unroll
        for (int l = 0; l < 8; ++l) {
            const int lv = l0 + l;
            if (lv >= J.nlev) break;
            const float u = (float)(U[l] + gx * zt[l]), v = (float)(V[l] + gy * zt[l]);
            const float w = (float)(W[l] + gz * zt[l]), p = (float)(P[l] + gp * zt[l]);
            L[1 + 4 * lv] = u; L[2 + 4 * lv] = v; L[3 + 4 * lv] = w; L[4 + 4 * lv] = p;
            if (lv % J.fstride == 0) {
                const int f = 3 + 3 * (lv / J.fstride);
                X[f] = u; X[f + 1] = v; X[f + 2] = w;
            }
        }
    }
    X[J.nx - 2] = (float)J.h;
    X[J.nx - 1] = (float)J.kp;
}

// X[q][col] = src[q] (row stride `stride`): a column of FP64 values into an FP32 matrix
__global__ void k_surr_set_column(int64_t n, const double* __restrict__ src, float* __restrict__ X, int stride,
                                  int col) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) X[q * stride + col] = (float)src[q];
}

// out = lin + pred elementwise (the surrogate's total fields)
__global__ void k_surr_add(int64_t n, const float* __restrict__ a, const float* __restrict__ b, float* __restrict__ o) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) o[i] = a[i] + b[i];
}

}  // namespace

int surrogate_linear(Sea& s, const SurrJob& J) {
    hipStream_t st = Engine::get().stream();
    const unsigned nb = (unsigned)((J.n + 255) / 256);
    if (J.n <= 0) return 0;
    if (s.prec == kFP32)
        hipLaunchKernelGGL((k_surr_linear<float>), dim3(nb), dim3(256), 0, st, s.dev<float>(), J);
    else
        hipLaunchKernelGGL((k_surr_linear<double>), dim3(nb), dim3(256), 0, st, s.dev<double>(), J);
    CFDWM_HIP(hipGetLastError());
    return 0;
}

int surrogate_set_column(int64_t n, const double* src, float* X, int stride, int col) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_surr_set_column, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, Engine::get().stream(),
                       n, src, X, stride, col);
    CFDWM_HIP(hipGetLastError());
    return 0;
}

int surrogate_add(int64_t n, const float* a, const float* b, float* o) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_surr_add, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, Engine::get().stream(), n, a, b, o);
    CFDWM_HIP(hipGetLastError());
    return 0;
}

}  // namespace cfdwm
