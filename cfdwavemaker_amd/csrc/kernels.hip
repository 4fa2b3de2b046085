// HIP kernels for CFDwavemaker's second-order irregular-wave kinematics on
// gfx950 (MI355X).  Work decomposition:
//   * one thread owns one column (x, y, t) and LC vertical levels of it;
//   * every lane of a wavefront walks the same component i and pair (i, m)
//     at the same time, so component/pair coefficients are wave-uniform and
//     arrive through scalar loads (SGPRs) -- no LDS traffic, no divergence;
//   * the column's own phases cos/sin(phi_m) are recomputed on the fly
//     (FP64 phase, range reduction, then FP32 or FP64 sincos), amortised over
//     the LC levels x 2 interaction branches that share them.
// Reference semantics: src/Irregular.cpp (eta1/eta2/u1..w1/uvw2/u2..dp2,
// phi1_d*dz), src/lsgrid_spline.cpp (column fill, FD gradients, query).
#include <hip/hip_runtime.h>

#include <cstdlib>

#include "device_math.hpp"
#include "engine.hpp"
#include "kernels.hpp"

namespace cfdwm {

namespace {

using dm::cs_of;
using dm::freq_ramp;
using dm::phase_of;

// ---------------------------------------------------------------- helpers

// Component records are read-only for a launch's whole duration: loads
// through the constant address space let the compiler use scalar loads for
// the wave-uniform component index (vector loads of one address otherwise).
template <typename T>
__device__ __forceinline__ T ld_const(const T* p, int64_t i) {
#if defined(__HIP_DEVICE_COMPILE__)
    return ((const __attribute__((address_space(4))) T*)p)[i];
#else
    return p[i];                              // host pass: never executed
#endif
}
// the same for the FP32 kernels; the FP64 ones keep vector loads (their
// records would take more scalar registers than they have left)
template <typename R, typename T>
__device__ __forceinline__ T ld_rec(const T* p, int64_t i) {
    if constexpr (sizeof(R) == 4) return ld_const(p, i);
    else return p[i];
}
// rowActive[i] from its 4-byte word (the table is padded to whole words;
// scalar loads take no single bytes)
__device__ __forceinline__ bool row_active(const uint8_t* a, int i) {
    return ((ld_const((const uint32_t*)a, i >> 2) >> (8 * (i & 3))) & 0xffu) != 0;
}

// depth profile pieces: E = e^{kz}, G = e^{-k(z+2h)}; cosh ratio = C(E+G), sinh ratio = C(E-G)
__device__ __forceinline__ void prof(double pk, double pg, double z, double& E, double& G) {
    E = exp(pk * z);
    G = exp(fma(-pk, z, pg));
}
__device__ __forceinline__ void prof(float pk, float pg, float z, float& E, float& G) {
    E = __builtin_amdgcn_exp2f(pk * z);
    G = __builtin_amdgcn_exp2f(fmaf(-pk, z, pg));
}

template <typename R>
__device__ __forceinline__ R fmaR(R a, R b, R c) { return a * b + c; }

// the exponential of prof(): natural base in FP64, base 2 (hardware) in FP32
__device__ __forceinline__ double pexp(double a) { return exp(a); }
__device__ __forceinline__ float pexp(float a) { return __builtin_amdgcn_exp2f(a); }

// Store N values (N a multiple of 4) at a 16-byte aligned address as 16-byte
// vectors.
template <int N, typename R>
__device__ __forceinline__ void store_rec(R* __restrict__ p, const R (&r)[N]) {
    if constexpr (sizeof(R) == 4) {
        float4* q = reinterpret_cast<float4*>(p);
#pragma unroll
        for (int k = 0; k < N / 4; ++k) q[k] = make_float4(r[4 * k], r[4 * k + 1], r[4 * k + 2], r[4 * k + 3]);
    } else {
        double2* q = reinterpret_cast<double2*>(p);
#pragma unroll
        for (int k = 0; k < N / 2; ++k) q[k] = make_double2(r[2 * k], r[2 * k + 1]);
    }
}

}  // namespace

// =================================================== thread -> column mapping
// Column kernel: one thread per (slot, active column) or per point.
template <typename R>
struct ColPos {
    int64_t q, col;
    int slot;
    double x, y, t;
};
template <typename R, bool GRID>
__device__ __forceinline__ ColPos<R> col_pos(const ColumnJob<R>& J, int64_t tid) {
    const int64_t total = J.nslot * J.ncol;
    ColPos<R> P;
    P.q = tid < total ? tid : total - 1;
    P.col = 0;
    P.slot = 0;
    if (GRID) {
        P.slot = (int)(P.q / J.ncol);
        P.col = J.cols ? J.cols[P.q % J.ncol] : P.q % J.ncol;
        const int i = (int)(P.col / J.g.ny), j = (int)(P.col % J.g.ny);
        P.x = J.g.x0 + J.g.dx * i;
        P.y = J.g.y0 + J.g.dy * j;
        P.t = J.slot_t[P.slot];
    } else {
        P.x = J.x[P.q]; P.y = J.y[P.q]; P.t = J.t[P.q];
    }
    return P;
}

// final column outputs (eta of the slot, or the point arrays)
template <typename R, bool GRID>
__device__ __forceinline__ void col_store(const SeaDev<R>& S, const ColumnJob<R>& J, const ColPos<R>& P,
                                          const double* v /* e1 e2 gx gy gz gp */) {
    if (GRID) {
        double e2 = v[1];
        if (J.q2) {
            double q = 0.;
            for (int k = 0; k < J.nq2; ++k) q += (double)J.q2[(int64_t)k * J.q2stride + P.q];
            e2 += J.q2scale[P.slot] * q;
        }
        J.eta_grid[(int64_t)J.slot_dst[P.slot] * J.ncell + P.col] = (R)(v[0] + e2 + S.swl);
    } else {
        double e2 = v[1];
        if (J.q2) e2 += freq_ramp(P.t, J.q2_rt, J.q2_ts, J.q2_du) * (double)J.q2[P.q];
        J.o_eta1[P.q] = v[0]; J.o_eta2[P.q] = e2;
        J.o_gx[P.q] = v[2]; J.o_gy[P.q] = v[3]; J.o_gz[P.q] = v[4]; J.o_gp[P.q] = v[5];
    }
}

// ============================================================ column kernel
// Per column: eta1, filtered eta2, Taylor gradients (Irregular.cpp:178-293, 1125-1192),
// over the pair records of segment blockIdx.y (rows it owns for the per-row terms).
template <typename R, bool GRID>
__global__ __launch_bounds__(256) void k_columns(SeaDev<R> S, ColumnJob<R> J) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = J.nslot * J.ncol;
    const bool live = tid < total;
    const ColPos<R> P = col_pos<R, GRID>(J, tid);
    const Segment sg = J.seg[blockIdx.y];
    const double dxp = P.x - S.x0, dyp = P.y - S.y0, t = P.t;
    double v[6] = {0, 0, 0, 0, 0, 0};     // e1 e2 gx gy gz gp: FP64 totals, rows in R
    const int ia = min(sg.rlo, sg.own_lo), ib = max(sg.rhi, sg.own_hi);
    // with the eta2 form (J.q2) every component's frequency ramp takes one
    // value at the column's time (checked on the host): one evaluation
    // serves all components, bit for bit
    R ri_q = 0;
    if (sizeof(R) == 4 && J.q2) {
        const PhaseRec p0 = ld_rec<R>(S.ph, 0);
        ri_q = (R)freq_ramp(t, p0.rt, p0.ts, p0.du);
    }
    for (int i = ia; i < ib; ++i) {
        const bool own = i >= sg.own_lo && i < sg.own_hi;
        const bool prs = J.want_eta2 && J.pairs && i >= sg.rlo && i < sg.rhi;
        if (!own && !prs) continue;
        const PhaseRec pi = ld_rec<R>(S.ph, i);
        const CompRec<R>& ci_ = S.cr[i];
        R ci, si;
        cs_of(phase_of(pi, dxp, dyp, t), ci, si);
        const R ri = sizeof(R) == 4 && J.q2 ? ri_q : (R)freq_ramp(t, pi.rt, pi.ts, pi.du);
        const R cr = ci * ri, sr = si * ri;
        if (own) {
            v[0] += (double)(ci_.A * cr);
            v[2] += (double)(ci_.gxz * cr);
            v[3] += (double)(ci_.gyz * cr);
            v[4] += (double)(ci_.gzz * sr);
            v[5] += (double)(ci_.gpz * cr);
            if (J.want_eta2) v[1] += (double)(ci_.e2d * (ci * ci - si * si) * ri);
        }
        if (!prs) continue;
        const int m0 = i == sg.rlo ? sg.mstart : i + 1;
        const int m1 = i == sg.rhi - 1 ? sg.mend : S.fmhi[i];
        const XRec<R>* __restrict__ xr = S.fx + S.frow[i] - (i + 1);
        R acc = 0;
        for (int m = m0; m < m1; ++m) {
            R cm, sm;
            cs_of(phase_of(S.ph[m], dxp, dyp, t), cm, sm);
            const XRec<R> rec = xr[m];
            const R cc = cr * cm, ss = sr * sm;
            acc = fmaR(rec.xP, cc - ss, acc);
            acc = fmaR(rec.xM, cc + ss, acc);
        }
        v[1] += (double)acc;
    }
    if (!live) return;
    if (J.partial) {
        double* o = J.partial + ((int64_t)blockIdx.y * total + tid) * 6;
#pragma unroll
        for (int k = 0; k < 6; ++k) o[k] = v[k];
    } else {
        col_store<R, GRID>(S, J, P, v);
    }
}

// sum of the segment partials of k_columns
template <typename R, bool GRID>
__global__ __launch_bounds__(256) void k_columns_reduce(SeaDev<R> S, ColumnJob<R> J) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = J.nslot * J.ncol;
    if (tid >= total) return;
    const ColPos<R> P = col_pos<R, GRID>(J, tid);
    double v[6] = {0, 0, 0, 0, 0, 0};
    for (int sgi = 0; sgi < J.nseg; ++sgi) {
        const double* o = J.partial + ((int64_t)sgi * total + tid) * 6;
#pragma unroll
        for (int k = 0; k < 6; ++k) v[k] += o[k];
    }
    col_store<R, GRID>(S, J, P, v);
}

// ============================================================= level kernel
// Thread -> (column, chunk of LC levels).  GRID: level chunks are the slowest
// index and every chunk is padded to whole wavefronts, so the chunk (hence
// the absolute level of each unrolled slot) is wave-uniform and the pruning
// tests are scalar branches.  POINTS: one level per point.
template <typename R, int LC, bool GRID>
struct LevPos {
    int64_t cq, col;
    int slot, l0;
    bool live;
    double x, y, t, e, z1;
    R zp[LC];
};
template <typename R, int LC, bool GRID>
__device__ __forceinline__ LevPos<R, LC, GRID> lev_pos(const SeaDev<R>& S, const LevelJob<R>& J, int64_t tid,
                                                       uint32_t fl) {
    LevPos<R, LC, GRID> P;
    const int64_t ncolt = (int64_t)J.nslot * J.ncol;
    const int64_t ncolp = GRID ? (ncolt + 63) / 64 * 64 : ncolt;
    const int nchunk = GRID ? (J.g.nl + LC - 1) / LC : 1;
    const int64_t chunk = tid / ncolp;
    const int64_t cq0 = tid - chunk * ncolp;
    P.live = chunk < nchunk && cq0 < ncolt;
    P.cq = cq0 < ncolt ? cq0 : ncolt - 1;
    P.col = 0; P.slot = 0; P.l0 = 0; P.e = 0.; P.z1 = 0.;
    if (GRID) {
        P.l0 = __builtin_amdgcn_readfirstlane((int)(chunk < nchunk ? chunk : nchunk - 1) * LC);
        P.slot = (int)(P.cq / J.ncol);
        P.col = J.cols ? J.cols[P.cq % J.ncol] : P.cq % J.ncol;
        const int i = (int)(P.col / J.g.ny), j = (int)(P.col % J.g.ny);
        P.x = J.g.x0 + J.g.dx * i;
        P.y = J.g.y0 + J.g.dy * j;
        P.t = J.slot_t[P.slot];
        P.e = (double)J.eta_grid[(int64_t)J.slot_dst[P.slot] * J.ncell + P.col] - J.g.swl;
#pragma unroll
        for (int l = 0; l < LC; ++l) {
            const int lv = P.l0 + l < J.g.nl ? P.l0 + l : J.g.nl - 1;
            const double z = P.e + J.sigma[lv] * (J.g.h + P.e);          // s2z (lsgrid_spline.cpp:85-87)
            P.zp[l] = (R)(z < 0. ? z : 0.);                                // min(0, z), no -swl (:711)
        }
    } else {
        P.x = J.x[P.cq]; P.y = J.y[P.cq]; P.t = J.t[P.cq];
        const double z = J.z[P.cq];
        double zz = z;
        if (fl & EV_ZCLAMP) zz = (z - S.swl) < 0. ? (z - S.swl) : 0.;
        else if (fl & EV_ZSHIFT) zz = z - S.swl;
#pragma unroll
        for (int l = 0; l < LC; ++l) P.zp[l] = (R)zz;
        P.z1 = (z - S.swl) > 0. ? (z - S.swl) : 0.;
    }
    return P;
}

// final level outputs.  a = {U[LC], V[LC], W[LC], gx, gy, gz, gp, P, P2}
template <typename R, int LC, bool GRID, bool WP>
__device__ __forceinline__ void lev_store(const SeaDev<R>& S, const LevelJob<R>& J, uint32_t fl,
                                          const LevPos<R, LC, GRID>& P, const double* a) {
    const double gx = a[3 * LC], gy = a[3 * LC + 1], gz = a[3 * LC + 2], gp = a[3 * LC + 3];
    if (GRID) {
        const int64_t o = ((int64_t)J.slot_dst[P.slot] * J.ncell + P.col) * J.g.nl + P.l0;
#pragma unroll
        for (int l = 0; l < LC; ++l) {
            if (P.l0 + l >= J.g.nl) break;
            const double z = P.e + J.sigma[P.l0 + l] * (J.g.h + P.e);
            const double zt = z - J.g.swl > 0. ? z - J.g.swl : 0.;      // max(0, z - swl)
            J.o_u[o + l] = (R)(a[l] + gx * zt);
            J.o_v[o + l] = (R)(a[LC + l] + gy * zt);
            J.o_w[o + l] = (R)(a[2 * LC + l] + gz * zt);
        }
    } else {
        const bool tay = (fl & EV_TAYLOR) != 0;
        const double z1 = P.z1;
        J.p_u[P.cq] = a[0] + (tay ? gx * z1 : 0.);
        J.p_v[P.cq] = a[LC] + (tay ? gy * z1 : 0.);
        J.p_w[P.cq] = a[2 * LC] + (tay ? gz * z1 : 0.);
        if (WP) J.p_p[P.cq] = a[3 * LC + 4] + S.rho * a[3 * LC + 5] + (tay ? gp * z1 : 0.);
    }
}

// The per-level evaluation of the linear-term-basis mode, once per column:
// the basis sums LU, LV, LW (all components), the form values and the Taylor
// gradients give u, v, w at the column's levels, written to the s-grid.
template <int LC>
__device__ __forceinline__ void lb_levels_out(const LevelJob<float>& J, const LevPos<float, LC, true>& PP,
                                              const float (&LU)[16], const float (&LV)[16], const float (&LW)[16],
                                              double gx, double gy, double gz) {
    using R = float;
    constexpr int LRM = 16, LQ = 16;
    const int64_t o = ((int64_t)J.slot_dst[PP.slot] * J.ncell + PP.col) * J.g.nl + PP.l0;
    const double sc = J.qf_scale[PP.slot];
    const int64_t cq = PP.cq;
    const int Rk = J.qf_rank, Rw = J.qf_rank_w;
    const int nlev = min(LC, J.g.nl - PP.l0);
    // four levels at a time: each form value and basis exponent is loaded
    // once for the four, and the outputs leave as 16-byte vectors when the
    // group is whole and aligned (level counts that are multiples of 4);
    // per level the operations and their order are those of one level alone
    static_assert(LC % 4 == 0, "level groups of four");
    const bool vec = (o & 3) == 0;
    for (int lg = 0; lg < LC; lg += 4) {
        double zz[4];
        R z[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int lv = min(PP.l0 + lg + k, J.g.nl - 1);
            zz[k] = PP.e + J.sigma[lv] * (J.g.h + PP.e);
            z[k] = (R)(zz[k] < 0. ? zz[k] : 0.);                      // min(0, z), as zp
        }
        R su[4] = {0, 0, 0, 0}, sv[4] = {0, 0, 0, 0}, sw[4] = {0, 0, 0, 0};
#pragma unroll
        for (int q = 0; q < LQ; ++q) {
            if (q < J.lin_R) {
                const R m = (R)ld_const(J.lin_mu, q), g = (R)ld_const(J.lin_mu, LRM + q);
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    R E, G;
                    prof(m, g, z[k], E, G);
                    su[k] = fmaR(LU[q], E + G, su[k]);
                    sv[k] = fmaR(LV[q], E + G, sv[k]);
                }
            }
            if (q < J.lin_Rw) {
                const R m = (R)ld_const(J.lin_mu, 2 * LRM + q), g = (R)ld_const(J.lin_mu, 3 * LRM + q);
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    R E, G;
                    prof(m, g, z[k], E, G);
                    sw[k] = fmaR(LW[q], E - G, sw[k]);
                }
            }
        }
        R fu[4] = {0, 0, 0, 0}, fv[4] = {0, 0, 0, 0}, fw[4] = {0, 0, 0, 0};
        for (int q = 0; q < Rk; ++q) {
            const R m2 = (R)J.qf_mu2[q], mg = (R)J.qf_mg[q];
            const R qu = (R)J.qf[(int64_t)q * J.qf_stride + cq], qv = (R)J.qf[(int64_t)(Rk + q) * J.qf_stride + cq];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                R E, G;
                prof(m2, mg, z[k], E, G);
                fu[k] = fmaR(qu, E + G, fu[k]);
                fv[k] = fmaR(qv, E + G, fv[k]);
            }
        }
        for (int q = 0; q < Rw; ++q) {
            const R m2 = (R)J.qf_mu2w[q], mg = (R)J.qf_mgw[q];
            const R qw = (R)J.qf[(int64_t)(2 * J.qf_rank + q) * J.qf_stride + cq];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                R E, G;
                prof(m2, mg, z[k], E, G);
                fw[k] = fmaR(qw, E - G, fw[k]);
            }
        }
        double u[4], v[4], w[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const bool in = lg + k < nlev;
            u[k] = in ? (double)su[k] + sc * (double)fu[k] : 0.;
            v[k] = in ? (double)sv[k] + sc * (double)fv[k] : 0.;
            w[k] = in ? (double)sw[k] + sc * (double)fw[k] : 0.;
        }
        R ou[4], ov[4], ow[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const double zt = zz[k] - J.g.swl > 0. ? zz[k] - J.g.swl : 0.;   // max(0, z - swl)
            ou[k] = (R)(u[k] + gx * zt);
            ov[k] = (R)(v[k] + gy * zt);
            ow[k] = (R)(w[k] + gz * zt);
        }
        if (vec && lg + 4 <= nlev) {
            store_rec<4>(J.o_u + o + lg, ou);
            store_rec<4>(J.o_v + o + lg, ov);
            store_rec<4>(J.o_w + o + lg, ow);
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (lg + k >= nlev) break;
                J.o_u[o + lg + k] = ou[k];
                J.o_v[o + lg + k] = ov[k];
                J.o_w[o + lg + k] = ow[k];
            }
        }
    }
}

// Component loop of the linear-term-basis mode (levels_body, F8 == 2): per
// component one phase, cos/sin, and the 16 x 6 products of the basis
// coefficients.  The component records (phase constants, coefficients and the
// 64-entry basis row, 368 bytes) are staged block-wide in LDS, 32 components
// at a time with the next chunk's global loads in flight meanwhile, and read
// back as wave-uniform 16-byte LDS loads; each component's phase and cos/sin
// are computed one component ahead.  Scalar loads would leave every wave
// waiting for its records once per component.  Same operations and order as
// the per-component form, so the sums are the same bit for bit.
constexpr int kLbChunk = 32, kLbPieces = 23;   // components per chunk, 16-byte pieces per component
constexpr int kLbRec = 52;                     // floats of a segment's partial record: LU, LV, LW, gx, gy, gz, pad
template <int LRM>
__device__ __forceinline__ void lb_component_sums(const SeaDev<float>& S, const LevelJob<float>& J, int lo, int hi,
                                                  double dxp, double dyp, double t, float ri, bool lin, bool secf,
                                                  float (&LU)[LRM], float (&LV)[LRM], float (&LW)[LRM],
                                                  float& gxr, float& gyr, float& gzr) {
    // piece k of chunk slot j: 0-2 PhaseRec (k, cth, sth, omega, ph0, rt),
    // 3-6 CompRec floats 0-15, 7-22 the basis row (4 x 16 floats)
    __shared__ uint4 lb[2][kLbChunk * kLbPieces];
    __shared__ int lsec[2][kLbChunk];
    const int tb = threadIdx.x;
    auto fetch = [&](int c0, uint4 (&r)[3], int& sf) {
#pragma unroll
        for (int u = 0; u < 3; ++u) {
            const int k = tb + u * 256;
            r[u] = make_uint4(0u, 0u, 0u, 0u);
            const int cc = k / kLbPieces, pc = k - cc * kLbPieces;
            const int i = c0 + cc;
            if (k < kLbChunk * kLbPieces && i < hi) {
                const uint4* src = pc < 3 ? reinterpret_cast<const uint4*>(S.ph + i) + pc
                                 : pc < 7 ? reinterpret_cast<const uint4*>(S.cr + i) + (pc - 3)
                                          : reinterpret_cast<const uint4*>(J.lin_c + (int64_t)i * (4 * LRM)) + (pc - 7);
                r[u] = *src;
            }
        }
        sf = 0;
        if (tb < kLbChunk && c0 + tb < hi) {
            const int i = c0 + tb;
            sf = secf && ((reinterpret_cast<const uint32_t*>(S.rowActive)[i >> 2] >> (8 * (i & 3))) & 0xffu) != 0;
        }
    };
    auto stash = [&](int b, const uint4 (&r)[3], int sf) {
#pragma unroll
        for (int u = 0; u < 3; ++u) {
            const int k = tb + u * 256;
            if (k < kLbChunk * kLbPieces) lb[b][k] = r[u];
        }
        if (tb < kLbChunk) lsec[b][tb] = sf;
    };
    uint4 rg[3];
    int sfr = 0;
    if (lo < hi) {
        fetch(lo, rg, sfr);
        stash(0, rg, sfr);
    }
    __syncthreads();
    int b = 0;
    for (int c0 = lo; c0 < hi; c0 += kLbChunk) {          // block-uniform
        const int c1 = c0 + kLbChunk;
        if (c1 < hi) fetch(c1, rg, sfr);                   // in flight during this chunk
        const int nc = min(kLbChunk, hi - c0);
        const uint4* L = lb[b];
        auto phase_cs = [&](int j, float& cj, float& sj) {
            const uint4 a0 = L[j * kLbPieces], a1 = L[j * kLbPieces + 1], a2 = L[j * kLbPieces + 2];
            PhaseRec p;
            p.k = __hiloint2double((int)a0.y, (int)a0.x);
            p.cth = __hiloint2double((int)a0.w, (int)a0.z);
            p.sth = __hiloint2double((int)a1.y, (int)a1.x);
            p.omega = __hiloint2double((int)a1.w, (int)a1.z);
            p.ph0 = __hiloint2double((int)a2.y, (int)a2.x);
            cs_of(phase_of(p, dxp, dyp, t), cj, sj);
        };
        float ci, si;
        phase_cs(0, ci, si);
        for (int j = 0; j < nc; ++j) {
            float cn = 0.f, sn = 0.f;
            if (j + 1 < nc) phase_cs(j + 1, cn, sn);       // next component, independent of this one's sums
            const uint4* P = L + j * kLbPieces;
            const uint4 r0 = P[3], r1 = P[4], r2 = P[5], r3 = P[6];
            const float lu = __uint_as_float(r0.x), lv = __uint_as_float(r0.y), lw = __uint_as_float(r0.z);
            const float gxz = __uint_as_float(r1.x), gyz = __uint_as_float(r1.y), gzz = __uint_as_float(r1.z);
            const float dfu = __uint_as_float(r2.z), dfv = __uint_as_float(r2.w), dfw = __uint_as_float(r3.x);
            const bool sec = lsec[b][j] != 0;
            const float cr = ci * ri, sr = si * ri;
            gxr = fmaR(gxz, cr, gxr);
            gyr = fmaR(gyz, cr, gyr);
            gzr = fmaR(gzz, sr, gzr);
            // first-order terms on the profile of k, diagonal on that of 2k
            const float xu = lin ? lu * cr : 0.f, xv = lin ? lv * cr : 0.f, xw = lin ? lw * sr : 0.f;
            const float c2 = sec ? (ci * ci - si * si) * ri : 0.f, s2 = sec ? 2.f * si * ci * ri : 0.f;
            const float du = dfu * c2, dv = dfv * c2, dw = dfw * s2;
            const float* cf = reinterpret_cast<const float*>(P + 7);
#pragma unroll
            for (int q = 0; q < LRM; ++q) {
                const float a = cf[q], a2 = cf[LRM + q], bq = cf[2 * LRM + q], b2 = cf[3 * LRM + q];
                LU[q] = fmaR(xu, a, fmaR(du, a2, LU[q]));
                LV[q] = fmaR(xv, a, fmaR(dv, a2, LV[q]));
                LW[q] = fmaR(xw, bq, fmaR(dw, b2, LW[q]));
            }
            ci = cn;
            si = sn;
        }
        if (c1 < hi) stash(b ^ 1, rg, sfr);                // the other buffer's readers passed the last barrier
        __syncthreads();
        b ^= 1;
    }
}

// Per (column, chunk of LC levels): first order + second order + Taylor over
// the pair records of segment blockIdx.y.
// MODE 0 (GRID, lsgrid_spline.cpp:708-721): z from the sigma map of the
//   column's eta, profiles at min(0, z), Taylor with max(0, z - swl), filtered
//   uvw2 terms.  MODE 1/2 (POINTS, one level per point, 2 = with pressure):
//   flags select the terms (EV_*).
// Accumulation: each row i (pairs (i, m), m > i) is summed in the working
// precision R and folded into FP64 totals, which keeps the FP32 error at the
// level of the per-term rounding instead of growing over ~N^2/2 terms.
// F8: FP32 grid fill whose pair sums come from the quadratic forms (J.qf);
// F8 == 2: its first-order and diagonal terms through the linear-term
// profile bases as well (J.lin_c): per component the basis coefficients are
// accumulated (LU, LV, LW), the profiles evaluated once per level at the end.
template <typename R, int LC, int MODE, int F8 = 0>
__device__ __forceinline__ void levels_body(const SeaDev<R>& S, const LevelJob<R>& J) {
    constexpr bool GRID = MODE == 0;
    constexpr bool WP = MODE == 2;
    constexpr int RS = 3 * LC + 6;
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t fl = GRID ? (uint32_t)(EV_LIN | EV_SEC | EV_TAYLOR) : J.flags;
    const LevPos<R, LC, GRID> PP = lev_pos<R, LC, GRID>(S, J, tid, fl);
    const int l0 = PP.l0;
    R zp[LC];
#pragma unroll
    for (int l = 0; l < LC; ++l) zp[l] = PP.zp[l];
    const bool direct = !GRID && (fl & EV_DIRECT) != 0;
    const bool lin = (fl & EV_LIN) != 0;
    const bool secf = (fl & EV_SEC) != 0;
    const double dxp = PP.x - S.x0, dyp = PP.y - S.y0, t = PP.t;
    const Segment sg = J.seg[blockIdx.y];
    double U[LC], V[LC], W[LC];
#pragma unroll
    for (int l = 0; l < LC; ++l) U[l] = V[l] = W[l] = 0.;
    double P = 0., P2 = 0.;   // pressure (MODE 2, LC == 1)
    double gx = 0, gy = 0, gz = 0, gp = 0;

    // dp2 kinetic terms (Irregular.cpp:755-756, 786-787): phase-free and
    // separable, -(Xx_i Xx_m + Xy_i Xy_m) px_i px_m + Z_i Z_m pz_i pz_m over
    // m > i, summed with running sums over descending i; plus the diagonal
    // "second term".  Done once per point, by segment 0.
    if (WP && direct && secf && blockIdx.y == 0) {
        double kx = 0., ky = 0., kz = 0.;
        for (int i = S.n - 1; i >= 0; --i) {
            const PhaseRec pi = S.ph[i];
            const CompRec<R> c = S.cr[i];
            const double ri = freq_ramp(t, pi.rt, pi.ts, pi.du);
            R E, G;
            prof(c.pk, c.pg, zp[0], E, G);
            const double dch = (double)(E + G), dsh = (double)(E - G);
            P2 += ri * (-dch * ((double)c.Xx * kx + (double)c.Xy * ky) + dsh * (double)c.Zk * kz);
            P2 += ri * (-(double)c.kd1 * dch + (double)c.kd2 * dsh);
            kx += (double)c.Xx * dch;
            ky += (double)c.Xy * dch;
            kz += (double)c.Zk * dsh;
        }
    }

    const UvwRec<R>* __restrict__ tab = direct ? S.duvw : S.fuvw;
    const int ia = min(sg.rlo, sg.own_lo), ib = max(sg.rhi, sg.own_hi);
    // The terms of a component meet in ur/vr/wr and are folded into the FP64
    // totals.  FP32 grid fills whose pair sums come from the quadratic forms
    // (F8: a component then adds its linear and diagonal terms only) fold
    // once per 8 components; elsewhere the FP32 row sums of the pair terms
    // are folded per component.  The component loop is wave-uniform.
    constexpr bool SQ = sizeof(R) == 4;
    constexpr int fold_n = (SQ && GRID && F8 != 0) ? 8 : 1;
    constexpr bool LB = F8 == 2 && SQ && GRID;      // linear-term bases
    constexpr int LRM = 16;                          // their largest rank
    constexpr int LQ = LB ? LRM : 1;
    R LU[LQ], LV[LQ], LW[LQ];
#pragma unroll
    for (int q = 0; q < LQ; ++q) LU[q] = LV[q] = LW[q] = 0;
    // LB: the Taylor gradients' component sums in the working precision
    R gxr = 0, gyr = 0, gzr = 0;
    R ur[LC], vr[LC], wr[LC];
#pragma unroll
    for (int l = 0; l < LC; ++l) ur[l] = vr[l] = wr[l] = 0;
    int nfold = 0;
    // form-assisted fills run only when every component's frequency ramp
    // takes one value at each slot time (checked on the host), so one
    // evaluation per thread serves all components, bit for bit
    R ri_f8 = 0;
    if constexpr (F8 != 0) {
        const PhaseRec p0 = ld_rec<R>(S.ph, 0);
        ri_f8 = (R)freq_ramp(t, p0.rt, p0.ts, p0.du);
    }
    if constexpr (LB) {
        lb_component_sums<LRM>(S, J, sg.own_lo, sg.own_hi, dxp, dyp, t, ri_f8, lin, secf, LU, LV, LW, gxr, gyr, gzr);
    }
    for (int i = ia; !LB && i < ib; ++i) {
        const bool own = i >= sg.own_lo && i < sg.own_hi;
        const bool sec = secf && (direct || (sizeof(R) == 4 ? row_active(S.rowActive, i) : S.rowActive[i] != 0));
        // (form-assisted fills, F8, launch only with the forms: no pair rows)
        const bool prs = F8 == 0 && sec && !(GRID && J.qf) && i >= sg.rlo && i < sg.rhi;
        if (!own && !prs) continue;
        const PhaseRec pi = ld_rec<R>(S.ph, i);
        const CompRec<R> c = ld_rec<R>(S.cr, i);
        R ci, si;
        cs_of(phase_of(pi, dxp, dyp, t), ci, si);
        const R ri = F8 != 0 ? ri_f8 : (R)freq_ramp(t, pi.rt, pi.ts, pi.du);
        const R cr = ci * ri, sr = si * ri;
        R pr = 0;
        if constexpr (fold_n == 1) {
#pragma unroll
            for (int l = 0; l < LC; ++l) ur[l] = vr[l] = wr[l] = 0;
        }
        if (own) {
            gx += (double)(c.gxz * cr);
            gy += (double)(c.gyz * cr);
            gz += (double)(c.gzz * sr);
            if (WP) gp += (double)(c.gpz * cr);
            // FP32: the diagonal's 2k profile is the square of the k profile
            // (p2k = 2 pk, p2g = 2 pg), so the k exponentials serve both
            R Ek[LC], Gk[LC];
            if (lin || (SQ && sec)) {
#pragma unroll
                for (int l = 0; l < LC; ++l) prof(c.pk, c.pg, zp[l], Ek[l], Gk[l]);
            }
            if (lin) {
#pragma unroll
                for (int l = 0; l < LC; ++l) {
                    const R E = Ek[l], G = Gk[l];
                    const R ch = E + G, sh = E - G;
                    if constexpr (SQ) {
                        ur[l] = fmaR(c.lu * cr, ch, ur[l]);
                        vr[l] = fmaR(c.lv * cr, ch, vr[l]);
                        wr[l] = fmaR(c.lw * sr, sh, wr[l]);
                    } else {
                        U[l] += (double)(c.lu * cr * ch);
                        V[l] += (double)(c.lv * cr * ch);
                        W[l] += (double)(c.lw * sr * sh);
                    }
                    if (WP) P += (double)(c.lp * cr * ch);
                }
            }
            if (sec) {
                // diagonal self-interaction (profile of 2k)
                const R c2 = (ci * ci - si * si) * ri, s2 = (R)2 * si * ci * ri;
                const R du = direct ? c.ddu : c.dfu, dv = direct ? c.ddv : c.dfv, dw = direct ? c.ddw : c.dfw;
#pragma unroll
                for (int l = 0; l < LC; ++l) {
                    R ch, sh;
                    if constexpr (SQ) {
                        // E^2 + G^2 = (E + G)^2 - 2 E G with E G = e^{-2kh} per
                        // component, E^2 - G^2 = (E + G)(E - G)
                        const R c1 = Ek[l] + Gk[l], s1 = Ek[l] - Gk[l];
                        ch = fmaR(c1, c1, -c.eg2);
                        sh = c1 * s1;
                    } else {
                        R E, G;
                        prof(c.p2k, c.p2g, zp[l], E, G);
                        ch = E + G;
                        sh = E - G;
                    }
                    if constexpr (SQ) {
                        ur[l] += du * c2 * ch;
                        vr[l] += dv * c2 * ch;
                        wr[l] += dw * s2 * sh;
                    } else {
                        ur[l] = du * c2 * ch;
                        vr[l] = dv * c2 * ch;
                        wr[l] = dw * s2 * sh;
                    }
                    if (WP && direct) pr = c.ddp * c2 * ch;
                }
            }
        }
        if (prs) {
            // pair terms, m in [m0, m1)
            const int m0 = i == sg.rlo ? sg.mstart : i + 1;
            const int m1 = i == sg.rhi - 1 ? sg.mend : (direct ? S.n : S.fmhi[i]);
            const int64_t base = (direct ? S.drow[i] : S.frow[i]) - (i + 1);
            const UvwRec<R>* __restrict__ rr = tab + base;
            const XRec<R>* __restrict__ xr = S.dx + base;
            const uint32_t* __restrict__ lvr = J.lvl + base;
            for (int m = m0; m < m1; ++m) {
                R cm, sm;
                cs_of(phase_of(S.ph[m], dxp, dyp, t), cm, sm);
                const UvwRec<R> r = rr[m];
                const R cc = cr * cm, ss = sr * sm, sc = sr * cm, cs = cr * sm;
                const R cP = cc - ss, cM = cc + ss, sP = sc + cs, sM = sc - cs;
                const R uP = r.cuP * cP, vP = r.cvP * cP, wP = r.cwP * sP;
                const R uM = r.cuM * cM, vM = r.cvM * cM, wM = r.cwM * sM;
                if (GRID && J.lvl) {
                    // levels at which e^{kz} / e^{-k(z+2h)} can exceed the pruning
                    // threshold (bounds from this launch's eta range)
                    const uint32_t lv = lvr[m];
                    const int lEP = (int)(lv & 255u), lGP = (int)((lv >> 8) & 255u);
                    const int lEM = (int)((lv >> 16) & 255u), lGM = (int)(lv >> 24);
#pragma unroll
                    for (int l = 0; l < LC; ++l) {
                        const int L = l0 + l;
                        // separate blocks for e^{kz} and e^{-k(z+2h)}; the empty asm keeps
                        // the compiler from speculating a skipped exponential into a select
                        if (L >= lEP) {
                            __asm__ volatile("");
                            const R E = pexp(r.pkP * zp[l]);
                            ur[l] = fmaR(uP, E, ur[l]); vr[l] = fmaR(vP, E, vr[l]); wr[l] = fmaR(wP, E, wr[l]);
                        }
                        if (L < lGP) {
                            __asm__ volatile("");
                            const R G = pexp(fmaR(-r.pkP, zp[l], r.pgP));
                            ur[l] = fmaR(uP, G, ur[l]); vr[l] = fmaR(vP, G, vr[l]); wr[l] = fmaR(-wP, G, wr[l]);
                        }
                        if (L >= lEM) {
                            __asm__ volatile("");
                            const R E = pexp(r.pkM * zp[l]);
                            ur[l] = fmaR(uM, E, ur[l]); vr[l] = fmaR(vM, E, vr[l]); wr[l] = fmaR(wM, E, wr[l]);
                        }
                        if (L < lGM) {
                            __asm__ volatile("");
                            const R G = pexp(fmaR(-r.pkM, zp[l], r.pgM));
                            ur[l] = fmaR(uM, G, ur[l]); vr[l] = fmaR(vM, G, vr[l]); wr[l] = fmaR(-wM, G, wr[l]);
                        }
                    }
                    continue;
                }
                R pP = 0, pM = 0;
                if (WP && direct) {
                    const XRec<R> xq = xr[m];
                    pP = xq.xP * cP;
                    pM = xq.xM * cM;
                }
#pragma unroll
                for (int l = 0; l < LC; ++l) {
                    R E, G;
                    prof(r.pkP, r.pgP, zp[l], E, G);
                    R ch = E + G, sh = E - G;
                    ur[l] = fmaR(uP, ch, ur[l]);
                    vr[l] = fmaR(vP, ch, vr[l]);
                    wr[l] = fmaR(wP, sh, wr[l]);
                    if (WP && direct) pr = fmaR(pP, ch, pr);
                    prof(r.pkM, r.pgM, zp[l], E, G);
                    ch = E + G; sh = E - G;
                    ur[l] = fmaR(uM, ch, ur[l]);
                    vr[l] = fmaR(vM, ch, vr[l]);
                    wr[l] = fmaR(wM, sh, wr[l]);
                    if (WP && direct) pr = fmaR(pM, ch, pr);
                }
            }
        }
        if (!LB && ++nfold == fold_n) {
#pragma unroll
            for (int l = 0; l < LC; ++l) {
                U[l] += (double)ur[l];
                V[l] += (double)vr[l];
                W[l] += (double)wr[l];
                if constexpr (fold_n > 1) ur[l] = vr[l] = wr[l] = 0;
            }
            nfold = 0;
        }
        if (WP && direct) P2 += (double)pr;
    }
    if (!LB && fold_n > 1 && nfold > 0) {
#pragma unroll
        for (int l = 0; l < LC; ++l) {
            U[l] += (double)ur[l];
            V[l] += (double)vr[l];
            W[l] += (double)wr[l];
        }
    }
    if constexpr (LB) {
        // all levels of the column in one thread: the segment's basis sums
        // (LU, LV, LW and the Taylor gradients, kLbRec floats) go to
        // k_levels_reduce_lb when the components are split over segments;
        // otherwise the levels are evaluated here
        if (!PP.live) return;
        if (J.partial) {
            const int64_t nthr = ((int64_t)J.nslot * J.ncol + 63) / 64 * 64 * ((J.g.nl + LC - 1) / LC);
            float* pf = reinterpret_cast<float*>(J.partial) + ((int64_t)blockIdx.y * nthr + tid) * kLbRec;
            store_rec<16>(pf, LU);
            store_rec<16>(pf + 16, LV);
            store_rec<16>(pf + 32, LW);
            const float g4[4] = {gxr, gyr, gzr, 0.f};
            store_rec<4>(pf + 48, g4);
            return;
        }
        lb_levels_out<LC>(J, PP, LU, LV, LW, (double)gxr, (double)gyr, (double)gzr);
        return;
    }
    if constexpr (GRID) {
        if (J.qf && blockIdx.y == 0) {
            // pair sums from the quadratic forms of the profile depth bases
            // (once per point: segment 0)
            R su[LC], sv[LC], sw[LC];
#pragma unroll
            for (int l = 0; l < LC; ++l) su[l] = sv[l] = sw[l] = 0;
            const int64_t cq = PP.cq;
            const int Rk = J.qf_rank, Rw = J.qf_rank_w;
            for (int q = 0; q < Rk; ++q) {            // cosh-type basis: u, v
                const R Qu = (R)J.qf[(int64_t)q * J.qf_stride + cq];
                const R Qv = (R)J.qf[(int64_t)(Rk + q) * J.qf_stride + cq];
                const R m2 = (R)J.qf_mu2[q], mg = (R)J.qf_mg[q];
#pragma unroll
                for (int l = 0; l < LC; ++l) {
                    R E, G;
                    prof(m2, mg, zp[l], E, G);
                    su[l] = fmaR(Qu, E + G, su[l]);
                    sv[l] = fmaR(Qv, E + G, sv[l]);
                }
            }
            for (int q = 0; q < Rw; ++q) {            // sinh-type basis: w
                const R Qw = (R)J.qf[(int64_t)(2 * Rk + q) * J.qf_stride + cq];
                const R m2 = (R)J.qf_mu2w[q], mg = (R)J.qf_mgw[q];
#pragma unroll
                for (int l = 0; l < LC; ++l) {
                    R E, G;
                    prof(m2, mg, zp[l], E, G);
                    sw[l] = fmaR(Qw, E - G, sw[l]);
                }
            }
            const double sc = J.qf_scale[PP.slot];
#pragma unroll
            for (int l = 0; l < LC; ++l) {
                U[l] += sc * (double)su[l];
                V[l] += sc * (double)sv[l];
                W[l] += sc * (double)sw[l];
            }
        }
    }
    if (!PP.live) return;
    double a[RS];
#pragma unroll
    for (int l = 0; l < LC; ++l) { a[l] = U[l]; a[LC + l] = V[l]; a[2 * LC + l] = W[l]; }
    a[3 * LC] = gx; a[3 * LC + 1] = gy; a[3 * LC + 2] = gz; a[3 * LC + 3] = gp;
    a[3 * LC + 4] = P; a[3 * LC + 5] = P2;
    if (J.partial) {
        const int64_t nthr = (GRID ? ((int64_t)J.nslot * J.ncol + 63) / 64 * 64 * ((J.g.nl + LC - 1) / LC)
                                   : (int64_t)J.nslot * J.ncol);
        double* o = J.partial + ((int64_t)blockIdx.y * nthr + tid) * RS;
#pragma unroll
        for (int k = 0; k < RS; ++k) o[k] = a[k];
    } else {
        lev_store<R, LC, GRID, WP>(S, J, fl, PP, a);
    }
}

template <typename R, int LC, int MODE>
__global__ __launch_bounds__(256) void k_levels(SeaDev<R> S, LevelJob<R> J) {
    levels_body<R, LC, MODE>(S, J);
}

// FP32 s-grid fill: the register budget is steered to 128 VGPRs so that four
// waves share each SIMD (the kernel is issue-bound; more waves hide the
// scalar-load and transcendental latencies).
template <int LC, int F8>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8)))
void k_levels_grid32(SeaDev<float> S, LevelJob<float> J) {
    levels_body<float, LC, 0, F8>(S, J);
}

// Linear-term-basis mode split over component segments: the sum of the
// segments' basis-sum records (segment order), then the per-level evaluation
// once per column.
template <int LC>
__global__ __launch_bounds__(256) void k_levels_reduce_lb(SeaDev<float> S, LevelJob<float> J) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nthr = ((int64_t)J.nslot * J.ncol + 63) / 64 * 64 * ((J.g.nl + LC - 1) / LC);
    if (tid >= nthr) return;
    const LevPos<float, LC, true> PP = lev_pos<float, LC, true>(S, J, tid, (uint32_t)(EV_LIN | EV_SEC | EV_TAYLOR));
    if (!PP.live) return;
    float LU[16], LV[16], LW[16], g[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int q = 0; q < 16; ++q) LU[q] = LV[q] = LW[q] = 0.f;
    for (int sgi = 0; sgi < J.nseg; ++sgi) {
        const float4* r = reinterpret_cast<const float4*>(reinterpret_cast<const float*>(J.partial) +
                                                          ((int64_t)sgi * nthr + tid) * kLbRec);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float4 a = r[k], b = r[4 + k], c = r[8 + k];
            LU[4 * k] += a.x; LU[4 * k + 1] += a.y; LU[4 * k + 2] += a.z; LU[4 * k + 3] += a.w;
            LV[4 * k] += b.x; LV[4 * k + 1] += b.y; LV[4 * k + 2] += b.z; LV[4 * k + 3] += b.w;
            LW[4 * k] += c.x; LW[4 * k + 1] += c.y; LW[4 * k + 2] += c.z; LW[4 * k + 3] += c.w;
        }
        const float4 d = r[12];
        g[0] += d.x; g[1] += d.y; g[2] += d.z;
    }
    lb_levels_out<LC>(J, PP, LU, LV, LW, (double)g[0], (double)g[1], (double)g[2]);
}

// sum of the segment partials of k_levels
template <typename R, int LC, int MODE>
__global__ __launch_bounds__(256) void k_levels_reduce(SeaDev<R> S, LevelJob<R> J) {
    constexpr bool GRID = MODE == 0;
    constexpr bool WP = MODE == 2;
    constexpr int RS = 3 * LC + 6;
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nthr = (GRID ? ((int64_t)J.nslot * J.ncol + 63) / 64 * 64 * ((J.g.nl + LC - 1) / LC)
                               : (int64_t)J.nslot * J.ncol);
    if (tid >= nthr) return;
    const uint32_t fl = GRID ? (uint32_t)(EV_LIN | EV_SEC | EV_TAYLOR) : J.flags;
    const LevPos<R, LC, GRID> PP = lev_pos<R, LC, GRID>(S, J, tid, fl);
    if (!PP.live) return;
    double a[RS];
#pragma unroll
    for (int k = 0; k < RS; ++k) a[k] = 0.;
    for (int sgi = 0; sgi < J.nseg; ++sgi) {
        const double* o = J.partial + ((int64_t)sgi * nthr + tid) * RS;
#pragma unroll
        for (int k = 0; k < RS; ++k) a[k] += o[k];
    }
    lev_store<R, LC, GRID, WP>(S, J, fl, PP, a);
}

// ======================================================== FD gradient kernels
// update_gradient_eta_dxdy / update_gradient_dxdydz (lsgrid_spline.cpp:163-366):
// central differences inside, one-sided at the edges, zero for a single point.
// Output slot o (0, 1) holds ring slot tids[o].
template <typename R>
__global__ __launch_bounds__(256) void k_grad_xyz(const R* __restrict__ D, R* __restrict__ GX,
                                                  R* __restrict__ GY, R* __restrict__ GS,
                                                  int nx, int ny, int L, int tid0, int tid1,
                                                  double dx, double dy, double ds) {
    const int64_t per = (int64_t)nx * ny * L;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 2 * per) return;
    const int o = (int)(e / per);
    const int64_t qq = e % per;
    const int m = (int)(qq % L);
    const int j = (int)((qq / L) % ny);
    const int i = (int)(qq / ((int64_t)L * ny));
    const R* d = D + (int64_t)(o ? tid1 : tid0) * per;
    const int64_t sx = (int64_t)ny * L, sy = L;
    R vx, vy;
    if (nx == 1) vx = 0;
    else if (i == 0) vx = (R)((d[qq + sx] - d[qq]) / dx);
    else if (i == nx - 1) vx = (R)((d[qq] - d[qq - sx]) / dx);
    else vx = (R)((d[qq + sx] - d[qq - sx]) / (2 * dx));
    if (ny == 1) vy = 0;
    else if (j == 0) vy = (R)((d[qq + sy] - d[qq]) / dy);
    else if (j == ny - 1) vy = (R)((d[qq] - d[qq - sy]) / dy);
    else vy = (R)((d[qq + sy] - d[qq - sy]) / (2 * dy));
    GX[o * per + qq] = vx;
    GY[o * per + qq] = vy;
    if (GS) {
        R vs;
        if (L == 1) vs = 0;
        else if (m == 0) vs = (R)((d[qq + 1] - d[qq]) / ds);
        else if (m == L - 1) vs = (R)((d[qq] - d[qq - 1]) / ds);
        else vs = (R)((d[qq + 1] - d[qq - 1]) / (2 * ds));
        GS[o * per + qq] = vs;
    }
}

// update_gradient_dt / update_gradient_eta_dt (lsgrid_spline.cpp:124-161)
template <typename R>
__global__ __launch_bounds__(256) void k_grad_dt(const R* __restrict__ D, R* __restrict__ GD,
                                                 int64_t per, int tstep, double dt) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= per) return;
    if (tstep > 0) {
        const int tid = (tstep + 2) % 4, t1 = (tstep + 3) % 4, t0 = (tstep + 1) % 4;
        GD[tid * per + e] = (R)((D[t1 * per + e] - D[t0 * per + e]) / (2 * dt));
    } else {
        GD[1 * per + e] = (R)((D[2 * per + e] - D[1 * per + e]) / dt);
        GD[2 * per + e] = (R)((D[3 * per + e] - D[1 * per + e]) / (2 * dt));
    }
}

// ================================================================ spline query
// The two time slots a query reads, (tstep+1)%4 and (tstep+2)%4, are packed
// after every fill/update by k_pack:
//   column record (8 R): eta(tid1) eta(tid2) eta_t(tid1) eta_t(tid2) eta_x(0) eta_x(1) eta_y(0) eta_y(1)
//   node records C and D (16 R each): for u, v, w in turn f, f_t, f_x, f_y, f_s,
//     C from time slot tid1 and derivative slot 0, D from tid2 and derivative slot 1.
// Levels are the fastest index, so the corners l and l+1 of a query share a
// cache line and one query touches 8 node lines instead of 8 x 30 arrays.
template <typename R>
__global__ __launch_bounds__(256) void k_pack(PackJob<R> J) {
    const int64_t per = (int64_t)J.nx * J.ny * J.nl;
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= per) return;
    // records are assembled in registers and stored as 16-byte vectors (the
    // packed arrays are 256-byte aligned allocations, records 16 R long)
    R c[16], d[16];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        c[5 * f + 0] = J.F[f][J.tid1 * per + n];
        c[5 * f + 1] = J.Ft[f][J.tid1 * per + n];
        c[5 * f + 2] = J.Fx[f][n];
        c[5 * f + 3] = J.Fy[f][n];
        c[5 * f + 4] = J.Fs[f][n];
        d[5 * f + 0] = J.F[f][J.tid2 * per + n];
        d[5 * f + 1] = J.Ft[f][J.tid2 * per + n];
        d[5 * f + 2] = J.Fx[f][per + n];
        d[5 * f + 3] = J.Fy[f][per + n];
        d[5 * f + 4] = J.Fs[f][per + n];
    }
    c[15] = 0;
    d[15] = 0;
    store_rec<16>(J.pk_C + 16 * n, c);
    store_rec<16>(J.pk_D + 16 * n, d);
    if (n % J.nl == 0) {
        const int64_t pc = (int64_t)J.nx * J.ny, col = n / J.nl;
        R e[8];
        e[0] = J.E[J.tid1 * pc + col];
        e[1] = J.E[J.tid2 * pc + col];
        e[2] = J.Et[J.tid1 * pc + col];
        e[3] = J.Et[J.tid2 * pc + col];
        e[4] = J.Ex[col];
        e[5] = J.Ex[pc + col];
        e[6] = J.Ey[col];
        e[7] = J.Ey[pc + col];
        store_rec<8>(J.pk_eta + 8 * col, e);
    }
}

namespace {
template <typename R>
__device__ __forceinline__ R herm(R f0, R f1, R a, R b, R u) {
    return f0 * ((R)1 - u) + f1 * u + u * ((R)1 - u) * (a * ((R)1 - u) + b * u);
}
template <typename R>
__device__ __forceinline__ R lin(R f0, R f1, R u) { return f0 * ((R)1 - u) + f1 * u; }

// cart2sigmaS (lsgrid_spline.cpp:107-112).  Points in (eta - swl, eta] give
// sigma > 0 when swl > 0, where the reference indexes with a NaN; they are
// mapped to the surface here.
__device__ __forceinline__ double cart2sigmaS(double zpt, double e, double h, double swl,
                                              double ta, double tb, double tanab) {
    const double z = zpt < e ? zpt : e;
    double sigma = (z - (e - swl)) / (h + (e - swl));
    sigma = sigma < -1. ? -1. : sigma;
    sigma = sigma > 0. ? 0. : sigma;
    return 1 - atan(pow(-sigma * tanab, 1. / tb)) / ta;
}

// The same map in FP32 for FP32 grids (k_query<float, .>): divisions by the
// grid constants become products with their reciprocals, pow and atan the
// FP32 library functions.  The sigma coordinate then carries a few 1e-7 of
// absolute error, far below the FP32 grid values' own.
struct QueryConst {
    double inv_dx, inv_dy, inv_dt;     // 1 / dx, 1 / dy, 1 / dt
    float h, swl, tanab;               // the FP64 map's depth, swl and tanab
    float ita, itb, inv_ds;            // 1 / tan_a, 1 / tan_b, 1 / ds
};
__device__ __forceinline__ float cart2sigmaS_f(float zpt, float e, const QueryConst& K) {
    const float z = zpt < e ? zpt : e;
    float sigma = (z - (e - K.swl)) / (K.h + (e - K.swl));
    sigma = sigma < -1.f ? -1.f : sigma;
    sigma = sigma > 0.f ? 0.f : sigma;
    return 1.f - atanf(powf(-sigma * K.tanab, K.itb)) * K.ita;
}

// A packed record of N values (N a multiple of 4) at a 16-byte aligned
// address, loaded as 16-byte vectors.  Node records hold 16 values (3 fields x
// {value, d/dt, d/dx, d/dy, d/ds} + pad), column surface records 8.
template <int N, typename R>
__device__ __forceinline__ void load_rec(R (&r)[N], const R* __restrict__ p) {
    if constexpr (sizeof(R) == 4) {
        const float4* q = reinterpret_cast<const float4*>(p);
#pragma unroll
        for (int k = 0; k < N / 4; ++k) {
            const float4 v = q[k];
            r[4 * k] = v.x; r[4 * k + 1] = v.y; r[4 * k + 2] = v.z; r[4 * k + 3] = v.w;
        }
    } else {
        const double2* q = reinterpret_cast<const double2*>(p);
#pragma unroll
        for (int k = 0; k < N / 2; ++k) {
            const double2 v = q[k];
            r[2 * k] = v.x; r[2 * k + 1] = v.y;
        }
    }
}

// x stage of spline_interp_velo (lsgrid_spline.cpp:386-502) for the three
// fields of the node pair (lo at i, hi at i+1): Hermite value, and linear
// d/dy, d/dt, d/ds.
template <typename R>
__device__ __forceinline__ void xstage(R (&xr)[3], R (&yr)[3], R (&tr)[3], R (&sr)[3],
                                       const R* __restrict__ plo, const R* __restrict__ phi, R xd, R dx) {
    R a[16], b[16];
    load_rec(a, plo);
    load_rec(b, phi);
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        const int o = 5 * f;
        const R ac = dx * a[o + 2] - (b[o] - a[o]);
        const R bc = -dx * b[o + 2] + (b[o] - a[o]);
        xr[f] = herm(a[o], b[o], ac, bc, xd);
        yr[f] = lin(a[o + 3], b[o + 3], xd);
        tr[f] = lin(a[o + 1], b[o + 1], xd);
        sr[f] = lin(a[o + 4], b[o + 4], xd);
    }
}

// x, y and s stages of spline_interp_velo (lsgrid_spline.cpp:386-502) on one
// time slot's records for all three fields: value v and d/dt vt at (xd, yd, sd)
// inside the cell of columns col = {(i,j), (i,j+1), (i+1,j), (i+1,j+1)} between
// levels l and l+1, both edge clamped (lsgrid_spline.cpp:375-384).  Each record
// is read once.
template <typename R>
__device__ __forceinline__ void slot_velo(R (&v)[3], R (&vt)[3], const R* __restrict__ pk, const int64_t (&col)[4],
                                          int nl, int l, R xd, R yd, R sd, R dx, R dy, R ds) {
    const int lp = min(l + 1, nl - 1);
    R yv[2][3], sv[2][3], tv[2][3];
#pragma unroll
    for (int lv = 0; lv < 2; ++lv) {
        const int ll = lv ? lp : l;
        R x0[3], y0[3], t0[3], s0[3], x1[3], y1[3], t1[3], s1[3];
        xstage(x0, y0, t0, s0, pk + 16 * (col[0] * nl + ll), pk + 16 * (col[2] * nl + ll), xd, dx);
        xstage(x1, y1, t1, s1, pk + 16 * (col[1] * nl + ll), pk + 16 * (col[3] * nl + ll), xd, dx);
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            const R ac = dy * y0[f] - (x1[f] - x0[f]);
            const R bc = -dy * y1[f] + (x1[f] - x0[f]);
            yv[lv][f] = herm(x0[f], x1[f], ac, bc, yd);
            sv[lv][f] = lin(s0[f], s1[f], yd);
            tv[lv][f] = lin(t0[f], t1[f], yd);
        }
    }
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        const R a = ds * sv[0][f] - (yv[1][f] - yv[0][f]);
        const R b = -ds * sv[1][f] + (yv[1][f] - yv[0][f]);
        v[f] = herm(yv[0][f], yv[1][f], a, b, sd);
        vt[f] = lin(tv[0][f], tv[1][f], sd);
    }
}
}  // namespace

// get_kinematics_at_point (lsgrid_spline.cpp:505-605).  The cell fractions
// follow std::modf(double, float*): the coordinate is rounded to float and the
// fraction taken in float.  TQ: coordinate/output type of the queries.  With
// J.srec thread k evaluates the k-th query in cell order; p is always the
// query's index in the caller's arrays.
template <typename R, typename TQ>
__global__ __launch_bounds__(256) void k_query(QueryJob<R, TQ> J, QueryConst K) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= J.n) return;
    int64_t p = k;
    double xpt, ypt, zpt;
    if (J.srec) {
        const QRec<TQ> r = J.srec[k];
        p = (int64_t)r.p;
        xpt = (double)r.x; ypt = (double)r.y; zpt = (double)r.z;
    } else {
        xpt = (double)J.x[p]; ypt = (double)J.y[p]; zpt = (double)J.z[p];
    }
    const GridGeom& g = J.g;
    const double tpt = J.t ? (double)J.t[p] : J.tconst;
    constexpr bool F = sizeof(R) == 4;
    double cx = F ? (xpt - g.x0) * K.inv_dx : (xpt - g.x0) / g.dx;
    cx = cx < (g.nx - 1.) ? cx : (g.nx - 1.);
    cx = cx > 0. ? cx : 0.;
    double cy = F ? (ypt - g.y0) * K.inv_dy : (ypt - g.y0) / g.dy;
    cy = cy < (g.ny - 1.) ? cy : (g.ny - 1.);
    cy = cy > 0. ? cy : 0.;
    const float fx = (float)cx, fy = (float)cy;
    const float ix = truncf(fx), iy = truncf(fy);
    const R xd = (R)(fx - ix), yd = (R)(fy - iy);
    double td = F ? (tpt - J.t0) * K.inv_dt : (tpt - J.t0) / J.dt;
    td = td > 0. ? td : 0.;
    td = td < 1. ? td : 1.;
    const int i = (int)ix, j = (int)iy;
    const int nx = g.nx, ny = g.ny, nl = g.nl;
    const int j1 = min(j + 1, ny - 1), i1 = min(i + 1, nx - 1);
    const int64_t ac[4] = {(int64_t)i * ny + j, (int64_t)i * ny + j1, (int64_t)i1 * ny + j, (int64_t)i1 * ny + j1};
    R C[4], D[4], Ct[4], Dt[4], Cx[4], Dx[4], Cy[4], Dy[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        R e[8];
        load_rec(e, J.pk_eta + 8 * ac[q]);
        C[q] = e[0]; D[q] = e[1]; Ct[q] = e[2]; Dt[q] = e[3];
        Cx[q] = e[4]; Dx[q] = e[5]; Cy[q] = e[6]; Dy[q] = e[7];
    }
    const R dx = (R)g.dx, dy = (R)g.dy, dt = (R)J.dt;
    const R a00 = dx * Cx[0] - (C[2] - C[0]), a01 = dx * Cx[1] - (C[3] - C[1]);
    const R b00 = -dx * Cx[2] + (C[2] - C[0]), b01 = -dx * Cx[3] + (C[3] - C[1]);
    const R a10 = dx * Dx[0] - (D[2] - D[0]), a11 = dx * Dx[1] - (D[3] - D[1]);
    const R b10 = -dx * Dx[2] + (D[2] - D[0]), b11 = -dx * Dx[3] + (D[3] - D[1]);
    const R C0 = herm(C[0], C[2], a00, b00, xd), C1 = herm(C[1], C[3], a01, b01, xd);
    const R D0 = herm(D[0], D[2], a10, b10, xd), D1 = herm(D[1], D[3], a11, b11, xd);
    const R C0y = lin(Cy[0], Cy[2], xd), C1y = lin(Cy[1], Cy[3], xd);
    const R D0y = lin(Dy[0], Dy[2], xd), D1y = lin(Dy[1], Dy[3], xd);
    const R C0t = lin(Ct[0], Ct[2], xd), C1t = lin(Ct[1], Ct[3], xd);
    const R D0t = lin(Dt[0], Dt[2], xd), D1t = lin(Dt[1], Dt[3], xd);
    const R ya0 = dy * C0y - (C1 - C0), yb0 = -dy * C1y + (C1 - C0);
    const R ya1 = dy * D0y - (D1 - D0), yb1 = -dy * D1y + (D1 - D0);
    const R e0 = herm(C0, C1, ya0, yb0, yd), e1 = herm(D0, D1, ya1, yb1, yd);
    const R CCt = lin(C0t, C1t, yd), DDt = lin(D0t, D1t, yd);
    const R ta = dt * CCt - (e1 - e0), tb = -dt * DDt + (e1 - e0);
    const R tdR = (R)td;
    const R eta = herm(e0, e1, ta, tb, tdR);
    float f0, f1;
    if constexpr (F) {
        f0 = cart2sigmaS_f((float)zpt, (float)e0, K) * K.inv_ds;
        f1 = cart2sigmaS_f((float)zpt, (float)e1, K) * K.inv_ds;
    } else {
        const double sp0 = cart2sigmaS(zpt, (double)e0, g.h, g.swl, J.tan_a, J.tan_b, J.tanab);
        const double sp1 = cart2sigmaS(zpt, (double)e1, g.h, g.swl, J.tan_a, J.tan_b, J.tanab);
        f0 = (float)(sp0 / g.ds);
        f1 = (float)(sp1 / g.ds);
    }
    const float k0 = truncf(f0), k1 = truncf(f1);
    const R sd0 = (R)(f0 - k0), sd1 = (R)(f1 - k1);
    int l0 = (int)k0, l1 = (int)k1;
    l0 = l0 < 0 ? 0 : (l0 > nl - 1 ? nl - 1 : l0);
    l1 = l1 < 0 ? 0 : (l1 > nl - 1 ? nl - 1 : l1);
    const R ds = (R)g.ds;
    const double rmp = J.ramp_scale ? (double)J.ramp_scale[p] : 1.0;
    const TQ oeta = (TQ)((double)eta * rmp);
    R VC[3], VCt[3], VD[3], VDt[3];
    slot_velo(VC, VCt, J.pk_C, ac, nl, l0, xd, yd, sd0, dx, dy, ds);
    slot_velo(VD, VDt, J.pk_D, ac, nl, l1, xd, yd, sd1, dx, dy, ds);
    TQ ov[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        const R a = dt * VCt[f] - (VD[f] - VC[f]), b = -dt * VDt[f] + (VD[f] - VC[f]);
        ov[f] = (TQ)((double)herm(VC[f], VD[f], a, b, tdR) * rmp);
    }
    if (J.o_aos) {
        // {eta, u, v, w} of the query in one vector store (16 B FP32, 2 x 16 B FP64)
        if constexpr (sizeof(TQ) == 4) {
            *reinterpret_cast<float4*>(J.o_aos + 4 * p) = make_float4(oeta, ov[0], ov[1], ov[2]);
        } else {
            double2* o = reinterpret_cast<double2*>(J.o_aos + 4 * p);
            o[0] = make_double2(oeta, ov[0]);
            o[1] = make_double2(ov[1], ov[2]);
        }
        return;
    }
    J.o_eta[p] = oeta;
    J.o_u[p] = ov[0];
    J.o_v[p] = ov[1];
    J.o_w[p] = ov[2];
}

// ===================================================== device point batches
namespace {
// RampCfg::r1d and operator() of the host entries, same operations and order
__device__ __forceinline__ double ramp_axis(double a, double xs, double xe, bool inv) {
    const double v = inv ? 1. - ((a - xs) / (xe - xs)) : ((a - xs) / (xe - xs));
    const double m = (1. < v) ? 1. : v;               // std::min(v, 1.)
    return (m < 0.) ? 0. : m;                         // std::max(., 0.)
}
__device__ __forceinline__ double ramp_factor(const RampDev& R, double t, double x, double y) {
    if (!R.init) return 1.0;
    const double a[6] = {t, t, x, x, y, y};
    double m = 1.0;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const double r = R.on[q] ? ramp_axis(a[q], R.se[2 * q], R.se[2 * q + 1], (q & 1) != 0) : 1.0;
        if (q == 0 || r < m) m = r;                   // std::min_element
    }
    return m;
}

__global__ __launch_bounds__(256) void k_points_finish(PointsFinish F) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= F.n) return;
    const double r = ramp_factor(F.ramp, F.t[q], F.x[q], F.y[q]);
    if (F.eta) {
        double e = F.ea[q];
        if (F.eadd) e = (F.eoff + e) + (F.eb ? F.eb[q] : 0.0);
        F.eta[q] = e * r;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c)
        if (F.dst[c]) F.dst[c][q] = F.src[c] ? F.src[c][q] * r : 0.0;
}

__global__ __launch_bounds__(256) void k_clamp_below(int64_t n, const double* __restrict__ z, double lo,
                                                     double* __restrict__ out) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const double v = z[q];
    out[q] = (lo < v) ? v : lo;                       // std::max(lo, v)
}
}  // namespace

// ==================================================== profile-term pruning
// eta range (minus swl) over the columns written by one fill
template <typename R>
__global__ __launch_bounds__(1024) void k_eta_bounds(const R* __restrict__ eta, int64_t ncell,
                                                     const int32_t* __restrict__ cols, int64_t ncol,
                                                     int nslot, int4 dst, double swl, double* __restrict__ eb) {
    __shared__ double smin[1024], smax[1024];
    double mn = 1e300, mx = -1e300;
    const int64_t total = (int64_t)nslot * ncol;
    for (int64_t q = threadIdx.x; q < total; q += blockDim.x) {
        const int s = (int)(q / ncol);
        const int64_t c = cols ? cols[q % ncol] : q % ncol;
        const int d = s == 0 ? dst.x : (s == 1 ? dst.y : (s == 2 ? dst.z : dst.w));
        const double v = (double)eta[(int64_t)d * ncell + c];
        mn = fmin(mn, v);
        mx = fmax(mx, v);
    }
    smin[threadIdx.x] = mn;
    smax[threadIdx.x] = mx;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + w]);
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        eb[0] = smin[0] - swl;
        eb[1] = smax[0] - swl;
    }
}

// Level bounds per pair record.  Profile z of level l is min(0, e + sigma_l (h + e)),
// increasing in e and in l, so over the fill's eta range:
//   e^{k z} can exceed e^{-T} only for l >= lE (first such level),
//   e^{-k(z+2h)} only for l < lG1.
// Exponents are in the units of prof() (log2 in FP32), T likewise.
template <typename R>
__global__ __launch_bounds__(256) void k_prune(const UvwRec<R>* __restrict__ rec, int64_t nrec,
                                               const double* __restrict__ eb, const double* __restrict__ sigma,
                                               int nl, double h, double T, uint32_t* __restrict__ lvl) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nrec) return;
    const double emin = eb[0], emax = eb[1];
    const UvwRec<R> r = rec[e];
    uint32_t out = 0;
    for (int b = 0; b < 2; ++b) {
        const double pk = (double)(b ? r.pkM : r.pkP), pg = (double)(b ? r.pgM : r.pgP);
        const bool present = b ? (r.cuM != 0 || r.cvM != 0 || r.cwM != 0) : (r.cuP != 0 || r.cvP != 0 || r.cwP != 0);
        int lE = 255, lG1 = 0;
        if (present) {
            lE = nl;
            for (int l = 0; l < nl; ++l) {
                const double zx = emax + sigma[l] * (h + emax), zn = emin + sigma[l] * (h + emin);
                const double zmx = zx < 0. ? zx : 0., zmn = zn < 0. ? zn : 0.;
                if (lE == nl && pk * zmx >= -T) lE = l;
                if (fma(-pk, zmn, pg) >= -T) lG1 = l + 1;
            }
            if (lE == nl) lE = 255;
        }
        out |= ((uint32_t)lE | ((uint32_t)lG1 << 8)) << (16 * b);
    }
    lvl[e] = out;
}

// ================================================================ launchers
namespace {
inline unsigned nblk(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }
}

int64_t column_threads(int64_t nslot, int64_t ncol) { return nslot * ncol; }
int64_t level_threads(int64_t nslot, int64_t ncol, int nl, int lc, bool grid) {
    if (!grid) return nslot * ncol;
    return (nslot * ncol + 63) / 64 * 64 * ((nl + lc - 1) / lc);
}

template <typename R>
hipError_t launch_columns(const SeaDev<R>& S, const ColumnJob<R>& J, bool grid, hipStream_t st) {
    const int64_t n = J.nslot * J.ncol;
    if (n <= 0) return hipSuccess;
    const dim3 g1(nblk(n), J.nseg);
    if (grid) hipLaunchKernelGGL((k_columns<R, true>), g1, dim3(256), 0, st, S, J);
    else hipLaunchKernelGGL((k_columns<R, false>), g1, dim3(256), 0, st, S, J);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess || !J.partial) return e;
    if (grid) hipLaunchKernelGGL((k_columns_reduce<R, true>), dim3(nblk(n)), dim3(256), 0, st, S, J);
    else hipLaunchKernelGGL((k_columns_reduce<R, false>), dim3(nblk(n)), dim3(256), 0, st, S, J);
    return hipGetLastError();
}

template <typename R, int LC, int MODE>
static hipError_t launch_levels_t(const SeaDev<R>& S, const LevelJob<R>& J, hipStream_t st) {
    const int64_t n = level_threads(J.nslot, J.ncol, J.g.nl, LC, MODE == 0);
    if (J.nslot * J.ncol <= 0) return hipSuccess;
    if constexpr (MODE == 0 && sizeof(R) == 4 && (LC == 8 || LC == 16)) {
        if (J.qf && J.lin_c && LC == 16) {
            hipLaunchKernelGGL((k_levels_grid32<LC, 2>), dim3(nblk(n), J.nseg), dim3(256), 0, st, S, J);
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess || !J.partial) return e;
            hipLaunchKernelGGL((k_levels_reduce_lb<LC>), dim3(nblk(n)), dim3(256), 0, st, S, J);
            return hipGetLastError();
        } else if (J.qf)
            hipLaunchKernelGGL((k_levels_grid32<LC, 1>), dim3(nblk(n), J.nseg), dim3(256), 0, st, S, J);
        else
            hipLaunchKernelGGL((k_levels_grid32<LC, 0>), dim3(nblk(n), J.nseg), dim3(256), 0, st, S, J);
    } else {
        hipLaunchKernelGGL((k_levels<R, LC, MODE>), dim3(nblk(n), J.nseg), dim3(256), 0, st, S, J);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess || !J.partial) return e;
    hipLaunchKernelGGL((k_levels_reduce<R, LC, MODE>), dim3(nblk(n)), dim3(256), 0, st, S, J);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_levels(const SeaDev<R>& S, const LevelJob<R>& J, bool grid, int lc, hipStream_t st) {
    if (!grid) return (J.flags & EV_P) ? launch_levels_t<R, 1, 2>(S, J, st) : launch_levels_t<R, 1, 1>(S, J, st);
    switch (lc) {
    case 1: return launch_levels_t<R, 1, 0>(S, J, st);
    case 2: return launch_levels_t<R, 2, 0>(S, J, st);
    case 4: return launch_levels_t<R, 4, 0>(S, J, st);
    case 8: return launch_levels_t<R, 8, 0>(S, J, st);
    default: return launch_levels_t<R, 16, 0>(S, J, st);
    }
}

template <typename R>
hipError_t launch_grad_xyz(const R* D, R* GX, R* GY, R* GS, int nx, int ny, int L, int tid0, int tid1,
                           double dx, double dy, double ds, hipStream_t st) {
    const int64_t n = 2LL * nx * ny * L;
    hipLaunchKernelGGL((k_grad_xyz<R>), dim3(nblk(n)), dim3(256), 0, st, D, GX, GY, GS, nx, ny, L,
                       tid0, tid1, dx, dy, ds);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_grad_dt(const R* D, R* GD, int64_t per, int tstep, double dt, hipStream_t st) {
    hipLaunchKernelGGL((k_grad_dt<R>), dim3(nblk(per)), dim3(256), 0, st, D, GD, per, tstep, dt);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_pack(const PackJob<R>& J, hipStream_t st) {
    const int64_t n = (int64_t)J.nx * J.ny * J.nl;
    hipLaunchKernelGGL((k_pack<R>), dim3(nblk(n)), dim3(256), 0, st, J);
    return hipGetLastError();
}
template hipError_t launch_pack<double>(const PackJob<double>&, hipStream_t);
template hipError_t launch_pack<float>(const PackJob<float>&, hipStream_t);

hipError_t launch_points_finish(const PointsFinish& F, hipStream_t st) {
    if (F.n <= 0) return hipSuccess;
    if (!F.t || !F.x || !F.y || (F.eta && !F.ea)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_points_finish, dim3(nblk(F.n)), dim3(256), 0, st, F);
    return hipGetLastError();
}

hipError_t launch_clamp_below(int64_t n, const double* z, double lo, double* out, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_clamp_below, dim3(nblk(n)), dim3(256), 0, st, n, z, lo, out);
    return hipGetLastError();
}

template <typename R, typename TQ>
hipError_t launch_query(const QueryJob<R, TQ>& J, hipStream_t st) {
    if (J.n <= 0) return hipSuccess;
    QueryConst K;
    K.inv_dx = 1.0 / J.g.dx; K.inv_dy = 1.0 / J.g.dy; K.inv_dt = 1.0 / J.dt;
    K.h = (float)J.g.h; K.swl = (float)J.g.swl; K.tanab = (float)J.tanab;
    K.ita = (float)(1.0 / J.tan_a); K.itb = (float)(1.0 / J.tan_b); K.inv_ds = (float)(1.0 / J.g.ds);
    hipLaunchKernelGGL((k_query<R, TQ>), dim3(nblk(J.n)), dim3(256), 0, st, J, K);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_prune(const R* eta, int64_t ncell, const int32_t* cols, int64_t ncol, int nslot,
                        const int* slot_dst, double swl, double* eb, const UvwRec<R>* rec, int64_t nrec,
                        const double* sigma, int nl, double h, uint32_t* lvl, hipStream_t st) {
    const int4 dst = make_int4(slot_dst[0], nslot > 1 ? slot_dst[1] : 0, nslot > 2 ? slot_dst[2] : 0,
                               nslot > 3 ? slot_dst[3] : 0);
    hipLaunchKernelGGL((k_eta_bounds<R>), dim3(1), dim3(1024), 0, st, eta, ncell, cols, ncol, nslot, dst, swl, eb);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (nrec <= 0) return hipSuccess;
    const double T = sizeof(R) == 4 ? 30.0 * 1.4426950408889634 : 40.0;
    hipLaunchKernelGGL((k_prune<R>), dim3(nblk(nrec)), dim3(256), 0, st, rec, nrec, eb, sigma, nl, h, T, lvl);
    return hipGetLastError();
}
template hipError_t launch_prune<double>(const double*, int64_t, const int32_t*, int64_t, int, const int*, double, double*,
                                         const UvwRec<double>*, int64_t, const double*, int, double, uint32_t*, hipStream_t);
template hipError_t launch_prune<float>(const float*, int64_t, const int32_t*, int64_t, int, const int*, double, double*,
                                        const UvwRec<float>*, int64_t, const double*, int, double, uint32_t*, hipStream_t);

template hipError_t launch_columns<double>(const SeaDev<double>&, const ColumnJob<double>&, bool, hipStream_t);
template hipError_t launch_columns<float>(const SeaDev<float>&, const ColumnJob<float>&, bool, hipStream_t);
template hipError_t launch_levels<double>(const SeaDev<double>&, const LevelJob<double>&, bool, int, hipStream_t);
template hipError_t launch_levels<float>(const SeaDev<float>&, const LevelJob<float>&, bool, int, hipStream_t);
template hipError_t launch_grad_xyz<double>(const double*, double*, double*, double*, int, int, int, int, int, double, double, double, hipStream_t);
template hipError_t launch_grad_xyz<float>(const float*, float*, float*, float*, int, int, int, int, int, double, double, double, hipStream_t);
template hipError_t launch_grad_dt<double>(const double*, double*, int64_t, int, double, hipStream_t);
template hipError_t launch_grad_dt<float>(const float*, float*, int64_t, int, double, hipStream_t);
template hipError_t launch_query<double, double>(const QueryJob<double, double>&, hipStream_t);
template hipError_t launch_query<float, double>(const QueryJob<float, double>&, hipStream_t);
template hipError_t launch_query<double, float>(const QueryJob<double, float>&, hipStream_t);
template hipError_t launch_query<float, float>(const QueryJob<float, float>&, hipStream_t);

}  // namespace cfdwm
