// Smit et al. (2017) s-coordinate second-order kinematics at points: the
// members u_smit, v_smit, w_smit, dp_smit of src/Irregular.cpp:1623-2027
// that Irregular::u/v/w/dp dispatch to when vertical_theory is "smit"
// (:25, :53, :77, :143).
//
// Per point, with eta the Sharma-Dean eta1 + eta2 (swl excluded) and
// zz = max(-h, z - swl), every depth profile is evaluated at the stretched
// argument sf = h (h + zz) / (h + eta) (Wheeler stretching, :118-141):
//   X(k) = cosh(k sf)/cosh(k h) = C_k (e^{k (sf-h)} + e^{-k (sf+h)}),
//   Z(k) = sinh(k sf)/cosh(k h) = C_k (e^{k (sf-h)} - e^{-k (sf+h)}),
// both finite for every depth (sf <= h below the surface).  The diagonal's
// X(2k), Z(2k) reuse the squares of the two exponentials of X(k).
//
// Two kernels: k_smit_pairs sums the pair terms over a balanced segment of
// the filtered pair table (blockIdx.y) into fixed-order partials;
// k_smit_final adds the first-order, diagonal and (u) chain terms, the
// partials in segment order, and writes u, v, w, p.  Points above the
// instantaneous surface get 0, as in the reference.
#include <hip/hip_runtime.h>

#include "device_math.hpp"
#include "kernels.hpp"

namespace cfdwm {

namespace {

using dm::cs_of;
using dm::freq_ramp;
using dm::phase_of;

__device__ __forceinline__ double sexp(double a) { return exp(a); }
__device__ __forceinline__ float sexp(float a) { return __builtin_amdgcn_exp2f(a * 1.4426950408889634f); }

// Point state shared by the two kernels.
template <typename R>
struct SPoint {
    double x, y, t, dx, dy;
    R sf, h, sp1, g1s;          // stretched depth, h, (1+s), g (1+s)
    double eta, zz;
    bool live, wet;
};

template <typename R>
__device__ __forceinline__ SPoint<R> smit_point(const SmitJob& J, int64_t p) {
    SPoint<R> P{};
    P.live = p < J.n;
    if (!P.live) return P;
    P.x = J.x[p]; P.y = J.y[p]; P.t = J.t[p];
    P.dx = P.x - J.x0; P.dy = P.y - J.y0;
    const double z = J.z[p];
    P.eta = J.eta1[p] + (J.order >= 1 ? J.eta2[p] : 0.0);
    P.wet = !(z > P.eta + J.swl);                          // Irregular.cpp:1627
    const double h = J.h;
    P.zz = fmax(-h, z - J.swl);
    const double sf = h * (h + P.zz) / (h + P.eta);
    const double sp1 = (h + P.zz) / (h + P.eta);
    P.sf = (R)sf; P.h = (R)h; P.sp1 = (R)sp1; P.g1s = (R)(J.g * sp1);
    return P;
}

// e^{k (sf-h)}, e^{-k (sf+h)}
template <typename R>
__device__ __forceinline__ void sexps(R k, const SPoint<R>& P, R& ea, R& eb) {
    ea = sexp(k * (P.sf - P.h));
    eb = sexp(-k * (P.sf + P.h));
}

// ------------------------------------------------------------- pair segment
template <typename R>
__global__ __launch_bounds__(256) void k_smit_pairs(SmitJob J) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const Segment S = J.seg[blockIdx.y];
    const SPoint<R> P = smit_point<R>(J, p);
    const bool act = P.live && P.wet && J.order >= 1;
    const R g = (R)J.g, rho = (R)J.rho, h = P.h;
    double au = 0, av = 0, aw = 0, ap = 0, ad = 0;
    if (act) {
        for (int i = S.rlo; i < S.rhi; ++i) {                // wave-uniform
            const int mb = i == S.rlo ? S.mstart : i + 1;
            const int me = i == S.rhi - 1 ? S.mend : J.fmhi[i];
            if (mb >= me) continue;
            const PhaseRec pi = J.ph[i];
            const SmitComp ci = J.sc[i];
            R c_i, s_i;
            cs_of(phase_of(pi, P.dx, P.dy, P.t), c_i, s_i);
            const R fr_i = (R)freq_ramp(P.t, pi.rt, pi.ts, pi.du);
            const R ki = (R)pi.k, oi = (R)pi.omega, cti = (R)pi.cth, sti = (R)pi.sth;
            R ea, eb;
            sexps(ki, P, ea, eb);
            const R Xi = (R)ci.Ck * (ea + eb), Zi = (R)ci.Ck * (ea - eb);
            const R ai = ki * Zi / oi, bi = ki * (ki * h * Xi) / oi;
            const R Ai = (R)ci.A;
            R ru = 0, rv = 0, rw = 0, rp = 0, rd = 0;
            const int64_t base = J.frow[i] - (i + 1);
            for (int m = mb; m < me; ++m) {
                const SmitPair q = J.sp[base + m];
                if (q.fP == 0.0 && q.fM == 0.0) continue;    // wave-uniform: filtered pair
                const PhaseRec pm = J.ph[m];
                R c_m, s_m;
                cs_of(phase_of(pm, P.dx, P.dy, P.t), c_m, s_m);
                const R fr_m = (R)freq_ramp(P.t, pm.rt, pm.ts, pm.du);
                const R km = (R)pm.k, om = (R)pm.omega, ctm = (R)pm.cth, stm = (R)pm.sth;
                R fa, fb;
                sexps(km, P, fa, fb);
                const R Ckm = (R)J.sc[m].Ck;
                const R Xm = Ckm * (fa + fb), Zm = Ckm * (fa - fb);
                const R am = km * Zm / om, bm = km * (km * h * Xm) / om;
                const R AA = Ai * (R)J.sc[m].A * fr_i * fr_m;
                const R k1dk2 = (R)q.k1dk2;
                if (q.fP != 0.0) {                            // sum interaction, :1718-1733
                    const R c = c_i * c_m - s_i * s_m, s = s_i * c_m + c_i * s_m;
                    const R kk = (R)q.kp, H = (R)q.Hp;
                    R e1, e2;
                    sexps(kk, P, e1, e2);
                    const R X12 = (R)q.CkP * (e1 + e2), Z12 = (R)q.CkP * (e1 - e2);
                    const R kx = ki * cti + km * ctm, ky = ki * sti + km * stm;
                    const R a = ai + am;
                    ru += AA * H * kx * (-c) * X12 + AA * kx * P.g1s * a * c;
                    rd += -AA * H * (kk * h * Z12) * s + AA * g * (a + P.sp1 * (bi + bm)) * s;
                    rv += AA * H * ky * (-c) * X12;
                    rw += AA * H * kk * (-s) * Z12;
                    rp += AA * rho * (oi + om) * H * (-c) * X12 - AA * rho * g * g * k1dk2 * (Xi * Xm - Zi * Zm) / (oi * om) * c;
                }
                if (q.fM != 0.0) {                            // difference interaction, :1736-1747
                    const R c = c_i * c_m + s_i * s_m, s = s_i * c_m - c_i * s_m;
                    const R kk = (R)q.km, H = (R)q.Hm;
                    R e1, e2;
                    sexps(kk, P, e1, e2);
                    const R X12 = (R)q.CkM * (e1 + e2), Z12 = (R)q.CkM * (e1 - e2);
                    const R kx = ki * cti - km * ctm, ky = ki * sti - km * stm;
                    const R a = ai - am;
                    ru += AA * H * kx * (-c) * X12 + AA * kx * P.g1s * a * c;
                    rd += -AA * H * (kk * h * Z12) * s + AA * g * (a + P.sp1 * (bi - bm)) * s;
                    rv += AA * H * ky * (-c) * X12;
                    rw += AA * H * kk * (-s) * Z12;
                    rp += AA * rho * (oi - om) * H * (-c) * X12 - AA * rho * g * g * k1dk2 * (Xi * Xm + Zi * Zm) / (oi * om) * c;
                }
            }
            au += (double)ru; av += (double)rv; aw += (double)rw; ap += (double)rp; ad += (double)rd;
        }
    }
    if (P.live) {
        double* o = J.part + (int64_t)blockIdx.y * 5 * J.n + p;
        o[0] = au; o[J.n] = av; o[2 * J.n] = aw; o[3 * J.n] = ap; o[4 * J.n] = ad;
    }
}

// ------------------------------------------------------- component terms
template <typename R>
__global__ __launch_bounds__(256) void k_smit_final(SmitJob J) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const SPoint<R> P = smit_point<R>(J, p);
    if (!P.live) return;
    double u = 0, v = 0, w = 0, pr = 0;
    if (P.wet) {
        const R g = (R)J.g, rho = (R)J.rho, h = P.h;
        const bool sec = J.order >= 1;
        double u1 = 0, v1 = 0, w1 = 0, p1 = 0, u2 = 0, v2 = 0, w2 = 0, p2 = 0, dd = 0, ex = 0, ps = 0;
        for (int i = 0; i < J.N; ++i) {                      // wave-uniform
            const PhaseRec pi = J.ph[i];
            const SmitComp ci = J.sc[i];
            R c, s;
            cs_of(phase_of(pi, P.dx, P.dy, P.t), c, s);
            const R f = (R)freq_ramp(P.t, pi.rt, pi.ts, pi.du);
            const R k = (R)pi.k, om = (R)pi.omega, ct = (R)pi.cth, st = (R)pi.sth, A = (R)ci.A;
            R ea, eb;
            sexps(k, P, ea, eb);
            const R X = (R)ci.Ck * (ea + eb), Z = (R)ci.Ck * (ea - eb);
            const R agk = A * g * k / om;
            // first order, :1633-1640 (u), :1781-1788 (v), :1868-1875 (w), :1941-1948 (dp)
            u1 += (double)(agk * X * ct * c * f);
            v1 += (double)(agk * X * st * c * f);
            w1 += (double)(agk * Z * s * f);
            p1 += (double)(A * rho * g * X * c * f);
            if (!sec) continue;
            // diagonal, :1657-1689, :1790-1800, :1877-1891, :1950-1968
            const R AA = A * A * f * f, H = (R)ci.Hd;
            const R c2 = c * c - s * s, s2 = (R)2 * s * c;
            const R X2 = (R)ci.C2k * (ea * ea + eb * eb), Z2 = (R)ci.C2k * (ea * ea - eb * eb);
            const R k2 = (R)2 * k, kx2 = k2 * ct, ky2 = k2 * st;
            const R a = k * Z / om;
            u2 += (double)((R)0.5 * AA * H * kx2 * (-c2) * X2 + (R)0.5 * AA * kx2 * P.g1s * a * c2);
            dd += (double)((R)-0.5 * AA * H * (k2 * h * Z2) * s2 +
                           (R)0.5 * AA * g * (a + P.sp1 * k * (k * h * X) / om) * s2);
            v2 += (double)((R)0.5 * AA * H * ky2 * (-c2) * X2);
            w2 += (double)((R)0.5 * AA * H * k2 * (-s2) * Z2);
            p2 += (double)((R)0.5 * AA * rho * ((R)2 * om) * H * (-c2) * X2 -
                           (R)0.5 * AA * rho * g * g * k * k * (X * X - Z * Z) / (om * om) * c2);
            // chain-term sums, :1752-1769
            ex += (double)(-A * k * ct * s * f);
            ps += (double)((g * A / om) * k * h * Z * s * f);
        }
        if (sec) {
            for (int sg = 0; sg < J.nseg; ++sg) {            // partials in segment order
                const double* o = J.part + (int64_t)sg * 5 * J.n + p;
                u2 += o[0]; v2 += o[J.n]; w2 += o[2 * J.n]; p2 += o[3 * J.n]; dd += o[4 * J.n];
            }
            const double hh = J.h;
            const double osd = (hh + P.zz) / ((hh + P.eta) * (hh + P.eta));
            u2 += -osd * ex * (ps + dd);
        }
        u = u1 + u2; v = v1 + v2; w = w1 + w2; pr = p1 + p2;
    }
    if (J.o_u) J.o_u[p] = u;
    if (J.o_v) J.o_v[p] = v;
    if (J.o_w) J.o_w[p] = w;
    if (J.o_p) J.o_p[p] = pr;
}

}  // namespace

template <typename R>
hipError_t launch_smit(const SmitJob& J, hipStream_t st) {
    if (J.n <= 0) return hipSuccess;
    if (J.nseg < 1 || !J.seg || !J.part) return hipErrorInvalidValue;
    const unsigned nb = (unsigned)((J.n + 255) / 256);
    hipLaunchKernelGGL(k_smit_pairs<R>, dim3(nb, (unsigned)J.nseg), dim3(256), 0, st, J);
    hipLaunchKernelGGL(k_smit_final<R>, dim3(nb), dim3(256), 0, st, J);
    return hipGetLastError();
}

template hipError_t launch_smit<double>(const SmitJob&, hipStream_t);
template hipError_t launch_smit<float>(const SmitJob&, hipStream_t);

}  // namespace cfdwm
