// Host-side coefficient tables for the second-order kinematics kernels.
// See sea_tables.hpp for the layout; reference expressions are cited inline.
#include "sea_tables.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <thread>

namespace cfdwm {

namespace {

double freq_default(const std::vector<double>& v, size_t i, double d) {
    return v.empty() ? d : v[i];
}

// Sharma & Dean D+- of the pair cache, Irregular.cpp:903-938 (cached Rs, R).
void cached_D(double ki, double km, double Rni, double Rnm, double Rsni, double Rsm,
              double gamma, double depth, double& kplus, double& kminus,
              double& Dp, double& Dm) {
    const double ki2 = ki * ki, km2 = km * km, kikm = ki * km;
    kplus = std::sqrt(ki2 + km2 + 2.0 * kikm * gamma);
    kminus = std::sqrt(ki2 + km2 - 2.0 * kikm * gamma);
    const double sp = Rsni + Rsm, sm = Rsni - Rsm;
    const double R2p = sp * sp, R2m = sm * sm;
    const double denp = R2p - kplus * std::tanh(kplus * depth);
    Dp = (2.0 * R2p * (kikm * gamma - Rni * Rnm)) / denp +
         (sp * (Rsni * (km2 - Rnm * Rnm) + Rsm * (ki2 - Rni * Rni))) / denp;
    const double denm = R2m - kminus * std::tanh(kminus * depth);
    Dm = (2.0 * R2m * (kikm * gamma + Rni * Rnm)) / denm +
         (sm * (-Rsni * (km2 - Rnm * Rnm) + Rsm * (ki2 - Rni * Rni))) / denm;
}

// D+- as recomputed per pair by u2/v2/w2/dp2, Irregular.cpp:466-486.
void direct_D(double ki, double km, double thi, double thm, double depth,
              double Rn, double Rsn, double& kp, double& kn, double& Dp, double& Dm) {
    const double gamma = std::cos(thi - thm);
    kp = std::sqrt(ki * ki + km * km + (2. * ki * km * gamma));
    kn = std::sqrt(ki * ki + km * km - (2. * ki * km * gamma));
    const double Rm = km * std::tanh(km * depth);
    const double Rsm = std::sqrt(Rm);
    const double R2p = std::pow(Rsn + Rsm, 2.);
    const double R2m = std::pow(Rsn - Rsm, 2.);
    Dp = (2. * R2p * (ki * km * gamma - Rn * Rm)) / (R2p - kp * std::tanh(kp * depth)) +
         ((Rsn + Rsm) * (Rsn * (km * km - Rm * Rm) + Rsm * (ki * ki - Rn * Rn))) /
             (R2p - kp * std::tanh(kp * depth));
    Dm = (2. * R2m * (ki * km * gamma + Rn * Rm)) / (R2m - kn * std::tanh(kn * depth)) +
         ((Rsn - Rsm) * (-Rsn * (km * km - Rm * Rm) + Rsm * (ki * ki - Rn * Rn))) /
             (R2m - kn * std::tanh(kn * depth));
}

// Runs body(i) for i in [0, n) on the host cores (rows have uneven cost:
// rows are handed out one at a time).
template <typename F>
void parallel_rows(int n, F body) {
    unsigned nt = std::thread::hardware_concurrency();
    nt = std::max(1u, std::min(nt, 16u));
    if (n < 64) nt = 1;
    std::atomic<int> next(0);
    auto work = [&]() {
        for (int i = next++; i < n; i = next++) body(i);
    };
    std::vector<std::thread> th;
    for (unsigned q = 1; q < nt; ++q) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
}

// 1 / (1 + exp(-2 k h)): folds cosh(k h) out of cosh(k (z+h)) / cosh(k h).
inline double cfac(double k, double h) { return 1.0 / (1.0 + std::exp(-2.0 * k * h)); }

}  // namespace

std::string build_tables(const SeaInput& in, SeaTables& out, bool with_direct) {
    const int N = (int)in.omega.size();
    if (N <= 0) return "no wave components";
    if ((int)in.A.size() != N || (int)in.k.size() != N || (int)in.phase.size() != N ||
        (int)in.theta.size() != N)
        return "component arrays differ in length";
    for (int i = 0; i < N; ++i)
        if (!(in.omega[i] > 0.0) || !(in.k[i] > 0.0))
            return "omega and k must be positive";
    out.in = in;
    const double h = in.depth, g = in.g, rho = in.rho;
    const double mt = in.mtheta * kRefPi / 180.0;
    CompTable& c = out.comp;
    c.n = N;
    auto R = [&](std::vector<double>& v) { v.assign(N, 0.0); };
    R(c.k); R(c.cth); R(c.sth); R(c.omega); R(c.ph0); R(c.rt); R(c.ts); R(c.du); R(c.Ck);
    R(c.lu); R(c.lv); R(c.lw); R(c.lp); R(c.gxz); R(c.gyz); R(c.gzz); R(c.gpz); R(c.A); R(c.e2d);
    R(c.C2k); R(c.dfu); R(c.dfv); R(c.dfw); R(c.ddu); R(c.ddv); R(c.ddw); R(c.ddp);
    R(c.Xx); R(c.Xy); R(c.Zk); R(c.kd1); R(c.kd2);
    c.rowActive.assign(N, 0);
    double kmax = 0.0;                                        // Irregular.cpp:834-836
    for (int i = 0; i < N; ++i) if (in.k[i] > kmax) kmax = in.k[i];
    out.kmax = kmax;
    std::vector<double> Rn(N), Rs(N), ag(N);
    out.diag_filtered = 0;
    for (int i = 0; i < N; ++i) {
        const double k = in.k[i], w = in.omega[i], A = in.A[i];
        const double th = in.theta[i] + mt;
        c.k[i] = k;
        c.cth[i] = std::cos(th);
        c.sth[i] = std::sin(th);
        c.omega[i] = w;
        c.ph0[i] = w * in.tofmax + in.phase[i];
        c.rt[i] = freq_default(in.ramptime, i, 0.0);
        c.ts[i] = freq_default(in.starttime, i, -kSmall);
        c.du[i] = freq_default(in.duration, i, kHuge);
        const double Ck = cfac(k, h), C2k = cfac(2.0 * k, h);
        c.Ck[i] = Ck;
        c.C2k[i] = C2k;
        const double agk = A * g * k / w;
        c.lu[i] = agk * c.cth[i] * Ck;
        c.lv[i] = agk * c.sth[i] * Ck;
        c.lw[i] = agk * Ck;
        c.lp[i] = A * rho * g * Ck;
        c.gxz[i] = c.cth[i] * A * w * k;
        c.gyz[i] = c.sth[i] * A * w * k;
        c.gzz[i] = A * w * k * (std::cosh(k * h) / std::sinh(k * h));
        c.gpz[i] = A * rho * g * k;
        c.A[i] = A;
        // component cache quantities (Irregular.cpp:822-830)
        Rn[i] = k * std::tanh(k * h);
        Rs[i] = std::sqrt(Rn[i]);
        ag[i] = A * g / w;
        const double Dnn = (6.0 * Rs[i] * Rs[i] * (k * k - Rn[i] * Rn[i])) /
                           (2.0 * Rs[i] * Rs[i] - k * std::tanh(2.0 * k * h));
        // filtered diagonal: uvw2 (Irregular.cpp:1008-1048), eta2 (:249-269)
        if (w <= in.cutoff) {
            c.rowActive[i] = 1;
            ++out.diag_filtered;
            const double AAd = 0.25 * ag[i] * ag[i] * (Dnn / (2.0 * w));
            c.dfu[i] = AAd * (2.0 * k * c.cth[i]) * C2k;
            c.dfv[i] = AAd * (2.0 * k * c.sth[i]) * C2k;
            c.dfw[i] = AAd * (2.0 * k) * C2k;
            const double ann = (Dnn - (k * k - Rn[i] * Rn[i])) / Rn[i] + 2.0 * Rn[i];
            if (2.0 * k <= kmax) c.e2d[i] = 0.25 * A * A * ann;
        }
        // unfiltered diagonal: u2/v2/w2/dp2 (Irregular.cpp:458-464 and peers)
        const double AAu = 0.25 * (A * g / w) * (A * g / w) * (Dnn / (2 * w));
        c.ddu[i] = AAu * (2 * k * c.cth[i]) * C2k;
        c.ddv[i] = AAu * (2 * k * c.sth[i]) * C2k;
        c.ddw[i] = AAu * 2 * k * C2k;
        c.ddp[i] = AAu * (2 * w) * C2k;
        // dp2 kinetic terms (Irregular.cpp:755-756, 786-787), profiles with C_k folded in
        c.Xx[i] = g * k * c.cth[i] / w * Ck;
        c.Xy[i] = g * k * c.sth[i] / w * Ck;
        c.Zk[i] = g * k / w * Ck;
        c.kd1[i] = (g * g * (2 * k * c.cth[i] + 2 * k * c.sth[i]) * 2) / (2 * w) * Ck;
        c.kd2[i] = (g * g * (2 * k) * 2) / (2 * w) * Ck;
    }

    // ---- filtered pair table (Irregular::rebuild_pair_cache, :840-967)
    {
        PairTable& P = out.filt;
        P.mhi.assign(N, 0);
        P.rowoff.assign(N + 1, 0);
        for (int i = 0; i < N; ++i) {
            int hi = i + 1;
            if (in.omega[i] <= in.cutoff) {
                for (int m = i + 1; m < N; ++m) {
                    if (in.omega[m] > in.cutoff) continue;
                    if (std::abs(in.omega[i] - in.omega[m]) > in.bandwidth) continue;
                    const bool hp = (in.k[i] + in.k[m]) <= kmax;
                    const bool hm = in.k[i] != in.k[m];
                    if (hp || hm) hi = m + 1;
                }
            }
            P.mhi[i] = hi;
            P.rowoff[i + 1] = P.rowoff[i] + (hi - i - 1);
        }
        P.rec.assign(P.rowoff[N], PairRec64{});
        std::atomic<int64_t> nterms(0);
        parallel_rows(N, [&](int i) {
            if (in.omega[i] > in.cutoff) return;
            int64_t nt = 0;
            const double ki = in.k[i], oi = in.omega[i];
            for (int m = i + 1; m < P.mhi[i]; ++m) {
                if (in.omega[m] > in.cutoff) continue;
                if (std::abs(oi - in.omega[m]) > in.bandwidth) continue;
                const double km = in.k[m], om = in.omega[m];
                const double gamma = c.cth[i] * c.cth[m] + c.sth[i] * c.sth[m];
                double kp, kn, Dp, Dm;
                cached_D(ki, km, Rn[i], Rn[m], Rs[i], Rs[m], gamma, h, kp, kn, Dp, Dm);
                PairRec64& r = P.rec[P.rowoff[i] + (m - i - 1)];
                const double AA = 0.5 * ag[i] * ag[m];
                const double amp = 0.5 * in.A[i] * in.A[m];
                if ((ki + km) <= kmax) {
                    const double ap = (Dp - (ki * km * gamma - Rn[i] * Rn[m])) / std::sqrt(Rn[i] * Rn[m]) + (Rn[i] + Rn[m]);
                    const double BB = Dp / (oi + om), C = cfac(kp, h);
                    r.cuP = AA * BB * (ki * c.cth[i] + km * c.cth[m]) * C;
                    r.cvP = AA * BB * (ki * c.sth[i] + km * c.sth[m]) * C;
                    r.cwP = AA * BB * kp * C;
                    r.kP = kp;
                    r.xP = amp * ap;
                    ++nt;
                }
                if (ki != km) {
                    const double am = (Dm - (ki * km * gamma + Rn[i] * Rn[m])) / std::sqrt(Rn[i] * Rn[m]) + (Rn[i] + Rn[m]);
                    const double CC = Dm / (oi - om), C = cfac(kn, h);
                    r.cuM = AA * CC * (ki * c.cth[i] - km * c.cth[m]) * C;
                    r.cvM = AA * CC * (ki * c.sth[i] - km * c.sth[m]) * C;
                    r.cwM = AA * CC * kn * C;
                    r.kM = kn;
                    r.xM = amp * am;
                    ++nt;
                }
            }
            nterms += nt;
        });
        P.nterms = nterms;
    }

    // ---- unfiltered direct table (u2/v2/w2/dp2)
    if (with_direct) {
        PairTable& P = out.direct;
        P.mhi.assign(N, N);
        P.rowoff.assign(N + 1, 0);
        for (int i = 0; i < N; ++i) P.rowoff[i + 1] = P.rowoff[i] + (N - i - 1);
        P.rec.assign(P.rowoff[N], PairRec64{});
        std::atomic<int64_t> nterms(0);
        parallel_rows(N, [&](int i) {
            int64_t nt = 0;
            const double ki = in.k[i], oi = in.omega[i], Ai = in.A[i];
            const double thi = in.theta[i] + (in.mtheta * kRefPi / 180.);
            const double Rni = ki * std::tanh(ki * h), Rsni = std::sqrt(Rni);
            for (int m = i + 1; m < N; ++m) {
                const double km = in.k[m], om = in.omega[m];
                const double thm = in.theta[m] + (in.mtheta * kRefPi / 180.);
                double kp, kn, Dp, Dm;
                direct_D(ki, km, thi, thm, h, Rni, Rsni, kp, kn, Dp, Dm);
                const double AA = 0.5 * (Ai * g / oi) * (in.A[m] * g / om);
                PairRec64& r = P.rec[P.rowoff[i] + (m - i - 1)];
                const double BB = (Dp / (oi + om)), Cp = cfac(kp, h);
                r.cuP = AA * BB * (ki * std::cos(thi) + km * std::cos(thm)) * Cp;
                r.cvP = AA * BB * (ki * std::sin(thi) + km * std::sin(thm)) * Cp;
                r.cwP = AA * BB * kp * Cp;
                r.kP = kp;
                r.xP = AA * BB * (oi + om) * Cp;
                ++nt;
                if (ki != km) {
                    const double CC = (Dm / (oi - om)), Cm = cfac(kn, h);
                    r.cuM = AA * CC * (ki * std::cos(thi) - km * std::cos(thm)) * Cm;
                    r.cvM = AA * CC * (ki * std::sin(thi) - km * std::sin(thm)) * Cm;
                    r.cwM = AA * CC * kn * Cm;
                    r.kM = kn;
                    r.xM = AA * CC * (oi - om) * Cm;
                    ++nt;
                }
            }
            nterms += nt;
        });
        P.nterms = nterms;
    }
    return std::string();
}

// Irregular::spectral_moments (Irregular.cpp:1336-1369): the loop stops at n-1.
void spectral_moments(const SeaInput& in, double m[3]) {
    m[0] = m[1] = m[2] = 0.;
    const int n = std::min((int)in.omega.size(), in.nfreq > 0 ? in.nfreq : (int)in.omega.size());
    for (int i = 0; i < n - 1; ++i) {
        m[0] += in.A[i] * in.A[i] / 2.;
        m[1] += in.omega[i] * in.A[i] * in.A[i] / 2.;
        m[2] += in.omega[i] * in.omega[i] * in.A[i] * in.A[i] / 2.;
    }
}

namespace {
// Irregular::interpolate (Irregular.cpp:1311-1334), extrapolating.
double interp(const std::vector<double>& xd, const std::vector<double>& yd, double x) {
    const int size = (int)xd.size();
    int i = 0;
    if (x >= xd[size - 2]) i = size - 2;
    else while (x > xd[i + 1]) i++;
    const double dydx = (yd[i + 1] - yd[i]) / (xd[i + 1] - xd[i]);
    return yd[i] + dydx * (x - xd[i]);
}
void periods(const SeaInput& in, double& t1, double& t2) {
    double m[3];
    spectral_moments(in, m);
    t1 = 2. * kRefPi * (m[0] / m[1]);
    t2 = 2. * kRefPi * std::sqrt(m[0] / m[2]);
}
}  // namespace

double mean_period(const SeaInput& in, int opt) {               // :1419-1436
    double t1, t2;
    periods(in, t1, t2);
    return opt == 1 ? t1 : (opt == 2 ? t2 : 0.);
}
double phase_speed(const SeaInput& in, int opt) {               // :1372-1392
    if (in.omega.size() < 2) return 0.;
    double t1, t2;
    periods(in, t1, t2);
    const double w1 = 2 * kRefPi / t1, w2 = 2 * kRefPi / t2;
    if (opt == 1) return w1 / interp(in.omega, in.k, w1);
    if (opt == 2) return w2 / interp(in.omega, in.k, w2);
    return 0.;
}
double mean_length(const SeaInput& in, int opt) {               // :1397-1415
    if (in.omega.size() < 2) return 0.;
    double t1, t2;
    periods(in, t1, t2);
    const double w1 = 2 * kRefPi / t1, w2 = 2 * kRefPi / t2;
    if (opt == 1) return (2. * kRefPi) / interp(in.omega, in.k, w1);
    if (opt == 2) return (2. * kRefPi) / interp(in.omega, in.k, w2);
    return 0.;
}

void build_smit(const SeaTables& t, std::vector<SmitComp>& comp, std::vector<SmitPair>& pair) {
    const SeaInput& in = t.in;
    const int N = (int)in.omega.size();
    const double h = in.depth, g = in.g;
    std::vector<double> Rn(N), Rs(N);
    comp.assign(N, SmitComp{});
    for (int i = 0; i < N; ++i) {
        const double k = in.k[i], w = in.omega[i];
        Rn[i] = k * std::tanh(k * h);                              // Irregular.cpp:826-828
        Rs[i] = std::sqrt(Rn[i]);
        const double D_nn = (6.0 * Rs[i] * Rs[i] * (k * k - Rn[i] * Rn[i])) /
                            (2.0 * Rs[i] * Rs[i] - k * std::tanh(2.0 * k * h));
        const double alpha_nn = (D_nn - (k * k - Rn[i] * Rn[i])) / Rn[i] + 2.0 * Rn[i];
        const double om2 = 2.0 * w;
        comp[i].A = in.A[i];
        comp[i].Ck = cfac(k, h);
        comp[i].C2k = cfac(2.0 * k, h);
        comp[i].Hd = -g / om2 * (alpha_nn / 2.0 - (3.0 * w * w) / (2.0 * g) + g * k * k / (2.0 * w * w));
    }
    const PairTable& P = t.filt;
    pair.assign(P.rec.size(), SmitPair{});
    const CompTable& c = t.comp;
    parallel_rows(N, [&](int i) {
        if (in.omega[i] > in.cutoff) return;
        const double ki = in.k[i], oi = in.omega[i];
        for (int m = i + 1; m < P.mhi[i]; ++m) {
            if (in.omega[m] > in.cutoff) continue;
            if (std::abs(oi - in.omega[m]) > in.bandwidth) continue;
            const double km = in.k[m], om = in.omega[m];
            const double gamma = c.cth[i] * c.cth[m] + c.sth[i] * c.sth[m];
            double kp, kn, Dp, Dm;
            cached_D(ki, km, Rn[i], Rn[m], Rs[i], Rs[m], gamma, h, kp, kn, Dp, Dm);
            SmitPair& r = pair[P.rowoff[i] + (m - i - 1)];
            const double k1dk2 = ki * km * gamma;
            r.k1dk2 = k1dk2;
            r.kp = kp;
            r.km = kn;
            r.CkP = cfac(kp, h);
            r.CkM = cfac(kn, h);
            if ((ki + km) <= t.kmax) {
                const double ap = (Dp - (ki * km * gamma - Rn[i] * Rn[m])) / std::sqrt(Rn[i] * Rn[m]) + (Rn[i] + Rn[m]);
                const double om12 = oi + om;
                r.Hp = -g / om12 * (ap / 2.0 - (oi * oi + om * om + oi * om) / (2.0 * g) + g * k1dk2 / (2.0 * oi * om));
                r.fP = 1.0;
            }
            if (ki != km) {
                const double am = (Dm - (ki * km * gamma + Rn[i] * Rn[m])) / std::sqrt(Rn[i] * Rn[m]) + (Rn[i] + Rn[m]);
                const double om12 = oi - om;
                if (std::abs(om12) > kSmall)
                    r.Hm = -g / om12 * (am / 2.0 - (oi * oi + om * om - oi * om) / (2.0 * g) + g * k1dk2 / (2.0 * oi * om));
                r.fM = 1.0;
            }
        }
    });
}

}  // namespace cfdwm
