// Skeleton-exponential basis of the depth profiles (see depth_basis.hpp).
#include "depth_basis.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>

namespace cfdwm {

namespace {

// Column-pivoted modified Gram-Schmidt: the first `want` pivot columns of the
// column set cols[j] (each of length len), in pick order; `first` >= 0 forces
// the first pick.
std::vector<int> pivot_columns(std::vector<std::vector<double>> cols, int want, int first = -1) {
    const int J = (int)cols.size();
    std::vector<double> nrm(J);
    for (int j = 0; j < J; ++j) {
        double s = 0;
        for (double v : cols[j]) s += v * v;
        nrm[j] = s;
    }
    std::vector<int> pick;
    std::vector<char> used(J, 0);
    for (int r = 0; r < want && r < J; ++r) {
        int best = -1;
        double bv = -1;
        for (int j = 0; j < J; ++j)
            if (!used[j] && nrm[j] > bv) { bv = nrm[j]; best = j; }
        if (r == 0 && first >= 0 && first < J) { best = first; bv = nrm[first]; }
        if (best < 0 || bv <= 0) break;
        used[best] = 1;
        pick.push_back(best);
        std::vector<double> q = cols[best];
        double s = 0;
        for (double v : q) s += v * v;
        s = std::sqrt(s);
        for (double& v : q) v /= s;
        for (int j = 0; j < J; ++j) {
            if (used[j]) continue;
            double d = 0;
            for (size_t k = 0; k < q.size(); ++k) d += q[k] * cols[j][k];
            double n2 = 0;
            for (size_t k = 0; k < q.size(); ++k) {
                cols[j][k] -= d * q[k];
                n2 += cols[j][k] * cols[j][k];
            }
            nrm[j] = n2;
        }
    }
    return pick;
}

// Gauss-Jordan inverse with partial pivoting (n x n, row-major); false if singular
bool invert(std::vector<double> a, int n, std::vector<double>& inv) {
    inv.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1.0;
    for (int c = 0; c < n; ++c) {
        int p = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(a[(size_t)r * n + c]) > std::fabs(a[(size_t)p * n + c])) p = r;
        if (a[(size_t)p * n + c] == 0.0) return false;
        if (p != c)
            for (int k = 0; k < n; ++k) {
                std::swap(a[(size_t)p * n + k], a[(size_t)c * n + k]);
                std::swap(inv[(size_t)p * n + k], inv[(size_t)c * n + k]);
            }
        const double d = a[(size_t)c * n + c];
        for (int k = 0; k < n; ++k) { a[(size_t)c * n + k] /= d; inv[(size_t)c * n + k] /= d; }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            const double f = a[(size_t)r * n + c];
            if (f == 0.0) continue;
            for (int k = 0; k < n; ++k) {
                a[(size_t)r * n + k] -= f * a[(size_t)c * n + k];
                inv[(size_t)r * n + k] -= f * inv[(size_t)c * n + k];
            }
        }
    }
    return true;
}

}  // namespace

void DepthBasis::coeffs(double kappa, double* c) const {
    double e[64];
    for (int s = 0; s < R; ++s) e[s] = std::exp(-kappa * tau[s]);
    for (int q = 0; q < R; ++q) {
        double v = 0;
        for (int s = 0; s < R; ++s) v += einv[(size_t)q * R + s] * e[s];
        c[q] = v;
    }
}

double DepthBasis::eval(double kappa, double t) const {
    double c[64];
    coeffs(kappa, c);
    double v = 0;
    for (int q = 0; q < R; ++q) v += c[q] * std::exp(-mu[q] * t);
    return v;
}

// e^{-kappa t} depends on kappa t only: a basis for kappa in [0, 1], t in [0, Pi]
// serves every (kmax, tmax) with kmax tmax <= Pi after scaling mu by kmax and
// tau by 1 / kmax.  Bases are cached for Pi on a geometric grid (ratio 1.1),
// so the sea states of a batch share a handful of them.
int build_depth_basis_cached(double kmax, double tmax, double tol, int rmax, DepthBasis& B) {
    kmax = std::max(kmax, 1e-9);
    const double prod = std::max(kmax * tmax, 1e-6);
    const int j = (int)std::ceil(std::log(prod / 0.1) / std::log(1.1));
    const double Pi = 0.1 * std::pow(1.1, j);
    static std::mutex mu;
    static std::map<std::tuple<int, double, int>, std::pair<int, DepthBasis>> cache;
    const auto key = std::make_tuple(j, tol, rmax);
    std::pair<int, DepthBasis> e;
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find(key);
        if (it != cache.end()) e = it->second;
    }
    if (e.second.R == 0) {
        e.first = build_depth_basis(1.0, Pi, tol, rmax, e.second);
        std::lock_guard<std::mutex> lk(mu);
        cache[key] = e;
    }
    B = e.second;
    B.kmax = kmax;
    B.tmax = Pi / kmax;
    for (double& m : B.mu) m *= kmax;
    for (double& t : B.tau) t /= kmax;
    return e.first;
}

int build_depth_basis(double kmax, double tmax, double tol, int rmax, DepthBasis& B) {
    rmax = std::min(std::max(rmax, 1), 64);
    kmax = std::max(kmax, 1e-9);
    const int J = 768, S = 1024;
    std::vector<double> kap(J), ts(S);
    for (int j = 0; j < J; ++j) kap[j] = kmax * j / (J - 1);
    for (int s = 0; s < S; ++s) ts[s] = tmax * s / (S - 1);
    std::vector<std::vector<double>> cols(J, std::vector<double>(S));
    for (int j = 0; j < J; ++j)
        for (int s = 0; s < S; ++s) cols[j][s] = std::exp(-kap[j] * ts[s]);
    const std::vector<int> kpick = pivot_columns(cols, rmax);
    // error check set: off-grid wavenumbers and depths
    const int JC = 181, SC = 263;
    int status = -1;
    for (int R = std::min(4, (int)kpick.size()); R <= (int)kpick.size(); ++R) {
        DepthBasis T;
        T.R = R; T.kmax = kmax; T.tmax = tmax;
        for (int q = 0; q < R; ++q) T.mu.push_back(kap[kpick[q]]);
        std::sort(T.mu.begin(), T.mu.end());
        std::vector<std::vector<double>> tc(S, std::vector<double>(R));
        for (int s = 0; s < S; ++s)
            for (int q = 0; q < R; ++q) tc[s][q] = std::exp(-T.mu[q] * ts[s]);
        const std::vector<int> tpick = pivot_columns(tc, R);
        if ((int)tpick.size() < R) continue;
        for (int s = 0; s < R; ++s) T.tau.push_back(ts[tpick[s]]);
        std::sort(T.tau.begin(), T.tau.end());
        std::vector<double> E((size_t)R * R);
        for (int s = 0; s < R; ++s)
            for (int q = 0; q < R; ++q) E[(size_t)s * R + q] = std::exp(-T.mu[q] * T.tau[s]);
        if (!invert(E, R, T.einv)) continue;
        double err = 0;
        for (int a = 0; a < JC; ++a) {
            const double k = kmax * (a + 0.37) / JC;
            double c[64];
            T.coeffs(k, c);
            for (int b = 0; b < SC; ++b) {
                const double t = tmax * (b + 0.61) / SC;
                double v = 0;
                for (int q = 0; q < R; ++q) v += c[q] * std::exp(-T.mu[q] * t);
                err = std::max(err, std::fabs(v - std::exp(-k * t)));
            }
        }
        T.err = err;
        B = T;
        if (err <= tol) { status = 0; break; }
    }
    return status;
}

double ProfileBasis::profile(double kappa, double z) const {
    return std::exp(kappa * z) + sign * std::exp(-kappa * (z + 2.0 * h));
}

void ProfileBasis::coeffs(double kappa, double* c) const {
    double e[64];
    for (int s = 0; s < R; ++s) e[s] = profile(kappa, zs[s]);
    for (int q = 0; q < R; ++q) {
        double v = 0;
        for (int s = 0; s < R; ++s) v += finv[(size_t)q * R + s] * e[s];
        c[q] = v;
    }
}

double ProfileBasis::eval(double kappa, double z) const {
    double c[64];
    coeffs(kappa, c);
    double v = 0;
    for (int q = 0; q < R; ++q) v += c[q] * profile(mu[q], z);
    return v;
}

int build_profile_basis(double kmax, double h, int sign, double tol, int rmax, ProfileBasis& B) {
    rmax = std::min(std::max(rmax, 1), 64);
    kmax = std::max(kmax, 1e-9);
    h = std::max(h, 1e-9);
    sign = sign < 0 ? -1 : 1;
    ProfileBasis P0;
    P0.sign = sign; P0.h = h; P0.kmax = kmax;
    B = ProfileBasis();
    const int J = 768, S = 1024;
    std::vector<double> kap(J), zz(S);
    for (int j = 0; j < J; ++j) kap[j] = kmax * j / (J - 1);
    // P_kappa(-2h - z) = sign P_kappa(z): sampling z in [-h, 0] covers [-2h, 0]
    // with the same error, and keeps mirror-image skeleton depths (rows of F
    // equal up to sign) out of the interpolation
    for (int s = 0; s < S; ++s) zz[s] = -h * s / (S - 1);
    std::vector<std::vector<double>> cols(J, std::vector<double>(S));
    for (int j = 0; j < J; ++j)
        for (int s = 0; s < S; ++s) cols[j][s] = P0.profile(kap[j], zz[s]);
    // Candidate wavenumber skeletons: the greedy pivot order, and the same
    // with the first pick forced to one of the smallest nonzero wavenumbers.
    // Interpolating in the sinh-type family near kappa = 0 (where P_kappa ~
    // 2 kappa (z + h)) is badly conditioned when the smallest skeleton
    // wavenumber is large, which amplifies the rounding of the bf16 form
    // tables; a small first pick fixes that at the same or a lower rank.
    const int firsts[] = {-1, 1, 2, 3, 5, 8};
    constexpr int NC = 6;
    std::vector<int> kpick[NC];
    // (the sampled profiles' numerical rank is below 24)
    for (int c = 0; c < NC; ++c) kpick[c] = pivot_columns(cols, std::min(rmax, 24), firsts[c]);
    const int JC = 181, SC = 263;            // error check: off-grid wavenumbers and depths
    int status = -1;
    for (int R = std::min(4, rmax); R <= std::min(rmax, 24); ++R) {
        ProfileBasis best;
        for (int c = 0; c < NC; ++c) {
            if ((int)kpick[c].size() < R) continue;
            ProfileBasis T = P0;
            T.R = R;
            for (int q = 0; q < R; ++q) T.mu.push_back(kap[kpick[c][q]]);
            std::sort(T.mu.begin(), T.mu.end());
            std::vector<std::vector<double>> zc(S, std::vector<double>(R));
            for (int s = 0; s < S; ++s)
                for (int q = 0; q < R; ++q) zc[s][q] = T.profile(T.mu[q], zz[s]);
            const std::vector<int> zpick = pivot_columns(zc, R);
            if ((int)zpick.size() < R) continue;
            for (int s = 0; s < R; ++s) T.zs.push_back(zz[zpick[s]]);
            std::sort(T.zs.begin(), T.zs.end());
            std::vector<double> F((size_t)R * R);
            for (int s = 0; s < R; ++s)
                for (int q = 0; q < R; ++q) F[(size_t)s * R + q] = T.profile(T.mu[q], T.zs[s]);
            if (!invert(F, R, T.finv)) continue;
            // max error, and the rounding amplification: per wavenumber, the
            // largest sqrt(sum_q c_q^2 P_{mu_q}(z)^2) over z against the
            // largest |P_kappa(z)| (floored at 0.05)
            std::vector<double> pqz((size_t)SC * R);
            for (int b = 0; b < SC; ++b)
                for (int q = 0; q < R; ++q) pqz[(size_t)b * R + q] = T.profile(T.mu[q], -h * (b + 0.61) / SC);
            double err = 0, amp = 0;
            for (int a = 0; a < JC; ++a) {
                const double k = kmax * (a + 0.37) / JC;
                double cq[64];
                T.coeffs(k, cq);
                double r2 = 0, pk = 0;
                for (int b = 0; b < SC; ++b) {
                    const double* pq = &pqz[(size_t)b * R];
                    double v = 0, w2 = 0;
                    for (int q = 0; q < R; ++q) {
                        v += cq[q] * pq[q];
                        w2 += cq[q] * cq[q] * pq[q] * pq[q];
                    }
                    const double pz = T.profile(k, -h * (b + 0.61) / SC);
                    err = std::max(err, std::fabs(v - pz));
                    r2 = std::max(r2, w2);
                    pk = std::max(pk, std::fabs(pz));
                }
                amp = std::max(amp, std::sqrt(r2) / std::max(pk, 0.05));
            }
            T.err = err;
            T.amp = amp;
            // the candidate meeting tol with the least amplification, else the most accurate
            const bool ok = err <= tol, bok = best.R > 0 && best.err <= tol;
            if (best.R == 0 || (ok && (!bok || amp < best.amp)) || (!ok && !bok && err < best.err)) best = T;
        }
        if (best.R == 0) continue;
        // past the sampled profiles' numerical rank (error ~1e-8) further
        // picks only worsen the conditioning of F: keep the best basis seen
        if (B.R == 0 || best.err < B.err || best.err <= tol) B = best;
        if (best.err <= tol) { status = 0; break; }
    }
    return status;
}

// P_kappa(z) depends on (kappa h, z / h) only: a basis for h = 1, kappa in
// [0, X] serves every (kmax, h) with kmax h <= X after scaling mu by 1 / h and
// z_s by h.  Cached for X on a geometric grid (ratio 1.1).
int build_profile_basis_cached(double kmax, double h, int sign, double tol, int rmax, ProfileBasis& B) {
    kmax = std::max(kmax, 1e-9);
    h = std::max(h, 1e-9);
    sign = sign < 0 ? -1 : 1;
    const double prod = std::max(kmax * h, 1e-6);
    const int j = (int)std::ceil(std::log(prod / 0.1) / std::log(1.1));
    const double X = 0.1 * std::pow(1.1, j);
    static std::mutex mu;
    static std::map<std::tuple<int, int, double, int>, std::pair<int, ProfileBasis>> cache;
    const auto key = std::make_tuple(j, sign, tol, rmax);
    std::pair<int, ProfileBasis> e;
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find(key);
        if (it != cache.end()) e = it->second;
    }
    if (e.second.R == 0) {
        e.first = build_profile_basis(X, 1.0, sign, tol, rmax, e.second);
        std::lock_guard<std::mutex> lk(mu);
        cache[key] = e;
    }
    B = e.second;
    B.h = h;
    B.kmax = X / h;
    for (double& m : B.mu) m /= h;
    for (double& z : B.zs) z *= h;
    return e.first;
}

}  // namespace cfdwm
