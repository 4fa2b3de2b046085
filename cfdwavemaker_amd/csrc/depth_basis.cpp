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
// column set cols[j] (each of length len), in pick order.
std::vector<int> pivot_columns(std::vector<std::vector<double>> cols, int want) {
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

}  // namespace cfdwm
