// Separable approximation of the depth-profile exponentials of the pair terms.
//
// Every second-order velocity pair term carries a profile e^{kappa z} and
// e^{-kappa (z + 2h)} (src/Irregular.cpp:1215-1290 through the cached
// cosh/sinh ratios), i.e. g(kappa, t) = e^{-kappa t} at t = -z in [0, h] and
// t = 2h + z in [h, 2h].  Over kappa in [0, kmax] and t in [0, 2h] this kernel
// has low numerical rank; with R skeleton wavenumbers mu_q and R skeleton
// depths tau_s (both picked by column-pivoted Gram-Schmidt on sampled kernel
// matrices),
//
//   e^{-kappa t} ~= sum_q c_q(kappa) e^{-mu_q t},   c(kappa) = E^{-1} [e^{-kappa tau_s}]_s,
//   E[s][q] = e^{-mu_q tau_s},
//
// is an interpolation at the skeleton depths whose coefficients stay O(1).
// R is the smallest rank whose measured maximum error over the range is
// within the requested tolerance.
#pragma once
#include <vector>

namespace cfdwm {

struct DepthBasis {
    int R = 0;
    double kmax = 0, tmax = 0;       // range of validity
    double err = 0;                  // measured max |approximation - e^{-kappa t}| over the range
    std::vector<double> mu, tau;     // skeleton wavenumbers and depths
    std::vector<double> einv;        // R x R, row-major: c = einv * e(kappa)
    void coeffs(double kappa, double* c) const;
    double eval(double kappa, double t) const;
};

// 0 on success, -1 when no rank <= rmax reaches tol (B then holds rank rmax).
int build_depth_basis(double kmax, double tmax, double tol, int rmax, DepthBasis& B);
// The same through a process-wide cache keyed by kmax * tmax (rounded up on a
// ratio-1.1 grid; the basis is scale invariant), thread-safe.
int build_depth_basis_cached(double kmax, double tmax, double tol, int rmax, DepthBasis& B);

}  // namespace cfdwm
