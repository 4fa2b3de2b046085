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

// The same construction on the pair terms' whole depth profiles,
//
//   P_kappa(z) = e^{kappa z} + s e^{-kappa (z + 2h)},   z in [-2h, 0],
//
// s = +1 for the cosh-type profile of u and v, s = -1 for the sinh-type
// profile of w (the 1 / (1 + e^{-2 kappa h}) of the cosh ratios is carried by
// the pair coefficients):  P_kappa(z) ~= sum_q c_q(kappa) P_{mu_q}(z) with
// c(kappa) = F^{-1} [P_kappa(z_s)]_s, F[s][q] = P_{mu_q}(z_s).  Since
// P_kappa(-2h - z) = s P_kappa(z), the construction samples z in [-h, 0] and
// the skeleton depths lie there.  One profile family needs fewer terms than
// the two exponentials of e^{-kappa t} on [0, 2h] (at kmax h = 140: 11 and 11
// against 14 at 1e-4, 15 and 15 against 18 at 1e-6), and the levels assemble
// each family's terms with the same two exponentials per skeleton wavenumber.
// The error floor of the construction is about 1e-8 (rank 17-18).
struct ProfileBasis {
    int R = 0, sign = 1;
    double kmax = 0, h = 0;
    double err = 0;                  // measured max |approximation - P_kappa(z)| over the range
    double amp = 0;                  // rounding amplification max_kappa max_z |c(kappa) P(z)|_2 / max_z |P_kappa|
    std::vector<double> mu, zs;      // skeleton wavenumbers and depths (-h <= z_s <= 0)
    std::vector<double> finv;        // R x R, row-major: c = finv * P_kappa(z_s)
    double profile(double kappa, double z) const;
    void coeffs(double kappa, double* c) const;
    double eval(double kappa, double z) const;
};
int build_profile_basis(double kmax, double h, int sign, double tol, int rmax, ProfileBasis& B);
// Through a process-wide cache keyed by kmax h (rounded up on a ratio-1.1
// grid; the basis depends on kmax h only after scaling), thread-safe.
int build_profile_basis_cached(double kmax, double h, int sign, double tol, int rmax, ProfileBasis& B);
// The same through a process-wide cache keyed by kmax * tmax (rounded up on a
// ratio-1.1 grid; the basis is scale invariant), thread-safe.
int build_depth_basis_cached(double kmax, double tmax, double tol, int rmax, DepthBasis& B);

}  // namespace cfdwm
