// Host-side construction of the per-component and per-pair coefficient tables
// that the HIP kernels stream.  Everything here is FP64 and follows the
// reference's expressions for the Sharma-Dean transfer coefficients:
//   filtered ("pair cache") set:  src/Irregular.cpp:805-967  (eta2, uvw2)
//   unfiltered (direct) set:      src/Irregular.cpp:450-803  (u2, v2, w2, dp2)
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace cfdwm {

constexpr double kRefPi = 3.1415926535897;   // truncated PI, src/Irregular.h:13-15
constexpr double kSmall = 1.e-12;            // src/Irregular.h:7-9
constexpr double kHuge = 1.e12;              // src/Irregular.h:10-12

// Inputs of one irregular sea, as Irregular holds them after parsing
// (components sorted by omega, amplitudes already normalised/amplified).
struct SeaInput {
    std::vector<double> omega, A, k, phase, theta;
    std::vector<double> ramptime, starttime, duration;   // empty -> reference defaults
    double depth = 300.0, g = 9.81, rho = 1025.0, swl = 0.0, mtheta = 0.0;
    double tofmax = 0.0, x0 = 0.0, y0 = 0.0;
    double bandwidth = 100.0, cutoff = 100000.0;
    int order = 2;
    // frequency rows of the input file (Irregular::nfreq): with ndir > 0 the
    // components are nfreq x ndir, and the spectral moments sum only the first
    // nfreq - 1 of them (Irregular.cpp:1356-1360); 0: the component count
    int nfreq = 0;
};

// Per-pair record.  Branch "P" is the sum interaction (phi_i + phi_m), "M"
// the difference interaction (phi_i - phi_m).  Velocity coefficients already
// carry 1/(1+exp(-2 k+- h)), so the depth profile needs only
// exp(k z) +- exp(-k (z + 2h)).  Absent branches have all-zero coefficients.
//   filtered table:   xP, xM = eta coefficients 0.5 A_i A_m alpha+-
//   unfiltered table: xP, xM = pressure coefficients AA BB (w_i+w_m) C+, AA CC (w_i-w_m) C-
struct PairRec64 {
    double cuP, cvP, cwP, kP, cuM, cvM, cwM, kM, xP, xM;
};
struct PairRec32 {           // FP32 form: k*log2(e) and -2 h k log2(e) replace k
    float cuP, cvP, cwP, klP, cuM, cvM, cwM, klM, xP, xM, g2P, g2M;
};

// Per-component table (FP64).  Profiles use the ratio form
// cosh(k(z+h))/cosh(kh) = C_k (e^{kz} + e^{-k(z+2h)}),  C_k = 1/(1+e^{-2kh}).
struct CompTable {
    int n = 0;
    // phase: k*(cth*(x-x0) + sth*(y-y0)) - omega*t + ph0   (Irregular.cpp:237-240)
    std::vector<double> k, cth, sth, omega, ph0;
    std::vector<double> rt, ts, du;                 // freq_ramp parameters
    std::vector<double> Ck;                         // 1/(1+exp(-2 k h))
    // linear fields (u1, v1, w1, dp1; Irregular.cpp:394-447, 1107-1122)
    std::vector<double> lu, lv, lw, lp;             // A g k/w cth C, .. sth C, A g k/w C, A rho g C
    // Taylor gradients at z=0 (Irregular.cpp:1125-1192)
    std::vector<double> gxz, gyz, gzz, gpz;         // cth A w k, sth A w k, A w k coth(kh), A rho g k
    // eta1 amplitude and the filtered eta2 diagonal (Irregular.cpp:249-269)
    std::vector<double> A, e2d;
    // second-order diagonal, profile of 2k: C_2k = 1/(1+exp(-4kh))
    std::vector<double> C2k;
    std::vector<double> dfu, dfv, dfw;              // filtered (uvw2) diag: AA 2k cth C2k, AA 2k sth C2k, AA 2k C2k
    std::vector<double> ddu, ddv, ddw, ddp;         // unfiltered (u2..dp2) diag, ddp = AA 2w C2k
    // dp2 kinetic terms (Irregular.cpp:755-756, 786-787): separable in i, m
    std::vector<double> Xx, Xy, Zk;                 // g k cth / w, g k sth / w, g k / w
    std::vector<double> kd1, kd2;                   // diagonal: g^2 (2k cth + 2k sth) 2/(2w) C_k, g^2 2k 2/(2w) C_k
    std::vector<uint8_t> rowActive;                 // filtered set: omega_i <= cutoff
};

struct PairTable {
    std::vector<int64_t> rowoff;     // record index of (i, i+1)
    std::vector<int32_t> mhi;        // row i covers m in [i+1, mhi[i])
    std::vector<PairRec64> rec;
    int64_t nterms = 0;              // number of present branches (+ diagonal terms counted separately)
};

struct SeaTables {
    SeaInput in;
    double kmax = 0.0;
    CompTable comp;
    PairTable filt;                  // pair-cache semantics (eta2, uvw2): s-grid and eta
    PairTable direct;                // unfiltered semantics (u2, v2, w2, dp2): direct mode
    int64_t diag_filtered = 0;       // diagonal terms present in uvw2
};

// Builds all tables.  Returns an empty string on success, else an error text.
std::string build_tables(const SeaInput& in, SeaTables& out, bool with_direct);

// Smit et al. (2017) s-coordinate coefficients (vertical_theory smit,
// src/Irregular.cpp:1623-2027).  Per component: A, the profile factors
// C_k = 1/(1+e^{-2kh}) and C_2k, and the diagonal H from the Sharma-Dean
// alpha_nn through Eq. 10 (:1672-1678).  Per record of the filtered pair
// table (same row layout): H of each branch (:919-927, :950-958), the
// combined wavenumbers and their profile factors, k_i k_m gamma, and 1/0
// branch-present flags (the branch terms not scaled by H, the P-derivative
// and the kinetic pressure, are gated by them).
struct SmitComp {
    double A, Ck, C2k, Hd;
};
struct SmitPair {
    double Hp, Hm, kp, km, CkP, CkM, k1dk2, fP, fM, pad;
};
void build_smit(const SeaTables& t, std::vector<SmitComp>& comp, std::vector<SmitPair>& pair);

// Spectral helpers (Irregular.cpp:1336-1453)
void spectral_moments(const SeaInput& in, double m[3]);
double mean_period(const SeaInput& in, int opt);
double phase_speed(const SeaInput& in, int opt);
double mean_length(const SeaInput& in, int opt);

}  // namespace cfdwm
