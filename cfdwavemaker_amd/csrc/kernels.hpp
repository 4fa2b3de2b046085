// Device data layout shared by the HIP kernels (kernels.hip) and the engine.
#pragma once
#include <cstdint>

#include "sea_tables.hpp"

namespace cfdwm {

// Per-component phase data, FP64 in both precisions: the phase
// k*(cth*(x-x0) + sth*(y-y0)) - omega*t + ph0 is always accumulated and
// range-reduced in FP64 (src/Irregular.cpp:237-240).
struct PhaseRec {
    double k, cth, sth, omega, ph0, rt, ts, du;
};

// Per-component coefficients in the working precision R.  Depth profiles
// are evaluated as E = e^{pk z}, G = e^{-pk z + pg} (FP64: natural base,
// FP32: base 2 with pk = k log2 e), so cosh-ratio = C (E + G) and
// sinh-ratio = C (E - G) with C folded into the coefficients.
template <typename R>
struct CompRec {
    R lu, lv, lw, lp;          // u1, v1, w1, dp1
    R gxz, gyz, gzz, gpz;      // Taylor gradients at z = 0
    R A, e2d;                  // eta1 amplitude, filtered eta2 diagonal
    R dfu, dfv, dfw;           // filtered (uvw2) diagonal
    R ddu, ddv, ddw, ddp;      // unfiltered (u2..dp2) diagonal
    R Xx, Xy, Zk, kd1, kd2;    // dp2 kinetic terms
    R pk, pg, p2k, p2g;        // profile exponents for k and 2k
    R eg2;                     // 2 E G = 2 e^{-2kh}: the 2k profile's E^2 + G^2 is (E + G)^2 - eg2
    R pad;
};

// Per-pair velocity record (sum branch P, difference branch M).
template <typename R>
struct UvwRec {
    R cuP, cvP, cwP, pkP, pgP, cuM, cvM, cwM, pkM, pgM;
};
// Per-pair scalar coefficients: eta (filtered) or pressure (unfiltered).
template <typename R>
struct XRec {
    R xP, xM;
};

template <typename R>
struct SeaDev {
    int n;
    double h, swl, x0, y0, rho;
    const PhaseRec* __restrict__ ph;
    const CompRec<R>* __restrict__ cr;
    const uint8_t* __restrict__ rowActive;
    const int64_t* __restrict__ frow;   // filtered table
    const int32_t* __restrict__ fmhi;
    const UvwRec<R>* __restrict__ fuvw;
    const XRec<R>* __restrict__ fx;
    const int64_t* __restrict__ drow;   // unfiltered table (rows cover m in [i+1, n))
    const UvwRec<R>* __restrict__ duvw;
    const XRec<R>* __restrict__ dx;
};

// Flags of the point evaluator.
enum : uint32_t {
    EV_LIN = 1u,        // first-order u1, v1, w1 (and dp1 with EV_P)
    EV_SEC = 2u,        // second-order pair sums
    EV_DIRECT = 4u,     // unfiltered u2/v2/w2/dp2 semantics (else filtered uvw2)
    EV_TAYLOR = 8u,     // + gradient at z=0 times max(0, z - swl)
    EV_P = 16u,         // pressure
    EV_ZCLAMP = 32u,    // profiles at min(0, z - swl) (order 1 and 2)
    EV_ZSHIFT = 64u,    // profiles at z - swl (order 0)
};

// A contiguous range of pair records (split-K over the pair sum).  Rows
// [rlo, rhi) have pairs in it (row rlo from m = mstart, row rhi-1 up to
// m = mend); rows [own_lo, own_hi) get their per-row terms (linear,
// diagonal, Taylor gradients) from this segment.
struct Segment {
    int rlo, rhi, mstart, mend, own_lo, own_hi, pad0, pad1;
};

// s-grid geometry and ring-buffer slot bookkeeping (lsGridSpline).
struct GridGeom {
    int nx, ny, nl;
    double x0, y0, dx, dy, ds, h, swl;
};

// Second-order pair sums as quadratic forms on MFMA (qform.hip).  Operands
// are bf16 hi/lo pairs in a tiled layout: rows in blocks of 32, k in blocks
// of 16, a 32 x 16 tile stored as 512 contiguous values in two 32 x 8 halves
// (k 0-7, then 8-15), each row's 8 values contiguous.
struct QfJob {
    int64_t p0 = 0, P = 0, ppad = 0;     // chunk of points [p0, p0 + P), padded to 128
    int n = 0, npad = 0;                  // components, padded to 128
    int grid = 1;                         // points: slot-major s-grid columns, else arrays
    int64_t ncol = 0;
    const int32_t* cols = nullptr;
    GridGeom g{};
    double slot_t[4] = {0, 0, 0, 0};
    const double *x = nullptr, *y = nullptr, *t = nullptr;
    double x0 = 0, y0 = 0;                // wave reference point
    const PhaseRec* ph = nullptr;
    uint16_t *zc_hi = nullptr, *zc_lo = nullptr, *zs_hi = nullptr, *zs_lo = nullptr;   // ppad x npad
    int64_t zn = 0;                       // elements between the four Z arrays (zc_lo = zc_hi + zn, ...)
    const uint16_t* b = nullptr;          // [nform][nbarr][npad x npad]: B1 hi, B1 lo, B2 hi, B2 lo (nbarr 4)
                                          // or B1 hi, B2 hi (nbarr 2)
    int nbarr = 4;
    int nlv = 3;                          // bf16 products per term: 3 (hi/lo split) or 1 (hi only, nbarr 2 or 4)
    const int8_t* ftype = nullptr;        // per form: 0 cos-type, 1 sin-type
    int nform = 0;
    const int32_t* kend = nullptr;        // per 64-wide tile of i: end of its m range (multiple of 16)
    float* out = nullptr;                 // [nform][ostride]
    int64_t ostride = 0;
    // optional clock measurement of the forms launch: [0] += shader-clock
    // cycles and [1] += constant-rate wall-clock ticks of every block
    unsigned long long* clk = nullptr;
};

// One tile of a batched forms launch: job (sea state), point tile, form.
struct QfTile {
    int job, ptile, form, pad;
};

// Device build of the form tables from the FP32 pair records (qform.hip).
struct QfBuild {
    int n = 0, npad = 0, nform = 0;
    int R = 0, Rw = 0;                    // depth-basis ranks: cosh-type (u, v) and sinh-type (w) profiles
    double h = 0;                         // sea depth of the profiles e^{kz} +- e^{-k (z + 2h)}
    const int64_t* frow = nullptr;        // filtered table rows
    const int32_t* fmhi = nullptr;
    const UvwRec<float>* uvw = nullptr;   // pk = kappa log2 e
    const XRec<float>* x = nullptr;
    const double* zs = nullptr;           // [R] skeleton depths of the cosh-type basis
    const double* finv = nullptr;         // [R][R]
    const double* zsw = nullptr;          // [Rw] skeleton depths of the sinh-type basis
    const double* finvw = nullptr;        // [Rw][Rw]
    uint16_t* b = nullptr;                // [nform][nbarr][npad x npad], zeroed by the caller
    int nbarr = 4;                        // 4: hi and lo parts, 2: hi parts only
    uint16_t* beta = nullptr;             // when set: form 0 (eta2) goes here as [4][npad x npad]
                                          // and form f > 0 to b at index f - 1
};

// Smit et al. s-coordinate point evaluation (smit.hip).  eta1/eta2 are the
// points' Sharma-Dean elevations (eta2 unused for order < 1); sc and sp are
// the Smit coefficients of the components and of the filtered pair table
// (frow/fmhi layout), split into nseg balanced segments; part holds
// [nseg][5][n] partial sums.
struct SmitJob {
    int64_t n = 0;
    const double *x = nullptr, *y = nullptr, *z = nullptr, *t = nullptr;
    const double *eta1 = nullptr, *eta2 = nullptr;
    int order = 2, N = 0, nseg = 1;
    double h = 0, g = 9.81, rho = 1025, swl = 0, x0 = 0, y0 = 0;
    const PhaseRec* ph = nullptr;
    const SmitComp* sc = nullptr;
    const int64_t* frow = nullptr;
    const int32_t* fmhi = nullptr;
    const SmitPair* sp = nullptr;
    const Segment* seg = nullptr;
    double* part = nullptr;
    double *o_u = nullptr, *o_v = nullptr, *o_w = nullptr, *o_p = nullptr;
};

}  // namespace cfdwm
