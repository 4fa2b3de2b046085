// Test-infrastructure driver around the reference CFDwavemaker classes.
//
// This file is NOT part of the product. It is compiled together with the
// reference's own sources (read in place, never copied) by oracle/Makefile into
// oracle/_ref/libcfdwm_refh.so, and is used only by tests/golden/make_golden.py
// (to write golden vectors) and by bench.py's cpu_baseline leg (to time the
// reference's OpenMP s-grid fill).  It drives:
//   Irregular        (src/Irregular.h:47-193, src/Irregular.cpp)
//   lsGridSpline     (src/lsgrid_spline.h:15-129, src/lsgrid_spline.cpp)
//   lsGrid           (src/lsgrid.h:15-114, src/lsgrid.cpp; the legacy two-slot grid)
//   Ramp             (src/Utils.h:37-56, src/Utils.cpp:26-68)
// through a plain C ABI so Python (ctypes) can call it.
#include <cstring>
#include <vector>
#include <chrono>
#include "Irregular.h"
#include "lsgrid_spline.h"
#include "lsgrid.h"
#include "Utils.h"
#include <cstdio>

static Irregular g_irr;     // static storage: zero-initialised like the reference's global
static lsGridSpline g_grid;
static Ramp g_ramp;
static bool g_grid_alloc = false;
static lsGrid g_lg;         // its destructor frees UX0 only (lsgrid.h:25-30)
static bool g_lg_alloc = false;

// Frees the legacy grid's arrays; with init_only the slot-1 pointers alias
// slot 0 (lsgrid.cpp:466-470, 665-667) and the original slot-1 arrays leaked.
static void lg_free() {
    if (!g_lg_alloc) return;
    const bool ea = g_lg.ETA1 == g_lg.ETA0, ua = g_lg.UX1 == g_lg.UX0;
    delete[] g_lg.ETA0; if (!ea) delete[] g_lg.ETA1;
    delete[] g_lg.UX0; delete[] g_lg.UY0; delete[] g_lg.UZ0;
    if (!ua) { delete[] g_lg.UX1; delete[] g_lg.UY1; delete[] g_lg.UZ1; }
    delete[] g_lg.IGNORE;
    g_lg.ETA0 = g_lg.ETA1 = nullptr;
    g_lg.UX0 = g_lg.UY0 = g_lg.UZ0 = g_lg.UX1 = g_lg.UY1 = g_lg.UZ1 = nullptr;
    g_lg.IGNORE = nullptr;
    g_lg_alloc = false;
}

static void assign(std::vector<double>& v, const double* p, int n, double dflt) {
    v.assign(n, dflt);
    if (p) for (int i = 0; i < n; ++i) v[i] = p[i];
}

extern "C" {

// Mirrors the irregular post-processing of process_inputdata
// (src/CFDwavemaker.cpp:1347-1377): copies parameters, normalize_data,
// then rebuild_second_order_cache when order == 2.  Components are taken in the
// order given (the caller sorts by omega as the parser does, :663-673).
int refh_setup(int n, const double* omega, const double* A, const double* k,
               const double* phase, const double* theta,
               const double* ramptime, const double* starttime, const double* duration,
               double depth, double g, double rho, double swl, double mtheta,
               double tofmax, double x0, double y0, double bandwidth, double cutoff,
               int order, int normalize, double ampl, int bw_auto)
{
    g_irr.omega.clear(); g_irr.A.clear(); g_irr.k.clear(); g_irr.phase.clear(); g_irr.theta.clear();
    assign(g_irr.omega, omega, n, 0.0);
    assign(g_irr.A, A, n, 0.0);
    assign(g_irr.k, k, n, 0.0);
    assign(g_irr.phase, phase, n, 0.0);
    assign(g_irr.theta, theta, n, 0.0);
    assign(g_irr.freq_ramptime, ramptime, n, 0.0);
    assign(g_irr.freq_starttime, starttime, n, -SMALL);
    assign(g_irr.freq_duration, duration, n, HUGE);
    g_irr.nfreq = n;
    g_irr.ndir = 1;
    g_irr.depth = depth;
    g_irr.g = g;
    g_irr.rho = rho;
    g_irr.swl = swl;
    g_irr.mtheta = mtheta;
    g_irr.tofmax = tofmax;
    g_irr.fpoint[0] = x0;
    g_irr.fpoint[1] = y0;
    g_irr.dw_bandwidth = bandwidth;
    g_irr.dw_cutoff = cutoff;
    g_irr.order = order;
    g_irr.vertical_theory = 0;
    g_irr.sloping_bottom = false;
    g_irr.normalize = normalize != 0;
    g_irr.ampl = ampl;
    if (bw_auto) g_irr.dw_bandwidth = g_irr.bandwidth_estimator();
    g_irr.normalize_data();
    g_irr.cth_cache.clear();
    if (order == 2) g_irr.rebuild_second_order_cache();
    return 0;
}

void refh_get_components(double* A, double* phase, double* bandwidth) {
    for (size_t i = 0; i < g_irr.A.size(); ++i) { A[i] = g_irr.A[i]; phase[i] = g_irr.phase[i]; }
    *bandwidth = g_irr.dw_bandwidth;
}

// Ramp (src/Utils.h:37-56).  flags bit0..5: time up/down, x up/down, y up/down.
void refh_set_ramp(int enabled, int flags, const double* se /* 12 values: start,end pairs */) {
    g_ramp.ramp_init = enabled != 0;
    g_ramp.ramp_init_time_up = flags & 1;   g_ramp.time_rampup_start = se[0];  g_ramp.time_rampup_end = se[1];
    g_ramp.ramp_init_time_down = flags & 2; g_ramp.time_rampdown_start = se[2]; g_ramp.time_rampdown_end = se[3];
    g_ramp.ramp_init_x_up = flags & 4;      g_ramp.x_rampup_start = se[4];     g_ramp.x_rampup_end = se[5];
    g_ramp.ramp_init_x_down = flags & 8;    g_ramp.x_rampdown_start = se[6];   g_ramp.x_rampdown_end = se[7];
    g_ramp.ramp_init_y_up = flags & 16;     g_ramp.y_rampup_start = se[8];     g_ramp.y_rampup_end = se[9];
    g_ramp.ramp_init_y_down = flags & 32;   g_ramp.y_rampdown_start = se[10];  g_ramp.y_rampdown_end = se[11];
}

double refh_ramp(double t, double x, double y) { return g_ramp.ramp(t, x, y); }
double refh_freq_ramp(double t, double rt, double ts, double du) { return g_irr.freq_ramp(t, rt, ts, du); }

// Point evaluation of one Irregular member over n points.
// kind: 0 eta1, 1 eta2, 2 eta, 3 u1, 4 v1, 5 w1, 6 dp1, 7 phi1_dxdz, 8 phi1_dydz,
//       9 phi1_dzdz, 10 dp1_dz, 11 u2, 12 v2, 13 w2, 14 dp2, 15 u, 16 v, 17 w, 18 dp,
//       19..22 uvw2 (u, v, w, p), 23..26 u_smit, v_smit, w_smit, dp_smit
void refh_eval(int kind, long n, const double* t, const double* x, const double* y,
               const double* z, double* out)
{
    #pragma omp parallel for schedule(dynamic, 1)
    for (long p = 0; p < n; ++p) {
        const double tt = t[p], xx = x[p], yy = y[p], zz = z ? z[p] : 0.0;
        double r = 0.0;
        switch (kind) {
        case 0: r = g_irr.eta1(tt, xx, yy); break;
        case 1: r = g_irr.eta2(tt, xx, yy); break;
        case 2: r = g_irr.eta(tt, xx, yy); break;
        case 3: r = g_irr.u1(tt, xx, yy, zz); break;
        case 4: r = g_irr.v1(tt, xx, yy, zz); break;
        case 5: r = g_irr.w1(tt, xx, yy, zz); break;
        case 6: r = g_irr.dp1(tt, xx, yy, zz); break;
        case 7: r = g_irr.phi1_dxdz(tt, xx, yy); break;
        case 8: r = g_irr.phi1_dydz(tt, xx, yy); break;
        case 9: r = g_irr.phi1_dzdz(tt, xx, yy); break;
        case 10: r = g_irr.dp1_dz(tt, xx, yy); break;
        case 11: r = g_irr.u2(tt, xx, yy, zz); break;
        case 12: r = g_irr.v2(tt, xx, yy, zz); break;
        case 13: r = g_irr.w2(tt, xx, yy, zz); break;
        case 14: r = g_irr.dp2(tt, xx, yy, zz); break;
        case 15: r = g_irr.u(tt, xx, yy, zz); break;
        case 16: r = g_irr.v(tt, xx, yy, zz); break;
        case 17: r = g_irr.w(tt, xx, yy, zz); break;
        case 18: r = g_irr.dp(tt, xx, yy, zz); break;
        case 19: case 20: case 21: case 22: {
            UVWP q = g_irr.uvw2(tt, xx, yy, zz);
            r = kind == 19 ? q.u : kind == 20 ? q.v : kind == 21 ? q.w : q.p;
            break;
        }
        case 23: r = g_irr.u_smit(tt, xx, yy, zz); break;
        case 24: r = g_irr.v_smit(tt, xx, yy, zz); break;
        case 25: r = g_irr.w_smit(tt, xx, yy, zz); break;
        case 26: r = g_irr.dp_smit(tt, xx, yy, zz); break;
        default: r = 0.0;
        }
        out[p] = r;
    }
}

// Pair cache export (src/Irregular.h:24-45).
long refh_pair_count() { return (long)g_irr.pair_cache.size(); }
void refh_pair_export(int* im, unsigned char* flags, double* vals /* 7 per entry */) {
    for (size_t e = 0; e < g_irr.pair_cache.size(); ++e) {
        const PairCacheEntry& p = g_irr.pair_cache[e];
        im[2 * e] = p.i; im[2 * e + 1] = p.m;
        flags[e] = (p.has_plus ? 1 : 0) | (p.has_minus ? 2 : 0);
        double* v = vals + 7 * e;
        v[0] = p.gamma_nm; v[1] = p.k_nm_plus; v[2] = p.k_nm_minus;
        v[3] = p.D_nm_plus; v[4] = p.D_nm_minus; v[5] = p.alpha_nm_plus; v[6] = p.alpha_nm_minus;
    }
}
double refh_kmax() { return g_irr.kmax_cache; }
// vertical_theory as the parser sets it (src/CFDwavemaker.cpp:571-585).  The
// Smit members build the second-order caches themselves when absent; build
// them here, before any parallel evaluation, so no two threads do it at once.
void refh_set_vertical_theory(int vt) {
    g_irr.vertical_theory = vt;
    if ((int)g_irr.cth_cache.size() != g_irr.nfreq * g_irr.ndir) g_irr.rebuild_second_order_cache();
}
void refh_pair_export_h(double* h2) {
    for (size_t e = 0; e < g_irr.pair_cache.size(); ++e) {
        h2[2 * e] = g_irr.pair_cache[e].H_nm_plus;
        h2[2 * e + 1] = g_irr.pair_cache[e].H_nm_minus;
    }
}

// lsGridSpline set-up as in process_inputdata (src/CFDwavemaker.cpp:1180-1196,
// 1426-1437).  Returns the wall time of initialize_kinematics[_with_ignore].
double refh_grid_init(const double* domain, int nx, int ny, int nl, double t0, double dt,
                      double tan_a, double tan_b, double water_depth, double swl,
                      const double* domain_ignore, int ignore_domain, int ignore_at_init)
{
    if (g_grid_alloc) {
        delete[] g_grid.UX; delete[] g_grid.UY; delete[] g_grid.UZ; delete[] g_grid.ETA;
        delete[] g_grid.IGNORE; delete[] g_grid.ETAdx; delete[] g_grid.ETAdy; delete[] g_grid.ETAdt;
        delete[] g_grid.Udt; delete[] g_grid.Udx; delete[] g_grid.Udy; delete[] g_grid.Uds;
        delete[] g_grid.Vdt; delete[] g_grid.Vdx; delete[] g_grid.Vdy; delete[] g_grid.Vds;
        delete[] g_grid.Wdt; delete[] g_grid.Wdx; delete[] g_grid.Wdy; delete[] g_grid.Wds;
    }
    for (int q = 0; q < 4; ++q) { g_grid.domain[q] = domain[q]; g_grid.domain_ignore[q] = domain_ignore[q]; }
    g_grid.nx = nx; g_grid.ny = ny; g_grid.nl = nl;
    g_grid.t0 = t0; g_grid.dt = dt; g_grid.tan_a = tan_a; g_grid.tan_b = tan_b;
    g_grid.ignore_domain = ignore_domain != 0;
    g_grid.ignore_at_init = ignore_at_init != 0;
    g_grid.init_only = false;
    g_grid.tstep = 0;
    g_grid.allocate();
    g_grid_alloc = true;
    // zero every array so cells that the reference never writes compare as 0
    const long c = (long)nx * ny, v = c * nl;
    std::memset(g_grid.ETA, 0, 4 * c * sizeof(double)); std::memset(g_grid.ETAdt, 0, 4 * c * sizeof(double));
    std::memset(g_grid.ETAdx, 0, 2 * c * sizeof(double)); std::memset(g_grid.ETAdy, 0, 2 * c * sizeof(double));
    double* a4[6] = {g_grid.UX, g_grid.UY, g_grid.UZ, g_grid.Udt, g_grid.Vdt, g_grid.Wdt};
    for (double* p : a4) std::memset(p, 0, 4 * v * sizeof(double));
    double* a2[9] = {g_grid.Udx, g_grid.Udy, g_grid.Uds, g_grid.Vdx, g_grid.Vdy, g_grid.Vds,
                     g_grid.Wdx, g_grid.Wdy, g_grid.Wds};
    for (double* p : a2) std::memset(p, 0, 2 * v * sizeof(double));
    g_grid.water_depth = water_depth;
    g_grid.swl = swl;
    g_grid.set_ignore();
    // initialize_kinematics calls compute_mpi_partition() from every OpenMP
    // thread inside its parallel region (lsgrid_spline.cpp:690), so the first
    // std::vector::assign on mpi_bounds races between threads and can corrupt
    // the heap.  Calling it once here first sizes the vector, after which the
    // in-region calls only rewrite the same two values.
    g_grid.compute_mpi_partition();
    auto t_start = std::chrono::steady_clock::now();
    if (g_grid.ignore_at_init) g_grid.initialize_kinematics_with_ignore(g_irr);
    else g_grid.initialize_kinematics(g_irr);
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
}

// wave_force_update for wavetype 5 (src/CFDwavemaker.cpp:2253-2261).
int refh_grid_force_update(double t) {
    int n = 0;
    while (g_grid.CheckTime(t)) { g_grid.update(g_irr, t); ++n; }
    return n;
}

void refh_grid_state(int* tstep, double* t0, double* dx, double* dy, double* ds) {
    *tstep = g_grid.tstep; *t0 = g_grid.t0; *dx = g_grid.dx; *dy = g_grid.dy; *ds = g_grid.ds;
}

// which: 0 ETA(4 slots), 1 ETAdx(2), 2 ETAdy(2), 3 ETAdt(4), 4 UX, 5 UY, 6 UZ (4),
//        7 Udt, 8 Vdt, 9 Wdt (4), 10 Udx 11 Udy 12 Uds 13 Vdx 14 Vdy 15 Vds 16 Wdx 17 Wdy 18 Wds (2),
//        19 IGNORE (as double)
long refh_grid_export(int which, double* out) {
    const long c = (long)g_grid.nx * g_grid.ny, v = c * g_grid.nl;
    const double* src = nullptr; long n = 0;
    switch (which) {
    case 0: src = g_grid.ETA; n = 4 * c; break;
    case 1: src = g_grid.ETAdx; n = 2 * c; break;
    case 2: src = g_grid.ETAdy; n = 2 * c; break;
    case 3: src = g_grid.ETAdt; n = 4 * c; break;
    case 4: src = g_grid.UX; n = 4 * v; break;
    case 5: src = g_grid.UY; n = 4 * v; break;
    case 6: src = g_grid.UZ; n = 4 * v; break;
    case 7: src = g_grid.Udt; n = 4 * v; break;
    case 8: src = g_grid.Vdt; n = 4 * v; break;
    case 9: src = g_grid.Wdt; n = 4 * v; break;
    case 10: src = g_grid.Udx; n = 2 * v; break;
    case 11: src = g_grid.Udy; n = 2 * v; break;
    case 12: src = g_grid.Uds; n = 2 * v; break;
    case 13: src = g_grid.Vdx; n = 2 * v; break;
    case 14: src = g_grid.Vdy; n = 2 * v; break;
    case 15: src = g_grid.Vds; n = 2 * v; break;
    case 16: src = g_grid.Wdx; n = 2 * v; break;
    case 17: src = g_grid.Wdy; n = 2 * v; break;
    case 18: src = g_grid.Wds; n = 2 * v; break;
    case 19:
        if (out) for (long q = 0; q < c; ++q) out[q] = g_grid.IGNORE[q];
        return c;
    default: return 0;
    }
    if (out) std::memcpy(out, src, n * sizeof(double));
    return n;
}

// get_kinematics_at_point (src/lsgrid_spline.cpp:505-605) over n points -> out[4n].
void refh_grid_query(long n, const double* t, const double* x, const double* y, const double* z,
                     double* out)
{
    #pragma omp parallel for schedule(static)
    for (long p = 0; p < n; ++p) {
        std::vector<double> r = g_grid.get_kinematics_at_point(t[p], x[p], y[p], z[p], g_grid.water_depth);
        for (int q = 0; q < 4; ++q) out[4 * p + q] = r[q];
    }
}

// Sigma-map helpers (src/lsgrid_spline.cpp:89-122).
double refh_sigmaS2cart(double s, double eta, double depth) { return g_grid.sigmaS2cart(s, eta, depth); }
double refh_cart2sigmaS(double z, double eta, double depth) { return g_grid.cart2sigmaS(z, eta, depth); }

// Legacy lsGrid as the unit tests drive it (unittests/lsgrid_tests.cpp:64-83):
// allocate, set_ignore, initialize_surface_elevation[_with_ignore](t0),
// initialize_kinematics[_with_ignore].  Returns the wall time of the two fills.
double refh_lsgrid_init(const double* domain, int nx, int ny, int nl, double t0, double dt,
                        double tan_a, double tan_b, double water_depth, double swl,
                        const double* domain_ignore, int ignore_at_init, int init_only)
{
    lg_free();
    for (int q = 0; q < 4; ++q) { g_lg.domain[q] = domain[q]; g_lg.domain_ignore[q] = domain_ignore[q]; }
    g_lg.nx = nx; g_lg.ny = ny; g_lg.nl = nl;
    g_lg.t0 = t0; g_lg.dt = dt; g_lg.tan_a = tan_a; g_lg.tan_b = tan_b;
    g_lg.water_depth = water_depth; g_lg.swl = swl;
    g_lg.ignore_at_init = ignore_at_init != 0;
    g_lg.init_only = init_only != 0;
    g_lg.update_count = 0;
    g_lg.dump_vtk = false;
    g_lg.allocate();
    g_lg_alloc = true;
    g_lg.set_ignore();
    auto t_start = std::chrono::steady_clock::now();
    if (g_lg.ignore_at_init) {
        g_lg.initialize_surface_elevation_with_ignore(g_irr, t0);
        g_lg.initialize_kinematics_with_ignore(g_irr);
    } else {
        g_lg.initialize_surface_elevation(g_irr, t0);
        g_lg.initialize_kinematics(g_irr);
    }
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
}

void refh_lsgrid_update(double t_target) { g_lg.update(g_irr, t_target); }
int refh_lsgrid_check_time(double t) { return g_lg.CheckTime(t) ? 1 : 0; }
double refh_lsgrid_t0() { return g_lg.t0; }

// which: 0 ETA, 4 UX, 5 UY, 6 UZ; slot 0 then slot 1
long refh_lsgrid_export(int which, double* out) {
    const long c = (long)g_lg.nx * g_lg.ny, v = c * g_lg.nl;
    const double *s0 = nullptr, *s1 = nullptr; long n = 0;
    switch (which) {
    case 0: s0 = g_lg.ETA0; s1 = g_lg.ETA1; n = c; break;
    case 4: s0 = g_lg.UX0; s1 = g_lg.UX1; n = v; break;
    case 5: s0 = g_lg.UY0; s1 = g_lg.UY1; n = v; break;
    case 6: s0 = g_lg.UZ0; s1 = g_lg.UZ1; n = v; break;
    default: return 0;
    }
    if (out) { std::memcpy(out, s0, n * sizeof(double)); std::memcpy(out + n, s1, n * sizeof(double)); }
    return 2 * n;
}

// lsGrid::eta / u / v / w (lsgrid.cpp:360-376) over n points -> out[4n].
void refh_lsgrid_query(long n, const double* t, const double* x, const double* y, const double* z,
                       double* out)
{
    #pragma omp parallel for schedule(static)
    for (long p = 0; p < n; ++p) {
        out[4 * p] = g_lg.eta(t[p], x[p], y[p]);
        out[4 * p + 1] = g_lg.u(t[p], x[p], y[p], z[p]);
        out[4 * p + 2] = g_lg.v(t[p], x[p], y[p], z[p]);
        out[4 * p + 3] = g_lg.w(t[p], x[p], y[p], z[p]);
    }
}

// lsGrid::export_vtu (lsgrid.cpp:1354-1552) into path.
int refh_lsgrid_write_vtu(const char* path, int last) {
    FILE* fp = std::fopen(path, "w");
    if (!fp) return -1;
    g_lg.export_vtu(fp, last != 0);
    std::fclose(fp);
    return 0;
}

} // extern "C"
