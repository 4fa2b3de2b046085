/* CPU oracle for the second-order irregular-wave kinematics path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is an FP64 restatement of the reference
 * CFDwavemaker algorithm (src/Irregular.cpp, src/lsgrid_spline.cpp,
 * src/lsgrid.cpp, src/Utils.cpp), used by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg as the checker.  The product library (libcfdwm) never links,
 * loads or calls it.
 *
 * Every function follows the reference's expression order so that, built
 * without -ffast-math against the same libm, it agrees with the compiled
 * reference to rounding level; tests/test_oracle_*.py pin it against golden
 * vectors written from the compiled reference.
 */
#ifndef CFDWM_ORACLE_H
#define CFDWM_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_sea or_sea;
typedef struct or_grid or_grid;

/* Evaluation kinds for or_eval (same numbering as the reference harness). */
enum {
    OR_ETA1 = 0, OR_ETA2, OR_ETA, OR_U1, OR_V1, OR_W1, OR_DP1, OR_PHI1_DXDZ, OR_PHI1_DYDZ,
    OR_PHI1_DZDZ, OR_DP1_DZ, OR_U2, OR_V2, OR_W2, OR_DP2, OR_U, OR_V, OR_W, OR_DP,
    OR_UVW2_U, OR_UVW2_V, OR_UVW2_W, OR_UVW2_P,
    OR_U_SMIT, OR_V_SMIT, OR_W_SMIT, OR_DP_SMIT
};

/* Create a sea state from component arrays (already sorted by omega, as the
 * reference parser does).  ramptime/starttime/duration may be NULL (reference
 * defaults 0, -1e-12, 1e12).  normalize/ampl follow Irregular::normalize_data;
 * bw_auto follows Irregular::bandwidth_estimator. */
or_sea* or_sea_create(int n, const double* omega, const double* A, const double* k,
                      const double* phase, const double* theta,
                      const double* ramptime, const double* starttime, const double* duration,
                      double depth, double g, double rho, double swl, double mtheta,
                      double tofmax, double x0, double y0, double bandwidth, double cutoff,
                      int order, int normalize, double ampl, int bw_auto);
void or_sea_destroy(or_sea* s);
/* Amplitudes/phases after normalize_data, and the effective bandwidth. */
void or_sea_components(const or_sea* s, double* A, double* phase, double* bandwidth);
double or_sea_kmax(const or_sea* s);
long or_pair_count(const or_sea* s);
/* vertical_theory (Irregular.h:83): 1 routes OR_U/OR_V/OR_W/OR_DP to the Smit
 * et al. s-coordinate members (Irregular.cpp:25, :53, :77, :143). */
void or_sea_set_vertical_theory(or_sea* s, int vt);
/* Smit H_nm_plus, H_nm_minus of each pair-cache entry (2 per entry). */
void or_pair_export_h(const or_sea* s, double* h2);
void or_pair_export(const or_sea* s, int* im, unsigned char* flags, double* vals7);

void or_eval(const or_sea* s, int kind, long n, const double* t, const double* x,
             const double* y, const double* z, double* out);

/* Irregular::freq_ramp and Ramp::ramp. */
double or_freq_ramp(double t, double ramptime, double tstart, double duration);
double or_ramp(int enabled, int flags, const double* se12, double t, double x, double y);

/* Spectral utilities (Irregular.cpp:1336-1453). out3 = m0, m1, m2. */
void or_spectral_moments(const or_sea* s, double* out3);
double or_mean_period(const or_sea* s, int opt);
double or_phase_speed(const or_sea* s, int opt);
double or_mean_length(const or_sea* s, int opt);

/* s-grid (lsGridSpline). */
or_grid* or_grid_create(const or_sea* s, const double* domain, int nx, int ny, int nl,
                        double t0, double dt, double tan_a, double tan_b,
                        double water_depth, double swl, const double* domain_ignore,
                        int ignore_at_init);
void or_grid_destroy(or_grid* g);
int or_grid_force_update(or_grid* g, double t);
void or_grid_state(const or_grid* g, int* tstep, double* t0, double* dx, double* dy, double* ds);
long or_grid_export(const or_grid* g, int which, double* out);
void or_grid_query(const or_grid* g, long n, const double* t, const double* x, const double* y,
                   const double* z, double* out4);
double or_sigmaS2cart(double s, double eta, double depth, double swl, double tan_a, double tan_b);
double or_cart2sigmaS(double z, double eta, double depth, double swl, double tan_a, double tan_b);

/* Legacy two-slot grid (lsGrid, src/lsgrid.h:15-114): slot 0 at t0, slot 1 at
 * t0 + dt, tri/bilinear queries.  Created as initialize_surface_elevation(t0)
 * + initialize_kinematics (the _with_ignore pair when ignore_at_init); with
 * init_only slot 1 aliases slot 0.  update returns 1 when it re-initialised
 * (a jump of more than one step), 0 after one step forward.  export: which 0
 * ETA, 4..6 UX/UY/UZ, slot 0 then slot 1.  query: out4 = {eta, u, v, w}. */
typedef struct or_lsgrid or_lsgrid;
or_lsgrid* or_lsgrid_create(const or_sea* s, const double* domain, int nx, int ny, int nl,
                            double t0, double dt, double tan_a, double tan_b,
                            double water_depth, double swl, const double* domain_ignore,
                            int ignore_at_init, int init_only);
/* A legacy grid holding given slots (eta [2][nx][ny], u/v/w [2][nx][ny][nl])
 * and no sea: for queries on fields filled elsewhere (no update). */
or_lsgrid* or_lsgrid_from_fields(const double* domain, int nx, int ny, int nl, double t0, double dt,
                                 double tan_a, double tan_b, double water_depth, double swl,
                                 const double* eta, const double* u, const double* v, const double* w);
void or_lsgrid_destroy(or_lsgrid* g);
int or_lsgrid_check_time(const or_lsgrid* g, double t);
int or_lsgrid_update(or_lsgrid* g, double t_target);
double or_lsgrid_t0(const or_lsgrid* g);
long or_lsgrid_export(const or_lsgrid* g, int which, double* out);
void or_lsgrid_query(const or_lsgrid* g, long n, const double* t, const double* x, const double* y,
                     const double* z, double* out4);

/* Fill only a subset of columns of one time level (used as the timed CPU
 * baseline "port" leg): evaluates ncol columns of the s-grid column body at
 * time t and writes eta[ncol], ux/uy/uz[ncol*nl].  Returns seconds. */
double or_column_fill(const or_sea* s, long ncol, const double* x, const double* y, double t,
                      int nl, double tan_a, double tan_b, double water_depth, double swl,
                      double* eta, double* ux, double* uy, double* uz);

#ifdef __cplusplus
}
#endif
#endif
