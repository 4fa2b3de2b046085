/* libcfdwm: MI355X (gfx950) second-order irregular-wave kinematics behind the
 * CFDwavemaker C ABI.
 *
 * Part 1 is the reference's exported API, same names, argument order and
 * meaning (reference: src/CFDwavemaker.h:25-69), served for wave types 1
 * (direct second-order irregular waves) and 5 (irregular waves on the
 * lsgrid s-grid).  Other wave types of the reference (regular Stokes/Fenton,
 * piston wavemaker, SWD, VTK, sloping water) answer 0.0, as the reference does
 * for unsupported builds.
 *
 * Part 2 adds batched entries with no counterpart in the reference: point
 * arrays evaluated in one launch, and explicit sea/grid objects for callers
 * that do not go through waveinput.dat.
 *
 * All values are SI doubles; argument order is (x, y, z, t).  Batched entries
 * return 0 on success and a negative code on error (never exit):
 *   -1 invalid argument / state, -2 no GPU, -3 HIP runtime error;
 * cfdwm_last_error() gives the message.
 */
#ifndef CFDWM_H
#define CFDWM_H
#include <stdint.h>

#if defined(__GNUC__)
#define CFDWM_EXPORT __attribute__((visibility("default")))
#else
#define CFDWM_EXPORT
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------ Part 1: reference API */
/* src/CFDwavemaker.h:25 -- 1 after wave_Initialize succeeded */
CFDWM_EXPORT int CFDwavemaker_is_initialized(void);
/* src/CFDwavemaker.h:27, CFDwavemaker.cpp:2065-2149 -- reads ./waveinput.dat
 * (or ./constant/, ../input_files/, ../); 0 ok, -1 file not found.  Input
 * errors print a message and exit the process, as the reference does. */
CFDWM_EXPORT int wave_Initialize(void);
/* src/CFDwavemaker.h:28 -- releases device memory; returns 0 */
CFDWM_EXPORT int wave_Cleanup(void);
/* src/CFDwavemaker.h:29-31, CFDwavemaker.cpp:1543-1797 -- velocity at (x,y,z,t),
 * z clamped to >= -depth, multiplied by the ramps */
CFDWM_EXPORT double wave_VeloX(double x, double y, double z, double t);
CFDWM_EXPORT double wave_VeloY(double x, double y, double z, double t);
CFDWM_EXPORT double wave_VeloZ(double x, double y, double z, double t);
/* src/CFDwavemaker.h:32, CFDwavemaker.cpp:1800-1813 -- wave type 1 only */
CFDWM_EXPORT double wave_DynPres(double x, double y, double z, double t);
/* src/CFDwavemaker.h:33, CFDwavemaker.cpp:1815-1832 -- dp - rho g eta, type 1 */
CFDWM_EXPORT double wave_nhPres(double x, double y, double z, double t);
/* src/CFDwavemaker.h:34, CFDwavemaker.cpp:1894-1969 */
CFDWM_EXPORT double wave_SurfElev(double x, double y, double t);
/* src/CFDwavemaker.h:35, CFDwavemaker.cpp:1990-2009 */
CFDWM_EXPORT double wave_VFrac(double x, double y, double z, double t, double delta_cell);
/* src/CFDwavemaker.h:36, CFDwavemaker.cpp:1971-1986 -- -depth */
CFDWM_EXPORT double wave_Seabed(double x, double y);
/* src/CFDwavemaker.h:37, CFDwavemaker.cpp:1623-1667 -- {eta, u, v, w} x ramp.
 * The pointer is thread-local storage valid until the calling thread's next
 * call (the reference returns a fresh new[] allocation that is never freed). */
CFDWM_EXPORT double* wave_Kinematics(double x, double y, double z, double t);
/* src/CFDwavemaker.h:38 -- piston wavemaker: 0.0 here */
CFDWM_EXPORT double wave_piston_position(double t);
/* src/CFDwavemaker.h:41, CFDwavemaker.cpp:2253-2281 -- advances the s-grid
 * one dt at a time while t > t0 + dt */
CFDWM_EXPORT void wave_force_update(double t);
/* src/CFDwavemaker.h:48 -- accepted and ignored (one GPU, no MPI) */
CFDWM_EXPORT void wave_set_mpi_comm(int fortran_comm);
/* src/CFDwavemaker.h:51 -- probes are never configured by the parser: no-op */
CFDWM_EXPORT void update_probes(double t);
/* src/CFDwavemaker.h:54-58, Irregular.cpp:1372-1436 */
CFDWM_EXPORT double wave_phase_speed(int opt);
CFDWM_EXPORT double wave_mean_period(int opt);
CFDWM_EXPORT double wave_mean_length(int opt);
CFDWM_EXPORT double wave_water_depth(void);
/* src/CFDwavemaker.h:61-69, CFDwavemaker.cpp:2168-2198 -- COMFLOW aliases */
CFDWM_EXPORT double VelocityX(int i, int j, int k, double x, double y, double z, double t);
CFDWM_EXPORT double VelocityY(int i, int j, int k, double x, double y, double z, double t);
CFDWM_EXPORT double VelocityZ(int i, int j, int k, double x, double y, double z, double t);
CFDWM_EXPORT double DynamicPressure(int i, int j, int k, double x, double y, double z, double t);
CFDWM_EXPORT double SurfaceElevation(int i, int j, double x, double y, double t);
CFDWM_EXPORT double VolumeFraction(double x, double y, double z, double t, double delta_cell);
CFDWM_EXPORT int Init(void* fptr);
CFDWM_EXPORT int Prepare(double t, int mode);
CFDWM_EXPORT int Cleanup(void);

/* ------------------------------------------------ Part 2: batched entries */
/* Output selection and precision flags for wave_EvalPoints. */
#define CFDWM_OUT_ETA   0x1u
#define CFDWM_OUT_U     0x2u
#define CFDWM_OUT_V     0x4u
#define CFDWM_OUT_W     0x8u
#define CFDWM_OUT_P     0x10u
#define CFDWM_FP32      0x100u   /* FP32 VALU sums with FP64 phase reduction */

/* Batched point evaluation of the initialised wave field: for i < n,
 * eta[i] = wave_SurfElev(x[i], y[i], t[i]), u[i] = wave_VeloX(x[i], y[i], z[i], t[i]),
 * v, w likewise and p[i] = wave_DynPres(...), in one launch.  Output pointers
 * may be NULL (not computed).  Host arrays.  Returns 0 or a negative code.
 * No counterpart in the reference (which has only the per-point calls above). */
CFDWM_EXPORT int wave_EvalPoints(int64_t n, const double* x, const double* y, const double* z,
                                 const double* t, double* eta, double* u, double* v, double* w,
                                 double* p, uint32_t flags);
/* Prefetch (no reference counterpart; a micro-batcher for per-point hosts):
 * evaluate the n points (outputs selected by CFDWM_OUT_* flags) in one batch
 * and keep the results; later wave_SurfElev / wave_VeloX/Y/Z / wave_DynPres /
 * wave_Kinematics calls whose arguments are bitwise equal to a prefetched
 * point return the kept value without a GPU round trip.  Wave type 5: the
 * kept value is the uncached result bit for bit.  Wave type 1: the batch sums
 * the same terms in another order (its split over pair segments, and for FP32
 * batches of 2048 points or more the eta2 quadratic form), so kept values
 * agree with uncached ones to rounding level.  A new prefetch replaces the
 * previous one.  Call from one thread, outside concurrent per-point calls
 * (like wave_force_update). */
CFDWM_EXPORT int wave_PrefetchPoints(int64_t n, const double* x, const double* y, const double* z,
                                     const double* t, uint32_t flags);
CFDWM_EXPORT int wave_PrefetchClear(void);
CFDWM_EXPORT int wave_PrefetchStats(int64_t* hits, int64_t* misses);
/* Per-point calls on the s-grid (wave type 5; no reference counterpart, no
 * host change needed): the points asked for at one time value and grid state
 * are remembered, and the first per-point call at a new time (or after
 * wave_force_update advanced the grid) evaluates all of them at that time in
 * one batched query; later calls for those points are served from the kept
 * values under a shared lock, concurrently from any number of host threads,
 * bit-equal to the uncached values.  On by default; enable 0 switches it off
 * (returns the previous setting). */
CFDWM_EXPORT int wave_PrefetchAuto(int enable);
CFDWM_EXPORT int wave_PrefetchAutoStats(int64_t* hits, int64_t* misses, int64_t* replays,
                                        int64_t* replayed_points);
/* As wave_EvalPoints, all pointers in device memory (the call synchronises
 * the library's stream before returning).  The points never leave the
 * device: the z clamp, the batch, the elevation sum and the ramps run as
 * device passes, with the same results as wave_EvalPoints bit for bit. */
CFDWM_EXPORT int wave_EvalPointsDevice(int64_t n, const double* x, const double* y, const double* z,
                                       const double* t, double* eta, double* u, double* v, double* w,
                                       double* p, uint32_t flags);

/* Explicit objects (no waveinput.dat, no process-global state). */
typedef struct cfdwm_sea cfdwm_sea;
typedef struct cfdwm_grid cfdwm_grid;

typedef struct {
    double depth, g, rho, swl, mtheta, tofmax, x0, y0, bandwidth, cutoff, ampl;
    int order;        /* 2: second order (Irregular::order default), 0: linear */
    int normalize;    /* Irregular::normalize_data: A = ampl A / sum(A), phases 0 */
    int bw_auto;      /* Irregular::bandwidth_estimator */
    int precision;    /* 0 FP64, 1 FP32 (FP64 phase reduction), 2 FP32 whose MFMA
                         velocity forms take one bf16 product per term instead
                         of three on coarser depth bases (tolerance 1e-3;
                         velocity errors ~3e-4 to 3e-3 of the field scale at
                         N=1024); the elevation form keeps three products
                         (error ~1e-6 Hs) */
    int with_direct;  /* also build the unfiltered table of u2/v2/w2/dp2 */
} cfdwm_sea_params;

/* Components as Irregular holds them (sorted by omega).  ramptime, starttime,
 * duration may be NULL (reference defaults 0, -1e-12, 1e12). */
CFDWM_EXPORT int cfdwm_sea_create(cfdwm_sea** out, int n, const double* omega, const double* A,
                                  const double* k, const double* phase, const double* theta,
                                  const double* ramptime, const double* starttime,
                                  const double* duration, const cfdwm_sea_params* params);
CFDWM_EXPORT int cfdwm_sea_destroy(cfdwm_sea* sea);
/* A second device copy of a sea at another precision (0, 1, 2 as in
 * cfdwm_sea_params), sharing nothing with src afterwards; the host
 * coefficient tables are copied, not recomputed. */
CFDWM_EXPORT int cfdwm_sea_clone(cfdwm_sea** out, const cfdwm_sea* src, int precision);
/* vertical_theory of the sea (Irregular.h:83, waveinput.dat key
 * "vertical_theory smit", src/CFDwavemaker.cpp:571-585): 1 routes the u, v, w
 * and p of every point entry (wave_VeloX..wave_DynPres, wave_EvalPoints,
 * cfdwm_eval kinds 15-18, cfdwm_eval_fields) to the Smit et al. (2017)
 * s-coordinate members u_smit/v_smit/w_smit/dp_smit (Irregular.cpp:1623-2027);
 * the elevation stays Sharma & Dean (Irregular::eta).  0 restores Sharma &
 * Dean.  Returns 0, or -1 for a bad argument. */
CFDWM_EXPORT int cfdwm_sea_set_vertical_theory(cfdwm_sea* sea, int vertical_theory);
/* Number of second-order terms kept by the pair filters (plus+minus branches
 * of the pair cache) and of diagonal terms: the algorithmic work per point. */
CFDWM_EXPORT int cfdwm_sea_terms(cfdwm_sea* sea, int64_t* pair_terms, int64_t* diag_terms,
                                 int64_t* direct_terms);

/* Evaluation kinds, one per reference Irregular member:
 *  0 eta1  1 eta2  2 eta  3 u1  4 v1  5 w1  6 dp1  7 phi1_dxdz  8 phi1_dydz  9 phi1_dzdz
 * 10 dp1_dz  11 u2  12 v2  13 w2  14 dp2  15 u  16 v  17 w  18 dp  19 uvw2.u  20 uvw2.v
 * 21 uvw2.w  (22 uvw2.p is computed by the reference but read by no API path: -1)
 * 23 u_smit  24 v_smit  25 w_smit  26 dp_smit (Irregular.cpp:1623-2027, whatever the
 * sea's vertical_theory).
 * Host arrays, n points; z ignored by the z-independent kinds. */
CFDWM_EXPORT int cfdwm_eval(cfdwm_sea* sea, int kind, int64_t n, const double* t, const double* x,
                            const double* y, const double* z, double* out);

typedef struct {
    double domain[4];            /* xmin xmax ymin ymax */
    int nx, ny, nl;
    double t0, dt, tan_a, tan_b, water_depth, swl;
    double domain_ignore[4];     /* columns inside are skipped by updates */
    int ignore_at_init;          /* initialize_kinematics_with_ignore */
    int defer_fill;              /* 1: allocate only; the first cfdwm_grid_refill or
                                    cfdwm_grids_refill_batched fills (queries, exports
                                    and updates before that fail with -1) */
} cfdwm_grid_desc;

/* lsGridSpline: allocate and fill (initialize_kinematics). */
/* Direct-mode fields of Irregular::eta/u/v/w/dp (src/Irregular.cpp:9-157)
 * for n points in one pass (no ramps, no z clamp): eta1, eta2 (elevation =
 * swl + eta1 + eta2), u, v, w, p; NULL outputs are skipped.  device != 0: all
 * pointers are device memory. */
CFDWM_EXPORT int cfdwm_eval_fields(cfdwm_sea* sea, int64_t n, const double* t, const double* x, const double* y,
                                   const double* z, double* eta1, double* eta2, double* u, double* v, double* w,
                                   double* p, int device);
CFDWM_EXPORT int cfdwm_grid_create(cfdwm_grid** out, cfdwm_sea* sea, const cfdwm_grid_desc* desc);
CFDWM_EXPORT int cfdwm_grid_destroy(cfdwm_grid* grid);
/* Re-run initialize_kinematics (time ring back to t0). */
CFDWM_EXPORT int cfdwm_grid_refill(cfdwm_grid* grid);
/* wave_force_update semantics; *nupdates = time slices computed. */
CFDWM_EXPORT int cfdwm_grid_force_update(cfdwm_grid* grid, double t, int* nupdates);
CFDWM_EXPORT int cfdwm_grid_state(cfdwm_grid* grid, int* tstep, double* t0, double* dx, double* dy,
                                  double* ds);
/* Copy a field to host doubles (field numbering of the reference members:
 * 0 ETA 1 ETAdx 2 ETAdy 3 ETAdt 4 UX 5 UY 6 UZ 7 Udt 8 Vdt 9 Wdt 10 Udx 11 Udy
 * 12 Uds 13 Vdx 14 Vdy 15 Vds 16 Wdx 17 Wdy 18 Wds 19 IGNORE).  Returns the
 * element count (out may be NULL to ask for it). */
CFDWM_EXPORT int64_t cfdwm_grid_export(cfdwm_grid* grid, int which, double* out);
/* get_kinematics_at_point for n host points -> out4[4n] = {eta, u, v, w} per point. */
CFDWM_EXPORT int cfdwm_grid_query(cfdwm_grid* grid, int64_t n, const double* t, const double* x,
                                  const double* y, const double* z, double* out4);
/* Device-pointer query: FP32 coordinates at one time t, FP32 outputs.
 * Enqueued on the library stream; cfdwm_sync() waits for it. */
CFDWM_EXPORT int cfdwm_grid_query_device_f32(cfdwm_grid* grid, int64_t n, double t, const float* x,
                                             const float* y, const float* z, float* eta, float* u,
                                             float* v, float* w);
/* Allocation-free variant of cfdwm_grid_query_device_f32 (no counterpart in
 * the reference): for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, kernels and results; the cell-ordering buffers of batches of
 * 65,536 queries or more live in a fixed 64 MiB device array per GPU (held
 * until the process ends) instead of the grid's own buffers, so the call makes
 * no device allocation.  A batch whose ordering buffers exceed the array
 * (about a million queries) returns -1 with the maximum in the message.  All
 * calls share the GPU's array and are serialised on the library stream; work
 * on other streams that overlaps one of them must not use it. */
CFDWM_EXPORT int cfdwm_grid_query_device_f32_noalloc(cfdwm_grid* grid, int64_t n, double t, const float* x,
                                                     const float* y, const float* z, float* eta, float* u,
                                                     float* v, float* w);
/* As cfdwm_grid_query_device_f32, with the results in the layout of
 * wave_Kinematics (src/CFDwavemaker.cpp:1623-1667): out4[4 i .. 4 i + 3] =
 * {eta, u, v, w} of query i (out4 16-byte aligned).  One 16-byte store per
 * query instead of four scattered 4-byte stores. */
CFDWM_EXPORT int cfdwm_grid_kinematics_device_f32(cfdwm_grid* grid, int64_t n, double t, const float* x,
                                                  const float* y, const float* z, float* out4);
/* Second-order quadratic forms of the last FP32 fill (no reference
 * counterpart: measurement hook).  Returns 1 when the fill ran the pair sums
 * as MFMA quadratic forms (then nform, the depth-basis rank, the executed MFMA
 * flops of that fill and the basis' max error are filled in), 0 when it ran
 * the pair loop. */
CFDWM_EXPORT int cfdwm_grid_qform_info(cfdwm_grid* grid, int* nform, int* rank, double* mfma_flop,
                                       double* basis_err);
/* Product mode of the last fill's quadratic forms (no reference
 * counterpart): bf16 products per term (3: hi/lo split, 1: hi only), the
 * depth-basis tolerance, and why the forms were or were not used: 0 used,
 * 1 FP64 sea, 2 switched off (CFDWM_QFORM=0), 3 no second-order pairs,
 * 4 frequency ramps differ between components at a slot time, 5 grid depth
 * beyond the depth basis, 6 no depth basis within the tolerance (the last
 * three print a one-time warning: the pair loop is much slower).  Returns 1
 * when that fill used the forms, else 0 (products and tolerance 0). */
CFDWM_EXPORT int cfdwm_grid_qform_mode(cfdwm_grid* grid, int* products, double* basis_tol, int* reason);
/* Shader clock of the MFMA form launches (no reference counterpart: a
 * measurement aid).  Writes the mean shader clock in MHz over the blocks of
 * the last fill's form launches, if that fill was measured (returns 0; else -1
 * and 0.0), then switches the measurement on (enable != 0) or off for later
 * fills of this grid.  Each block adds its shader-clock cycles and wall-clock
 * ticks to two device counters; results of the fill do not change. */
CFDWM_EXPORT int cfdwm_grid_qform_clock(cfdwm_grid* grid, int enable, double* mhz);
/* Batched refill of many s-grids (no reference counterpart: the batched
 * workload of many sea states).  Equivalent to cfdwm_grid_refill on each grid;
 * FP32 grids have their second-order forms evaluated in persistent launches
 * over all grids of a group (groups bounded by budget_gb of device memory),
 * with the launch shape and tile order picked by an autotuner on the first
 * call per workload size (stats->tuned = 1; forms_ms then times the last
 * candidate run).  release != 0 frees each sea's form tables and each grid's
 * operand workspace after its group (they are rebuilt on the next fill). */
typedef struct {
    int groups, batched, single, tiles, blocks, order, tuned;
    double forms_ms, mfma_flop;
    int group_forms, group_ptiles;   /* tile grouping of the tuned launch (order 1) */
    double tables_ms;                /* device time of the form-table builds of the call */
    double setup_ms;                 /* host time of the groups' set-up */
} cfdwm_batch_stats;
/* Write the s-grid as an ASCII VTK XML UnstructuredGrid (.vtu) for QA in
 * ParaView (lsGridSpline::export_vtu, src/lsgrid_spline.cpp:1222-1403): the
 * slot the spline query uses as its base (last != 0: the following slot), its
 * time as field data named time_label (NULL: "TimeValue"), velocity vectors at
 * the nodes, node positions from the sigma map, hexahedral cells. */
CFDWM_EXPORT int cfdwm_grid_write_vtu(cfdwm_grid* grid, const char* path, int last, const char* time_label);
/* Legacy two-slot s-grid (lsGrid, src/lsgrid.h:15-114, src/lsgrid.cpp; the
 * reference compiles it only into its unit tests, unittests/lsgrid_tests.cpp).
 * Slot 0 holds t0, slot 1 t0 + dt: ETA (with swl) and UX/UY/UZ on the grid of
 * desc, filled by the same column body as cfdwm_grid (the sea's precision and
 * forms), queried by tri/bilinear interpolation with the swl-free sigma maps
 * (z2s / s2z, lsgrid.cpp:82-89).  desc.defer_fill is ignored. */
typedef struct cfdwm_lsgrid cfdwm_lsgrid;
/* allocate (lsgrid.cpp:23-56, every array zero), set_ignore (:1554-1573),
 * initialize_surface_elevation(t0) (:622-680) and initialize_kinematics
 * (:419-516); with desc.ignore_at_init the _with_ignore pair (:682-747,
 * :518-620: ignored columns keep eta = swl and zero velocities).  init_only:
 * slot 1 is slot 0 (the reference aliases the arrays, :466-470, :665-667), so
 * both always hold the same values. */
CFDWM_EXPORT int cfdwm_lsgrid_create(cfdwm_lsgrid** out, cfdwm_sea* sea, const cfdwm_grid_desc* desc,
                                     int init_only);
/* Frees the grid; with VTK dumps on, first writes the last dump (slot 1 at
 * t0 + dt) as the reference's destructor does (lsgrid.h:25-30). */
CFDWM_EXPORT int cfdwm_lsgrid_destroy(cfdwm_lsgrid* grid);
/* CheckTime (lsgrid.cpp:378-386): 1 while t <= t0 + dt, 0 when an update is due. */
CFDWM_EXPORT int cfdwm_lsgrid_check_time(cfdwm_lsgrid* grid, double t);
/* update (lsgrid.cpp:749-827): when t_target / dt - (t0 + 2 dt) / dt > 0,
 * every column is re-initialised at dt floor(t_target / dt) (*reinit = 1);
 * otherwise t0 += dt, and for the columns outside the ignore region slot 1
 * moves to slot 0 and slot 1 is filled at the new t0 + dt (*reinit = 0). */
CFDWM_EXPORT int cfdwm_lsgrid_update(cfdwm_lsgrid* grid, double t_target, int* reinit);
CFDWM_EXPORT int cfdwm_lsgrid_state(cfdwm_lsgrid* grid, double* t0, int* update_count);
/* which: 0 ETA, 4 UX, 5 UY, 6 UZ; slot 0 then slot 1 ([2][nx][ny] or
 * [2][nx][ny][nl]).  Returns the element count (out may be NULL). */
CFDWM_EXPORT int64_t cfdwm_lsgrid_export(cfdwm_lsgrid* grid, int which, double* out);
/* lsGrid::eta (bilinear_interpolation2, lsgrid.cpp:322-357) and u, v, w
 * (trilinear_interpolation2, :192-287) for n host points -> out4[4n] =
 * {eta, u, v, w} per point.  Fractions of a cell through float as the
 * reference's std::modf(double, float*); coordinates outside the grid clamp
 * to its edge cells, t outside [t0, t0 + dt] to the nearer slot. */
CFDWM_EXPORT int cfdwm_lsgrid_query(cfdwm_lsgrid* grid, int64_t n, const double* t, const double* x,
                                    const double* y, const double* z, double* out4);
/* The same with all six arrays in device memory (synchronises before returning). */
CFDWM_EXPORT int cfdwm_lsgrid_query_device(cfdwm_lsgrid* grid, int64_t n, const double* t, const double* x,
                                           const double* y, const double* z, double* out4);
/* export_vtu (lsgrid.cpp:1354-1552): slot 0 at t0 (last != 0: slot 1 at
 * t0 + dt), node z = eta + s (h + eta) with eta including swl as the
 * reference writes it. */
CFDWM_EXPORT int cfdwm_lsgrid_write_vtu(cfdwm_lsgrid* grid, const char* path, int last, const char* time_label);
/* dump_vtk (lsgrid.h:43-46): on != 0 writes dir + prefix + %05d + ".vtu"
 * after every fill and update (write_vtk, lsgrid.cpp:1329-1351, counter
 * update_count) and the last dump at destroy; NULL strings keep "./",
 * "kinematics", "TimeValue".  Switching dumps on writes the current state at
 * once, as the fill does when dump_vtk is set before it (:507-510), so
 * switching on right after cfdwm_lsgrid_create gives the reference's files. */
CFDWM_EXPORT int cfdwm_lsgrid_set_vtk(cfdwm_lsgrid* grid, int on, const char* dir, const char* prefix,
                                      const char* time_label);
/* Free the sea's form tables, the grid's operand workspace (rebuilt on the
 * next FP32 fill) and its query-ordering buffers (rebuilt by the next batch of
 * 65,536 or more queries); the grid's fields are kept. */
CFDWM_EXPORT int cfdwm_grid_release_workspace(cfdwm_grid* grid);
CFDWM_EXPORT int cfdwm_grids_refill_batched(cfdwm_grid** grids, int n, double budget_gb, int release,
                                            cfdwm_batch_stats* stats);
/* Skeleton-exponential depth basis e^{-kappa t} ~ sum_q c_q(kappa) e^{-mu_q t}
 * over kappa in [0, kmax], t in [0, tmax] (host only, no GPU needed): the
 * smallest rank <= rmax whose max error is <= tol.  mu and tau (capacity
 * rmax) receive the skeleton wavenumbers and depths, err the measured max
 * error.  Returns the rank, or -(rank) when tol was not reached. */
CFDWM_EXPORT int cfdwm_depth_basis(double kmax, double tmax, double tol, int rmax, double* mu, double* tau,
                                   double* err);
/* The depth bases the s-grid fills use: the pair terms' whole profiles
 * P_kappa(z) = e^{kappa z} + sign e^{-kappa (z + 2h)} (sign +1: cosh-type, u and
 * v; -1: sinh-type, w) over kappa in [0, kmax], z in [-2h, 0], approximated as
 * sum_q c_q(kappa) P_{mu_q}(z) with c = F^{-1} [P_kappa(z_s)], F[s][q] =
 * P_{mu_q}(z_s) (host only).  P_kappa(-2h - z) = sign P_kappa(z), so the
 * construction samples z in [-h, 0].  mu and zs (capacity rmax) receive the
 * skeleton wavenumbers and depths, err the measured max error.  Returns the
 * rank, -(rank) when tol was not reached (the best basis found is returned),
 * 0 for invalid arguments. */
CFDWM_EXPORT int cfdwm_profile_basis(double kmax, double h, int sign, double tol, int rmax, double* mu, double* zs,
                                     double* err);
/* ---- Learned second-order surrogate: per-column MLP (no reference
 * counterpart; the exact kernels above stay the fallback and the check).
 * L = nhidden + 1 linear layers (nin -> width -> ... -> nout), GELU (erf
 * form) after all but the last; FP32 master parameters, bf16 products with
 * FP32 accumulation on the matrix cores; inputs / outputs standardised with
 * the per-feature mean and std set by cfdwm_mlp_set_norm.  Parameter layout:
 * per layer W (dout x din, row-major) then b (dout). */
typedef struct cfdwm_mlp cfdwm_mlp;
CFDWM_EXPORT int cfdwm_mlp_create(cfdwm_mlp** out, int nin, int width, int nhidden, int nout, uint64_t seed);
CFDWM_EXPORT void cfdwm_mlp_destroy(cfdwm_mlp* mlp);
CFDWM_EXPORT int64_t cfdwm_mlp_nparams(cfdwm_mlp* mlp);
CFDWM_EXPORT int cfdwm_mlp_get_params(cfdwm_mlp* mlp, float* out);
CFDWM_EXPORT int cfdwm_mlp_set_params(cfdwm_mlp* mlp, const float* in);
CFDWM_EXPORT int cfdwm_mlp_set_norm(cfdwm_mlp* mlp, const float* in_mean, const float* in_std,
                                    const float* out_mean, const float* out_std);
/* out[2 nin + 2 nout] = in_mean, in_std, out_mean, out_std */
CFDWM_EXPORT int cfdwm_mlp_get_norm(cfdwm_mlp* mlp, float* out);
/* Y[n][nout] = net(X[n][nin]) (device pointers, raw units) */
CFDWM_EXPORT int cfdwm_mlp_forward_device(cfdwm_mlp* mlp, int64_t n, const float* X, float* Y);
/* mean squared error of the standardised outputs over the first n <= 65536
 * rows and its gradient (grad_host: nparams floats, may be NULL) */
CFDWM_EXPORT int cfdwm_mlp_loss_grad_device(cfdwm_mlp* mlp, int64_t n, const float* X, const float* Y,
                                            float* loss, float* grad_host);
/* Adam on seeded minibatches of the n rows of X / Y (device pointers);
 * cosine learning-rate decay to lr / 20; loss_curve[steps] (host, may be
 * NULL) gets the minibatch loss every `report` steps. */
CFDWM_EXPORT int cfdwm_mlp_train_device(cfdwm_mlp* mlp, int64_t n, const float* X, const float* Y, int steps,
                                        int batch, float lr, uint64_t seed, int report, float* loss_curve);
/* Linear-field features of the surrogate for n columns (x, y: device
 * doubles) at time t: X[n][3 + 3 ceil(nlev / fstride) + 2] = eta1, eta1_x,
 * eta1_y, (u1, v1, w1) at every fstride-th level, h, kp; Lin[n][1 + 4 nlev] =
 * eta1, (u1, v1, w1, p1) at the levels z_l = eta1 + sigma_l (h + eta1)
 * (nlev <= 32).  Enqueued on the library stream. */
CFDWM_EXPORT int cfdwm_surrogate_linear(cfdwm_sea* sea, int64_t n, const double* x, const double* y, double t,
                                        double h, int nlev, const double* sigma, int fstride, double kp,
                                        float* X, float* Lin);
/* As cfdwm_surrogate_linear, with X rows of xstride floats (xstride >=
 * nx + 3, nx = 3 + 3 ceil(nlev / fstride) + 2): the nx features above, then
 * the quadrature (Hilbert pair) of eta1, sum A r sin(theta), and its x and y
 * slopes; columns past nx + 3 are left to the caller. */
CFDWM_EXPORT int cfdwm_surrogate_features(cfdwm_sea* sea, int64_t n, const double* x, const double* y, double t,
                                          double h, int nlev, const double* sigma, int fstride, double kp,
                                          int xstride, float* X, float* Lin);
/* X[i * stride + col] = (float)src[i] for i < n (device pointers; enqueued on
 * the library stream): an FP64 column (the exact eta2) into the surrogate's
 * feature or output matrix. */
CFDWM_EXPORT int cfdwm_surrogate_set_column(int64_t n, const double* src, float* X, int stride, int col);
/* out[i] = a[i] + b[i] over n device floats (assembly of the surrogate fields). */
CFDWM_EXPORT int cfdwm_surrogate_add(int64_t n, const float* a, const float* b, float* out);
CFDWM_EXPORT int cfdwm_sync(void);
CFDWM_EXPORT int cfdwm_last_error(char* buf, int len);
/* Kernel timing with HIP events on the library stream. */
CFDWM_EXPORT int cfdwm_prof_enable(int on);
/* Fills up to cap entries: name (64 bytes each), total ms, launch count. */
CFDWM_EXPORT int cfdwm_prof_read(char* names, double* total_ms, int64_t* count, int cap, int reset);

#ifdef __cplusplus
}
#endif
#endif
