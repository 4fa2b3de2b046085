// Device engine: sea states and s-grids resident in HBM, kernel launch jobs.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "sea_tables.hpp"

namespace cfdwm {

// ---------------------------------------------------------------- jobs
template <typename R>
struct ColumnJob {
    int64_t nslot = 1, ncol = 0, ncell = 0;
    const int32_t* cols = nullptr;        // GRID: active columns (null: all)
    GridGeom g{};
    double slot_t[4] = {0, 0, 0, 0};
    int slot_dst[4] = {0, 0, 0, 0};
    R* eta_grid = nullptr;
    const double *x = nullptr, *y = nullptr, *t = nullptr;
    int want_eta2 = 1;
    int pairs = 1;                        // 0: eta2 pair sum comes from the MFMA partials q2
    const float* q2 = nullptr;            // [nq2][q2stride] per-point partials (GRID)
    int nq2 = 0;
    int64_t q2stride = 0;
    double q2scale[4] = {1, 1, 1, 1};     // per slot ramp factor of the partials (GRID)
    double q2_rt = 0, q2_ts = 0, q2_du = 0;  // POINTS: the ramp triple shared by all components
    double *o_eta1 = nullptr, *o_eta2 = nullptr, *o_gx = nullptr, *o_gy = nullptr, *o_gz = nullptr, *o_gp = nullptr;
    const Segment* seg = nullptr;         // nseg segments of the filtered table (blockIdx.y)
    int nseg = 1;
    double* partial = nullptr;            // nseg > 1: [seg][thread][6] partial sums
};

template <typename R>
struct LevelJob {
    int64_t nslot = 1, ncol = 0, ncell = 0;
    const int32_t* cols = nullptr;
    GridGeom g{};
    double slot_t[4] = {0, 0, 0, 0};
    int slot_dst[4] = {0, 0, 0, 0};
    const R* eta_grid = nullptr;
    const double* sigma = nullptr;
    R *o_u = nullptr, *o_v = nullptr, *o_w = nullptr;
    const double *x = nullptr, *y = nullptr, *t = nullptr, *z = nullptr;
    double *p_u = nullptr, *p_v = nullptr, *p_w = nullptr, *p_p = nullptr;
    uint32_t flags = 0;
    const uint32_t* lvl = nullptr;        // GRID: per filtered pair record, pruning level bounds
    // GRID: pair sums from the MFMA quadratic forms instead of the pair loop:
    // u, v += scale * sum_q (e^{mu_q z} + e^{-mu_q (z + 2h)}) Q_{f,q}(column) (cosh-type
    // basis, rank qf_rank), w += scale * sum_q (e^{nu_q z} - e^{-nu_q (z + 2h)}) Q_{w,q}
    // (sinh-type basis, rank qf_rank_w)
    const float* qf = nullptr;            // [2 qf_rank + qf_rank_w][qf_stride]: u, v, w forms
    int qf_rank = 0, qf_rank_w = 0;
    int64_t qf_stride = 0;
    float qf_mu2[32] = {};                // mu_q log2 e
    float qf_mg[32] = {};                 // -2 h mu_q log2 e
    float qf_mu2w[32] = {};               // nu_q log2 e
    float qf_mgw[32] = {};                // -2 h nu_q log2 e
    double qf_scale[4] = {1, 1, 1, 1};
    // GRID, form-assisted: the first-order and diagonal terms through two
    // profile bases instead of per-level profiles (lin_c null: per level):
    // u += sum_q (e^{m_q z} + e^{-m_q (z + 2h)}) sum_i [(lu_i c_i r) a_q(k_i)
    // + (dfu_i cos 2theta_i r) a_q(2 k_i)], v likewise, w with the sinh-type
    // basis, lw_i s_i r and dfw_i sin 2theta_i r
    const float* lin_c = nullptr;         // [n][4 x 16]: a(k), a(2k), b(k), b(2k), zeros past the ranks
    const float* lin_mu = nullptr;        // [4][16]: m_q log2 e, -2 h m_q log2 e (cosh-type), the same (sinh-type)
    int lin_R = 0, lin_Rw = 0;
    const Segment* seg = nullptr;         // nseg segments of the table in use (blockIdx.y)
    int nseg = 1;
    double* partial = nullptr;            // nseg > 1: [seg][thread][3 LC + 6] partial sums
};

template <typename R>
struct PackJob {
    int nx, ny, nl, tid1, tid2;
    const R *E, *Et, *Ex, *Ey;
    const R *F[3], *Ft[3], *Fx[3], *Fy[3], *Fs[3];
    R *pk_eta, *pk_C, *pk_D;
};

// Ramp of the reference's Ramp class (Utils.cpp:26-68) for device point
// batches: axis q (time in, time out, x in, x out, y in, y out) is on when
// on[q], a linear 0..1 ramp from se[2q] to se[2q + 1] (reversed for the odd,
// "out" axes); the factor is the smallest over the axes, 1 when init is 0.
struct RampDev {
    int init = 0;
    int on[6] = {0, 0, 0, 0, 0, 0};
    double se[12] = {};
};

// Last pass of a device point batch (wave_EvalPointsDevice): with r the ramp
// factor of (t, x, y)[q],
//   eta[q] = (eadd ? (eoff + ea[q]) + (eb ? eb[q] : 0) : ea[q]) * r   (eta != null)
//   dst[c][q] = src[c] ? src[c][q] * r : 0                             (dst[c] != null)
struct PointsFinish {
    int64_t n = 0;
    RampDev ramp{};
    const double *t = nullptr, *x = nullptr, *y = nullptr;
    int eadd = 0;
    double eoff = 0;
    const double *ea = nullptr, *eb = nullptr;
    double* eta = nullptr;
    const double* src[4] = {nullptr, nullptr, nullptr, nullptr};
    double* dst[4] = {nullptr, nullptr, nullptr, nullptr};
};
hipError_t launch_points_finish(const PointsFinish& F, hipStream_t st);
// out[q] = max(lo, z[q]) in the host's std::max semantics
hipError_t launch_clamp_below(int64_t n, const double* z, double lo, double* out, hipStream_t st);

// A query in cell order: its coordinates and its index in the caller's arrays.
template <typename TQ>
struct QRec;
template <>
struct QRec<float> {
    float x, y, z;
    uint32_t p;
};
template <>
struct QRec<double> {
    double x, y, z;
    uint64_t p;
};

template <typename R, typename TQ>
struct QueryJob {
    int64_t n = 0;
    GridGeom g{};
    int tstep = 0;
    double t0 = 0, dt = 1, tan_a = 0, tan_b = 1, tanab = 1;
    const R *pk_eta = nullptr, *pk_C = nullptr, *pk_D = nullptr;
    const TQ *t = nullptr, *x = nullptr, *y = nullptr, *z = nullptr;
    double tconst = 0;
    const double* ramp_scale = nullptr;
    TQ *o_eta = nullptr, *o_u = nullptr, *o_v = nullptr, *o_w = nullptr;
    TQ* o_aos = nullptr;                  // when set: {eta, u, v, w} of query p at o_aos[4 p .. 4 p + 3]
    // when set: thread p evaluates the query srec[p] (cell order) instead of
    // (x, y, z)[p] and writes its outputs at the query's own index
    const QRec<TQ>* srec = nullptr;
};

template <typename R>
hipError_t launch_columns(const SeaDev<R>& S, const ColumnJob<R>& J, bool grid, hipStream_t st);
template <typename R>
hipError_t launch_levels(const SeaDev<R>& S, const LevelJob<R>& J, bool grid, int lc, hipStream_t st);
template <typename R>
hipError_t launch_smit(const SmitJob& J, hipStream_t st);
// thread counts of one launch (partial buffers hold nseg x threads records)
int64_t column_threads(int64_t nslot, int64_t ncol);
int64_t level_threads(int64_t nslot, int64_t ncol, int nl, int lc, bool grid);
template <typename R>
hipError_t launch_grad_xyz(const R* D, R* GX, R* GY, R* GS, int nx, int ny, int L, int tid0, int tid1,
                           double dx, double dy, double ds, hipStream_t st);
template <typename R>
hipError_t launch_grad_dt(const R* D, R* GD, int64_t per, int tstep, double dt, hipStream_t st);
template <typename R, typename TQ>
hipError_t launch_query(const QueryJob<R, TQ>& J, hipStream_t st);
template <typename R>
hipError_t launch_pack(const PackJob<R>& J, hipStream_t st);
// Cell order of a query batch (qsort.hip): key[p] = column of query p's cell,
// rec[p] = (x, y, z, p), then a radix sort of (key, rec) over key_bits bits.
// keys and recs hold 2 n entries (input half, sorted half), tmp tmp_bytes
// (query_sort_tmp_bytes).  On return recs + n holds the cell-ordered records.
template <typename TQ>
hipError_t query_sort_tmp_bytes(int64_t n, int key_bits, size_t* tmp_bytes);
template <typename TQ>
hipError_t launch_query_sort(int64_t n, const GridGeom& g, const TQ* x, const TQ* y, const TQ* z, int key_bits,
                             uint32_t* keys, QRec<TQ>* recs, void* tmp, size_t tmp_bytes, hipStream_t st);
// Z operands, then the MFMA quadratic form of the filtered eta2 pair sum
hipError_t launch_qf_zgen(const QfJob& J, hipStream_t st);
hipError_t launch_qf_forms(const QfJob& J, hipStream_t st);
hipError_t launch_qf_build(const QfBuild& B, hipStream_t st);
hipError_t launch_qf_persist(const QfJob* jobs, const QfTile* tiles, int ntiles, int blocks, int nlv, hipStream_t st);
// eta range of the columns of a fill (minus swl) -> eb[2]; then per filtered
// pair record the levels where its profile exponentials exceed e^{-T}
template <typename R>
hipError_t launch_prune(const R* eta, int64_t ncell, const int32_t* cols, int64_t ncol, int nslot,
                        const int* slot_dst, double swl, double* eb, const UvwRec<R>* rec, int64_t nrec,
                        const double* sigma, int nl, double h, uint32_t* lvl, hipStream_t st);

// Linear-field features of the learned surrogate (surrogate.hip)
struct SurrJob {
    int64_t n = 0;
    const double *x = nullptr, *y = nullptr;   // device column coordinates
    double t = 0, h = 0, kp = 0;
    int nlev = 0, fstride = 2, nx = 0;         // nx = 3 + 3 ceil(nlev / fstride) + 2
    int xstride = 0;                           // row stride of X (>= nx, >= nx + 3 with quad)
    int quad = 0;                              // also X[nx .. nx + 2] = quadrature of eta1, its x and y slopes
    double sigma[32] = {};
    float* X = nullptr;                        // [n][xstride]
    float* Lin = nullptr;                      // [n][1 + 4 nlev]
};

// ---------------------------------------------------------------- errors
struct Error {
    int code;
    std::string msg;
};
void set_error(const std::string& msg);
std::string last_error();

#define CFDWM_HIP(expr)                                                                  \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) {                                                           \
            ::cfdwm::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));        \
            return -3;                                                                    \
        }                                                                                 \
    } while (0)

// --------------------------------------------------------- device buffers
class DevBuf {
public:
    DevBuf() = default;
    ~DevBuf();
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    int alloc(size_t bytes);                 // keeps contents when large enough
    int upload(const void* src, size_t bytes, hipStream_t st);
    void* get() const { return p_; }
    size_t size() const { return n_; }
    void release();
private:
    void* p_ = nullptr;
    size_t n_ = 0;
};

// Kernel timing with HIP events on the engine stream (enabled on request).
struct KernelTimes {
    double total_ms = 0;
    int64_t count = 0;
};

class Engine {
public:
    static Engine& get();
    int init();                              // device + stream, idempotent
    hipStream_t stream() const { return stream_; }
    // a second stream (default priority) for work that overlaps the main
    // stream's, forked and joined with events; created on first use
    hipStream_t side_stream();
    // timing: record an event pair around a launch when enabled
    void prof_enable(bool on) { prof_ = on; }
    bool prof_on() const { return prof_; }
    void prof_begin(hipStream_t st = nullptr);
    void prof_end(const char* name, hipStream_t st = nullptr);
    int prof_flush();                        // resolves pending events
    std::map<std::string, KernelTimes> prof_read(bool reset);
private:
    Engine() = default;
    hipStream_t stream_ = nullptr, side_ = nullptr;
    bool ready_ = false, prof_ = false;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending_;
    hipEvent_t cur_ = nullptr;
    std::map<std::string, KernelTimes> times_;
    std::vector<hipEvent_t> pool_;
    hipEvent_t take_event();
};

// A launch timed with an event pair when timing is on; errors return -3.
#define PROF(name, call)                    \
    do {                                    \
        Engine::get().prof_begin();         \
        hipError_t pe_ = (call);            \
        Engine::get().prof_end(name);       \
        CFDWM_HIP(pe_);                     \
    } while (0)
// the same for a launch on another stream
#define PROF_ON(st, name, call)             \
    do {                                    \
        Engine::get().prof_begin(st);       \
        hipError_t pe_ = (call);            \
        Engine::get().prof_end(name, st);   \
        CFDWM_HIP(pe_);                     \
    } while (0)

enum Precision { kFP64 = 0, kFP32 = 1 };

// A sea state resident on the device.
class Sea;
int surrogate_linear(Sea& s, const SurrJob& J);
int surrogate_add(int64_t n, const float* a, const float* b, float* o);
int surrogate_set_column(int64_t n, const double* src, float* X, int stride, int col);

class Sea {
public:
    SeaTables tab;
    int prec = kFP64;
    int64_t nfilt() const { return (int64_t)tab.filt.rec.size(); }
    int upload(bool with_direct);
    template <typename R>
    const SeaDev<R>& dev() const;
    // balanced split of a pair table into nseg segments (cached on the device)
    const Segment* segments(bool direct, int nseg);
    // nseg segments of the components alone (no pair rows), for passes whose
    // pair sums come from the quadratic forms
    const Segment* comp_segments(int nseg);
    int64_t table_records(bool direct) const {
        return direct ? (int64_t)tab.direct.rec.size() : (int64_t)tab.filt.rec.size();
    }
    // bf16 hi/lo tables of the second-order quadratic forms (built on first
    // use): form 0 the filtered eta2, then the u and v forms (cos-type) of the
    // cosh-type profile basis and the w forms (sin-type) of the sinh-type one
    struct QfTables {
        int npad = 0, nform = 0, rank = 0, rank_w = 0;   // ranks of the cosh- and sinh-type profile bases
        int nlv = 3, nbarr = 4;              // bf16 products per term, table arrays per form
        double basis_err = 0, basis_tol = 0;
        std::vector<double> mu, mu_w;        // skeleton wavenumbers of the two profile bases
        // the tables: [nform][nbarr][npad x npad] tiled; with one product per
        // term (nlv 1) bp holds the 2 R + R_w velocity forms (hi parts) and betap the
        // eta2 form (hi and lo parts: the elevation keeps three products per
        // term).  They live in b / beta, or in caller memory (a batched
        // refill's arena); null while not built.
        const uint16_t* bp = nullptr;
        const uint16_t* betap = nullptr;
        DevBuf b, beta;
        size_t table_bytes = 0, eta_bytes = 0;   // device bytes of bp and betap
        // kept while the tables come and go: sizes, basis, tile ends
        DevBuf ftype;                        // [nform] int8
        DevBuf kend;                         // per 64-wide tile of i
        DevBuf zs, finv, zsw, finvw;         // the profile bases' skeleton depths and interpolation matrices
        // first-order and diagonal terms of the level pass through two more
        // profile bases over [0, 2 kmax] (lin_rank 0: per-level profiles):
        // per component a_q(k), a_q(2k) (cosh-type), b_q(k), b_q(2k) (sinh-type)
        int lin_rank = 0, lin_rank_w = 0;
        double lin_err = 0, lin_tol = 0;
        std::vector<double> lin_mu, lin_mu_w;
        DevBuf lin_c;                        // [n][4 x 16] float (zeros past the ranks)
        DevBuf lin_exp;                      // [4][16] float: the bases' exponents for the level pass
        std::vector<int32_t> kend_h;
        double cost = 0;                     // sum over i tiles of their m range (tile ordering)
    };
    // 1 and *out set: tables built (or cached); 0: the forms do not apply to
    // this sea (FP64, or no depth basis within the tolerance at rank <= 32:
    // the fills then run the pair loop); < 0: device error
    int qform_tables(const QfTables** out);
    // The sizes and basis of the tables without building them (1, 0 or < 0 as above).
    int qform_meta(const QfTables** out);
    // Build the tables into caller device memory (table_bytes + eta_bytes,
    // 256-byte aligned parts) on the library stream, without waiting; they
    // stay valid until release_qform or the memory's reuse.
    int qform_tables_into(void* mem, const QfTables** out);
    void release_qform();
    // bf16 products per term of the forms (3: hi/lo split, 1: hi only) and the
    // depth-basis tolerance (0: 1e-6 for 3 products, 1e-4 for 1 product)
    int qf_products = 3;
    double qf_tol = 0;
    bool qf_fallback = false;                // the last qform_tables call found the forms not applicable
    // vertical_theory (Irregular.h:83): 1 = Smit et al. s-coordinate
    // kinematics for u, v, w, p at points (elevation stays Sharma-Dean)
    int vertical_theory = 0;
    struct SmitDev {
        DevBuf comp, pair;                   // SmitComp[n], SmitPair[filtered records]
    };
    const SmitDev* smit_tables();            // built on first use
    DevBuf partial;                          // split-K partial sums of point batches
    DevBuf qz, qo;                           // quadratic-form operands / values of point batches
    std::mutex mu;                           // serialises the per-point scratch use
    // scratch for point batches (inputs 4n doubles, outputs 6n doubles)
    DevBuf pts_in, pts_out;
    int eval_points(int64_t n, const double* x, const double* y, const double* z, const double* t,
                    uint32_t flags, bool want_eta2, double* eta1, double* eta2, double* grads4,
                    double* u, double* v, double* w, double* p, bool device_ptrs, bool smit_always = false);
private:
    DevBuf ph_, cr_, act_, frow_, fmhi_, fuvw_, fx_, drow_, duvw_, dx_;
    std::map<std::pair<bool, int>, std::unique_ptr<DevBuf>> seg_cache_;
    std::map<int, std::unique_ptr<DevBuf>> cseg_cache_;
    std::unique_ptr<QfTables> qf_;
    int qform_build(QfTables& T, uint16_t* b, uint16_t* beta);
    bool qf_ext_ = false;                    // the tables live in caller memory
public:
    bool qform_external() const { return qf_ext_; }
private:
    std::unique_ptr<SmitDev> smit_;
    SeaDev<double> d64_{};
    SeaDev<float> d32_{};
    template <typename R>
    int upload_t(bool with_direct);
};

struct GridDesc {
    double domain[4];
    int nx, ny, nl;
    double t0, dt, tan_a, tan_b, water_depth, swl;
    double domain_ignore[4];
    int ignore_at_init;
};

// lsGridSpline resident on the device (4-slot time ring + derivatives).
class Grid {
public:
    Grid(Sea* sea, const GridDesc& d);
    ~Grid();
    // storage of the legacy two-slot grid (LsGrid): ETA and UX/UY/UZ in two
    // slots, no derivative fields; set before init
    bool legacy = false;
    int init(bool fill = true);              // allocate (+ initialize_kinematics)
    bool filled = false;                     // a fill has run (queries need the packed slots)
    int refill();                            // initialize_kinematics again (tstep back to 0)
    int force_update(double t, int* n);      // wave_force_update
    int64_t export_field(int which, double* out);
    // VTK XML unstructured grid of one time slot (lsgrid_spline.cpp:1200-1403 role)
    int write_vtu(const char* path, bool last, const char* time_label);
    // the writer of one slot at time label tv; swl_free: node z = eta +
    // sigma (h + eta) as the legacy grid's export_vtu (lsgrid.cpp:1404-1428)
    int write_vtu_slot(const char* path, int slot, double tv, const char* time_label, bool swl_free);
    // [lsgrid2vtk] dumps (lsgrid_spline.h:58-61): after every fill and update,
    // and of the following slot at the end (last)
    struct VtkDump {
        bool on = false;
        std::string dir = "./", prefix = "kinematics", label = "TimeValue";
    } vtk;
    int write_vtk_dump(bool last);
    int query(int64_t n, const double* t, const double* x, const double* y, const double* z,
              double* out4, const int* ramp_flags, const double* ramp_se);
    // SoA outputs (eta, u, v, w) or, with out4, {eta, u, v, w} per query;
    // ws (ws_bytes): caller workspace for the cell-ordering buffers instead of
    // the grid's own (query_ws_bytes gives the size)
    int query_device_f32(int64_t n, double t, const float* x, const float* y, const float* z,
                         float* eta, float* u, float* v, float* w, float* out4 = nullptr,
                         void* ws = nullptr, size_t ws_bytes = 0);
    int query_device_f64(int64_t n, const double* t, const double* x, const double* y, const double* z,
                         double* eta, double* u, double* v, double* w);
    Sea* sea;
    GridDesc d;
    GridGeom g{};
    int tstep = 0;
    double t0 = 0;
    int level_chunk = 16;
    int64_t points_per_fill() const { return 4LL * g.nx * g.ny * g.nl; }
    // quadratic forms of the last fill (measurement)
    bool qf_used = false;
    // why the last fill did or did not use the forms (QF_USED when it did)
    enum QfSkip { QF_USED = 0, QF_FP64 = 1, QF_OFF = 2, QF_NO_PAIRS = 3, QF_RAMPS = 4, QF_DEPTH = 5, QF_BASIS = 6 };
    int qf_skip = QF_FP64;
    // an FP32 fill that takes the pair loop for a reason other than FP64 or an
    // explicit switch runs about 14x slower: say so once per grid on stderr
    void warn_pair_loop(const char* why);
    bool qf_warned_ = false;
    int qf_nform = 0, qf_rank = 0, qf_rank_w = 0, qf_nlv = 3;
    double qf_flop = 0, qf_basis_err = 0;
    // shader clock of the forms launches of later fills (diagnostic)
    void qform_clock_enable(bool on) { qf_clock_on_ = on; }
    int qform_clock(double* mhz);
    // the column body (eta1 + eta2 + swl, Taylor gradients, levels) of nslot
    // time levels into the given slots, all columns or (use_ignore) those
    // outside the ignore region; no gradients, no query records
    int fill_at(int nslot, const double* slot_t, const int* slot_dst, bool use_ignore);
    void* field(int f) const { return f_[f].get(); }
    const int32_t* active_cols() const { return (const int32_t*)cols_act_d_.get(); }
    int64_t active_count() const { return nact_; }
    int precision() const { return prec; }
private:
    int prec;
    std::vector<int32_t> ignore_;
    std::vector<double> sigma_;
    DevBuf sigma_d_, cols_all_d_, cols_act_d_;
    int64_t nact_ = 0;
    DevBuf lvl_, eb_, partial_;
    bool prune_ = true;
    bool qform_ = true;                       // FP32: eta2 pair sum as an MFMA quadratic form
    bool qf_overlap_ = true;                  // velocity forms beside the column pass (CFDWM_QF_OVERLAP=0: off)
    DevBuf zbuf_;                             // its Z operands (one chunk of points)
    uint16_t* zext_ = nullptr;                // or caller memory (a batched refill's arena)
    size_t zext_bytes_ = 0;
public:
    // Z operand bytes of a full refill whose forms fit one launch (0: they do not)
    size_t refill_z_bytes() const;
    void set_z_external(uint16_t* p, size_t bytes) { zext_ = p; zext_bytes_ = bytes; }
private:
    DevBuf qout_;                             // form values [nform][points of the fill]
    QfJob qj_{};                              // prepared forms launch (first chunk): all forms, or
                                              // the eta2 form alone when qj2_ holds the velocity forms
    QfJob qj2_{};                             // velocity forms at one product per term (split fills)
    bool qf_split_ = false;
    double qf_fpp1_ = 0;                      // executed flops per point and velocity form (split fills)
    int64_t qf_chunk_ = 0, qf_P_ = 0;
    bool qf_clock_on_ = false, qf_clock_valid_ = false;
    DevBuf qf_clk_;                           // 2 counters of the clock measurement
    double qf_fpp_ = 0, qf_scale_[4] = {1, 1, 1, 1};
    bool batch_pending_ = false;
    // the velocity forms of a one-chunk split fill run on the side stream,
    // beside the column pass; the level pass joins them
    hipEvent_t qf_fork_ = nullptr, qf_join_ = nullptr;
    bool qf_join_pending_ = false;
    void qform_join();
    int qform_prepare(int nslot, const double* slot_t, int64_t ncol, const int32_t* cols);
    int qform_run();
    void qform_attach(ColumnJob<float>& C, LevelJob<float>& L);
    int qform_fill(int nslot, const double* slot_t, int64_t ncol, const int32_t* cols, ColumnJob<float>& C,
                   LevelJob<float>& L);
    void init_slot_times(double* tt) const;
public:
    // batched refill over many grids (refill_batched)
    int batch_begin();
    int batch_end();
    const QfJob& batch_job() const { return qj_; }
    bool batch_split() const { return qf_split_; }
    const QfJob& batch_job2() const { return qj2_; }
    void release_workspace();
private:
    // fields: 0 ETA(4) 1 ETAdx(2) 2 ETAdy(2) 3 ETAdt(4) 4..6 U/V/W(4) 7..9 dt(4) 10..18 dx/dy/ds(2)
    DevBuf f_[19];
    DevBuf q_in_, q_out_;
    DevBuf pk_eta_, pk_C_, pk_D_;             // query records of the two active slots
    DevBuf qs_key_, qs_rec_, qs_tmp_;         // cell-ordered query batches
    template <typename R>
    int pack();
    template <typename R>
    int fill_slots(int nslot, const double* slot_t, const int* slot_dst, bool use_ignore);
    template <typename R>
    int gradients();
    template <typename R, typename TQ>
    int run_query(int64_t n, const TQ* t, double tconst, const TQ* x, const TQ* y, const TQ* z,
                  TQ* eta, TQ* u, TQ* v, TQ* w, const double* ramp, TQ* aos = nullptr,
                  void* ws = nullptr, size_t ws_bytes = 0);
    int64_t field_count(int which) const;
};

// The legacy linear s-grid (lsGrid, src/lsgrid.h:15-114) resident on the
// device: slot 0 at t0 and slot 1 at t0 + dt of ETA (incl. swl) and UX/UY/UZ,
// filled by the same column body as the spline grid, queried by
// tri/bilinear interpolation with the swl-free sigma maps.  With init_only
// the two slots always hold the same values (the reference aliases them).
class LsGrid {
public:
    LsGrid(Sea* sea, const GridDesc& d, bool init_only);
    // initialize_surface_elevation(t0) + initialize_kinematics, the
    // _with_ignore pair when d.ignore_at_init (lsgrid.cpp:419-747)
    int init();
    // CheckTime (lsgrid.cpp:378-386): true while t <= t0 + dt
    bool check_time(double t) const { return !(t > t0 + d.dt); }
    // update (lsgrid.cpp:749-827): *reinit = 1 when t_target is more than a
    // step ahead (every column re-initialised at dt floor(t_target / dt)),
    // else one step forward for the columns outside the ignore region
    int update(double t_target, int* reinit);
    // eta (bilinear_interpolation2, :322-357) and u, v, w
    // (trilinear_interpolation2, :192-287) -> out4 = {eta, u, v, w} per point;
    // device: the five arrays are device pointers
    int query(int64_t n, const double* t, const double* x, const double* y, const double* z, double* out4,
              bool device);
    // which: 0 ETA, 4..6 UX/UY/UZ; slot 0 then slot 1 (2 x cells or nodes)
    int64_t export_field(int which, double* out) { return g->export_field(which, out); }
    // export_vtu (lsgrid.cpp:1354-1552): slot 0 at t0, or slot 1 at t0 + dt (last)
    int write_vtu(const char* path, bool last, const char* time_label);
    // write_vtk (lsgrid.cpp:1329-1351): after every fill and step while on
    Grid::VtkDump vtk;
    int write_vtk_dump(bool last);
    double t0 = 0;
    int update_count = 0;
    bool init_only;
    GridDesc d;
    std::unique_ptr<Grid> g;
private:
    int fill_both(bool use_ignore);
    int copy_slot(int src, int dst, bool active_only);
    DevBuf qin_, qout_;
    DevBuf okey_, oidx_, otmp_;              // cell-ordering buffers of large query batches
};

// Refill many s-grids (sea states) with their second-order forms evaluated
// in persistent launches over all of them (groups bounded by budget_bytes of
// device memory); launch shape and tile order picked by an autotuner and
// cached per workload size.
struct BatchStats {
    int groups = 0, batched = 0, single = 0, tiles = 0, blocks = 0, order = 0, tuned = 0;
    double forms_ms = 0, mfma_flop = 0;
    int group_forms = 0, group_ptiles = 0;   // tile grouping of the tuned launch (order 1)
    double tables_ms = 0;                    // device time of the form-table builds of the call
    double setup_ms = 0;                     // host time of the groups' set-up (incl. waiting for the builds)
};
int refill_batched(Grid** grids, int n, int64_t budget_bytes, bool release, BatchStats* stats);
// Device bytes of the cell-ordering buffers of an n-query batch on a grid of
// ncol columns (0 below the ordering threshold); 0 or a negative code.
template <typename TQ>
int query_ws_bytes(int64_t n, int ncol, size_t* bytes);
// Grid::query_device_f32 with the ordering buffers in a fixed device array
// (noalloc.hip): allocates no device memory; too large a batch gives -1.
int query_device_f32_noalloc(Grid& grid, int64_t n, double t, const float* x, const float* y, const float* z,
                             float* eta, float* u, float* v, float* w);

}  // namespace cfdwm
