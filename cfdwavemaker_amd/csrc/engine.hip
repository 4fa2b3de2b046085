// Device engine: HIP stream, device-resident sea tables and s-grids,
// launch orchestration for the column / level / gradient / query kernels.
#include "engine.hpp"
#include "depth_basis.hpp"

#include <algorithm>
#include <cerrno>
#include <chrono>
#include <iostream>
#include <sys/stat.h>
#include <sys/types.h>
#include <cmath>
#include <cstring>
#include <cstdio>
#include <mutex>
#include <map>
#include <thread>

namespace cfdwm {

// ------------------------------------------------------------------ errors
namespace {
std::mutex g_err_mu;
std::string g_err;
}  // namespace
void set_error(const std::string& msg) {
    std::lock_guard<std::mutex> lk(g_err_mu);
    g_err = msg;
}
std::string last_error() {
    std::lock_guard<std::mutex> lk(g_err_mu);
    return g_err;
}

// ------------------------------------------------------------------ DevBuf
DevBuf::~DevBuf() { release(); }
void DevBuf::release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
}
int DevBuf::alloc(size_t bytes) {
    if (bytes <= n_ && p_) return 0;
    release();
    if (bytes == 0) return 0;
    CFDWM_HIP(hipMalloc(&p_, bytes));
    n_ = bytes;
    return 0;
}
int DevBuf::upload(const void* src, size_t bytes, hipStream_t st) {
    if (alloc(bytes)) return -3;
    if (bytes) CFDWM_HIP(hipMemcpyAsync(p_, src, bytes, hipMemcpyHostToDevice, st));
    return 0;
}

// ------------------------------------------------------------------ Engine
Engine& Engine::get() {
    static Engine e;
    return e;
}
int Engine::init() {
    if (ready_) return 0;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        set_error("no HIP device available (hipGetDeviceCount: " + std::string(hipGetErrorString(e)) + ")");
        return -2;
    }
    CFDWM_HIP(hipSetDevice(0));
    CFDWM_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    ready_ = true;
    return 0;
}
hipEvent_t Engine::take_event() {
    if (!pool_.empty()) {
        hipEvent_t e = pool_.back();
        pool_.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}
hipStream_t Engine::side_stream() {
    if (!side_ && hipStreamCreateWithFlags(&side_, hipStreamNonBlocking) != hipSuccess) side_ = nullptr;
    return side_;
}
void Engine::prof_begin(hipStream_t st) {
    if (!prof_) return;
    cur_ = take_event();
    if (cur_) (void)hipEventRecord(cur_, st ? st : stream_);
}
void Engine::prof_end(const char* name, hipStream_t st) {
    if (!prof_ || !cur_) return;
    hipEvent_t e = take_event();
    if (!e) return;
    (void)hipEventRecord(e, st ? st : stream_);
    pending_.push_back({name, {cur_, e}});
    cur_ = nullptr;
}
int Engine::prof_flush() {
    for (auto& p : pending_) {
        float ms = 0;
        if (hipEventSynchronize(p.second.second) == hipSuccess &&
            hipEventElapsedTime(&ms, p.second.first, p.second.second) == hipSuccess) {
            KernelTimes& k = times_[p.first];
            k.total_ms += ms;
            k.count += 1;
        }
        pool_.push_back(p.second.first);
        pool_.push_back(p.second.second);
    }
    pending_.clear();
    return 0;
}
std::map<std::string, KernelTimes> Engine::prof_read(bool reset) {
    prof_flush();
    auto r = times_;
    if (reset) times_.clear();
    return r;
}

// ------------------------------------------------------------------ Sea
namespace {
constexpr double kLog2e = 1.4426950408889634074;

template <typename R>
CompRec<R> comp_rec(const CompTable& c, int i, double h) {
    CompRec<R> r{};
    r.lu = (R)c.lu[i]; r.lv = (R)c.lv[i]; r.lw = (R)c.lw[i]; r.lp = (R)c.lp[i];
    r.gxz = (R)c.gxz[i]; r.gyz = (R)c.gyz[i]; r.gzz = (R)c.gzz[i]; r.gpz = (R)c.gpz[i];
    r.A = (R)c.A[i]; r.e2d = (R)c.e2d[i];
    r.dfu = (R)c.dfu[i]; r.dfv = (R)c.dfv[i]; r.dfw = (R)c.dfw[i];
    r.ddu = (R)c.ddu[i]; r.ddv = (R)c.ddv[i]; r.ddw = (R)c.ddw[i]; r.ddp = (R)c.ddp[i];
    r.Xx = (R)c.Xx[i]; r.Xy = (R)c.Xy[i]; r.Zk = (R)c.Zk[i]; r.kd1 = (R)c.kd1[i]; r.kd2 = (R)c.kd2[i];
    const double s = sizeof(R) == 4 ? kLog2e : 1.0;
    r.pk = (R)(c.k[i] * s);
    r.pg = (R)(-2.0 * h * c.k[i] * s);
    r.p2k = (R)(2.0 * c.k[i] * s);
    r.p2g = (R)(-4.0 * h * c.k[i] * s);
    r.eg2 = (R)(2.0 * std::exp(-2.0 * h * c.k[i]));
    return r;
}

template <typename R>
UvwRec<R> uvw_rec(const PairRec64& p, double h) {
    const double s = sizeof(R) == 4 ? kLog2e : 1.0;
    UvwRec<R> r{};
    r.cuP = (R)p.cuP; r.cvP = (R)p.cvP; r.cwP = (R)p.cwP;
    r.pkP = (R)(p.kP * s); r.pgP = (R)(-2.0 * h * p.kP * s);
    r.cuM = (R)p.cuM; r.cvM = (R)p.cvM; r.cwM = (R)p.cwM;
    r.pkM = (R)(p.kM * s); r.pgM = (R)(-2.0 * h * p.kM * s);
    return r;
}
}  // namespace

template <typename R>
int Sea::upload_t(bool with_direct) {
    Engine& E = Engine::get();
    hipStream_t st = E.stream();
    const CompTable& c = tab.comp;
    const int N = c.n;
    const double h = tab.in.depth;
    std::vector<PhaseRec> ph(N);
    std::vector<CompRec<R>> cr(N);
    std::vector<uint8_t> act((N + 3) / 4 * 4, 0);     // whole 4-byte words (read as such)
    for (int i = 0; i < N; ++i) {
        ph[i] = PhaseRec{c.k[i], c.cth[i], c.sth[i], c.omega[i], c.ph0[i], c.rt[i], c.ts[i], c.du[i]};
        cr[i] = comp_rec<R>(c, i, h);
        act[i] = c.rowActive[i];
    }
    if (ph_.upload(ph.data(), ph.size() * sizeof(PhaseRec), st)) return -3;
    if (cr_.upload(cr.data(), cr.size() * sizeof(CompRec<R>), st)) return -3;
    if (act_.upload(act.data(), act.size(), st)) return -3;
    auto put_table = [&](const PairTable& P, bool eta, DevBuf& row, DevBuf* mhi, DevBuf& uvw, DevBuf& xr) -> int {
        const int64_t ne = (int64_t)P.rec.size();
        std::vector<UvwRec<R>> u(ne > 0 ? ne : 1);
        std::vector<XRec<R>> x(ne > 0 ? ne : 1);
        for (int64_t e = 0; e < ne; ++e) {
            u[e] = uvw_rec<R>(P.rec[e], h);
            x[e] = XRec<R>{(R)P.rec[e].xP, (R)P.rec[e].xM};
        }
        (void)eta;
        if (row.upload(P.rowoff.data(), P.rowoff.size() * sizeof(int64_t), st)) return -3;
        if (mhi && mhi->upload(P.mhi.data(), P.mhi.size() * sizeof(int32_t), st)) return -3;
        if (uvw.upload(u.data(), u.size() * sizeof(UvwRec<R>), st)) return -3;
        if (xr.upload(x.data(), x.size() * sizeof(XRec<R>), st)) return -3;
        CFDWM_HIP(hipStreamSynchronize(st));   // host staging vectors go out of scope
        return 0;
    };
    if (put_table(tab.filt, true, frow_, &fmhi_, fuvw_, fx_)) return -3;
    if (with_direct && put_table(tab.direct, false, drow_, nullptr, duvw_, dx_)) return -3;
    SeaDev<R> d{};
    d.n = N;
    d.h = h; d.swl = tab.in.swl; d.x0 = tab.in.x0; d.y0 = tab.in.y0; d.rho = tab.in.rho;
    d.ph = (const PhaseRec*)ph_.get();
    d.cr = (const CompRec<R>*)cr_.get();
    d.rowActive = (const uint8_t*)act_.get();
    d.frow = (const int64_t*)frow_.get();
    d.fmhi = (const int32_t*)fmhi_.get();
    d.fuvw = (const UvwRec<R>*)fuvw_.get();
    d.fx = (const XRec<R>*)fx_.get();
    d.drow = (const int64_t*)drow_.get();
    d.duvw = (const UvwRec<R>*)duvw_.get();
    d.dx = (const XRec<R>*)dx_.get();
    if constexpr (sizeof(R) == 8) d64_ = d; else d32_ = d;
    CFDWM_HIP(hipStreamSynchronize(st));
    return 0;
}

int Sea::upload(bool with_direct) {
    if (int e = Engine::get().init()) return e;
    return prec == kFP32 ? upload_t<float>(with_direct) : upload_t<double>(with_direct);
}

const Segment* Sea::segments(bool direct, int nseg) {
    const auto key = std::make_pair(direct, nseg);
    auto it = seg_cache_.find(key);
    if (it != seg_cache_.end()) return (const Segment*)it->second->get();
    const PairTable& P = direct ? tab.direct : tab.filt;
    const int N = tab.comp.n;
    const std::vector<int64_t>& ro = P.rowoff;      // N + 1 entries, ro[N] = total
    const int64_t total = ro[N];
    // row holding record e < total: the last i with ro[i] <= e
    auto row_of = [&](int64_t e) { return (int)(std::upper_bound(ro.begin(), ro.begin() + N + 1, e) - ro.begin()) - 1; };
    auto first_at_least = [&](int64_t e) { return (int)(std::lower_bound(ro.begin(), ro.begin() + N, e) - ro.begin()); };
    std::vector<Segment> sg(nseg);
    for (int s = 0; s < nseg; ++s) {
        const int64_t e0 = total * s / nseg, e1 = total * (s + 1) / nseg;
        Segment& g = sg[s];
        g.own_lo = s == 0 ? 0 : first_at_least(e0);
        g.own_hi = s == nseg - 1 ? N : first_at_least(e1);
        if (e1 > e0) {
            g.rlo = row_of(e0);
            g.mstart = g.rlo + 1 + (int)(e0 - ro[g.rlo]);
            const int rl = row_of(e1 - 1);
            g.rhi = rl + 1;
            g.mend = rl + 1 + (int)(e1 - ro[rl]);
        } else {
            g.rlo = g.rhi = g.own_lo;
            g.mstart = g.mend = 0;
        }
        g.pad0 = g.pad1 = 0;
    }
    auto buf = std::make_unique<DevBuf>();
    if (buf->upload(sg.data(), sg.size() * sizeof(Segment), Engine::get().stream())) return nullptr;
    if (hipStreamSynchronize(Engine::get().stream()) != hipSuccess) return nullptr;
    const Segment* r = (const Segment*)buf->get();
    seg_cache_[key] = std::move(buf);
    return r;
}

const Segment* Sea::comp_segments(int nseg) {
    auto it = cseg_cache_.find(nseg);
    if (it != cseg_cache_.end()) return (const Segment*)it->second->get();
    const int N = tab.comp.n;
    std::vector<Segment> sg(nseg);
    for (int s = 0; s < nseg; ++s) {
        Segment& g = sg[s];
        g.own_lo = (int)((int64_t)N * s / nseg);
        g.own_hi = (int)((int64_t)N * (s + 1) / nseg);
        g.rlo = g.rhi = g.own_lo;
        g.mstart = g.mend = 0;
        g.pad0 = g.pad1 = 0;
    }
    auto buf = std::make_unique<DevBuf>();
    if (buf->upload(sg.data(), sg.size() * sizeof(Segment), Engine::get().stream())) return nullptr;
    if (hipStreamSynchronize(Engine::get().stream()) != hipSuccess) return nullptr;
    const Segment* r = (const Segment*)buf->get();
    cseg_cache_[nseg] = std::move(buf);
    return r;
}

namespace {
// Irregular::freq_ramp (Irregular.cpp:200-204), host side
double freq_ramp_host(double t, double rt, double ts, double du) {
    const double a = 1. - ((t - ts - du + rt) / (rt + 1e-12));
    const double b = (t - ts) / (rt + 1e-12);
    const double down = a < 1. ? (a < 0. ? 0. : a) : 1.;
    const double up = b < 1. ? (b < 0. ? 0. : b) : 1.;
    return up * down;
}
}  // namespace

// Tables of the quadratic forms over the filtered pairs (i, m), m in
// [i+1, mhi[i]), split into bf16 hi + lo in the tiled operand layout (rows i,
// k = m):  form 0 (eta2):  B1 = xP + xM, B2 = xM - xP;  per rank q of the
// cosh-type profile basis (coefficients aP = a_q(kP), aM = a_q(kM)):
// u: B1 = cuP aP + cuM aM, B2 = cuM aM - cuP aP;  v likewise with cv;  per
// rank q of the sinh-type basis (bP, bM):  w (sin-type): B1 = cwP bP + cwM bM,
// B2 = cwP bP - cwM bM.
int Sea::qform_meta(const QfTables** out) {
    *out = nullptr;
    if (qf_) { *out = qf_.get(); return 1; }
    if (prec != kFP32) return 0;
    if (qf_fallback) return 0;
    const PairTable& P = tab.filt;
    const int N = tab.comp.n;
    const double h = tab.in.depth;
    double kmx = 0;
    for (const PairRec64& r : P.rec) kmx = std::max(kmx, std::max(std::fabs(r.kP), std::fabs(r.kM)));
    int nlv = qf_products == 1 ? 1 : 3;
    if (const char* e = std::getenv("CFDWM_QF_PRODUCTS")) nlv = std::atoi(e) == 1 ? 1 : 3;
    // The bases' error reaches the fields at about a fifth of the tolerance
    // (relative to the field scale): 1e-6 keeps the FP32 mode's ~1e-6; at one
    // bf16 product per term the rounding gives ~1e-3 and 1e-3 adds ~2e-4 to it
    double tol = qf_tol > 0 ? qf_tol : (nlv == 1 ? 1e-3 : 1e-6);
    if (const char* e = std::getenv("CFDWM_QF_TOL")) tol = std::atof(e);
    ProfileBasis BC, BS;
    const double kb = kmx * (1.0 + 1e-9) + 1e-12;
    const int st_c = build_profile_basis_cached(kb, h, +1, tol, 32, BC);
    const int st_s = build_profile_basis_cached(kb, h, -1, tol, 32, BS);
    const int R = BC.R, Rw = BS.R;
    // no basis within the tolerance at rank <= 32: the fills take the pair loop
    qf_fallback = st_c != 0 || st_s != 0 || !(BC.err <= tol) || !(BS.err <= tol) || R < 1 || R > 32 || Rw < 1 ||
                  Rw > 32;
    if (qf_fallback) return 0;
    const int nform = 1 + 2 * R + Rw;
    const int nbarr = nlv == 1 ? 2 : 4;
    const int npad = (N + 127) / 128 * 128;
    const size_t nn = (size_t)npad * npad;
    std::vector<int8_t> ftype(nform, 0);
    for (int q = 0; q < Rw; ++q) ftype[1 + 2 * R + q] = 1;
    std::vector<int32_t> kend(npad / 64, 0);
    for (int i = 0; i < N; ++i) {
        int& ke = kend[i / 64];
        ke = std::max(ke, std::min(npad, (P.mhi[i] + 15) / 16 * 16));
    }
    auto T = std::make_unique<QfTables>();
    T->npad = npad; T->nform = nform; T->rank = R; T->rank_w = Rw; T->mu = BC.mu; T->mu_w = BS.mu;
    T->basis_err = std::max(BC.err, BS.err);
    // The level pass's first-order and diagonal terms through profile bases
    // over [0, 2 kmax] (the diagonal's profile is that of 2 k_i): by default in
    // the one-product (bf16) mode, whose velocity error (~1e-3 of the field
    // scale, from the rounding of the pair forms) dwarfs the 1e-5 they add;
    // CFDWM_LIN_TOL sets the tolerance (0: off, per-level profiles).
    double lin_tol = nlv == 1 ? 1e-5 : 0.0;
    if (const char* e = std::getenv("CFDWM_LIN_TOL")) lin_tol = std::atof(e);
    if (lin_tol > 0) {
        double kl = 0;
        for (int i = 0; i < N; ++i) kl = std::max(kl, 2.0 * std::fabs(tab.comp.k[i]));
        ProfileBasis LC, LS;
        const double klb = kl * (1.0 + 1e-9) + 1e-12;
        const int sc = build_profile_basis_cached(klb, h, +1, lin_tol, 16, LC);
        const int ss = build_profile_basis_cached(klb, h, -1, lin_tol, 16, LS);
        if (sc == 0 && ss == 0 && LC.R >= 1 && LC.R <= 16 && LS.R >= 1 && LS.R <= 16) {
            // per component 4 x 16 floats: a(k), a(2k), b(k), b(2k), zeros past the ranks
            const int R1 = LC.R, R2 = LS.R, ls = 4 * 16;
            std::vector<float> tabc((size_t)N * ls, 0.0f);
            double ca[64], cb[64];
            for (int i = 0; i < N; ++i) {
                const double k = tab.comp.k[i];
                float* o = &tabc[(size_t)i * ls];
                LC.coeffs(k, ca);
                LC.coeffs(2.0 * k, cb);
                for (int q = 0; q < R1; ++q) { o[q] = (float)ca[q]; o[16 + q] = (float)cb[q]; }
                LS.coeffs(k, ca);
                LS.coeffs(2.0 * k, cb);
                for (int q = 0; q < R2; ++q) { o[32 + q] = (float)ca[q]; o[48 + q] = (float)cb[q]; }
            }
            constexpr double kL2e = 1.4426950408889634074;
            float mut[64] = {};
            for (int q = 0; q < R1; ++q) { mut[q] = (float)(LC.mu[q] * kL2e); mut[16 + q] = (float)(-2.0 * h * LC.mu[q] * kL2e); }
            for (int q = 0; q < R2; ++q) { mut[32 + q] = (float)(LS.mu[q] * kL2e); mut[48 + q] = (float)(-2.0 * h * LS.mu[q] * kL2e); }
            if (T->lin_c.upload(tabc.data(), tabc.size() * sizeof(float), Engine::get().stream()) ||
                T->lin_exp.upload(mut, sizeof(mut), Engine::get().stream()))
                return -3;
            CFDWM_HIP(hipStreamSynchronize(Engine::get().stream()));   // the staging vector goes out of scope
            T->lin_rank = R1; T->lin_rank_w = R2;
            T->lin_mu = LC.mu; T->lin_mu_w = LS.mu;
            T->lin_err = std::max(LC.err, LS.err);
            T->lin_tol = lin_tol;
        }
    }
    T->basis_tol = tol; T->nlv = nlv; T->nbarr = nbarr;
    T->kend_h = kend;
    for (size_t t = 0; t < kend.size(); ++t) T->cost += std::max(0, kend[t] - (int)(64 * t));
    // one product per term: the velocity forms (hi parts) in one buffer, the
    // eta2 form (hi and lo parts, three products) in another
    const bool split = nlv == 1;
    T->table_bytes = (size_t)(split ? nform - 1 : nform) * nbarr * nn * sizeof(uint16_t);
    T->eta_bytes = split ? 4 * nn * sizeof(uint16_t) : 0;
    hipStream_t st = Engine::get().stream();
    if (T->zs.upload(BC.zs.data(), R * sizeof(double), st) ||
        T->finv.upload(BC.finv.data(), (size_t)R * R * sizeof(double), st) ||
        T->zsw.upload(BS.zs.data(), Rw * sizeof(double), st) ||
        T->finvw.upload(BS.finv.data(), (size_t)Rw * Rw * sizeof(double), st) ||
        T->ftype.upload(ftype.data(), ftype.size(), st) ||
        T->kend.upload(kend.data(), kend.size() * sizeof(int32_t), st))
        return -3;
    CFDWM_HIP(hipStreamSynchronize(st));       // host staging vectors go out of scope
    qf_ = std::move(T);
    *out = qf_.get();
    return 1;
}

// Zero the table memory and launch the table build into it (library stream).
int Sea::qform_build(QfTables& T, uint16_t* b, uint16_t* beta) {
    hipStream_t st = Engine::get().stream();
    const bool split = T.nlv == 1;
    if (split) CFDWM_HIP(hipMemsetAsync(beta, 0, T.eta_bytes, st));
    CFDWM_HIP(hipMemsetAsync(b, 0, T.table_bytes, st));
    QfBuild Bq;
    Bq.n = tab.comp.n; Bq.npad = T.npad; Bq.R = T.rank; Bq.Rw = T.rank_w; Bq.nform = T.nform; Bq.nbarr = T.nbarr;
    Bq.h = tab.in.depth;
    Bq.frow = d32_.frow; Bq.fmhi = d32_.fmhi; Bq.uvw = d32_.fuvw; Bq.x = d32_.fx;
    Bq.zs = (const double*)T.zs.get(); Bq.finv = (const double*)T.finv.get();
    Bq.zsw = (const double*)T.zsw.get(); Bq.finvw = (const double*)T.finvw.get();
    Bq.b = b;
    Bq.beta = split ? beta : nullptr;
    if (Bq.n > 0) PROF("k_qf_build", launch_qf_build(Bq, st));
    T.bp = b;
    T.betap = split ? beta : nullptr;
    return 0;
}

int Sea::qform_tables(const QfTables** out) {
    const QfTables* M = nullptr;
    const int rc = qform_meta(&M);
    *out = nullptr;
    if (rc != 1) return rc;
    QfTables& T = *qf_;
    if (!T.bp) {
        if (T.b.alloc(T.table_bytes) || (T.eta_bytes && T.beta.alloc(T.eta_bytes))) return -3;
        if (int e = qform_build(T, (uint16_t*)T.b.get(), (uint16_t*)T.beta.get())) return e;
        CFDWM_HIP(hipStreamSynchronize(Engine::get().stream()));
    }
    *out = &T;
    return 1;
}

int Sea::qform_tables_into(void* mem, const QfTables** out) {
    const QfTables* M = nullptr;
    const int rc = qform_meta(&M);
    *out = nullptr;
    if (rc != 1) return rc;
    QfTables& T = *qf_;
    if (!T.bp) {
        char* m = (char*)mem;
        const size_t tb = (T.table_bytes + 255) / 256 * 256;
        if (int e = qform_build(T, (uint16_t*)m, (uint16_t*)(m + tb))) return e;
        qf_ext_ = true;
    }
    *out = &T;
    return 1;
}

// Drop the tables (owned or in caller memory); the sizes and basis stay.
void Sea::release_qform() {
    if (!qf_) return;
    qf_->b.release();
    qf_->beta.release();
    qf_->bp = nullptr;
    qf_->betap = nullptr;
    qf_ext_ = false;
}

const Sea::SmitDev* Sea::smit_tables() {
    if (smit_) return smit_.get();
    std::vector<SmitComp> comp;
    std::vector<SmitPair> pair;
    build_smit(tab, comp, pair);
    auto T = std::make_unique<SmitDev>();
    hipStream_t st = Engine::get().stream();
    if (T->comp.upload(comp.data(), std::max<size_t>(1, comp.size()) * sizeof(SmitComp), st)) return nullptr;
    if (T->pair.upload(pair.data(), std::max<size_t>(1, pair.size()) * sizeof(SmitPair), st)) return nullptr;
    if (hipStreamSynchronize(st) != hipSuccess) return nullptr;
    smit_ = std::move(T);
    return smit_.get();
}

namespace {
// Segments so that about 256 CUs x 4 SIMDs x 4 waves of threads are busy,
// with at least 64 pair records per segment.
int pick_nseg(int64_t threads, int64_t nrec) {
    if (const char* e = std::getenv("CFDWM_NSEG")) return std::max(1, std::atoi(e));
    const int64_t target = 256LL * 4 * 4 * 64;
    int64_t s = (target + threads - 1) / std::max<int64_t>(threads, 1);
    s = std::min<int64_t>(s, std::max<int64_t>(1, nrec / 64));
    s = std::min<int64_t>(s, 4096);
    return (int)std::max<int64_t>(1, s);
}
}  // namespace

template <>
const SeaDev<double>& Sea::dev<double>() const { return d64_; }
template <>
const SeaDev<float>& Sea::dev<float>() const { return d32_; }

namespace {
// FP32 point batches: the filtered eta2 pair sum as the MFMA quadratic form
// (form 0 of the sea's tables) when the batch is large and every component
// has the same frequency-ramp triple (the ramp factor is then one number per
// point, applied by the column kernel).  Returns 1 when used.
int points_eta2_form(Sea& s, int64_t n, const double* dx, const double* dy, const double* dt, bool want_eta2,
                     ColumnJob<float>& J) {
    const char* e = std::getenv("CFDWM_QFORM");
    if (!want_eta2 || n < 2048 || (e && std::atoi(e) == 0) || s.nfilt() == 0) return 0;
    const CompTable& c = s.tab.comp;
    for (int i = 1; i < c.n; ++i)
        if (c.rt[i] != c.rt[0] || c.ts[i] != c.ts[0] || c.du[i] != c.du[0]) return 0;
    const Sea::QfTables* T = nullptr;
    if (const int rc = s.qform_tables(&T); rc != 1) return rc;
    const int npad = T->npad;
    const int64_t ppad_all = (n + 127) / 128 * 128;
    const int64_t cap = std::max<int64_t>(128, ((int64_t)1 << 31) / ((int64_t)npad * 8) / 128 * 128);
    const int64_t chunk = std::min(ppad_all, cap);
    const size_t zn = (size_t)chunk * npad;
    if (s.qz.alloc(zn * 4 * sizeof(uint16_t)) || s.qo.alloc((size_t)ppad_all * sizeof(float))) return -3;
    QfJob Q;
    Q.n = c.n; Q.npad = npad;
    Q.grid = 0; Q.x = dx; Q.y = dy; Q.t = dt;
    Q.x0 = s.tab.in.x0; Q.y0 = s.tab.in.y0;
    Q.ph = s.dev<float>().ph;
    uint16_t* z = (uint16_t*)s.qz.get();
    Q.zc_hi = z; Q.zc_lo = z + zn; Q.zs_hi = z + 2 * zn; Q.zs_lo = z + 3 * zn;
    Q.zn = (int64_t)zn;
    // eta2 only, with three products per term in every mode
    Q.b = T->nlv == 1 ? T->betap : T->bp;
    Q.nbarr = 4; Q.nlv = 3;
    Q.ftype = (const int8_t*)T->ftype.get();
    Q.nform = 1;
    Q.kend = (const int32_t*)T->kend.get();
    Q.out = (float*)s.qo.get();
    Q.ostride = ppad_all;
    hipStream_t st = Engine::get().stream();
    for (int64_t p0 = 0; p0 < n; p0 += chunk) {
        Q.p0 = p0;
        Q.P = std::min(chunk, n - p0);
        Q.ppad = (Q.P + 127) / 128 * 128;
        PROF("k_qf_zgen", launch_qf_zgen(Q, st));
        PROF("k_qf_forms", launch_qf_forms(Q, st));
    }
    J.pairs = 0;
    J.q2 = (const float*)s.qo.get();
    J.nq2 = 1;
    J.q2stride = ppad_all;
    J.q2_rt = c.rt[0]; J.q2_ts = c.ts[0]; J.q2_du = c.du[0];
    return 1;
}

template <typename R>
int run_points(Sea& s, int64_t n, const double* dx, const double* dy, const double* dz, const double* dt,
               uint32_t flags, bool want_cols, bool want_eta2, bool want_lev, double* o) {
    hipStream_t st = Engine::get().stream();
    const SeaDev<R>& S = s.dev<R>();
    if (want_cols) {
        ColumnJob<R> J;
        J.nslot = 1; J.ncol = n;
        J.x = dx; J.y = dy; J.t = dt;
        J.want_eta2 = want_eta2 ? 1 : 0;
        J.o_eta1 = o; J.o_eta2 = o + n; J.o_gx = o + 2 * n; J.o_gy = o + 3 * n; J.o_gz = o + 4 * n; J.o_gp = o + 5 * n;
        J.nseg = want_eta2 ? pick_nseg(n, s.table_records(false)) : 1;
        if constexpr (sizeof(R) == 4) {
            const int rc = points_eta2_form(s, n, dx, dy, dt, want_eta2, J);
            if (rc < 0) return rc;
            if (rc == 1) J.nseg = 1;
        }
        J.seg = s.segments(false, J.nseg);
        if (!J.seg) return -3;
        if (J.nseg > 1) {
            if (s.partial.alloc((size_t)J.nseg * n * 6 * sizeof(double))) return -3;
            J.partial = (double*)s.partial.get();
        }
        PROF("k_columns_points", launch_columns<R>(S, J, false, st));
    }
    if (want_lev) {
        LevelJob<R> J;
        J.nslot = 1; J.ncol = n;
        J.x = dx; J.y = dy; J.z = dz; J.t = dt;
        J.flags = flags;
        J.p_u = o + 6 * n; J.p_v = o + 7 * n; J.p_w = o + 8 * n; J.p_p = o + 9 * n;
        const bool direct = (flags & EV_DIRECT) != 0;
        J.nseg = (flags & EV_SEC) ? pick_nseg(n, s.table_records(direct)) : 1;
        J.seg = s.segments(direct, J.nseg);
        if (!J.seg) return -3;
        if (J.nseg > 1) {
            if (s.partial.alloc((size_t)J.nseg * n * 9 * sizeof(double))) return -3;
            J.partial = (double*)s.partial.get();
        }
        PROF("k_levels_points", launch_levels<R>(S, J, false, 1, st));
    }
    return 0;
}

// Smit et al. u, v, w, p at points (smit.hip); o[0..n) and o[n..2n) hold the
// points' eta1 and eta2, the outputs go to o[6n..10n).
template <typename R>
int run_smit(Sea& s, int64_t n, const double* dx, const double* dy, const double* dz, const double* dt, double* o) {
    hipStream_t st = Engine::get().stream();
    const Sea::SmitDev* T = s.smit_tables();
    if (!T) { set_error("Smit coefficient tables: device allocation failed"); return -3; }
    const SeaDev<R>& S = s.dev<R>();
    const SeaInput& in = s.tab.in;
    SmitJob J;
    J.n = n; J.x = dx; J.y = dy; J.z = dz; J.t = dt;
    J.eta1 = o; J.eta2 = o + n;
    J.order = in.order; J.N = S.n;
    J.h = in.depth; J.g = in.g; J.rho = in.rho; J.swl = in.swl; J.x0 = in.x0; J.y0 = in.y0;
    J.ph = S.ph; J.frow = S.frow; J.fmhi = S.fmhi;
    J.sc = (const SmitComp*)T->comp.get();
    J.sp = (const SmitPair*)T->pair.get();
    J.nseg = in.order >= 1 ? pick_nseg(n, s.table_records(false)) : 1;
    J.seg = s.segments(false, J.nseg);
    if (!J.seg) return -3;
    if (s.partial.alloc((size_t)J.nseg * 5 * n * sizeof(double))) return -3;
    J.part = (double*)s.partial.get();
    J.o_u = o + 6 * n; J.o_v = o + 7 * n; J.o_w = o + 8 * n; J.o_p = o + 9 * n;
    PROF("k_smit", launch_smit<R>(J, st));
    return 0;
}
}  // namespace

int Sea::eval_points(int64_t n, const double* x, const double* y, const double* z, const double* t,
                     uint32_t flags, bool want_eta2, double* eta1, double* eta2, double* grads4,
                     double* u, double* v, double* w, double* p, bool device_ptrs, bool smit_always) {
    if (n <= 0) return 0;
    std::lock_guard<std::mutex> lk(mu);
    hipStream_t st = Engine::get().stream();
    const bool want_cols = eta1 || eta2 || grads4;
    const bool want_lev = u || v || w || p;
    const bool smit = (vertical_theory == 1 || smit_always) && want_lev;
    if (want_lev && !smit && (flags & EV_DIRECT) && tab.direct.rowoff.empty()) {
        set_error("the unfiltered (direct) pair table was not built for this sea");
        return -1;
    }
    const double *dx = x, *dy = y, *dz = z, *dt = t;
    if (!device_ptrs) {
        if (pts_in.alloc(4 * n * sizeof(double))) return -3;
        double* b = (double*)pts_in.get();
        CFDWM_HIP(hipMemcpyAsync(b, x, n * sizeof(double), hipMemcpyHostToDevice, st));
        CFDWM_HIP(hipMemcpyAsync(b + n, y, n * sizeof(double), hipMemcpyHostToDevice, st));
        CFDWM_HIP(hipMemcpyAsync(b + 3 * n, t, n * sizeof(double), hipMemcpyHostToDevice, st));
        if (z) CFDWM_HIP(hipMemcpyAsync(b + 2 * n, z, n * sizeof(double), hipMemcpyHostToDevice, st));
        else CFDWM_HIP(hipMemsetAsync(b + 2 * n, 0, n * sizeof(double), st));
        dx = b; dy = b + n; dz = b + 2 * n; dt = b + 3 * n;
    }
    if (pts_out.alloc(10 * n * sizeof(double))) return -3;
    double* o = (double*)pts_out.get();
    int rc;
    if (smit) {
        // the stretch height eta1 + eta2 of every point, then the Smit kernels
        const bool e2 = tab.in.order >= 1;
        rc = prec == kFP32 ? run_points<float>(*this, n, dx, dy, dz, dt, 0, true, e2, false, o)
                           : run_points<double>(*this, n, dx, dy, dz, dt, 0, true, e2, false, o);
        if (!rc)
            rc = prec == kFP32 ? run_smit<float>(*this, n, dx, dy, dz, dt, o)
                               : run_smit<double>(*this, n, dx, dy, dz, dt, o);
    } else {
        rc = prec == kFP32 ? run_points<float>(*this, n, dx, dy, dz, dt, flags, want_cols, want_eta2, want_lev, o)
                           : run_points<double>(*this, n, dx, dy, dz, dt, flags, want_cols, want_eta2, want_lev, o);
    }
    if (rc) return rc;
    const hipMemcpyKind kind = device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    auto get = [&](double* dst, int slot) -> int {
        if (dst) CFDWM_HIP(hipMemcpyAsync(dst, o + slot * n, n * sizeof(double), kind, st));
        return 0;
    };
    if (get(eta1, 0) || get(eta2, 1) || get(u, 6) || get(v, 7) || get(w, 8) || get(p, 9)) return -3;
    if (grads4)
        for (int q = 0; q < 4; ++q)
            if (get(grads4 + q * n, 2 + q)) return -3;
    CFDWM_HIP(hipStreamSynchronize(st));
    return 0;
}

// ------------------------------------------------------------------ Grid
Grid::Grid(Sea* s, const GridDesc& desc) : sea(s), d(desc), prec(s->prec) {}
Grid::~Grid() {
    qform_join();
    if (qf_fork_) (void)hipEventDestroy(qf_fork_);
    if (qf_join_) (void)hipEventDestroy(qf_join_);
}

// the main stream waits for the velocity forms on the side stream
void Grid::qform_join() {
    if (!qf_join_pending_) return;
    (void)hipStreamWaitEvent(Engine::get().stream(), qf_join_, 0);
    qf_join_pending_ = false;
}

int64_t Grid::field_count(int which) const {
    const int64_t c = (int64_t)g.nx * g.ny, v = c * g.nl;
    if (legacy) return which == 0 ? 2 * c : (which >= 4 && which <= 6 ? 2 * v : 0);
    switch (which) {
    case 0: case 3: return 4 * c;
    case 1: case 2: return 2 * c;
    case 4: case 5: case 6: case 7: case 8: case 9: return 4 * v;
    default: return which >= 10 && which <= 18 ? 2 * v : 0;
    }
}

int Grid::init(bool fill) {
    if (d.nx < 1 || d.ny < 1 || d.nl < 1) {
        set_error("s-grid needs nx, ny, nl >= 1");
        return -1;
    }
    // set_ignore / initialize_kinematics geometry (lsgrid_spline.cpp:1406-1425, 672-675)
    g.nx = d.nx; g.ny = d.ny; g.nl = d.nl;
    g.x0 = d.domain[0]; g.y0 = d.domain[2];
    g.dx = (d.domain[1] - d.domain[0]) / std::max(1., double(d.nx - 1));
    g.dy = (d.domain[3] - d.domain[2]) / std::max(1., double(d.ny - 1));
    g.ds = 1. / std::max(1., double(d.nl - 1));
    g.h = d.water_depth;
    g.swl = d.swl;
    ignore_.assign((size_t)d.nx * d.ny, 0);
    std::vector<int32_t> act;
    for (int i = 0; i < d.nx; ++i) {
        const double xp = d.domain[0] + g.dx * i;
        for (int j = 0; j < d.ny; ++j) {
            const double yp = d.domain[2] + g.dy * j;
            const bool ig = xp >= d.domain_ignore[0] && xp <= d.domain_ignore[1] &&
                            yp >= d.domain_ignore[2] && yp <= d.domain_ignore[3];
            ignore_[(size_t)i * d.ny + j] = ig ? 1 : 0;
            if (!ig) act.push_back(i * d.ny + j);
        }
    }
    nact_ = (int64_t)act.size();
    // sigma of each level: sigmaS2cart (lsgrid_spline.cpp:114-119), host libm
    sigma_.assign(d.nl, 0.0);
    for (int m = 0; m < d.nl; ++m) {
        const double s = g.ds * m;
        sigma_[m] = -std::pow(std::tan(-(s - 1) * d.tan_a), d.tan_b) / std::pow(std::tan(d.tan_a), d.tan_b);
    }
    hipStream_t st = Engine::get().stream();
    if (sigma_d_.upload(sigma_.data(), sigma_.size() * sizeof(double), st)) return -3;
    if (cols_act_d_.upload(act.data(), std::max<size_t>(1, act.size()) * sizeof(int32_t), st)) return -3;
    const size_t esz = prec == kFP32 ? 4 : 8;
    for (int f = 0; f < 19; ++f) {
        const size_t b = (size_t)field_count(f) * esz;
        if (b == 0) continue;
        if (f_[f].alloc(b)) return -3;
        CFDWM_HIP(hipMemsetAsync(f_[f].get(), 0, b, st));
    }
    // levels per thread: FP32 8 (124 VGPRs, 4 waves/SIMD), FP64 4 (139 VGPRs, 3 waves/SIMD)
    const int cap = prec == kFP32 ? 8 : 4;
    level_chunk = 1;
    while (level_chunk < d.nl && level_chunk < cap) level_chunk *= 2;
    if (const char* e = std::getenv("CFDWM_LEVEL_CHUNK")) {
        // only the compiled chunk sizes (buffers are sized for the chunk the kernel uses)
        const int v = std::atoi(e);
        if (v == 1 || v == 2 || v == 4 || v == 8 || v == 16) level_chunk = v;
    }
    if (const char* e = std::getenv("CFDWM_PRUNE")) prune_ = std::atoi(e) != 0;
    if (const char* e = std::getenv("CFDWM_QFORM")) qform_ = std::atoi(e) != 0;
    if (const char* e = std::getenv("CFDWM_QF_OVERLAP")) qf_overlap_ = std::atoi(e) != 0;
    return fill ? refill() : 0;
}

int Grid::fill_at(int nslot, const double* slot_t, const int* slot_dst, bool use_ignore) {
    const int maxslot = legacy ? 2 : 4;
    if (nslot < 1 || nslot > maxslot) { set_error("fill_at: 1 to " + std::to_string(maxslot) + " slots"); return -1; }
    for (int s = 0; s < nslot; ++s)
        if (slot_dst[s] < 0 || slot_dst[s] >= maxslot) { set_error("fill_at: slot out of range"); return -1; }
    return prec == kFP32 ? fill_slots<float>(nslot, slot_t, slot_dst, use_ignore)
                         : fill_slots<double>(nslot, slot_t, slot_dst, use_ignore);
}

template <typename R>
int Grid::fill_slots(int nslot, const double* slot_t, const int* slot_dst, bool use_ignore) {
    const SeaDev<R>& S = sea->dev<R>();
    hipStream_t st = Engine::get().stream();
    const int64_t ncell = (int64_t)g.nx * g.ny;
    ColumnJob<R> C;
    C.nslot = nslot;
    C.ncol = use_ignore ? nact_ : ncell;
    C.ncell = ncell;
    C.cols = use_ignore ? (const int32_t*)cols_act_d_.get() : nullptr;
    C.g = g;
    for (int s = 0; s < nslot; ++s) { C.slot_t[s] = slot_t[s]; C.slot_dst[s] = slot_dst[s]; }
    C.eta_grid = (R*)f_[0].get();
    C.want_eta2 = 1;
    const int64_t nrec = sea->nfilt();
    C.nseg = pick_nseg(C.ncol * nslot, nrec);
    LevelJob<R> L;
    bool qf_used = false;
    // every return below joins the side stream's velocity forms first
    struct Join {
        Grid* g;
        ~Join() { g->qform_join(); }
    } join{this};
    if constexpr (sizeof(R) == 4) {
        const int rc = qform_fill(nslot, slot_t, C.ncol, C.cols, C, L);
        if (rc < 0) return rc;
        qf_used = rc == 1;
    }
    // with the pair sums on the matrix cores the column and level passes are
    // O(N) per column: split them over component segments when the grid alone
    // does not fill the GPU
    const int64_t ncomp = sea->tab.comp.n;
    if (qf_used) C.nseg = pick_nseg(C.ncol * nslot, 64 * ncomp);
    C.seg = qf_used ? sea->comp_segments(C.nseg) : sea->segments(false, C.nseg);
    if (!C.seg) return -3;
    // the linear-term bases (L.lin_c, set with the forms) take all 16 levels
    // of a column in one thread: their per-component work no longer depends
    // on the level count
    const bool lin16 = qf_used && L.lin_c && level_chunk == 8 && g.nl <= 16;
    const int lchunk = lin16 ? 16 : level_chunk;
    const int64_t lthr = level_threads(nslot, C.ncol, g.nl, lchunk, true);
    const int lseg = pick_nseg(lthr, qf_used ? 64 * ncomp : nrec);
    if constexpr (sizeof(R) == 4) {
        if (!lin16) L.lin_c = nullptr;
    }
    {
        const size_t need = std::max((size_t)C.nseg * nslot * C.ncol * 6, (size_t)lseg * lthr * (3 * lchunk + 6));
        if ((C.nseg > 1 || lseg > 1) && partial_.alloc(need * sizeof(double))) return -3;
    }
    if (C.nseg > 1) C.partial = (double*)partial_.get();
    // initialize_kinematics: ETA = eta1 + eta2 + swl (lsgrid_spline.cpp:700)
    PROF("k_columns_grid", launch_columns<R>(S, C, true, st));
    qform_join();
    L.nslot = nslot; L.ncol = C.ncol; L.ncell = ncell; L.cols = C.cols; L.g = g;
    for (int s = 0; s < nslot; ++s) { L.slot_t[s] = slot_t[s]; L.slot_dst[s] = slot_dst[s]; }
    L.eta_grid = (const R*)f_[0].get();
    L.sigma = (const double*)sigma_d_.get();
    L.o_u = (R*)f_[4].get(); L.o_v = (R*)f_[5].get(); L.o_w = (R*)f_[6].get();
    L.flags = EV_LIN | EV_SEC | EV_TAYLOR;
    L.nseg = lseg;
    L.seg = qf_used ? sea->comp_segments(lseg) : sea->segments(false, lseg);
    if (!L.seg) return -3;
    if (lseg > 1) L.partial = (double*)partial_.get();
    if (prune_ && !qf_used) {
        if (lvl_.alloc(std::max<int64_t>(1, nrec) * sizeof(uint32_t)) || eb_.alloc(2 * sizeof(double))) return -3;
        PROF("k_prune", launch_prune<R>((const R*)f_[0].get(), ncell, C.cols, C.ncol, nslot, slot_dst, g.swl,
                                        (double*)eb_.get(), S.fuvw, nrec, (const double*)sigma_d_.get(), g.nl, g.h,
                                        (uint32_t*)lvl_.get(), st));
        L.lvl = (const uint32_t*)lvl_.get();
    }
    PROF("k_levels_grid", launch_levels<R>(S, L, true, lchunk, st));
    return 0;
}

// Second-order pair sums of an FP32 fill as MFMA quadratic forms: the
// filtered eta2 for the column pass and, through the two profile depth
// bases, the u/v/w pair sums for the level pass.  qform_prepare checks
// applicability (switched on, pairs present, frequency ramps equal for all
// components at each slot time, levels inside the basis range), builds the
// sea's tables once and sizes the workspace; returns 1 when the forms apply,
// 0 when not, < 0 on error.
int Grid::qform_prepare(int nslot, const double* slot_t, int64_t ncol, const int32_t* cols) {
    qf_used = false;
    qf_clock_valid_ = false;                  // a measurement belongs to one fill
    qf_skip = QF_FP64;
    if (sea->prec != kFP32) return 0;
    qf_skip = QF_OFF;
    if (!qform_) return 0;
    qf_skip = QF_NO_PAIRS;
    if (sea->nfilt() == 0) return 0;
    const CompTable& c = sea->tab.comp;
    for (int s = 0; s < 4; ++s) qf_scale_[s] = 1.0;
    for (int s = 0; s < nslot; ++s) {
        const double r0 = freq_ramp_host(slot_t[s], c.rt[0], c.ts[0], c.du[0]);
        for (int i = 1; i < c.n; ++i)
            if (freq_ramp_host(slot_t[s], c.rt[i], c.ts[i], c.du[i]) != r0) {
                qf_skip = QF_RAMPS;
                warn_pair_loop("the components' frequency ramps differ at a slot time");
                return 0;
            }
        qf_scale_[s] = r0;
    }
    const double hs = sea->tab.in.depth;
    // the profile bases cover z in [-2 hs, 0]: the grid's levels must lie above -2 hs
    qf_skip = QF_DEPTH;
    if (g.h > 2.0 * hs) {
        warn_pair_loop("the grid's water depth exceeds twice the sea's");
        return 0;
    }
    const Sea::QfTables* T = nullptr;
    if (const int rc = sea->qform_tables(&T); rc != 1) {
        qf_skip = QF_BASIS;
        if (rc == 0) warn_pair_loop("no depth basis within the tolerance at rank <= 32");
        return rc;
    }
    qf_skip = QF_USED;
    const int npad = T->npad, R = T->rank;
    const int64_t P = (int64_t)nslot * ncol;
    const int64_t ppad_all = (P + 127) / 128 * 128;
    // Z operands of one chunk of points: 8 bytes per (point, component), at most 2 GiB
    const int64_t cap = std::max<int64_t>(128, ((int64_t)1 << 31) / ((int64_t)npad * 8) / 128 * 128);
    qf_chunk_ = std::min(ppad_all, cap);
    qf_P_ = P;
    const size_t zn = (size_t)qf_chunk_ * npad;
    uint16_t* z = nullptr;
    if (zext_ && zext_bytes_ >= zn * 4 * sizeof(uint16_t)) {
        z = zext_;                            // a batched refill's arena
    } else {
        if (zbuf_.alloc(zn * 4 * sizeof(uint16_t))) return -3;
        z = (uint16_t*)zbuf_.get();
    }
    if (qout_.alloc((size_t)T->nform * ppad_all * sizeof(float))) return -3;
    QfJob& Q = qj_;
    Q = QfJob{};
    Q.n = c.n; Q.npad = npad;
    Q.grid = 1; Q.ncol = ncol; Q.cols = cols; Q.g = g;
    for (int s = 0; s < nslot; ++s) Q.slot_t[s] = slot_t[s];
    Q.x0 = sea->tab.in.x0; Q.y0 = sea->tab.in.y0;
    Q.ph = sea->dev<float>().ph;
    Q.zc_hi = z; Q.zc_lo = z + zn; Q.zs_hi = z + 2 * zn; Q.zs_lo = z + 3 * zn;
    Q.zn = (int64_t)zn;
    Q.ftype = (const int8_t*)T->ftype.get();
    Q.kend = (const int32_t*)T->kend.get();
    Q.out = (float*)qout_.get();
    Q.ostride = ppad_all;
    Q.p0 = 0;
    Q.P = std::min(qf_chunk_, P);
    Q.ppad = (Q.P + 127) / 128 * 128;
    qf_split_ = T->nlv == 1;
    if (!qf_split_) {                             // every form, three products per term
        Q.b = T->bp;
        Q.nbarr = 4; Q.nlv = 3;
        Q.nform = T->nform;
    } else {                                      // eta2 at three products; velocity forms at one
        Q.b = T->betap;
        Q.nbarr = 4; Q.nlv = 3;
        Q.nform = 1;
        QfJob& V = qj2_;
        V = Q;
        V.b = T->bp;
        V.nbarr = T->nbarr; V.nlv = 1;
        V.ftype = Q.ftype + 1;
        V.nform = T->nform - 1;
        V.out = Q.out + ppad_all;
    }
    // executed MFMA flops per padded point and form: a 128-wide i tile runs
    // both 64-row halves from its first row to the larger of their two kend,
    // 128 x (kend - i0) products, two operand pairs (c and s), 3 or 1 bf16 products
    double tri = 0;
    for (size_t t = 0; t + 1 < T->kend_h.size(); t += 2) {
        const int ke = std::max(T->kend_h[t], T->kend_h[t + 1]);
        tri += 128.0 * std::max(0, ke - (int)(64 * t)) * 2 * 2;
    }
    qf_fpp_ = 3 * tri;
    qf_fpp1_ = tri;
    qf_nlv = T->nlv;
    qf_nform = T->nform;
    qf_rank = R;
    qf_rank_w = T->rank_w;
    qf_basis_err = T->basis_err;
    qf_flop = 0;
    return 1;
}

void Grid::warn_pair_loop(const char* why) {
    if (qf_warned_ || std::getenv("CFDWM_QUIET")) return;
    qf_warned_ = true;
    std::fprintf(stderr, "cfdwm: FP32 s-grid fill runs the second-order pair loop instead of the matrix-core "
                         "forms (%s); expect it to be much slower\n", why);
}

// Z operands and forms of every chunk of the prepared fill (one grid).
int Grid::qform_run() {
    hipStream_t st = Engine::get().stream();
    QfJob Q = qj_;
    qf_clock_valid_ = false;
    if (qf_clock_on_) {
        if (qf_clk_.alloc(2 * sizeof(unsigned long long))) return -3;
        CFDWM_HIP(hipMemsetAsync(qf_clk_.get(), 0, 2 * sizeof(unsigned long long), st));
        Q.clk = (unsigned long long*)qf_clk_.get();
        qf_clock_valid_ = true;
    }
    QfJob V = qj2_;
    V.clk = Q.clk;
    for (int64_t p0 = 0; p0 < qf_P_; p0 += qf_chunk_) {
        Q.p0 = p0;
        Q.P = std::min(qf_chunk_, qf_P_ - p0);
        Q.ppad = (Q.P + 127) / 128 * 128;
        PROF("k_qf_zgen", launch_qf_zgen(Q, st));
        if (qf_split_) {
            // eta2 at three products, then the velocity forms at one.  With
            // one chunk the velocity forms go to the side stream (they need
            // only the Z operands), so that the column pass, which needs the
            // eta2 form, runs beside them; the level pass joins them
            // (qform_join), and so does anything later on the main stream.
            V.p0 = Q.p0; V.P = Q.P; V.ppad = Q.ppad;
            PROF("k_qf_forms_eta", launch_qf_forms(Q, st));
            hipStream_t s2 = qf_overlap_ && qf_P_ <= qf_chunk_ ? Engine::get().side_stream() : nullptr;
            if (s2 && !qf_fork_ && (hipEventCreateWithFlags(&qf_fork_, hipEventDisableTiming) != hipSuccess ||
                                    hipEventCreateWithFlags(&qf_join_, hipEventDisableTiming) != hipSuccess))
                s2 = nullptr;
            if (s2 && qf_fork_ && qf_join_) {
                CFDWM_HIP(hipEventRecord(qf_fork_, st));
                CFDWM_HIP(hipStreamWaitEvent(s2, qf_fork_, 0));
                PROF_ON(s2, "k_qf_forms", launch_qf_forms(V, s2));
                CFDWM_HIP(hipEventRecord(qf_join_, s2));
                qf_join_pending_ = true;
            } else {
                PROF("k_qf_forms", launch_qf_forms(V, st));
            }
            qf_flop += (qf_fpp_ * Q.nform + qf_fpp1_ * V.nform) * (double)Q.ppad;
        } else {
            PROF("k_qf_forms", launch_qf_forms(Q, st));
            qf_flop += qf_fpp_ * (double)Q.ppad * Q.nform;
        }
    }
    return 0;
}

// Mean shader clock over the blocks of the last fill's forms launches:
// cycles / wall-clock ticks x the wall-clock rate.
int Grid::qform_clock(double* mhz) {
    if (mhz) *mhz = 0;
    if (!qf_clock_valid_ || !qf_clk_.get()) return -1;
    unsigned long long c[2] = {0, 0};
    CFDWM_HIP(hipStreamSynchronize(Engine::get().stream()));
    CFDWM_HIP(hipMemcpy(c, qf_clk_.get(), sizeof(c), hipMemcpyDeviceToHost));
    int khz = 0;
    CFDWM_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, 0));
    if (c[1] == 0 || khz <= 0) return -1;
    if (mhz) *mhz = (double)c[0] / (double)c[1] * khz * 1e-3;
    return 0;
}

// Point the column and level passes of the fill at the form values.
void Grid::qform_attach(ColumnJob<float>& C, LevelJob<float>& L) {
    const Sea::QfTables* T = nullptr;
    (void)sea->qform_tables(&T);              // cached: qform_prepare built them
    const int R = qf_rank;
    const double hs = sea->tab.in.depth;
    qf_used = true;
    C.pairs = 0;
    C.q2 = (const float*)qout_.get();
    C.nq2 = 1;
    C.q2stride = qj_.ostride;
    for (int s = 0; s < 4; ++s) C.q2scale[s] = qf_scale_[s];
    L.qf = (const float*)qout_.get() + qj_.ostride;
    L.qf_rank = R;
    L.qf_rank_w = qf_rank_w;
    L.qf_stride = qj_.ostride;
    constexpr double kL2e = 1.4426950408889634074;
    for (int q = 0; q < R; ++q) {
        L.qf_mu2[q] = (float)(T->mu[q] * kL2e);
        L.qf_mg[q] = (float)(-2.0 * hs * T->mu[q] * kL2e);
    }
    for (int q = 0; q < qf_rank_w; ++q) {
        L.qf_mu2w[q] = (float)(T->mu_w[q] * kL2e);
        L.qf_mgw[q] = (float)(-2.0 * hs * T->mu_w[q] * kL2e);
    }
    for (int s = 0; s < 4; ++s) L.qf_scale[s] = qf_scale_[s];
    L.lin_c = T->lin_rank > 0 ? (const float*)T->lin_c.get() : nullptr;
    L.lin_mu = (const float*)T->lin_exp.get();
    L.lin_R = T->lin_rank;
    L.lin_Rw = T->lin_rank_w;
}

int Grid::qform_fill(int nslot, const double* slot_t, int64_t ncol, const int32_t* cols, ColumnJob<float>& C,
                     LevelJob<float>& L) {
    if (batch_pending_) {                     // forms already evaluated by a batched launch
        batch_pending_ = false;
        qform_attach(C, L);
        return 1;
    }
    const int rc = qform_prepare(nslot, slot_t, ncol, cols);
    if (rc != 1) return rc;
    if (int e = qform_run()) return e;
    qform_attach(C, L);
    return 1;
}

// initialize_kinematics slot times (lsgrid_spline.cpp:678, :767)
void Grid::init_slot_times(double* tt) const {
    const double t0i = d.t0;
    const bool ign = d.ignore_at_init != 0;
    tt[0] = ign ? t0i : t0i - d.dt;
    tt[1] = t0i;
    tt[2] = t0i + d.dt;
    tt[3] = t0i + 2 * d.dt;
}

// Batched refill, part 1: prepare the forms of a full refill and launch its Z
// operands.  Returns 1 when this grid joins the batched forms launch (one
// chunk), 0 when it must be refilled on its own.
int Grid::batch_begin() {
    batch_pending_ = false;
    if (prec != kFP32) return 0;
    double tt[4];
    init_slot_times(tt);
    const bool ign = d.ignore_at_init != 0;
    const int64_t ncol = ign ? nact_ : (int64_t)g.nx * g.ny;
    const int32_t* cols = ign ? (const int32_t*)cols_act_d_.get() : nullptr;
    const int rc = qform_prepare(4, tt, ncol, cols);
    if (rc != 1) return rc;
    if (qf_chunk_ < qf_P_) return 0;          // needs chunking: not batched
    PROF("k_qf_zgen", launch_qf_zgen(qj_, Engine::get().stream()));
    qf_flop = (qf_fpp_ * qj_.nform + (qf_split_ ? qf_fpp1_ * qj2_.nform : 0.0)) * (double)qj_.ppad;
    batch_pending_ = true;
    return 1;
}

// Batched refill, part 2 (after the batched forms launch): the rest of the refill.
int Grid::batch_end() {
    if (!batch_pending_) return refill();
    tstep = 0;
    t0 = d.t0;
    double tt[4];
    init_slot_times(tt);
    const int dst[4] = {0, 1, 2, 3};
    int rc = fill_slots<float>(4, tt, dst, d.ignore_at_init != 0);
    batch_pending_ = false;
    if (rc) return rc;
    if ((rc = gradients<float>())) return rc;
    if ((rc = pack<float>())) return rc;
    filled = true;
    return 0;
}

size_t Grid::refill_z_bytes() const {
    const Sea::QfTables* T = nullptr;
    if (sea->qform_meta(&T) != 1) return 0;
    const int64_t ncol = d.ignore_at_init ? nact_ : (int64_t)g.nx * g.ny;
    const int64_t ppad = (4 * ncol + 127) / 128 * 128;
    const int64_t cap = std::max<int64_t>(128, ((int64_t)1 << 31) / ((int64_t)T->npad * 8) / 128 * 128);
    if (ppad > cap) return 0;
    return (size_t)ppad * T->npad * 4 * sizeof(uint16_t);
}

void Grid::release_workspace() {
    zbuf_.release();
    qs_key_.release();
    qs_rec_.release();
    qs_tmp_.release();
}

namespace {
// Shape of a batched persistent launch: tile order, grid size, and (order 1)
// the group of gf forms x gp point tiles that consecutive tiles walk.
struct BatchShape {
    int order, blocks, gf, gp;
};

// Tile lists of a batched launch.  order 0: all tiles by decreasing cost
// (point-tile cost ~ the job's i-range work), so round-robin blocks get equal
// shares; order 1: job by job (decreasing cost), inside a job groups of gf
// forms x gp point tiles (cache locality of the job's tables and operands:
// consecutive blocks share a group's B tables and Z tiles).
std::vector<QfTile> batch_tiles(const std::vector<QfJob>& jobs, const std::vector<double>& cost, const BatchShape& b) {
    std::vector<int> idx(jobs.size());
    for (size_t j = 0; j < jobs.size(); ++j) idx[j] = (int)j;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int c) { return cost[a] > cost[c]; });
    std::vector<QfTile> t;
    if (b.order == 0) {
        for (int j : idx) {
            const int npt = (int)(jobs[j].ppad / 128);
            for (int f = 0; f < jobs[j].nform; ++f)
                for (int p = 0; p < npt; ++p) t.push_back(QfTile{j, p, f, 0});
        }
        std::stable_sort(t.begin(), t.end(), [&](const QfTile& a, const QfTile& c) { return cost[a.job] > cost[c.job]; });
    } else {
        const int gf = std::max(1, b.gf), gp = std::max(1, b.gp);
        for (int j : idx) {
            const int npt = (int)(jobs[j].ppad / 128), nf = jobs[j].nform;
            for (int f0 = 0; f0 < nf; f0 += gf)
                for (int p0 = 0; p0 < npt; p0 += gp)
                    for (int f = f0; f < std::min(nf, f0 + gf); ++f)
                        for (int p = p0; p < std::min(npt, p0 + gp); ++p) t.push_back(QfTile{j, p, f, 0});
        }
    }
    return t;
}

// Candidates of the batched-launch tuner: cost-ordered tiles, or job-ordered
// tiles in groups of 8 x 32, 4 x 64 or 16 x 16 (forms x point tiles), each at
// one, two or four blocks per CU of the 256 CUs.
const std::vector<BatchShape>& batch_candidates() {
    static const std::vector<BatchShape> c = {
        {0, 256, 0, 0}, {0, 1024, 0, 0},
        {1, 256, 8, 32}, {1, 512, 8, 32}, {1, 1024, 8, 32},
        {1, 256, 4, 64}, {1, 1024, 4, 64},
        {1, 256, 16, 16}, {1, 1024, 16, 16},
    };
    return c;
}

// Device arena of the batched refill's form tables and Z operands (one
// group at a time); freed at the end of a call with release, else kept.
DevBuf& batch_arena() {
    static DevBuf a;
    return a;
}

struct TuneKey {
    int nlv, jobs_log2, tiles_log2;
    bool operator<(const TuneKey& o) const {
        if (nlv != o.nlv) return nlv < o.nlv;
        return jobs_log2 != o.jobs_log2 ? jobs_log2 < o.jobs_log2 : tiles_log2 < o.tiles_log2;
    }
};
std::map<TuneKey, BatchShape>& tune_cache() {
    static std::map<TuneKey, BatchShape> m;
    return m;
}
std::mutex& tune_mutex() {
    static std::mutex m;
    return m;
}
int ilog2(int64_t v) {
    int r = 0;
    while (v > 1) { v >>= 1; ++r; }
    return r;
}
}  // namespace

int refill_batched(Grid** grids, int n, int64_t budget_bytes, bool release, BatchStats* stats) {
    BatchStats S;
    hipStream_t st = Engine::get().stream();
    int i = 0;
    DevBuf djobs, dtiles;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    CFDWM_HIP(hipEventCreate(&e0));
    CFDWM_HIP(hipEventCreate(&e1));
    auto cleanup = [&]() { (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); };
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    // device bytes of a grid's share of a group: its sea's tables (unless
    // shared with an earlier grid or resident) and its Z operands, or an
    // estimate for grids whose forms do not apply
    auto share = [&](Grid* g, size_t* tbytes, size_t* zbytes) -> int {
        *tbytes = *zbytes = 0;
        const Sea::QfTables* T = nullptr;
        const int rc = g->sea->qform_meta(&T);
        if (rc < 0) return rc;
        if (rc == 1) {
            if (!T->bp) *tbytes = al(T->table_bytes) + al(T->eta_bytes);
            *zbytes = al(g->refill_z_bytes());
        }
        return 0;
    };
    hipEvent_t tb0 = nullptr, tb1 = nullptr;
    CFDWM_HIP(hipEventCreate(&tb0));
    CFDWM_HIP(hipEventCreate(&tb1));
    auto cleanup_all = [&]() {
        cleanup();
        (void)hipEventDestroy(tb0);
        (void)hipEventDestroy(tb1);
        // on an error no sea or grid may keep pointers into the arena
        for (int k = 0; k < n; ++k) {
            if (grids[k]->sea->qform_external()) grids[k]->sea->release_qform();
            grids[k]->set_z_external(nullptr, 0);
        }
    };
    while (i < n) {
        // group of grids whose tables and Z operands fit the budget
        std::vector<Grid*> grp;
        std::vector<size_t> tbytes, zbytes;
        int64_t bytes = 0;
        const auto h0 = std::chrono::steady_clock::now();
        while (i < n) {
            Grid* g = grids[i];
            size_t tb = 0, zb = 0;
            if (g->sea->prec == kFP32)
                if (int rc = share(g, &tb, &zb)) { cleanup_all(); return rc; }
            // a sea shared by grids of this group is built once
            for (Grid* o : grp)
                if (o->sea == g->sea) tb = 0;
            const int64_t est = (int64_t)(tb + zb);
            if (!grp.empty() && bytes + est > budget_bytes) break;
            grp.push_back(g);
            tbytes.push_back(tb);
            zbytes.push_back(zb);
            bytes += est;
            ++i;
        }
        ++S.groups;
        // release != 0: the group's tables and Z operands live in one arena,
        // reused by the next group (no allocation per sea or grid); else each
        // sea keeps its own tables resident for later fills
        CFDWM_HIP(hipEventRecord(tb0, st));
        if (release) {
            if (bytes > 0 && batch_arena().alloc((size_t)bytes)) { cleanup_all(); return -3; }
            char* a = (char*)batch_arena().get();
            size_t off = 0;
            for (size_t k = 0; k < grp.size(); ++k) {
                if (tbytes[k]) {
                    const Sea::QfTables* T = nullptr;
                    if (grp[k]->sea->qform_tables_into(a + off, &T) < 0) { cleanup_all(); return -3; }
                    off += tbytes[k];
                }
                if (zbytes[k]) {
                    grp[k]->set_z_external((uint16_t*)(a + off), zbytes[k]);
                    off += zbytes[k];
                }
            }
        } else {
            for (Grid* g : grp) {
                const Sea::QfTables* T = nullptr;
                if (g->sea->prec == kFP32 && g->sea->qform_tables(&T) < 0) { cleanup_all(); return -3; }
            }
        }
        CFDWM_HIP(hipEventRecord(tb1, st));
        // per product mode (3: hi/lo split, 1: hi only) the jobs of the group
        // and their costs; a grid whose velocity forms take one product adds
        // its eta2 form to the three-product list and its velocity forms to
        // the one-product list
        std::vector<QfJob> jobs3, jobs1;
        std::vector<double> cost3, cost1;
        std::vector<char> joined(grp.size(), 0);
        for (size_t k = 0; k < grp.size(); ++k) {
            const int rc = grp[k]->batch_begin();
            if (rc < 0) { cleanup_all(); return rc; }
            if (rc == 1) {
                joined[k] = 1;
                const Sea::QfTables* T = nullptr;
                (void)grp[k]->sea->qform_meta(&T);
                jobs3.push_back(grp[k]->batch_job());
                cost3.push_back(T->cost);
                if (grp[k]->batch_split()) {
                    jobs1.push_back(grp[k]->batch_job2());
                    cost1.push_back(T->cost);
                }
                S.mfma_flop += grp[k]->qf_flop;
            }
        }
        {
            // host time of the group's set-up plus the device time of its table builds
            float ms = 0;
            CFDWM_HIP(hipEventSynchronize(tb1));
            CFDWM_HIP(hipEventElapsedTime(&ms, tb0, tb1));
            S.tables_ms += ms;
            S.setup_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count();
        }
        // one persistent launch per product mode, its shape from the tuner
        auto launch_list = [&](const std::vector<QfJob>& jobs, const std::vector<double>& cost, int nlv) -> int {
            if (jobs.empty()) return 0;
            if (djobs.upload(jobs.data(), jobs.size() * sizeof(QfJob), st)) return -3;
            int64_t ntile = 0;
            for (const QfJob& J : jobs) ntile += (J.ppad / 128) * J.nform;
            const TuneKey key{nlv, ilog2((int64_t)jobs.size()), ilog2(ntile)};
            BatchShape choice{-1, 0, 0, 0};
            {
                std::lock_guard<std::mutex> lk(tune_mutex());
                auto it = tune_cache().find(key);
                if (it != tune_cache().end()) choice = it->second;
            }
            if (const char* e = std::getenv("CFDWM_BATCH_SHAPE")) {
                BatchShape b{0, 256, 8, 32};
                if (std::sscanf(e, "%d,%d,%d,%d", &b.order, &b.blocks, &b.gf, &b.gp) >= 2) choice = b;
            }
            auto run = [&](const BatchShape& b, float* ms) -> int {
                const std::vector<QfTile> tl = batch_tiles(jobs, cost, b);
                if (dtiles.upload(tl.data(), tl.size() * sizeof(QfTile), st)) return -3;
                CFDWM_HIP(hipEventRecord(e0, st));
                CFDWM_HIP(launch_qf_persist((const QfJob*)djobs.get(), (const QfTile*)dtiles.get(), (int)tl.size(),
                                            b.blocks, nlv, st));
                CFDWM_HIP(hipEventRecord(e1, st));
                CFDWM_HIP(hipEventSynchronize(e1));
                CFDWM_HIP(hipEventElapsedTime(ms, e0, e1));
                S.tiles += (int)tl.size();
                return 0;
            };
            float ms = 0;
            if (choice.order < 0) {
                // autotune on this launch: every candidate computes the same
                // values; keep the fastest
                float best = 1e30f;
                for (const BatchShape& c : batch_candidates()) {
                    float t = 0;
                    if (int rc = run(c, &t)) return rc;
                    if (t < best) { best = t; choice = c; }
                }
                ms = best;
                S.tuned = 1;
                std::lock_guard<std::mutex> lk(tune_mutex());
                tune_cache()[key] = choice;
            } else {
                if (int rc = run(choice, &ms)) return rc;
            }
            S.forms_ms += ms;
            S.order = choice.order;
            S.blocks = choice.blocks;
            S.group_forms = choice.gf;
            S.group_ptiles = choice.gp;
            return 0;
        };
        if (int rc = launch_list(jobs3, cost3, 3)) { cleanup_all(); return rc; }
        if (int rc = launch_list(jobs1, cost1, 1)) { cleanup_all(); return rc; }
        for (size_t k = 0; k < grp.size(); ++k) {
            const int rc = joined[k] ? grp[k]->batch_end() : grp[k]->refill();
            if (rc) { cleanup_all(); return rc; }
            if (joined[k]) ++S.batched; else ++S.single;
        }
        if (hipStreamSynchronize(st) != hipSuccess) { cleanup_all(); return -3; }
        for (Grid* g : grp) {
            if (release) { g->sea->release_qform(); g->release_workspace(); }
            g->set_z_external(nullptr, 0);
        }
    }
    if (release) batch_arena().release();
    cleanup_all();
    if (stats) *stats = S;
    return 0;
}

template <typename R>
int Grid::gradients() {
    hipStream_t st = Engine::get().stream();
    const int tid0 = (tstep + 1) % 4, tid1 = (tstep + 2) % 4;
    const int64_t c = (int64_t)g.nx * g.ny, v = c * g.nl;
    auto P = [&](int f) { return (R*)f_[f].get(); };
    PROF("k_grad_xyz", launch_grad_xyz<R>(P(0), P(1), P(2), (R*)nullptr, g.nx, g.ny, 1, tid0, tid1, g.dx, g.dy, g.ds, st));
    PROF("k_grad_dt", launch_grad_dt<R>(P(0), P(3), c, tstep, d.dt, st));
    for (int f = 0; f < 3; ++f) {
        PROF("k_grad_xyz", launch_grad_xyz<R>(P(4 + f), P(10 + 3 * f), P(11 + 3 * f), P(12 + 3 * f), g.nx, g.ny, g.nl,
                                              tid0, tid1, g.dx, g.dy, g.ds, st));
        PROF("k_grad_dt", launch_grad_dt<R>(P(4 + f), P(7 + f), v, tstep, d.dt, st));
    }
    return 0;
}

template <typename R>
int Grid::pack() {
    const size_t esz = sizeof(R);
    const int64_t c = (int64_t)g.nx * g.ny, v = c * g.nl;
    if (pk_eta_.alloc(8 * c * esz) || pk_C_.alloc(16 * v * esz) || pk_D_.alloc(16 * v * esz)) return -3;
    PackJob<R> J;
    J.nx = g.nx; J.ny = g.ny; J.nl = g.nl;
    J.tid1 = (tstep + 1) % 4; J.tid2 = (tstep + 2) % 4;
    auto P = [&](int f) { return (const R*)f_[f].get(); };
    J.E = P(0); J.Ex = P(1); J.Ey = P(2); J.Et = P(3);
    for (int f = 0; f < 3; ++f) {
        J.F[f] = P(4 + f); J.Ft[f] = P(7 + f);
        J.Fx[f] = P(10 + 3 * f); J.Fy[f] = P(11 + 3 * f); J.Fs[f] = P(12 + 3 * f);
    }
    J.pk_eta = (R*)pk_eta_.get(); J.pk_C = (R*)pk_C_.get(); J.pk_D = (R*)pk_D_.get();
    PROF("k_pack", launch_pack<R>(J, Engine::get().stream()));
    return 0;
}

int Grid::refill() {
    if (legacy) { set_error("a legacy two-slot grid fills through LsGrid"); return -1; }
    tstep = 0;
    t0 = d.t0;
    // initialize_kinematics (:678) / initialize_kinematics_with_ignore (:767)
    const double ta[4] = {t0 - d.dt, t0, t0 + d.dt, t0 + 2 * d.dt};
    const double tb[4] = {t0, t0, t0 + d.dt, t0 + 2 * d.dt};
    const int dst[4] = {0, 1, 2, 3};
    const bool ign = d.ignore_at_init != 0;
    int rc = prec == kFP32 ? fill_slots<float>(4, ign ? tb : ta, dst, ign)
                           : fill_slots<double>(4, ign ? tb : ta, dst, ign);
    if (rc) return rc;
    rc = prec == kFP32 ? gradients<float>() : gradients<double>();
    if (rc) return rc;
    rc = prec == kFP32 ? pack<float>() : pack<double>();
    if (rc) return rc;
    CFDWM_HIP(hipStreamSynchronize(Engine::get().stream()));
    filled = true;
    return write_vtk_dump(false);                 // initialize_kinematics (lsgrid_spline.cpp:746-748)
}

int Grid::force_update(double t, int* n) {
    if (!filled) { set_error("the s-grid has not been filled yet (created with defer_fill)"); return -1; }
    int k = 0;
    // CheckTime / update (lsgrid_spline.cpp:610-617, 847-931)
    while (t > t0 + d.dt) {
        tstep++;
        t0 += d.dt;
        const double t2 = t0 + 2 * d.dt;
        const int slot = (tstep + 3) % 4;
        int rc = prec == kFP32 ? fill_slots<float>(1, &t2, &slot, true) : fill_slots<double>(1, &t2, &slot, true);
        if (rc) return rc;
        rc = prec == kFP32 ? gradients<float>() : gradients<double>();
        if (rc) return rc;
        rc = prec == kFP32 ? pack<float>() : pack<double>();
        if (rc) return rc;
        ++k;
        if (vtk.on) {                             // update (lsgrid_spline.cpp:924-926)
            CFDWM_HIP(hipStreamSynchronize(Engine::get().stream()));
            if ((rc = write_vtk_dump(false))) return rc;
        }
    }
    if (k) CFDWM_HIP(hipStreamSynchronize(Engine::get().stream()));
    if (n) *n = k;
    return 0;
}

int64_t Grid::export_field(int which, double* out) {
    if (which == 19) {
        if (out) for (size_t q = 0; q < ignore_.size(); ++q) out[q] = ignore_[q];
        return (int64_t)ignore_.size();
    }
    const int64_t n = field_count(which);
    if (n == 0 || !out) return n;
    hipStream_t st = Engine::get().stream();
    if (prec == kFP32) {
        std::vector<float> tmp(n);
        if (hipMemcpyAsync(tmp.data(), f_[which].get(), n * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) return -3;
        for (int64_t q = 0; q < n; ++q) out[q] = tmp[q];
    } else {
        if (hipMemcpyAsync(out, f_[which].get(), n * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) return -3;
    }
    return n;
}

// ASCII VTK XML UnstructuredGrid of the slot that the query uses as its base
// (last: the following slot), at time t0 (last: t0 + dt), in the layout and
// number formats of the reference's lsGridSpline::export_vtu
// (lsgrid_spline.cpp:1223-1403), so that tools reading its files read these:
// velocity vectors at the nodes, node positions from the sigma map of the
// slot's elevation, hexahedra between neighbouring nodes (quadrilaterals for
// grids one node wide in x, y or the vertical; no cells for grids one node
// wide in two directions, whose cell count is still given as the product of
// max(n - 1, 1)), and the time as field data.
int Grid::write_vtu(const char* path, bool last, const char* time_label) {
    if (!filled) { set_error("the s-grid has not been filled yet (created with defer_fill)"); return -1; }
    const int tt = last ? (tstep + 2) % 4 : (tstep + 1) % 4;
    return write_vtu_slot(path, tt, last ? t0 + d.dt : t0, time_label, false);
}

int Grid::write_vtu_slot(const char* path, int tt, double tv, const char* time_label, bool swl_free) {
    const int nx = g.nx, ny = g.ny, nl = g.nl;
    const int64_t nc = (int64_t)nx * ny, nv = nc * nl;
    const int ns = legacy ? 2 : 4;
    if (tt < 0 || tt >= ns) { set_error("write_vtu: slot out of range"); return -1; }
    std::vector<double> eta(ns * nc), U(ns * nv), V(ns * nv), W(ns * nv);
    if (export_field(0, eta.data()) < 0 || export_field(4, U.data()) < 0 || export_field(5, V.data()) < 0 ||
        export_field(6, W.data()) < 0) return -3;
    FILE* fp = std::fopen(path, "w");
    if (!fp) { set_error(std::string("cannot open ") + path); return -1; }
    const char* lab = time_label && time_label[0] ? time_label : "TimeValue";
    std::fputs("<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"1.0\" "
               "byte_order=\"LittleEndian\" header_type=\"UInt64\">\n", fp);
    std::fputs("\t <UnstructuredGrid>\n\t\t <FieldData> \n", fp);
    std::fprintf(fp, "\t\t\t <DataArray type = \"Float64\" Name = \"%s\" NumberOfTuples = \"1\" format = \"ascii\" "
                     "RangeMin = \"%.3f\" RangeMax = \"%.3f\"> \n", lab, tv, tv);
    std::fprintf(fp, "\t\t\t %.3f \n\t\t\t </DataArray > \n\t\t </FieldData> \n", tv);
    const long long ncells = (long long)std::max(nx - 1, 1) * std::max(ny - 1, 1) * std::max(nl - 1, 1);
    std::fprintf(fp, "\t\t <Piece NumberOfPoints=\"%lld\" NumberOfCells=\"%lld\">\n", (long long)nv, ncells);
    std::fputs("\t\t\t <PointData Scalars=\"scalars\">\n", fp);
    std::fputs("\t\t\t\t <DataArray type=\"Float64\" NumberOfComponents=\"3\" Name=\"Velocity\" format=\"ascii\">\n",
               fp);
    const int64_t so = (int64_t)tt * nv;
    for (int64_t q = 0; q < nv; ++q) std::fprintf(fp, "%g %g %g\n", U[so + q], V[so + q], W[so + q]);
    std::fputs("\t\t\t\t </DataArray>\n\t\t\t </PointData>\n\t\t\t <Points>\n", fp);
    std::fputs("\t\t\t\t <DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n", fp);
    for (int i = 0; i < nx; ++i) {
        const double xp = g.x0 + g.dx * i;
        for (int j = 0; j < ny; ++j) {
            const double yp = g.y0 + g.dy * j;
            const double e = eta[(int64_t)tt * nc + (int64_t)i * ny + j];
            for (int m = 0; m < nl; ++m) {
                // s2z (lsgrid_spline.cpp:96-98); the legacy writer passes eta
                // with its swl (lsgrid.cpp:1420-1424, 87-89)
                const double z = swl_free ? e + sigma_[m] * (g.h + e)
                                          : (e - g.swl) + sigma_[m] * (g.h + (e - g.swl));
                std::fprintf(fp, "%12.4f %12.4f %12.4f\n", xp, yp, z);
            }
        }
    }
    std::fputs("\t\t\t\t </DataArray>\n\t\t\t </Points>\n\t\t\t <Cells>\n", fp);
    std::fputs("\t\t\t\t <DataArray type=\"Int64\" Name=\"connectivity\" format=\"ascii\">\n", fp);
    auto tail = [&](long long nce, int per, int type) {
        std::fputs("\t\t\t\t </DataArray>\n\t\t\t\t <DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\">\n",
                   fp);
        for (long long c = 1; c <= nce; ++c) std::fprintf(fp, "%lld \n", c * per);
        std::fputs("\t\t\t\t </DataArray>\n\t\t\t\t <DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n", fp);
        for (long long c = 1; c <= nce; ++c) std::fprintf(fp, "%d \n", type);
    };
    if (nx > 1 && ny > 1 && nl > 1) {
        for (int i = 0; i < nx - 1; ++i)
            for (int j = 0; j < ny - 1; ++j)
                for (int m = 0; m < nl - 1; ++m) {
                    const long long a = (long long)nl * ny * i + (long long)nl * j + m;
                    const long long b = a + (long long)nl * ny, c = b + nl, e = a + nl;
                    std::fprintf(fp, "%lld %lld %lld %lld %lld %lld %lld %lld\n", a, b, c, e, a + 1, b + 1, c + 1, e + 1);
                }
        tail((long long)(nx - 1) * (ny - 1) * (nl - 1), 8, 12);
    } else if (nx > 1 && ny == 1 && nl > 1) {
        for (int i = 0; i < nx - 1; ++i)
            for (int m = 0; m < nl - 1; ++m) {
                const long long a = (long long)nl * i + m, b = (long long)nl * (i + 1) + m;
                std::fprintf(fp, "%lld %lld %lld %lld\n", a, b, b + 1, a + 1);
            }
        tail((long long)(nx - 1) * (nl - 1), 4, 9);
    } else if (nx == 1 && ny > 1 && nl > 1) {
        for (int j = 0; j < ny - 1; ++j)
            for (int m = 0; m < nl - 1; ++m) {
                const long long a = (long long)nl * j + m, b = (long long)nl * (j + 1) + m;
                std::fprintf(fp, "%lld %lld %lld %lld\n", a, b, b + 1, a + 1);
            }
        tail((long long)(ny - 1) * (nl - 1), 4, 9);
    } else if (nx > 1 && ny > 1 && nl == 1) {
        for (int i = 0; i < nx - 1; ++i)
            for (int j = 0; j < ny - 1; ++j) {
                const long long a = (long long)ny * i + j, b = (long long)ny * (i + 1) + j;
                std::fprintf(fp, "%lld %lld %lld %lld\n", a, b, b + 1, a + 1);
            }
        tail((long long)(nx - 1) * (ny - 1), 4, 9);
    }
    std::fputs("\t\t\t\t </DataArray>\n\t\t\t </Cells>\n\t\t </Piece>\n\t </UnstructuredGrid>\n</VTKFile>\n", fp);
    const bool ok = std::ferror(fp) == 0;
    std::fclose(fp);
    if (!ok) { set_error(std::string("write error on ") + path); return -1; }
    return 0;
}

// lsGridSpline::write_vtk (lsgrid_spline.cpp:1200-1221): directory (created if
// missing) + prefix + the time step as %05d + ".vtu".
int Grid::write_vtk_dump(bool last) {
    if (!vtk.on) return 0;
    if (!vtk.dir.empty()) {
        struct stat sb;
        if (stat(vtk.dir.c_str(), &sb) != 0 || !S_ISDIR(sb.st_mode)) {
            std::cout << "WARNING: Specified directory for storage of VTK files does not exist. Directory will be "
                         "created at the following path:  " << vtk.dir << std::endl;
            if (mkdir(vtk.dir.c_str(), 0755) != 0 && errno != EEXIST) {
                set_error("cannot create " + vtk.dir);
                return -1;
            }
        }
    }
    char buf[32];
    std::snprintf(buf, sizeof(buf), "%05d", tstep);
    const std::string path = vtk.dir + vtk.prefix + buf + ".vtu";
    std::cout << path << std::endl;
    if (int e = write_vtu(path.c_str(), last, vtk.label.c_str())) return e;
    std::cout << "wrote kinematics to: " << path << std::endl;
    return 0;
}

// Query batches from this size on are evaluated in cell order.
constexpr int64_t kQuerySortMin = 65536;

// Device bytes of the cell-ordering buffers of an n-query batch (keys and
// records, input and sorted halves, and the radix sort's scratch, each part
// 256-byte aligned); 0 below the ordering threshold.
template <typename TQ>
int query_ws_bytes(int64_t n, int ncol, size_t* bytes) {
    *bytes = 0;
    if (n < kQuerySortMin || n > INT32_MAX) return 0;
    int bits = 1;
    while (((int64_t)1 << bits) < ncol) ++bits;
    size_t tmp = 0;
    CFDWM_HIP(query_sort_tmp_bytes<TQ>(n, bits, &tmp));
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    *bytes = up(2 * n * sizeof(uint32_t)) + up(2 * n * sizeof(QRec<TQ>)) + up(std::max<size_t>(tmp, 256));
    return 0;
}
template int query_ws_bytes<float>(int64_t, int, size_t*);
template int query_ws_bytes<double>(int64_t, int, size_t*);

template <typename R, typename TQ>
int Grid::run_query(int64_t n, const TQ* t, double tconst, const TQ* x, const TQ* y, const TQ* z,
                    TQ* eta, TQ* u, TQ* v, TQ* w, const double* ramp, TQ* aos, void* ws, size_t ws_bytes) {
    if (!filled) { set_error("the s-grid has not been filled yet (created with defer_fill)"); return -1; }
    QueryJob<R, TQ> J;
    J.n = n; J.g = g; J.tstep = tstep; J.t0 = t0; J.dt = d.dt;
    J.tan_a = d.tan_a; J.tan_b = d.tan_b; J.tanab = std::pow(std::tan(d.tan_a), d.tan_b);
    J.pk_eta = (const R*)pk_eta_.get();
    J.pk_C = (const R*)pk_C_.get();
    J.pk_D = (const R*)pk_D_.get();
    J.t = t; J.tconst = tconst; J.x = x; J.y = y; J.z = z;
    J.ramp_scale = ramp;
    J.o_eta = eta; J.o_u = u; J.o_v = v; J.o_w = w;
    J.o_aos = aos;
    hipStream_t st = Engine::get().stream();
    const int64_t ncol = (int64_t)g.nx * g.ny;
    if (n >= kQuerySortMin && n <= INT32_MAX) {
        // large batches run in cell order (qsort.hip)
        int bits = 1;
        while (((int64_t)1 << bits) < ncol) ++bits;
        size_t tmp = 0;
        CFDWM_HIP(query_sort_tmp_bytes<TQ>(n, bits, &tmp));
        uint32_t* keys = nullptr;
        QRec<TQ>* rec = nullptr;
        void* tbuf = nullptr;
        size_t tlen = 0;
        if (ws) {
            // caller's workspace (the allocation-free entry): keys, records, sort scratch
            size_t need = 0;
            if (int e = query_ws_bytes<TQ>(n, (int)ncol, &need)) return e;
            if (need > ws_bytes) { set_error("query workspace too small"); return -1; }
            auto up = [](size_t b) { return (b + 255) / 256 * 256; };
            char* b = (char*)ws;
            keys = (uint32_t*)b;
            rec = (QRec<TQ>*)(b + up(2 * n * sizeof(uint32_t)));
            tbuf = b + up(2 * n * sizeof(uint32_t)) + up(2 * n * sizeof(QRec<TQ>));
            tlen = std::max<size_t>(tmp, 256);
        } else {
            if (qs_key_.alloc(2 * n * sizeof(uint32_t)) || qs_rec_.alloc(2 * n * sizeof(QRec<TQ>)) ||
                qs_tmp_.alloc(std::max<size_t>(tmp, 256)))
                return -3;
            keys = (uint32_t*)qs_key_.get();
            rec = (QRec<TQ>*)qs_rec_.get();
            tbuf = qs_tmp_.get();
            tlen = qs_tmp_.size();
        }
        PROF("k_query_sort", (launch_query_sort<TQ>(n, g, x, y, z, bits, keys, rec, tbuf, tlen, st)));
        J.srec = rec + n;
    }
    PROF("k_query", (launch_query<R, TQ>(J, st)));
    return 0;
}

int Grid::query(int64_t n, const double* t, const double* x, const double* y, const double* z,
                double* out4, const int* ramp_flags, const double* ramp_se) {
    (void)ramp_flags; (void)ramp_se;
    if (n <= 0) return 0;
    hipStream_t st = Engine::get().stream();
    std::lock_guard<std::mutex> lk(sea->mu);
    if (q_in_.alloc(4 * n * sizeof(double)) || q_out_.alloc(4 * n * sizeof(double))) return -3;
    double* b = (double*)q_in_.get();
    double* o = (double*)q_out_.get();
    CFDWM_HIP(hipMemcpyAsync(b, t, n * 8, hipMemcpyHostToDevice, st));
    CFDWM_HIP(hipMemcpyAsync(b + n, x, n * 8, hipMemcpyHostToDevice, st));
    CFDWM_HIP(hipMemcpyAsync(b + 2 * n, y, n * 8, hipMemcpyHostToDevice, st));
    CFDWM_HIP(hipMemcpyAsync(b + 3 * n, z, n * 8, hipMemcpyHostToDevice, st));
    // {eta, u, v, w} per point, the layout of out4
    int rc = prec == kFP32
        ? run_query<float, double>(n, b, 0.0, b + n, b + 2 * n, b + 3 * n, nullptr, nullptr, nullptr, nullptr, nullptr, o)
        : run_query<double, double>(n, b, 0.0, b + n, b + 2 * n, b + 3 * n, nullptr, nullptr, nullptr, nullptr, nullptr, o);
    if (rc) return rc;
    CFDWM_HIP(hipMemcpyAsync(out4, o, 4 * n * 8, hipMemcpyDeviceToHost, st));
    CFDWM_HIP(hipStreamSynchronize(st));
    return 0;
}

int Grid::query_device_f32(int64_t n, double t, const float* x, const float* y, const float* z,
                           float* eta, float* u, float* v, float* w, float* out4, void* ws, size_t ws_bytes) {
    std::lock_guard<std::mutex> lk(sea->mu);          // the grid's query-ordering buffers
    if (prec == kFP32) return run_query<float, float>(n, nullptr, t, x, y, z, eta, u, v, w, nullptr, out4, ws, ws_bytes);
    return run_query<double, float>(n, nullptr, t, x, y, z, eta, u, v, w, nullptr, out4, ws, ws_bytes);
}

int Grid::query_device_f64(int64_t n, const double* t, const double* x, const double* y, const double* z,
                           double* eta, double* u, double* v, double* w) {
    std::lock_guard<std::mutex> lk(sea->mu);          // the grid's query-ordering buffers
    if (prec == kFP32) return run_query<float, double>(n, t, 0.0, x, y, z, eta, u, v, w, nullptr);
    return run_query<double, double>(n, t, 0.0, x, y, z, eta, u, v, w, nullptr);
}

}  // namespace cfdwm
