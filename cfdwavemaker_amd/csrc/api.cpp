// C ABI of libcfdwm: the reference's wave_* entry points (process-global wave
// field configured from waveinput.dat) and the batched / object entries.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <array>
#include <atomic>
#include <unordered_map>
#include <mutex>
#include <shared_mutex>
#include <fstream>
#include <iostream>
#include <memory>
#include <numeric>
#include <sstream>
#include <string>
#include <unistd.h>
#include <vector>

#include "../../include/cfdwm.h"
#include "engine.hpp"
#include "depth_basis.hpp"
#include "sea_tables.hpp"

struct cfdwm_sea {
    std::unique_ptr<cfdwm::Sea> s;
    bool with_direct = false;
};
struct cfdwm_grid {
    std::unique_ptr<cfdwm::Grid> g;
};
struct cfdwm_lsgrid {
    std::unique_ptr<cfdwm::LsGrid> g;
};

namespace cfdwm {
namespace {

// ------------------------------------------------------------------ Ramp
// Ramp::ramp (src/Utils.cpp:26-68): smallest of the active linear ramps.
struct RampCfg {
    bool init = false;
    bool on[6] = {false, false, false, false, false, false};
    double se[12] = {0};
    static double r1d(double x, double xs, double xe, bool inv) {
        const double v = inv ? 1. - ((x - xs) / (xe - xs)) : ((x - xs) / (xe - xs));
        return std::max(std::min(v, 1.), 0.);
    }
    double operator()(double t, double x, double y) const {
        if (!init) return 1.0;
        double r[6] = {1.0, 1.0, 1.0, 1.0, 1.0, 1.0};
        const double a[6] = {t, t, x, x, y, y};
        for (int q = 0; q < 6; ++q)
            if (on[q]) r[q] = r1d(a[q], se[2 * q], se[2 * q + 1], q % 2 == 1);
        return *std::min_element(r, r + 6);
    }
};

// ---------------------------------------------------- process-global state
// Mirrors the singletons of src/CFDwavemaker.cpp:69-163.
struct Global {
    bool initialized = false;
    int wavetype = 0;                 // 1 direct irregular, 5 s-grid irregular, others unsupported
    double depth = 0, x_pos = 0, y_pos = 0, tofmax = 0, mtheta = 0, swl = 0;
    double rho = 1., gravity = 9.81, ampl = 1.;
    bool bw_auto = false, property_read = false;
    // irregular
    SeaInput sea;
    int nfreq = 0, ndir = 1, extrapolation = 0, order = 2;
    bool normalize = false, individual_ramp = false;
    int vertical_theory = 0;
    // lsgrid
    bool has_grid = false;
    double lsg_domain[4] = {0, 0, 0, 0};
    int lsg_nx = 5, lsg_ny = 0, lsg_nl = 4;
    double lsg_t0 = 0., lsg_dt = 0.1, lsg_tan_a = 0., lsg_tan_b = 0.;
    double lsg_ignore[4] = {1.E12, -1.E12, 1.E12, -1.E12};
    bool lsg_ignore_domain = false, lsg_ignore_at_init = false, lsg_init_only = false;
    // [lsgrid2vtk] (lsgrid_spline.h:58-61 defaults)
    bool vtk_on = false, vtk_last_done = false;
    std::string vtk_dir = "./", vtk_prefix = "kinematics", vtk_label = "TimeValue";
    RampCfg ramp;
    int precision = kFP64;
    cfdwm_sea* csea = nullptr;
    cfdwm_grid* cgrid = nullptr;
};
Global G;

// Device scratch of wave_EvalPointsDevice (kept between calls; one caller at
// a time).
struct DevPoints {
    std::mutex mu;
    DevBuf z, e1, e2, z0, q4, e4;
};
DevPoints g_dev;

// Device staging of wave_EvalPoints on the s-grid (inputs, outputs).
struct StagePoints {
    std::mutex mu;
    DevBuf in, out;
};
StagePoints g_stage;

RampDev ramp_dev(const RampCfg& r) {
    RampDev d;
    d.init = r.init ? 1 : 0;
    for (int q = 0; q < 6; ++q) d.on[q] = r.on[q] ? 1 : 0;
    for (int q = 0; q < 12; ++q) d.se[q] = r.se[q];
    return d;
}

void trim(std::string& s) {
    auto ns = [](int ch) { return !std::isspace(ch); };
    s.erase(s.begin(), std::find_if(s.begin(), s.end(), ns));
    s.erase(std::find_if(s.rbegin(), s.rend(), ns).base(), s.end());
}
bool starts(const std::string& s, const char* key) { return s.compare(0, std::strlen(key), key) == 0; }

[[noreturn]] void input_error(const std::string& msg) {
    std::cerr << "INPUTFILE ERROR: " << msg << std::endl;
    std::exit(-1);
}

template <typename T>
void read_after_key(const std::string& line, T& v) {
    std::istringstream b(line);
    std::string key;
    b >> key >> v;
}

// waveinput.dat sections for irregular waves and the s-grid
// (src/CFDwavemaker.cpp:232-1203).  Line handling follows the reference:
// a section loop reads trimmed lines until a line starting with '[', which is
// then examined by the following section tests without reading a new line.
void parse(const std::string& res) {
    std::istringstream f(res);
    std::string lineA;
    bool skip = false;
    while (!f.eof()) {
        if (skip) skip = false;
        else { std::getline(f, lineA); trim(lineA); }
        if (lineA == "[wave type]") {
            std::getline(f, lineA);
            trim(lineA);
            if (lineA == "irregular" || lineA == "1") G.wavetype = 1;
            else if (lineA == "piston") G.wavetype = 11;
            else if (lineA == "regular") G.wavetype = 21;
            else if (lineA == "swd") G.wavetype = 31;
            else if (lineA == "vtk") G.wavetype = 41;
            else if (lineA == "slope") G.wavetype = 51;
            else input_error("Unknown wave type specified. Valid alternatives are: irregular regular piston swd vtk slope");
            if (G.wavetype != 1)
                std::cerr << "cfdwm: wave type '" << lineA << "' is not supported in this build; "
                          << "kinematics calls return 0.0" << std::endl;
        }
        if (lineA == "[general input data]") {
            while (!f.eof()) {
                std::getline(f, lineA);
                trim(lineA);
                if (starts(lineA, "depth")) read_after_key(lineA, G.depth);
                if (starts(lineA, "mtheta")) read_after_key(lineA, G.mtheta);
                if (starts(lineA, "normalize")) read_after_key(lineA, G.normalize);
                if (starts(lineA, "amplify")) read_after_key(lineA, G.ampl);
                if (starts(lineA, "swl")) read_after_key(lineA, G.swl);
                if (starts(lineA, "gravity")) read_after_key(lineA, G.gravity);
                if (starts(lineA, "rho")) read_after_key(lineA, G.rho);
                if (starts(lineA, "[")) { skip = true; break; }
            }
        }
        if (lineA == "[wave reference point]") {
            while (!f.eof()) {
                std::getline(f, lineA);
                trim(lineA);
                if (starts(lineA, "time")) read_after_key(lineA, G.tofmax);
                if (starts(lineA, "x")) read_after_key(lineA, G.x_pos);
                if (starts(lineA, "y")) read_after_key(lineA, G.y_pos);
                if (starts(lineA, "[")) { skip = true; break; }
            }
        }
        if (lineA == "[ramps]") {
            G.ramp.init = true;
            static const char* keys[6] = {"time_rampup", "time_rampdown", "x_rampup", "x_rampdown", "y_rampup", "y_rampdown"};
            while (!f.eof()) {
                std::getline(f, lineA);
                trim(lineA);
                for (int q = 0; q < 6; ++q) {
                    if (starts(lineA, keys[q])) {
                        std::istringstream b(lineA);
                        std::string key;
                        b >> key >> G.ramp.se[2 * q] >> G.ramp.se[2 * q + 1];
                        G.ramp.on[q] = true;
                    }
                }
                if (starts(lineA, "[")) { skip = true; break; }
            }
        }
        if (lineA == "[wave input irregular]") {
            if (G.wavetype != 1) input_error("irregular wave components does not match the specified wave type. Check inputfile");
            while (!f.eof()) {
                std::getline(f, lineA);
                trim(lineA);
                if (starts(lineA, "kinematics_surf_extrap")) {
                    read_after_key(lineA, G.extrapolation);
                    if (G.extrapolation == 2) G.order = 2;
                    else if (G.extrapolation == 1 || G.extrapolation == 0) G.order = 0;
                    else { std::cout << "Unknown extrapolation method. try a value between 0 and 2." << std::endl; std::exit(-1); }
                }
                if (starts(lineA, "cutoff")) read_after_key(lineA, G.sea.cutoff);
                if (starts(lineA, "bandwidth")) {
                    std::string v;
                    read_after_key(lineA, v);
                    if (!v.compare(0, 3, "off")) {}
                    else if (!v.compare(0, 4, "auto")) G.bw_auto = true;
                    else G.sea.bandwidth = std::atof(v.c_str());
                }
                if (starts(lineA, "individual_ramp")) {
                    std::string v;
                    read_after_key(lineA, v);
                    if (!v.compare(0, 4, "true")) G.individual_ramp = true;
                }
                if (starts(lineA, "vertical_theory")) {
                    std::string v;
                    read_after_key(lineA, v);
                    G.vertical_theory = !v.compare(0, 4, "smit") ? 1 : 0;
                    if (G.vertical_theory == 1)
                        std::cout << "Smit et al. (2017) s-coordinate vertical theory selected (experimental in "
                                     "the reference: off-diagonal second-order terms not yet validated)." << std::endl;
                }
                if (starts(lineA, "nfreq")) break;
            }
            if (starts(lineA, "nfreq")) read_after_key(lineA, G.nfreq);
            else input_error("parameter nfreq missing or not specified correctly");
            std::getline(f, lineA);
            trim(lineA);
            if (starts(lineA, "ndir")) read_after_key(lineA, G.ndir);
            else input_error("parameter ndir missing or not specified correctly");
            std::vector<double> om, A, k, ph, th, rt, ts, du;
            if (G.ndir == 0) {
                // one direction per frequency row (:596-705), optional ramp triplet
                G.ndir = 1;
                for (int i = 0; i < G.nfreq; ++i) {
                    std::getline(f, lineA);
                    if (starts(lineA, "["))
                        input_error("Not enough frequency components specified. please check to make sure that the number of lines matches the specified number of frequency components");
                    std::istringstream b(lineA);
                    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
                    const int nv = G.individual_ramp ? 8 : 5;
                    for (int q = 0; q < nv; ++q) b >> v[q];
                    om.push_back(v[0]); A.push_back(v[1]); k.push_back(v[2]); ph.push_back(v[3]); th.push_back(v[4]);
                    if (G.individual_ramp) { rt.push_back(v[5]); ts.push_back(v[6]); du.push_back(v[7]); }
                    else { rt.push_back(0.); du.push_back(kHuge); ts.push_back(-kSmall); }
                }
            } else {
                // frequency rows then direction rows, expanded frequency-major (:706-797)
                std::vector<double> fo(G.nfreq), fa(G.nfreq), fk(G.nfreq), fp(G.nfreq);
                for (int i = 0; i < G.nfreq; ++i) {
                    std::getline(f, lineA);
                    std::istringstream b(lineA);
                    b >> fo[i] >> fa[i] >> fk[i] >> fp[i];
                }
                std::vector<double> dt_(G.ndir), dd(G.ndir);
                for (int j = 0; j < G.ndir; ++j) {
                    std::getline(f, lineA);
                    if (starts(lineA, "[")) input_error("too few directional components..");
                    std::istringstream b(lineA);
                    b >> dt_[j] >> dd[j];
                }
                for (int i = 0; i < G.nfreq; ++i)
                    for (int j = 0; j < G.ndir; ++j) {
                        om.push_back(fo[i]); k.push_back(fk[i]); A.push_back(fa[i] * dd[j]);
                        ph.push_back(fp[i]); th.push_back(dt_[j]);
                        rt.push_back(0.); du.push_back(kHuge); ts.push_back(-kSmall);
                    }
            }
            // stable sort by omega (:663-673)
            std::vector<size_t> idx(om.size());
            std::iota(idx.begin(), idx.end(), 0);
            std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return om[a] < om[b]; });
            SeaInput& S = G.sea;
            for (size_t q : idx) {
                S.omega.push_back(om[q]); S.A.push_back(A[q]); S.k.push_back(k[q]);
                S.phase.push_back(ph[q]); S.theta.push_back(th[q]);
                S.ramptime.push_back(rt[q]); S.starttime.push_back(ts[q]); S.duration.push_back(du[q]);
            }
            G.property_read = true;
        }
        if (lineA == "[lsgrid]") {
            while (!f.eof()) {
                std::getline(f, lineA);
                trim(lineA);
                if (starts(lineA, "bounds")) {
                    std::istringstream b(lineA);
                    std::string key;
                    b >> key >> G.lsg_domain[0] >> G.lsg_domain[1] >> G.lsg_domain[2] >> G.lsg_domain[3];
                }
                if (starts(lineA, "nx")) read_after_key(lineA, G.lsg_nx);
                if (starts(lineA, "ny")) read_after_key(lineA, G.lsg_ny);
                if (starts(lineA, "nl")) read_after_key(lineA, G.lsg_nl);
                if (starts(lineA, "t0")) read_after_key(lineA, G.lsg_t0);
                if (starts(lineA, "dt")) read_after_key(lineA, G.lsg_dt);
                if (starts(lineA, "stretch_params")) {
                    std::istringstream b(lineA);
                    std::string key;
                    b >> key >> G.lsg_tan_a >> G.lsg_tan_b;
                }
                if (starts(lineA, "ignore_subdomain")) {
                    std::istringstream b(lineA);
                    std::string key;
                    b >> key >> G.lsg_ignore[0] >> G.lsg_ignore[1] >> G.lsg_ignore[2] >> G.lsg_ignore[3];
                    G.lsg_ignore_domain = true;
                }
                if (starts(lineA, "ignore_at_init")) read_after_key(lineA, G.lsg_ignore_at_init);
                if (starts(lineA, "init_only")) read_after_key(lineA, G.lsg_init_only);
                if (starts(lineA, "[")) { skip = true; break; }
            }
            if (G.wavetype == 1) G.wavetype = 5;
            else if (G.wavetype == 31) G.wavetype = 34;
            else input_error("LS grid may currently only be used with irregular second order wave theory or swd");
            G.has_grid = true;
        }
        if (lineA == "[lsgrid2vtk]") {                // optional (src/CFDwavemaker.cpp:1205-1244)
            std::cout << "--------------------" << std::endl << "VTK output settings:" << std::endl
                      << "--------------------" << std::endl;
            if (G.wavetype != 5) {
                std::cout << "InputError: VTK output only works in combination with LSgrid at the moment." << std::endl;
                std::exit(-1);
            }
            G.vtk_on = true;
            while (!f.eof()) {
                std::getline(f, lineA);
                trim(lineA);
                std::istringstream b(lineA);
                std::string key;
                if (starts(lineA, "storage_path")) {
                    b >> key >> G.vtk_dir;
                    std::cout << "Directory for storage of vtk files: " << G.vtk_dir << std::endl;
                }
                if (starts(lineA, "filename")) {
                    b >> key >> G.vtk_prefix;
                    std::cout << "filename prefix: " << G.vtk_prefix << std::endl;
                }
                if (starts(lineA, "vtk_timelabel")) {
                    b >> key >> G.vtk_label;
                    std::cout << "time label to use in vtk specified as: " << G.vtk_label << std::endl;
                }
                if (starts(lineA, "[")) { skip = true; break; }
            }
        }
    }
}

// Irregular::dumpSpectralComponents (src/Irregular.cpp:1456-1468)
void dump_spectral_components() {
    FILE* fp = std::fopen("./spectral_components.dat", "w");
    if (!fp) return;
    std::fprintf(fp, "# spectral wave data output\n# [irregular wave components]\n# nfreq %d\n# ndir 0\n", G.nfreq);
    std::fprintf(fp, "# omega [rad/s]    a[m]       k[rad/m]        phi[rad]    theta[rad]\n");
    const SeaInput& S = G.sea;
    for (int i = 0; i < G.nfreq && i < (int)S.omega.size(); ++i)
        std::fprintf(fp, "%12.5E  %12.5E  %12.5E  %12.5E  %12.5E\n", S.omega[i], S.A[i], S.k[i], S.phase[i], S.theta[i]);
    std::fclose(fp);
}

// CFDWM_PRECISION: fp64 (default), fp32, or bf16 (FP32 with the
// one-product bf16 quadratic forms); the values of cfdwm_sea_params.precision
int env_precision() {
    const char* e = std::getenv("CFDWM_PRECISION");
    if (!e) return 0;
    if (!std::strcmp(e, "fp32") || !std::strcmp(e, "FP32")) return 1;
    if (!std::strcmp(e, "bf16") || !std::strcmp(e, "BF16")) return 2;
    return 0;
}

[[noreturn]] void gpu_fatal(const char* where) {
    std::cerr << "cfdwm: " << where << " failed: " << last_error() << std::endl;
    std::abort();
}

// post-processing of process_inputdata (src/CFDwavemaker.cpp:1334-1437)
// The end-of-run .vtu of the slot after the base slot (the reference writes it
// from its grid's destructor): once, at wave_Cleanup or at process exit.
void vtk_final_dump() {
    if (G.vtk_on && !G.vtk_last_done && G.cgrid) {
        G.vtk_last_done = true;
        (void)G.cgrid->g->write_vtk_dump(true);
    }
}

int setup_after_parse() {
    if (G.wavetype == 5 && G.order != 2)
        input_error("LSgrid interpolation uses strictly second order theory. Set kinematics_surf_extrap to 2 under [wave input irregular], or remove [lsgrid].");
    if (G.wavetype != 1 && G.wavetype != 5) return 0;
    if (!G.property_read) input_error("Irregular wave selected, but no wave components/wave properties found in input file.");
    SeaInput& S = G.sea;
    S.depth = G.depth; S.mtheta = G.mtheta; S.tofmax = G.tofmax; S.x0 = G.x_pos; S.y0 = G.y_pos;
    S.swl = G.swl; S.g = G.gravity; S.rho = G.rho; S.order = G.order;
    S.nfreq = G.nfreq;
    if (G.bw_auto) {                                                    // bandwidth_estimator
        if (G.ndir > 1) input_error("Automatic calculation of bandwidth parameter only supported for 2D wave spectra");
        double m[3];
        spectral_moments(S, m);
        S.bandwidth = 0.7 * (m[1] / m[0]);
    }
    const double sumA = std::accumulate(S.A.begin(), S.A.end(), 0.);   // normalize_data
    for (size_t i = 0; i < S.A.size(); ++i) {
        if (G.normalize) { S.A[i] = G.ampl * S.A[i] / sumA; S.phase[i] = 0.; }
        else S.A[i] = G.ampl * S.A[i];
    }
    dump_spectral_components();
    {
        double m[3];
        spectral_moments(S, m);
        std::cout << "cfdwm: " << S.omega.size() << " components, Hs " << 4 * std::sqrt(m[0])
                  << " m, Tm " << mean_period(S, 1) << " s, Tz " << mean_period(S, 2) << " s" << std::endl;
    }
    G.precision = env_precision();
    cfdwm_sea_params p{};
    p.depth = S.depth; p.g = S.g; p.rho = S.rho; p.swl = S.swl; p.mtheta = S.mtheta;
    p.tofmax = S.tofmax; p.x0 = S.x0; p.y0 = S.y0; p.bandwidth = S.bandwidth; p.cutoff = S.cutoff;
    p.ampl = 1.0; p.order = S.order; p.normalize = 0; p.bw_auto = 0; p.precision = G.precision;
    p.with_direct = G.wavetype == 1 ? 1 : 0;
    const int n = (int)S.omega.size();
    int rc = cfdwm_sea_create(&G.csea, n, S.omega.data(), S.A.data(), S.k.data(), S.phase.data(), S.theta.data(),
                              S.ramptime.data(), S.starttime.data(), S.duration.data(), &p);
    if (rc) gpu_fatal("sea set-up");
    G.csea->s->vertical_theory = G.vertical_theory;        // Irregular::u/v/w/dp dispatch
    if (G.wavetype == 5) {
        if (G.lsg_ny < 1 || G.lsg_nx < 1 || G.lsg_nl < 1)
            input_error("lsgrid needs nx, ny and nl of at least 1");
        cfdwm_grid_desc d{};
        std::memcpy(d.domain, G.lsg_domain, sizeof(d.domain));
        d.nx = G.lsg_nx; d.ny = G.lsg_ny; d.nl = G.lsg_nl; d.t0 = G.lsg_t0; d.dt = G.lsg_dt;
        d.tan_a = G.lsg_tan_a; d.tan_b = G.lsg_tan_b; d.water_depth = G.depth; d.swl = G.swl;
        std::memcpy(d.domain_ignore, G.lsg_ignore, sizeof(d.domain_ignore));
        d.ignore_at_init = G.lsg_ignore_at_init ? 1 : 0;
        d.defer_fill = 1;                             // the dump settings first, then the fill
        rc = cfdwm_grid_create(&G.cgrid, G.csea, &d);
        if (rc) gpu_fatal("s-grid initialisation");
        if (G.vtk_on) {
            Grid& gr = *G.cgrid->g;
            gr.vtk.on = true;
            gr.vtk.dir = G.vtk_dir;
            gr.vtk.prefix = G.vtk_prefix;
            gr.vtk.label = G.vtk_label;
            // the reference writes the following slot once more when its grid
            // is destroyed at process exit (lsgrid_spline.h:25-28)
            static bool registered = false;
            if (!registered) {
                std::atexit([] { vtk_final_dump(); });
                registered = true;
            }
        }
        if (cfdwm_grid_refill(G.cgrid)) gpu_fatal("s-grid initialisation");
    }
    return 0;
}

// ---------------------------------------------------- direct point helpers
uint32_t order_flags(int order) {
    if (order == 2) return EV_LIN | EV_SEC | EV_DIRECT | EV_TAYLOR | EV_ZCLAMP;
    if (order == 1) return EV_LIN | EV_SEC | EV_DIRECT | EV_ZCLAMP;
    return EV_LIN | EV_ZSHIFT;
}

// One direct-mode point: eta, u, v, w (and p) of Irregular::eta/u/v/w/dp.
void direct_point(double x, double y, double z, double t, double* eta, double* u, double* v, double* w, double* p) {
    Sea& s = *G.csea->s;
    double e1 = 0, e2 = 0;
    const bool we = eta != nullptr;
    uint32_t fl = order_flags(G.order) | (p ? EV_P : 0u);
    int rc = s.eval_points(1, &x, &y, &z, &t, fl, G.order == 2, we ? &e1 : nullptr, we && G.order == 2 ? &e2 : nullptr,
                           nullptr, u, v, w, p, false);
    if (rc) gpu_fatal("point evaluation");
    if (eta) *eta = G.swl + e1 + (G.order == 2 ? e2 : 0.0);
}


// ------------------------------------------------- prefetched point values
// A CFD host knows the points it will ask for in a time step: evaluating
// them in one batch (wave_PrefetchPoints) and serving the per-point calls
// from the results turns one GPU round trip per call into one per batch.
// Keys are the bit patterns of the call arguments; values are the final API
// values (z clamp and ramps applied), so a hit returns exactly what the
// uncached call returns.
struct PKey {
    uint64_t a, b, c, d;
    bool operator==(const PKey& o) const { return a == o.a && b == o.b && c == o.c && d == o.d; }
};
struct PKeyHash {
    size_t operator()(const PKey& k) const noexcept {
        uint64_t h = k.a * 0x9E3779B97F4A7C15ull;
        h ^= (k.b + 0x632BE59BD9B4E019ull + (h << 6) + (h >> 2));
        h ^= (k.c * 0xBF58476D1CE4E5B9ull + (h << 6) + (h >> 2));
        h ^= (k.d * 0x94D049BB133111EBull + (h << 6) + (h >> 2));
        return (size_t)h;
    }
};
struct PVals {
    double u, v, w, p;
};
struct PointCache {
    std::shared_mutex mu;
    std::unordered_map<PKey, PVals, PKeyHash> vel;     // (x, y, z, t)
    std::unordered_map<PKey, double, PKeyHash> eta;    // (x, y, 0, t)
    bool has_vel = false, has_p = false, has_eta = false;
    std::atomic<int64_t> hits{0}, misses{0};
};
PointCache PC;
inline uint64_t bits(double v) {
    uint64_t u;
    std::memcpy(&u, &v, 8);
    return u;
}
bool cached_vel(double x, double y, double z, double t, PVals& out) {
    std::shared_lock<std::shared_mutex> lk(PC.mu);
    if (!PC.has_vel) return false;
    auto it = PC.vel.find(PKey{bits(x), bits(y), bits(z), bits(t)});
    if (it == PC.vel.end()) { ++PC.misses; return false; }
    ++PC.hits;
    out = it->second;
    return true;
}
bool cached_eta(double x, double y, double t, double& out) {
    std::shared_lock<std::shared_mutex> lk(PC.mu);
    if (!PC.has_eta) return false;
    auto it = PC.eta.find(PKey{bits(x), bits(y), 0, bits(t)});
    if (it == PC.eta.end()) { ++PC.misses; return false; }
    ++PC.hits;
    out = it->second;
    return true;
}

// ------------------------------------------ per-point calls on the s-grid
// A CFD host that calls wave_SurfElev / wave_VeloX .. per cell asks for the
// same points every time step.  On the s-grid each per-point call is one
// spline query; the calls of one time value and grid state form a generation
// whose points are remembered (the cache keys).  The first call at a new
// time (or after wave_force_update advanced the grid) evaluates every point
// of the previous generation at the new time in one batched query, and later
// calls for those points are served from the cache under a shared lock, from
// any number of host threads.  A point not in the cache is queried on its
// own, as without the cache, and joins it.  The spline query evaluates every
// point identically whatever the batch, so served values equal the uncached
// ones bit for bit.  wave_PrefetchAuto(0) switches this off.
struct K3 {
    uint64_t a, b, c;
    bool operator==(const K3& o) const { return a == o.a && b == o.b && c == o.c; }
};
struct K3Hash {
    size_t operator()(const K3& k) const noexcept {
        uint64_t h = k.a * 0x9E3779B97F4A7C15ull;
        h ^= (k.b + 0x632BE59BD9B4E019ull + (h << 6) + (h >> 2));
        h ^= (k.c * 0xBF58476D1CE4E5B9ull + (h << 6) + (h >> 2));
        return (size_t)h;
    }
};
struct AutoCache {
    std::shared_mutex mu;
    std::atomic<bool> on{true};
    bool valid = false;                   // entries belong to (t_bits, version)
    uint64_t t_bits = 0, version = 0;
    std::unordered_map<K3, std::array<double, 4>, K3Hash> m;   // (x, y, z of the query) -> {eta, u, v, w} x ramp
    std::atomic<int64_t> hits{0}, misses{0}, replays{0}, replayed{0};
};
AutoCache AC;
std::atomic<uint64_t> g_grid_version{0};
constexpr size_t kAutoCacheMax = size_t(1) << 23;

// {eta, u, v, w} x ramp of the s-grid query at (x, y, zq, t), through the cache.
void grid_kin(double x, double y, double zq, double t, double* o4) {
    auto single = [&]() {
        if (G.cgrid->g->query(1, &t, &x, &y, &zq, o4, nullptr, nullptr)) gpu_fatal("s-grid query");
        const double r = G.ramp(t, x, y);
        for (int q = 0; q < 4; ++q) o4[q] *= r;
    };
    if (!AC.on.load(std::memory_order_relaxed)) { single(); return; }
    const K3 key{bits(x), bits(y), bits(zq)};
    const uint64_t tb = bits(t), ver = g_grid_version.load();
    {
        std::shared_lock<std::shared_mutex> lk(AC.mu);
        if (AC.valid && AC.t_bits == tb && AC.version == ver) {
            auto it = AC.m.find(key);
            if (it != AC.m.end()) {
                for (int q = 0; q < 4; ++q) o4[q] = it->second[q];
                ++AC.hits;
                return;
            }
        }
    }
    std::unique_lock<std::shared_mutex> lk(AC.mu);
    if (!(AC.valid && AC.t_bits == tb && AC.version == ver)) {
        // a new generation: the previous generation's points at time t, one batch
        const size_t n = AC.m.size() <= kAutoCacheMax ? AC.m.size() : 0;
        std::vector<double> px(n), py(n), pz(n), pt(n, t), out(4 * n);
        size_t k = 0;
        for (const auto& e : AC.m) {
            std::memcpy(&px[k], &e.first.a, 8);
            std::memcpy(&py[k], &e.first.b, 8);
            std::memcpy(&pz[k], &e.first.c, 8);
            ++k;
        }
        AC.m.clear();
        if (n) {
            if (G.cgrid->g->query((int64_t)n, pt.data(), px.data(), py.data(), pz.data(), out.data(), nullptr, nullptr))
                gpu_fatal("s-grid query");
            AC.m.reserve(n);
            for (size_t q = 0; q < n; ++q) {
                const double r = G.ramp(t, px[q], py[q]);
                AC.m[K3{bits(px[q]), bits(py[q]), bits(pz[q])}] = {out[4 * q] * r, out[4 * q + 1] * r,
                                                                    out[4 * q + 2] * r, out[4 * q + 3] * r};
            }
            ++AC.replays;
            AC.replayed += (int64_t)n;
        }
        AC.t_bits = tb;
        AC.version = ver;
        AC.valid = true;
        auto it = AC.m.find(key);
        if (it != AC.m.end()) {
            for (int q = 0; q < 4; ++q) o4[q] = it->second[q];
            ++AC.hits;
            return;
        }
    }
    single();
    AC.m[key] = {o4[0], o4[1], o4[2], o4[3]};
    ++AC.misses;
}

void auto_cache_reset() {
    std::unique_lock<std::shared_mutex> lk(AC.mu);
    AC.m.clear();
    AC.valid = false;
}

double velo(double x, double y, double z, double t, int comp) {
    if (!G.initialized) return 0.0;
    {
        PVals c;
        if (cached_vel(x, y, z, t, c)) return comp == 1 ? c.u : (comp == 2 ? c.v : c.w);
    }
    z = std::max(-G.depth, z);                       // CFDwavemaker.cpp:1547
    if (G.wavetype == 5) {
        double o4[4];
        grid_kin(x, y, z, t, o4);
        return o4[comp];
    }
    if (G.wavetype == 1) {
        double r = 0;
        direct_point(x, y, z, t, nullptr, comp == 1 ? &r : nullptr, comp == 2 ? &r : nullptr, comp == 3 ? &r : nullptr, nullptr);
        return G.ramp(t, x, y) * r;
    }
    return 0.0;
}

}  // namespace
}  // namespace cfdwm

using namespace cfdwm;

// =================================================================== Part 1
extern "C" {

int CFDwavemaker_is_initialized(void) { return G.initialized ? 1 : 0; }

int wave_Initialize(void) {
    std::cout << "CFDWAVEMAKER (cfdwm, MI355X) " << std::endl;
    const char* paths[4] = {"./waveinput.dat", "./constant/waveinput.dat", "../input_files/waveinput.dat", "../waveinput.dat"};
    std::ifstream fid;
    for (const char* p : paths) {
        fid.open(p);
        if (!fid.fail()) break;
        fid.clear();
    }
    if (!fid.is_open() || fid.fail()) {
        std::cerr << "Could not open file (is it really there?) " << std::endl;
        return -1;
    }
    std::string res, line;
    while (fid.good()) {                              // strip '#' comments, drop empty lines
        std::getline(fid, line);
        line.erase(std::find(line.begin(), line.end(), '#'), line.end());
        if (line.length() > 0) res += line + "\n";
    }
    fid.close();
    std::istringstream f(res);
    std::getline(f, line);
    trim(line);
    if (line.substr(0, 2) != "@v") {
        std::cout << "Please specify the input file version on the first line of the input file (hint: @v310)" << std::endl;
        std::exit(0);
    }
    if (std::stoi(line.substr(2, 3)) < 300) {
        std::cout << "your waveinput.dat file is outdated and does not match the current version of CFDwavemaker." << std::endl;
        std::exit(0);
    }
    if (G.csea || G.cgrid) wave_Cleanup();
    auto_cache_reset();
    G = Global();
    parse(res);
    setup_after_parse();
    G.initialized = true;
    return 0;
}

int wave_Cleanup(void) {
    vtk_final_dump();
    wave_PrefetchClear();
    auto_cache_reset();
    if (G.cgrid) { cfdwm_grid_destroy(G.cgrid); G.cgrid = nullptr; }
    if (G.csea) { cfdwm_sea_destroy(G.csea); G.csea = nullptr; }
    return 0;
}

double wave_VeloX(double x, double y, double z, double t) { return velo(x, y, z, t, 1); }
double wave_VeloY(double x, double y, double z, double t) { return velo(x, y, z, t, 2); }
double wave_VeloZ(double x, double y, double z, double t) { return velo(x, y, z, t, 3); }

double wave_DynPres(double x, double y, double z, double t) {
    if (!G.initialized || G.wavetype != 1) return 0.0;
    {
        PVals c;
        if (PC.has_p && cached_vel(x, y, z, t, c)) return c.p;
    }
    z = std::max(-G.depth, z);
    double p = 0;
    direct_point(x, y, z, t, nullptr, nullptr, nullptr, nullptr, &p);
    return G.ramp(t, x, y) * p;
}

double wave_nhPres(double x, double y, double z, double t) {
    if (!G.initialized || G.wavetype != 1) return 0.0;
    z = std::max(-G.depth, z);
    double p = 0, eta = 0;
    direct_point(x, y, z, t, &eta, nullptr, nullptr, nullptr, &p);
    return G.ramp(t, x, y) * (p - G.rho * G.gravity * eta);
}

double wave_SurfElev(double x, double y, double t) {
    if (!G.initialized) return 0.0;
    {
        double e;
        if (cached_eta(x, y, t, e)) return e;
    }
    if (G.wavetype == 1) {
        double eta = 0;
        direct_point(x, y, 0.0, t, &eta, nullptr, nullptr, nullptr, nullptr);
        return G.ramp(t, x, y) * eta;
    }
    if (G.wavetype == 5) {
        double o4[4];
        grid_kin(x, y, 0.0, t, o4);
        return o4[0];
    }
    return 0.0;
}

double wave_VFrac(double x, double y, double z, double t, double dc) {
    const double e = wave_SurfElev(x, y, t);
    if (e < z - (dc / 2.)) return 0.0;
    if (e > z + (dc / 2.)) return 1.0;
    return (e - (z - (dc / 2.))) / dc;
}

double wave_Seabed(double, double) { return -G.depth; }

double* wave_Kinematics(double x, double y, double z, double t) {
    thread_local double res[4];
    res[0] = res[1] = res[2] = res[3] = 0.0;
    if (!G.initialized) return res;
    {
        PVals c;
        double e;
        if (cached_eta(x, y, t, e) && cached_vel(x, y, z, t, c)) {
            res[0] = e; res[1] = c.u; res[2] = c.v; res[3] = c.w;
            return res;
        }
    }
    if (G.wavetype == 5) {
        grid_kin(x, y, z, t, res);
    } else if (G.wavetype == 1) {
        const double zc = std::max(-G.depth, z);
        direct_point(x, y, zc, t, &res[0], &res[1], &res[2], &res[3], nullptr);
        const double r = G.ramp(t, x, y);
        for (int q = 0; q < 4; ++q) res[q] *= r;
    }
    return res;
}

double wave_piston_position(double) { return 0.0; }

void wave_force_update(double t) {
    if (!G.initialized || G.wavetype != 5) return;
    int n = 0;
    if (cfdwm_grid_force_update(G.cgrid, t, &n)) gpu_fatal("s-grid update");
    if (n > 0) {
        wave_PrefetchClear();                         // kept values belong to the previous grid state
        ++g_grid_version;                             // the per-point cache starts a new generation
    }
}

void wave_set_mpi_comm(int) {}
void update_probes(double) {}

double wave_phase_speed(int opt) { return G.wavetype == 1 ? phase_speed(G.sea, opt) : 0.; }
double wave_mean_period(int opt) { return G.wavetype == 1 ? mean_period(G.sea, opt) : 0.; }
double wave_mean_length(int opt) { return G.wavetype == 1 ? mean_length(G.sea, opt) : 0.; }
double wave_water_depth(void) { return G.depth; }

double VelocityX(int, int, int, double x, double y, double z, double t) { return wave_VeloX(x, y, z, t); }
double VelocityY(int, int, int, double x, double y, double z, double t) { return wave_VeloY(x, y, z, t); }
double VelocityZ(int, int, int, double x, double y, double z, double t) { return wave_VeloZ(x, y, z, t); }
double DynamicPressure(int, int, int, double x, double y, double z, double t) { return wave_DynPres(x, y, z, t); }
double SurfaceElevation(int, int, double x, double y, double t) { return wave_SurfElev(x, y, t); }
double VolumeFraction(double x, double y, double z, double t, double d) { return wave_VFrac(x, y, z, t, d); }
int Init(void*) { return wave_Initialize(); }
int Prepare(double, int) { return 0; }
int Cleanup(void) { return wave_Cleanup(); }

// =================================================================== Part 2
static int eval_points_impl(int64_t n, const double* x, const double* y, const double* z, const double* t,
                            double* eta, double* u, double* v, double* w, double* p, uint32_t flags, bool dev) {
    if (!G.initialized) { set_error("wave_Initialize has not been called"); return -1; }
    if (n < 0 || !x || !y || !t || ((u || v || w || p) && !z)) { set_error("bad arguments"); return -1; }
    if (n == 0) return 0;
    if (G.wavetype == 1 && dev) {
        // device-resident: z clamp, the batch, then elevation sum and ramps
        // in one pass (the host branch's operations, in its order)
        Sea& s = *G.csea->s;
        hipStream_t st = Engine::get().stream();
        std::lock_guard<std::mutex> lk(g_dev.mu);
        const double* zz = z;
        if (z) {
            if (g_dev.z.alloc(n * sizeof(double))) return -3;
            CFDWM_HIP(launch_clamp_below(n, z, -G.depth, (double*)g_dev.z.get(), st));
            zz = (const double*)g_dev.z.get();
        }
        double *e1 = nullptr, *e2 = nullptr;
        if (eta) {
            if (g_dev.e1.alloc(n * sizeof(double)) || g_dev.e2.alloc(n * sizeof(double))) return -3;
            e1 = (double*)g_dev.e1.get();
            e2 = G.order == 2 ? (double*)g_dev.e2.get() : nullptr;
        }
        const uint32_t fl = order_flags(G.order) | (p ? EV_P : 0u);
        if (int rc = s.eval_points(n, x, y, zz, t, fl, G.order == 2, e1, e2, nullptr, u, v, w, p, true)) return rc;
        PointsFinish F;
        F.n = n; F.ramp = ramp_dev(G.ramp); F.t = t; F.x = x; F.y = y;
        F.eadd = 1; F.eoff = G.swl; F.ea = e1; F.eb = e2; F.eta = eta;
        double* outs[4] = {u, v, w, p};
        for (int c = 0; c < 4; ++c) { F.src[c] = outs[c]; F.dst[c] = outs[c]; }
        CFDWM_HIP(launch_points_finish(F, st));
        CFDWM_HIP(hipStreamSynchronize(st));
        return 0;
    }
    if (G.wavetype == 1) {
        Sea& s = *G.csea->s;
        std::vector<double> zc, e1, e2;
        const double* zz = z;
        // z >= -depth (CFDwavemaker.cpp:1547) and the ramps are applied on the host
        std::vector<double> hx, hy, ht, hz;
        if (dev) {
            hx.resize(n); hy.resize(n); ht.resize(n);
            CFDWM_HIP(hipMemcpy(hx.data(), x, n * 8, hipMemcpyDeviceToHost));
            CFDWM_HIP(hipMemcpy(hy.data(), y, n * 8, hipMemcpyDeviceToHost));
            CFDWM_HIP(hipMemcpy(ht.data(), t, n * 8, hipMemcpyDeviceToHost));
            if (z) { hz.resize(n); CFDWM_HIP(hipMemcpy(hz.data(), z, n * 8, hipMemcpyDeviceToHost)); }
            x = hx.data(); y = hy.data(); t = ht.data(); z = z ? hz.data() : nullptr;
        }
        if (z) {
            zc.resize(n);
            for (int64_t q = 0; q < n; ++q) zc[q] = std::max(-G.depth, z[q]);
            zz = zc.data();
        }
        std::vector<double> hu, hv, hw, hp, he;
        auto want = [&](double* o, std::vector<double>& b) -> double* { if (!o) return nullptr; b.resize(n); return b.data(); };
        double *ou = want(u, hu), *ov = want(v, hv), *ow = want(w, hw), *op = want(p, hp);
        if (eta) { e1.resize(n); e2.resize(n); }
        const uint32_t fl = order_flags(G.order) | (p ? EV_P : 0u);
        int rc = s.eval_points(n, x, y, zz, t, fl, G.order == 2, eta ? e1.data() : nullptr,
                               eta && G.order == 2 ? e2.data() : nullptr, nullptr, ou, ov, ow, op, false);
        if (rc) return rc;
        if (eta) { he.resize(n); for (int64_t q = 0; q < n; ++q) he[q] = G.swl + e1[q] + (G.order == 2 ? e2[q] : 0.0); }
        auto fin = [&](double* dst, std::vector<double>& src) -> int {
            if (!dst) return 0;
            for (int64_t q = 0; q < n; ++q) src[q] *= G.ramp(t[q], x[q], y[q]);
            if (dev) CFDWM_HIP(hipMemcpy(dst, src.data(), n * 8, hipMemcpyHostToDevice));
            else std::memcpy(dst, src.data(), n * 8);
            return 0;
        };
        if (fin(eta, he) || fin(u, hu) || fin(v, hv) || fin(w, hw) || fin(p, hp)) return -3;
        return 0;
    }
    if (G.wavetype == 5 && !dev) {
        // host arrays: one copy up, the device path below, one copy down
        // (the device path's results equal the host computation bit for bit)
        std::lock_guard<std::mutex> lk(g_stage.mu);
        const size_t nb = (size_t)n * sizeof(double);
        if (g_stage.in.alloc(4 * nb) || g_stage.out.alloc(5 * nb)) return -3;
        double* din = (double*)g_stage.in.get();
        double* dout = (double*)g_stage.out.get();
        double *dx = din, *dy = din + n, *dt = din + 2 * n, *dz = z ? din + 3 * n : nullptr;
        CFDWM_HIP(hipMemcpy(dx, x, nb, hipMemcpyHostToDevice));
        CFDWM_HIP(hipMemcpy(dy, y, nb, hipMemcpyHostToDevice));
        CFDWM_HIP(hipMemcpy(dt, t, nb, hipMemcpyHostToDevice));
        if (z) CFDWM_HIP(hipMemcpy(dz, z, nb, hipMemcpyHostToDevice));
        double* host_out[5] = {eta, u, v, w, p};
        double* dev_out[5];
        for (int c = 0; c < 5; ++c) dev_out[c] = host_out[c] ? dout + c * n : nullptr;
        if (int rc = eval_points_impl(n, dx, dy, dz, dt, dev_out[0], dev_out[1], dev_out[2], dev_out[3], dev_out[4],
                                      flags, true))
            return rc;
        for (int c = 0; c < 5; ++c)
            if (host_out[c]) CFDWM_HIP(hipMemcpy(host_out[c], dev_out[c], nb, hipMemcpyDeviceToHost));
        return 0;
    }
    if (G.wavetype == 5 && dev) {
        // device-resident s-grid queries: velocities at the clamped z, the
        // elevation from a query at z = 0, then the ramps
        Grid& g = *G.cgrid->g;
        hipStream_t st = Engine::get().stream();
        std::lock_guard<std::mutex> lk(g_dev.mu);
        double *q4 = nullptr, *e4 = nullptr;
        if (u || v || w) {
            if (g_dev.z.alloc(n * sizeof(double)) || g_dev.q4.alloc(4 * n * sizeof(double))) return -3;
            double* zc = (double*)g_dev.z.get();
            q4 = (double*)g_dev.q4.get();
            CFDWM_HIP(launch_clamp_below(n, z, -G.depth, zc, st));
            if (int rc = g.query_device_f64(n, t, x, y, zc, q4, q4 + n, q4 + 2 * n, q4 + 3 * n)) return rc;
        }
        if (eta) {
            if (g_dev.z0.alloc(n * sizeof(double)) || g_dev.e4.alloc(4 * n * sizeof(double))) return -3;
            double* z0 = (double*)g_dev.z0.get();
            e4 = (double*)g_dev.e4.get();
            CFDWM_HIP(hipMemsetAsync(z0, 0, n * sizeof(double), st));
            if (int rc = g.query_device_f64(n, t, x, y, z0, e4, e4 + n, e4 + 2 * n, e4 + 3 * n)) return rc;
        }
        if (p) { set_error("dynamic pressure is defined for wave type 1 only (CFDwavemaker.cpp:1800-1813)"); }
        PointsFinish F;
        F.n = n; F.ramp = ramp_dev(G.ramp); F.t = t; F.x = x; F.y = y;
        F.eadd = 0; F.ea = e4; F.eta = eta;
        double* outs[4] = {u, v, w, p};
        for (int c = 0; c < 4; ++c) {
            F.dst[c] = outs[c];
            F.src[c] = (c < 3 && outs[c]) ? q4 + (c + 1) * n : nullptr;    // p: 0
        }
        CFDWM_HIP(launch_points_finish(F, st));
        CFDWM_HIP(hipStreamSynchronize(st));
        return 0;
    }
    set_error("wave type not supported by this build");
    return -1;
}

int wave_PrefetchPoints(int64_t n, const double* x, const double* y, const double* z, const double* t,
                        uint32_t flags) {
    if (!G.initialized) { set_error("wave_Initialize has not been called"); return -1; }
    if (n < 0 || (n && (!x || !y || !t))) { set_error("bad arguments"); return -1; }
    const bool we = (flags & CFDWM_OUT_ETA) != 0;
    const bool wv = (flags & (CFDWM_OUT_U | CFDWM_OUT_V | CFDWM_OUT_W | CFDWM_OUT_P)) != 0;
    const bool wp = (flags & CFDWM_OUT_P) != 0 && G.wavetype == 1;
    if (wv && !z) { set_error("z required for velocities"); return -1; }
    std::vector<double> eta(we ? n : 0), u(wv ? n : 0), v(wv ? n : 0), w(wv ? n : 0), p(wp ? n : 0);
    if (n) {
        if (int rc = eval_points_impl(n, x, y, z, t, we ? eta.data() : nullptr, wv ? u.data() : nullptr,
                                      wv ? v.data() : nullptr, wv ? w.data() : nullptr, wp ? p.data() : nullptr, flags,
                                      false))
            return rc;
    }
    std::unique_lock<std::shared_mutex> lk(PC.mu);
    PC.vel.clear();
    PC.eta.clear();
    PC.has_vel = wv && n > 0;
    PC.has_p = wp && n > 0;
    PC.has_eta = we && n > 0;
    if (PC.has_vel) PC.vel.reserve((size_t)n);
    if (PC.has_eta) PC.eta.reserve((size_t)n);
    for (int64_t q = 0; q < n; ++q) {
        if (wv) PC.vel[PKey{bits(x[q]), bits(y[q]), bits(z[q]), bits(t[q])}] = PVals{u[q], v[q], w[q], wp ? p[q] : 0.0};
        if (we) PC.eta[PKey{bits(x[q]), bits(y[q]), 0, bits(t[q])}] = eta[q];
    }
    return 0;
}
int wave_PrefetchClear(void) {
    std::unique_lock<std::shared_mutex> lk(PC.mu);
    PC.vel.clear();
    PC.eta.clear();
    PC.has_vel = PC.has_p = PC.has_eta = false;
    return 0;
}
int wave_PrefetchAuto(int enable) {
    // the remembered points stay: kept values are only served for the time
    // value and grid state they were computed at
    const bool was = AC.on.exchange(enable != 0);
    return was ? 1 : 0;
}
int wave_PrefetchAutoStats(int64_t* hits, int64_t* misses, int64_t* replays, int64_t* replayed_points) {
    if (hits) *hits = AC.hits.load();
    if (misses) *misses = AC.misses.load();
    if (replays) *replays = AC.replays.load();
    if (replayed_points) *replayed_points = AC.replayed.load();
    return 0;
}
int wave_PrefetchStats(int64_t* hits, int64_t* misses) {
    if (hits) *hits = PC.hits.load();
    if (misses) *misses = PC.misses.load();
    return 0;
}
int wave_EvalPoints(int64_t n, const double* x, const double* y, const double* z, const double* t,
                    double* eta, double* u, double* v, double* w, double* p, uint32_t flags) {
    return eval_points_impl(n, x, y, z, t, eta, u, v, w, p, flags, false);
}
int wave_EvalPointsDevice(int64_t n, const double* x, const double* y, const double* z, const double* t,
                          double* eta, double* u, double* v, double* w, double* p, uint32_t flags) {
    return eval_points_impl(n, x, y, z, t, eta, u, v, w, p, flags, true);
}

int cfdwm_sea_create(cfdwm_sea** out, int n, const double* omega, const double* A, const double* k,
                     const double* phase, const double* theta, const double* ramptime,
                     const double* starttime, const double* duration, const cfdwm_sea_params* P) {
    if (!out || n <= 0 || !omega || !A || !k || !phase || !theta || !P) { set_error("bad arguments"); return -1; }
    SeaInput in;
    in.omega.assign(omega, omega + n); in.A.assign(A, A + n); in.k.assign(k, k + n);
    in.phase.assign(phase, phase + n); in.theta.assign(theta, theta + n);
    if (ramptime) in.ramptime.assign(ramptime, ramptime + n);
    if (starttime) in.starttime.assign(starttime, starttime + n);
    if (duration) in.duration.assign(duration, duration + n);
    in.depth = P->depth; in.g = P->g; in.rho = P->rho; in.swl = P->swl; in.mtheta = P->mtheta;
    in.tofmax = P->tofmax; in.x0 = P->x0; in.y0 = P->y0; in.bandwidth = P->bandwidth; in.cutoff = P->cutoff;
    in.order = P->order;
    if (P->bw_auto) {
        double m[3];
        spectral_moments(in, m);
        in.bandwidth = 0.7 * (m[1] / m[0]);
    }
    const double ampl = P->ampl == 0.0 ? 1.0 : P->ampl;
    const double sumA = std::accumulate(in.A.begin(), in.A.end(), 0.);
    for (int i = 0; i < n; ++i) {
        if (P->normalize) { in.A[i] = ampl * in.A[i] / sumA; in.phase[i] = 0.; }
        else in.A[i] = ampl * in.A[i];
    }
    auto cs = std::make_unique<cfdwm_sea>();
    cs->s = std::make_unique<Sea>();
    if (P->precision < 0 || P->precision > 2) { set_error("precision must be 0 (FP64), 1 (FP32) or 2 (BF16 forms)"); return -1; }
    cs->s->prec = P->precision >= 1 ? kFP32 : kFP64;
    cs->s->qf_products = P->precision == 2 ? 1 : 3;
    cs->with_direct = P->with_direct != 0;
    std::string err = build_tables(in, cs->s->tab, cs->with_direct);
    if (!err.empty()) { set_error(err); return -1; }
    if (int rc = cs->s->upload(cs->with_direct)) return rc;
    *out = cs.release();
    return 0;
}

int cfdwm_sea_clone(cfdwm_sea** out, const cfdwm_sea* src, int precision) {
    if (!out || !src || precision < 0 || precision > 2) { set_error("bad arguments"); return -1; }
    auto cs = std::make_unique<cfdwm_sea>();
    cs->s = std::make_unique<Sea>();
    cs->s->tab = src->s->tab;                 // host tables: no second O(N^2) coefficient pass
    cs->s->prec = precision >= 1 ? kFP32 : kFP64;
    cs->s->qf_products = precision == 2 ? 1 : 3;
    cs->s->vertical_theory = src->s->vertical_theory;
    cs->with_direct = src->with_direct;
    if (int rc = cs->s->upload(cs->with_direct)) return rc;
    *out = cs.release();
    return 0;
}

int cfdwm_sea_destroy(cfdwm_sea* s) { delete s; return 0; }

int cfdwm_sea_set_vertical_theory(cfdwm_sea* s, int vt) {
    if (!s || (vt != 0 && vt != 1)) { set_error("bad arguments"); return -1; }
    s->s->vertical_theory = vt;
    return 0;
}

int cfdwm_sea_terms(cfdwm_sea* s, int64_t* pair_terms, int64_t* diag_terms, int64_t* direct_terms) {
    if (!s) return -1;
    if (pair_terms) *pair_terms = s->s->tab.filt.nterms;
    if (diag_terms) *diag_terms = s->s->tab.diag_filtered;
    if (direct_terms) *direct_terms = s->s->tab.direct.nterms;
    return 0;
}

int cfdwm_eval(cfdwm_sea* cs, int kind, int64_t n, const double* t, const double* x, const double* y,
               const double* z, double* out) {
    if (!cs || n < 0 || !t || !x || !y || !out) { set_error("bad arguments"); return -1; }
    Sea& s = *cs->s;
    const int order = s.tab.in.order;
    std::vector<double> a(n), b(n), gr(4 * n);
    auto needz = [&]() { return z != nullptr; };
    switch (kind) {
    case 0: case 1: case 2: {
        const bool e2 = kind != 0 && (kind == 1 || order == 2);
        if (int rc = s.eval_points(n, x, y, z, t, 0, e2, a.data(), e2 ? b.data() : nullptr, nullptr,
                                   nullptr, nullptr, nullptr, nullptr, false)) return rc;
        for (int64_t q = 0; q < n; ++q)
            out[q] = kind == 0 ? a[q] : (kind == 1 ? b[q] : s.tab.in.swl + a[q] + (e2 ? b[q] : 0.0));
        return 0;
    }
    case 7: case 8: case 9: case 10: {
        if (int rc = s.eval_points(n, x, y, z, t, 0, false, nullptr, nullptr, gr.data(), nullptr, nullptr,
                                   nullptr, nullptr, false)) return rc;
        std::memcpy(out, gr.data() + (kind - 7) * n, n * 8);
        return 0;
    }
    default: break;
    }
    if (!needz()) { set_error("z required"); return -1; }
    uint32_t fl = 0;
    int comp = 0;   // 0 u 1 v 2 w 3 p
    switch (kind) {
    case 3: case 4: case 5: fl = EV_LIN; comp = kind - 3; break;
    case 6: fl = EV_LIN | EV_P; comp = 3; break;
    case 11: case 12: case 13: fl = EV_SEC | EV_DIRECT; comp = kind - 11; break;
    case 14: fl = EV_SEC | EV_DIRECT | EV_P; comp = 3; break;
    case 15: case 16: case 17: fl = order_flags(order); comp = kind - 15; break;
    case 18: fl = order_flags(order) | EV_P; comp = 3; break;
    case 19: case 20: case 21: fl = EV_SEC; comp = kind - 19; break;
    case 23: case 24: case 25: case 26: {                  // u_smit, v_smit, w_smit, dp_smit
        double* so[4] = {nullptr, nullptr, nullptr, nullptr};
        so[kind - 23] = out;
        return s.eval_points(n, x, y, z, t, 0, false, nullptr, nullptr, nullptr, so[0], so[1], so[2], so[3],
                             false, true);
    }
    default:
        set_error("evaluation kind not supported (uvw2.p is read by no API path of the reference)");
        return -1;
    }
    if ((fl & EV_DIRECT) && !cs->with_direct && s.vertical_theory != 1) {
        set_error("sea created without the unfiltered table");
        return -1;
    }
    double* o[4] = {nullptr, nullptr, nullptr, nullptr};
    o[comp] = out;
    return s.eval_points(n, x, y, z, t, fl, false, nullptr, nullptr, nullptr, o[0], o[1], o[2], o[3], false);
}

int cfdwm_eval_fields(cfdwm_sea* cs, int64_t n, const double* t, const double* x, const double* y, const double* z,
                      double* eta1, double* eta2, double* u, double* v, double* w, double* p, int device) {
    if (!cs || n < 0 || (n && (!t || !x || !y))) { set_error("bad arguments"); return -1; }
    Sea& s = *cs->s;
    const int order = s.tab.in.order;
    const bool lev = u || v || w || p;
    if (lev && !z) { set_error("z required"); return -1; }
    if (lev && order >= 1 && !cs->with_direct && s.vertical_theory != 1) {
        set_error("sea created without the unfiltered table");
        return -1;
    }
    const uint32_t fl = order_flags(order) | (p ? EV_P : 0u);
    return s.eval_points(n, x, y, z, t, fl, order == 2 && eta2 != nullptr, eta1, order == 2 ? eta2 : nullptr, nullptr,
                         u, v, w, p, device != 0);
}
int cfdwm_grid_create(cfdwm_grid** out, cfdwm_sea* sea, const cfdwm_grid_desc* d) {
    if (!out || !sea || !d) { set_error("bad arguments"); return -1; }
    GridDesc gd{};
    std::memcpy(gd.domain, d->domain, sizeof(gd.domain));
    gd.nx = d->nx; gd.ny = d->ny; gd.nl = d->nl; gd.t0 = d->t0; gd.dt = d->dt;
    gd.tan_a = d->tan_a; gd.tan_b = d->tan_b; gd.water_depth = d->water_depth; gd.swl = d->swl;
    std::memcpy(gd.domain_ignore, d->domain_ignore, sizeof(gd.domain_ignore));
    gd.ignore_at_init = d->ignore_at_init;
    auto cg = std::make_unique<cfdwm_grid>();
    cg->g = std::make_unique<Grid>(sea->s.get(), gd);
    if (int rc = cg->g->init(d->defer_fill == 0)) return rc;
    *out = cg.release();
    return 0;
}
int cfdwm_grid_destroy(cfdwm_grid* g) { delete g; return 0; }
int cfdwm_grid_refill(cfdwm_grid* g) { return g ? g->g->refill() : -1; }
int cfdwm_grid_force_update(cfdwm_grid* g, double t, int* n) { return g ? g->g->force_update(t, n) : -1; }
int cfdwm_grid_state(cfdwm_grid* g, int* tstep, double* t0, double* dx, double* dy, double* ds) {
    if (!g) return -1;
    if (tstep) *tstep = g->g->tstep;
    if (t0) *t0 = g->g->t0;
    if (dx) *dx = g->g->g.dx;
    if (dy) *dy = g->g->g.dy;
    if (ds) *ds = g->g->g.ds;
    return 0;
}
int64_t cfdwm_grid_export(cfdwm_grid* g, int which, double* out) { return g ? g->g->export_field(which, out) : -1; }
int cfdwm_grid_query(cfdwm_grid* g, int64_t n, const double* t, const double* x, const double* y,
                     const double* z, double* out4) {
    if (!g || n < 0 || !t || !x || !y || !z || !out4) { set_error("bad arguments"); return -1; }
    return g->g->query(n, t, x, y, z, out4, nullptr, nullptr);
}
int cfdwm_grid_query_device_f32(cfdwm_grid* g, int64_t n, double t, const float* x, const float* y,
                                const float* z, float* eta, float* u, float* v, float* w) {
    if (!g || n < 0 || !x || !y || !z || !eta || !u || !v || !w) { set_error("bad arguments"); return -1; }
    return g->g->query_device_f32(n, t, x, y, z, eta, u, v, w);
}
int cfdwm_grid_query_device_f32_noalloc(cfdwm_grid* g, int64_t n, double t, const float* x, const float* y,
                                        const float* z, float* eta, float* u, float* v, float* w) {
    if (!g || n < 0 || !x || !y || !z || !eta || !u || !v || !w) { set_error("bad arguments"); return -1; }
    return query_device_f32_noalloc(*g->g, n, t, x, y, z, eta, u, v, w);
}
int cfdwm_grid_kinematics_device_f32(cfdwm_grid* g, int64_t n, double t, const float* x, const float* y,
                                     const float* z, float* out4) {
    if (!g || n < 0 || !x || !y || !z || !out4) { set_error("bad arguments"); return -1; }
    if ((reinterpret_cast<uintptr_t>(out4) & 15) != 0) { set_error("out4 must be 16-byte aligned"); return -1; }
    return g->g->query_device_f32(n, t, x, y, z, nullptr, nullptr, nullptr, nullptr, out4);
}
int cfdwm_grid_qform_info(cfdwm_grid* grid, int* nform, int* rank, double* mfma_flop, double* basis_err) {
    if (!grid) return -1;
    const Grid& g = *grid->g;
    if (nform) *nform = g.qf_nform;
    if (rank) *rank = g.qf_rank;
    if (mfma_flop) *mfma_flop = g.qf_flop;
    if (basis_err) *basis_err = g.qf_basis_err;
    return g.qf_used ? 1 : 0;
}
int cfdwm_grid_qform_mode(cfdwm_grid* grid, int* products, double* basis_tol, int* reason) {
    if (!grid) { set_error("bad arguments"); return -1; }
    const Grid& g = *grid->g;
    if (reason) *reason = g.qf_skip;
    const Sea::QfTables* T = nullptr;
    if (g.qf_used && g.sea->qform_tables(&T) == 1) {
        if (products) *products = T->nlv;
        if (basis_tol) *basis_tol = T->basis_tol;
        return 1;
    }
    if (products) *products = 0;
    if (basis_tol) *basis_tol = 0;
    return 0;
}
int cfdwm_grid_qform_clock(cfdwm_grid* grid, int enable, double* mhz) {
    if (!grid) { set_error("bad arguments"); return -1; }
    const int r = grid->g->qform_clock(mhz);
    grid->g->qform_clock_enable(enable != 0);
    return r;
}
int cfdwm_grid_write_vtu(cfdwm_grid* grid, const char* path, int last, const char* time_label) {
    if (!grid || !path) { set_error("bad arguments"); return -1; }
    return grid->g->write_vtu(path, last != 0, time_label);
}
int cfdwm_lsgrid_create(cfdwm_lsgrid** out, cfdwm_sea* sea, const cfdwm_grid_desc* d, int init_only) {
    if (!out || !sea || !d) { set_error("bad arguments"); return -1; }
    GridDesc gd{};
    std::memcpy(gd.domain, d->domain, sizeof(gd.domain));
    gd.nx = d->nx; gd.ny = d->ny; gd.nl = d->nl; gd.t0 = d->t0; gd.dt = d->dt;
    gd.tan_a = d->tan_a; gd.tan_b = d->tan_b; gd.water_depth = d->water_depth; gd.swl = d->swl;
    std::memcpy(gd.domain_ignore, d->domain_ignore, sizeof(gd.domain_ignore));
    gd.ignore_at_init = d->ignore_at_init;
    auto cg = std::make_unique<cfdwm_lsgrid>();
    cg->g = std::make_unique<LsGrid>(sea->s.get(), gd, init_only != 0);
    if (int rc = cg->g->init()) return rc;
    *out = cg.release();
    return 0;
}
int cfdwm_lsgrid_destroy(cfdwm_lsgrid* g) {
    if (!g) return 0;
    int rc = 0;
    if (g->g->vtk.on) rc = g->g->write_vtk_dump(true);     // ~lsGrid (lsgrid.h:25-30)
    delete g;
    return rc;
}
int cfdwm_lsgrid_check_time(cfdwm_lsgrid* g, double t) {
    if (!g) { set_error("bad arguments"); return -1; }
    return g->g->check_time(t) ? 1 : 0;
}
int cfdwm_lsgrid_update(cfdwm_lsgrid* g, double t_target, int* reinit) {
    if (!g) { set_error("bad arguments"); return -1; }
    return g->g->update(t_target, reinit);
}
int cfdwm_lsgrid_state(cfdwm_lsgrid* g, double* t0, int* update_count) {
    if (!g) { set_error("bad arguments"); return -1; }
    if (t0) *t0 = g->g->t0;
    if (update_count) *update_count = g->g->update_count;
    return 0;
}
int64_t cfdwm_lsgrid_export(cfdwm_lsgrid* g, int which, double* out) {
    if (!g) { set_error("bad arguments"); return -1; }
    return g->g->export_field(which, out);
}
int cfdwm_lsgrid_query(cfdwm_lsgrid* g, int64_t n, const double* t, const double* x, const double* y,
                       const double* z, double* out4) {
    if (!g || n < 0 || (n > 0 && (!t || !x || !y || !z || !out4))) { set_error("bad arguments"); return -1; }
    return g->g->query(n, t, x, y, z, out4, false);
}
int cfdwm_lsgrid_query_device(cfdwm_lsgrid* g, int64_t n, const double* t, const double* x, const double* y,
                              const double* z, double* out4) {
    if (!g || n < 0 || (n > 0 && (!t || !x || !y || !z || !out4))) { set_error("bad arguments"); return -1; }
    int rc = g->g->query(n, t, x, y, z, out4, true);
    if (rc) return rc;
    CFDWM_HIP(hipStreamSynchronize(Engine::get().stream()));
    return 0;
}
int cfdwm_lsgrid_write_vtu(cfdwm_lsgrid* g, const char* path, int last, const char* time_label) {
    if (!g || !path) { set_error("bad arguments"); return -1; }
    return g->g->write_vtu(path, last != 0, time_label);
}
int cfdwm_lsgrid_set_vtk(cfdwm_lsgrid* g, int on, const char* dir, const char* prefix, const char* time_label) {
    if (!g) { set_error("bad arguments"); return -1; }
    const bool was = g->g->vtk.on;
    g->g->vtk.on = on != 0;
    if (dir) g->g->vtk.dir = dir;
    if (prefix) g->g->vtk.prefix = prefix;
    if (time_label) g->g->vtk.label = time_label;
    if (on && !was) {
        // the dump the fill writes when dump_vtk is set before it (lsgrid.cpp:507-510)
        if (int rc = g->g->write_vtk_dump(false)) return rc;
        ++g->g->update_count;
    }
    return 0;
}
int cfdwm_grid_release_workspace(cfdwm_grid* grid) {
    if (!grid) return -1;
    grid->g->sea->release_qform();
    grid->g->release_workspace();
    return 0;
}
int cfdwm_grids_refill_batched(cfdwm_grid** grids, int n, double budget_gb, int release, cfdwm_batch_stats* out) {
    if (n < 0 || (n > 0 && !grids)) { set_error("bad arguments"); return -1; }
    std::vector<Grid*> gs(n);
    for (int k = 0; k < n; ++k) {
        if (!grids[k]) { set_error("bad arguments"); return -1; }
        gs[k] = grids[k]->g.get();
    }
    BatchStats S;
    const int64_t budget = (int64_t)(std::max(1.0, budget_gb) * 1073741824.0);
    if (int rc = refill_batched(gs.data(), n, budget, release != 0, &S)) return rc;
    if (out) {
        out->groups = S.groups; out->batched = S.batched; out->single = S.single; out->tiles = S.tiles;
        out->blocks = S.blocks; out->order = S.order; out->tuned = S.tuned;
        out->forms_ms = S.forms_ms; out->mfma_flop = S.mfma_flop;
        out->group_forms = S.group_forms; out->group_ptiles = S.group_ptiles;
        out->tables_ms = S.tables_ms;
        out->setup_ms = S.setup_ms;
    }
    return 0;
}
int cfdwm_depth_basis(double kmax, double tmax, double tol, int rmax, double* mu, double* tau, double* err) {
    if (rmax < 1 || rmax > 64 || !(kmax > 0) || !(tmax > 0)) return 0;
    DepthBasis B;
    const int rc = build_depth_basis(kmax, tmax, tol, rmax, B);
    for (int q = 0; q < B.R; ++q) {
        if (mu) mu[q] = B.mu[q];
        if (tau) tau[q] = B.tau[q];
    }
    if (err) *err = B.err;
    return rc == 0 ? B.R : -B.R;
}
int cfdwm_profile_basis(double kmax, double h, int sign, double tol, int rmax, double* mu, double* zs,
                        double* err) {
    if (rmax < 1 || rmax > 64 || !(kmax > 0) || !(h > 0) || (sign != 1 && sign != -1)) return 0;
    ProfileBasis B;
    const int rc = build_profile_basis(kmax, h, sign, tol, rmax, B);
    for (int q = 0; q < B.R; ++q) {
        if (mu) mu[q] = B.mu[q];
        if (zs) zs[q] = B.zs[q];
    }
    if (err) *err = B.err;
    return rc == 0 ? B.R : -B.R;
}
int cfdwm_surrogate_linear(cfdwm_sea* sea, int64_t n, const double* x, const double* y, double t, double h,
                           int nlev, const double* sigma, int fstride, double kp, float* X, float* Lin) {
    if (!sea || n < 0 || nlev < 1 || nlev > 32 || fstride < 1 || !sigma || (n && (!x || !y || !X || !Lin))) {
        set_error("bad arguments");
        return -1;
    }
    SurrJob J;
    J.n = n; J.x = x; J.y = y; J.t = t; J.h = h; J.kp = kp;
    J.nlev = nlev; J.fstride = fstride;
    J.nx = 3 + 3 * ((nlev + fstride - 1) / fstride) + 2;
    J.xstride = J.nx;
    for (int l = 0; l < nlev; ++l) J.sigma[l] = sigma[l];
    J.X = X; J.Lin = Lin;
    return surrogate_linear(*sea->s, J);
}
int cfdwm_surrogate_features(cfdwm_sea* sea, int64_t n, const double* x, const double* y, double t, double h,
                             int nlev, const double* sigma, int fstride, double kp, int xstride, float* X,
                             float* Lin) {
    const int nx = 3 + 3 * ((std::max(nlev, 1) + std::max(fstride, 1) - 1) / std::max(fstride, 1)) + 2;
    if (!sea || n < 0 || nlev < 1 || nlev > 32 || fstride < 1 || !sigma || xstride < nx + 3 ||
        (n && (!x || !y || !X || !Lin))) {
        set_error("bad arguments");
        return -1;
    }
    SurrJob J;
    J.n = n; J.x = x; J.y = y; J.t = t; J.h = h; J.kp = kp;
    J.nlev = nlev; J.fstride = fstride;
    J.nx = nx;
    J.xstride = xstride;
    J.quad = 1;
    for (int l = 0; l < nlev; ++l) J.sigma[l] = sigma[l];
    J.X = X; J.Lin = Lin;
    return surrogate_linear(*sea->s, J);
}
int cfdwm_surrogate_add(int64_t n, const float* a, const float* b, float* out) {
    if (n < 0 || (n && (!a || !b || !out))) { set_error("bad arguments"); return -1; }
    if (int e = Engine::get().init()) return e;
    return surrogate_add(n, a, b, out);
}
int cfdwm_surrogate_set_column(int64_t n, const double* src, float* X, int stride, int col) {
    if (n < 0 || stride < 1 || col < 0 || col >= stride || (n && (!src || !X))) { set_error("bad arguments"); return -1; }
    if (int e = Engine::get().init()) return e;
    return surrogate_set_column(n, src, X, stride, col);
}
int cfdwm_sync(void) {
    if (Engine::get().init()) return -2;
    CFDWM_HIP(hipStreamSynchronize(Engine::get().stream()));
    return 0;
}
int cfdwm_last_error(char* buf, int len) {
    const std::string e = last_error();
    if (buf && len > 0) {
        std::strncpy(buf, e.c_str(), (size_t)len - 1);
        buf[len - 1] = 0;
    }
    return (int)e.size();
}
int cfdwm_prof_enable(int on) {
    if (int e = Engine::get().init()) return e;
    Engine::get().prof_enable(on != 0);
    return 0;
}
int cfdwm_prof_read(char* names, double* total_ms, int64_t* count, int cap, int reset) {
    auto m = Engine::get().prof_read(reset != 0);
    int q = 0;
    for (auto& kv : m) {
        if (q >= cap) break;
        if (names) { std::strncpy(names + 64 * q, kv.first.c_str(), 63); names[64 * q + 63] = 0; }
        if (total_ms) total_ms[q] = kv.second.total_ms;
        if (count) count[q] = kv.second.count;
        ++q;
    }
    return q;
}

}  // extern "C"
