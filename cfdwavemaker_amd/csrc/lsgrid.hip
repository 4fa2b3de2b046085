// The legacy linear s-grid (lsGrid, src/lsgrid.h:15-114, src/lsgrid.cpp) on
// the device: two time slots filled by the spline grid's column body
// (Grid::fill_at), one-step updates that move slot 1 into slot 0, and
// tri/bilinear queries.
//
// Query: one thread per point, reading the 4 surface values of its cell in
// both slots (8 loads, shared by the elevation and the sigma maps) and the
// 8 corner values of each velocity component in both slots (48 loads).
// Indices are clamped into the grid for every input (a NaN coordinate maps to
// cell 0 as std::max(0, NaN) does in the reference's clip), so no input can
// read outside the fields.
#include <hip/hip_runtime.h>

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <iostream>
#include <sys/stat.h>

#include <hipcub/device/device_radix_sort.hpp>

#include "engine.hpp"

namespace cfdwm {

namespace {


struct LsQueryJob {
    int64_t n = 0;
    int nx = 1, ny = 1, nl = 1;
    double x0 = 0, y0 = 0, dx = 1, dy = 1, ds = 1;
    double t0 = 0, dt = 1, swl = 0, depth = 1;
    double tan_a = 1, inv_b = 1, tanab = 1;    // tan2s: -atan((-t tanab)^(1/b)) / a, tanab = tan(a)^b
    const void *eta = nullptr, *u = nullptr, *v = nullptr, *w = nullptr;
    const double *t = nullptr, *x = nullptr, *y = nullptr, *z = nullptr;
    double* out = nullptr;                     // [n][4]
    // when set: thread q evaluates the query sidx[q] (cell order): its inputs
    // are gathered, its outputs written at its own index
    const uint32_t* sidx = nullptr;
};

// std::min / std::max / clip (lsgrid.cpp:187-190) with the reference's
// comparison order, so NaN handling matches
__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ int iclip(int n, int lo, int hi) {
    const int m = (hi < n) ? hi : n;
    return (lo < m) ? m : lo;
}

// std::modf(double, float*) resolves to the float overload: the argument is
// rounded to float, the fraction has float precision.  Returns the fraction,
// the integer part goes to *ip (clamped to [0, hi] for the index).
__device__ __forceinline__ double fmodf_split(double v, int hi, int* ip) {
    const float f = (float)v;
    const float tr = truncf(f);
    int k = (tr >= 0.f && tr <= (float)hi) ? (int)tr : (tr > (float)hi ? hi : 0);
    *ip = k;
    return (double)(f - tr);
}

template <typename R>
__global__ void __launch_bounds__(256) k_lsquery(LsQueryJob J) {
    // the reference's expression order, without fused multiply-adds
#pragma clang fp contract(off)
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= J.n) return;
    const int64_t p = J.sidx ? (int64_t)J.sidx[q] : q;
    const double tpt = J.t[p], xpt = J.x[p], ypt = J.y[p], zpt = J.z[p];
    const int nx = J.nx, ny = J.ny, nl = J.nl;
    int i, j;
    const double xd = fmodf_split(smax(0., smin((xpt - J.x0) / J.dx, nx - 1.)), nx - 1, &i);
    const double yd = fmodf_split(smax(0., smin((ypt - J.y0) / J.dy, ny - 1.)), ny - 1, &j);
    const int i1 = iclip(i + 1, 0, nx - 1), j1 = iclip(j + 1, 0, ny - 1);
    const int64_t nc = (int64_t)nx * ny;
    const R* E = (const R*)J.eta;
    const double C00 = E[(int64_t)i * ny + j], C01 = E[(int64_t)i * ny + j1];
    const double C10 = E[(int64_t)i1 * ny + j], C11 = E[(int64_t)i1 * ny + j1];
    const double D00 = E[nc + (int64_t)i * ny + j], D01 = E[nc + (int64_t)i * ny + j1];
    const double D10 = E[nc + (int64_t)i1 * ny + j], D11 = E[nc + (int64_t)i1 * ny + j1];
    const double td = smin(1., smax(0., (tpt - J.t0) / J.dt));
    const double C0 = C00 * (1. - xd) + C10 * xd, C1 = C01 * (1. - xd) + C11 * xd;
    const double D0 = D00 * (1. - xd) + D10 * xd, D1 = D01 * (1. - xd) + D11 * xd;
    const double we0 = C0 * (1. - yd) + C1 * yd, we1 = D0 * (1. - yd) + D1 * yd;
    double* o = J.out + 4 * p;
    o[0] = we0 * (1. - td) + we1 * td;
    // z -> tan-stretched s -> level (z2s, tan2s: lsgrid.cpp:82-101)
    const double e0 = we0 - J.swl, e1 = we1 - J.swl, zs = zpt - J.swl;
    const double q0 = smax((smin(zs, e0) - e0) / (J.depth + e0), -1.);
    const double q1 = smax((smin(zs, e1) - e1) / (J.depth + e1), -1.);
    const double spt0 = -atan(pow(-q0 * J.tanab, J.inv_b)) / J.tan_a;
    const double spt1 = -atan(pow(-q1 * J.tanab, J.inv_b)) / J.tan_a;
    int l0, l1;
    const double sd0 = fmodf_split((spt0 + 1.) / J.ds, nl - 1, &l0);
    const double sd1 = fmodf_split((spt1 + 1.) / J.ds, nl - 1, &l1);
    const int l0p = iclip(l0 + 1, 0, nl - 1), l1p = iclip(l1 + 1, 0, nl - 1);
    const int64_t nv = nc * nl;
    const int64_t b00 = ((int64_t)i * ny + j) * nl, b01 = ((int64_t)i * ny + j1) * nl;
    const int64_t b10 = ((int64_t)i1 * ny + j) * nl, b11 = ((int64_t)i1 * ny + j1) * nl;
    const void* F[3] = {J.u, J.v, J.w};
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        const R* V0 = (const R*)F[f];
        const R* V1 = V0 + nv;
        const double c000 = V0[b00 + l0], c001 = V0[b00 + l0p], c010 = V0[b01 + l0], c011 = V0[b01 + l0p];
        const double c100 = V0[b10 + l0], c101 = V0[b10 + l0p], c110 = V0[b11 + l0], c111 = V0[b11 + l0p];
        const double d000 = V1[b00 + l1], d001 = V1[b00 + l1p], d010 = V1[b01 + l1], d011 = V1[b01 + l1p];
        const double d100 = V1[b10 + l1], d101 = V1[b10 + l1p], d110 = V1[b11 + l1], d111 = V1[b11 + l1p];
        // x, then y, then s, then t (lsgrid.cpp:266-286)
        const double c00 = c000 * (1. - xd) + c100 * xd, c01 = c001 * (1. - xd) + c101 * xd;
        const double c10 = c010 * (1. - xd) + c110 * xd, c11 = c011 * (1. - xd) + c111 * xd;
        const double d00 = d000 * (1. - xd) + d100 * xd, d01 = d001 * (1. - xd) + d101 * xd;
        const double d10 = d010 * (1. - xd) + d110 * xd, d11 = d011 * (1. - xd) + d111 * xd;
        const double cs0 = c00 * (1. - yd) + c10 * yd, cs1 = c01 * (1. - yd) + c11 * yd;
        const double ds0 = d00 * (1. - yd) + d10 * yd, ds1 = d01 * (1. - yd) + d11 * yd;
        o[1 + f] = (cs0 * (1. - sd0) + cs1 * sd0) * (1. - td) + (ds0 * (1. - sd1) + ds1 * sd1) * td;
    }
}

// key = column i ny + j of the query's cell (the query kernel's clamps; the
// order only affects locality, not results) and the query's index
__global__ void __launch_bounds__(256) k_ls_keys(int64_t n, int nx, int ny, double x0, double y0, double dx,
                                                 double dy, const double* __restrict__ x, const double* __restrict__ y,
                                                 uint32_t* __restrict__ key, uint32_t* __restrict__ idx) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int i, j;
    fmodf_split(smax(0., smin((x[p] - x0) / dx, nx - 1.)), nx - 1, &i);
    fmodf_split(smax(0., smin((y[p] - y0) / dy, ny - 1.)), ny - 1, &j);
    key[p] = (uint32_t)((int64_t)i * ny + j);
    idx[p] = (uint32_t)p;
}

// slot dst := slot src for ETA and UX/UY/UZ of the listed columns (all when cols is null)
template <typename R>
__global__ void __launch_bounds__(256) k_ls_copy(R* eta, R* u, R* v, R* w, const int32_t* cols, int64_t ncol,
                                                 int64_t nc, int nl, int src, int dst) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per = (int64_t)nl + 1;
    if (q >= ncol * per) return;
    const int64_t k = q / per;
    const int m = (int)(q - k * per);
    const int64_t c = cols ? (int64_t)cols[k] : k;
    if (m == nl) {
        eta[dst * nc + c] = eta[src * nc + c];
        return;
    }
    const int64_t nv = nc * nl, e = c * nl + m;
    u[dst * nv + e] = u[src * nv + e];
    v[dst * nv + e] = v[src * nv + e];
    w[dst * nv + e] = w[src * nv + e];
}

template <typename R>
__global__ void __launch_bounds__(256) k_ls_set(R* a, int64_t n, double value) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) a[q] = (R)value;
}

unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

// batches of this many queries or more run in cell order
constexpr int64_t kLsOrderMin = 65536;

// keys and indices of the batch J, then their radix sort by cell into keys + n, idx + n
hipError_t launch_ls_order(const LsQueryJob& J, uint32_t* keys, uint32_t* idx, void* tmp, size_t tmp_bytes, int bits,
                           hipStream_t st) {
    const int64_t n = J.n;
    hipLaunchKernelGGL(k_ls_keys, dim3(blocks_for(n)), dim3(256), 0, st, n, J.nx, J.ny, J.x0, J.y0, J.dx, J.dy, J.x,
                       J.y, keys, idx);
    if (hipError_t e = hipGetLastError()) return e;
    return hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, keys, keys + n, idx, idx + n, (int)n, 0, bits, st);
}

template <typename R>
hipError_t launch_lsquery(const LsQueryJob& J, hipStream_t st) {
    hipLaunchKernelGGL(k_lsquery<R>, dim3(blocks_for(J.n)), dim3(256), 0, st, J);
    return hipGetLastError();
}

}  // namespace

LsGrid::LsGrid(Sea* sea, const GridDesc& desc, bool io) : init_only(io), d(desc) {
    g = std::make_unique<Grid>(sea, desc);
    g->legacy = true;
}

// slot dst := slot src (ETA, UX/UY/UZ) for the columns outside the ignore
// region (active_only) or for every column
int LsGrid::copy_slot(int src, int dst, bool active_only) {
    hipStream_t st = Engine::get().stream();
    const int64_t nc = (int64_t)g->g.nx * g->g.ny;
    const int64_t ncol = active_only ? g->active_count() : nc;
    if (ncol == 0) return 0;
    const int32_t* cols = active_only ? g->active_cols() : nullptr;
    const int64_t nthr = ncol * (g->g.nl + 1);
    if (g->precision() == kFP32)
        hipLaunchKernelGGL(k_ls_copy<float>, dim3(blocks_for(nthr)), dim3(256), 0, st, (float*)g->field(0),
                           (float*)g->field(4), (float*)g->field(5), (float*)g->field(6), cols, ncol, nc, g->g.nl,
                           src, dst);
    else
        hipLaunchKernelGGL(k_ls_copy<double>, dim3(blocks_for(nthr)), dim3(256), 0, st, (double*)g->field(0),
                           (double*)g->field(4), (double*)g->field(5), (double*)g->field(6), cols, ncol, nc,
                           g->g.nl, src, dst);
    CFDWM_HIP(hipGetLastError());
    return 0;
}

// initialize_surface_elevation[_with_ignore](t0) + initialize_kinematics[_with_ignore]
// (lsgrid.cpp:419-747): slot 0 at t0, slot 1 at t0 + dt (init_only: slot 1 := slot 0)
int LsGrid::fill_both(bool use_ignore) {
    const double ts[2] = {t0, t0 + d.dt};
    const int dst[2] = {0, 1};
    int rc = g->fill_at(init_only ? 1 : 2, ts, dst, use_ignore);
    if (rc) return rc;
    if (init_only && (rc = copy_slot(0, 1, false))) return rc;
    CFDWM_HIP(hipStreamSynchronize(Engine::get().stream()));
    return 0;
}

int LsGrid::init() {
    if (d.dt <= 0) { set_error("lsGrid: dt must be positive"); return -1; }
    int rc = g->init(false);
    if (rc) return rc;
    t0 = d.t0;
    update_count = 0;
    if (d.ignore_at_init) {
        // _with_ignore: the ignored columns' surface is swl in both slots, their
        // velocities keep the zeros of allocate (lsgrid.cpp:23-56, 701-709)
        const int64_t n = 2LL * g->g.nx * g->g.ny;
        hipStream_t st = Engine::get().stream();
        if (g->precision() == kFP32)
            hipLaunchKernelGGL(k_ls_set<float>, dim3(blocks_for(n)), dim3(256), 0, st, (float*)g->field(0), n, d.swl);
        else
            hipLaunchKernelGGL(k_ls_set<double>, dim3(blocks_for(n)), dim3(256), 0, st, (double*)g->field(0), n,
                               d.swl);
        CFDWM_HIP(hipGetLastError());
    }
    if ((rc = fill_both(d.ignore_at_init != 0))) return rc;
    if (vtk.on) {                                // initialize_kinematics (lsgrid.cpp:507-510)
        if ((rc = write_vtk_dump(false))) return rc;
        ++update_count;
    }
    return 0;
}

int LsGrid::update(double t_target, int* reinit) {
    int rc;
    const double dt = d.dt;
    if ((t_target / dt - (t0 + 2 * dt) / dt) > 0.) {
        // initialize_surface_elevation + initialize_kinematics at the new time, every column
        t0 = dt * std::floor(t_target / dt);
        if ((rc = fill_both(false))) return rc;
        if (reinit) *reinit = 1;
    } else {
        t0 += dt;
        const double t1 = t0 + dt;
        const int one = 1;
        if (!init_only && (rc = copy_slot(1, 0, true))) return rc;
        if ((rc = g->fill_at(1, &t1, &one, true))) return rc;
        if (init_only && (rc = copy_slot(1, 0, true))) return rc;
        CFDWM_HIP(hipStreamSynchronize(Engine::get().stream()));
        if (reinit) *reinit = 0;
    }
    if (vtk.on) {                                // lsgrid.cpp:507-510, 815-818
        if ((rc = write_vtk_dump(false))) return rc;
        ++update_count;
    }
    return 0;
}

int LsGrid::query(int64_t n, const double* t, const double* x, const double* y, const double* z, double* out4,
                  bool device) {
    if (n < 0) { set_error("lsGrid query: negative count"); return -1; }
    if (n == 0) return 0;
    hipStream_t st = Engine::get().stream();
    LsQueryJob J;
    J.n = n;
    J.nx = g->g.nx; J.ny = g->g.ny; J.nl = g->g.nl;
    J.x0 = g->g.x0; J.y0 = g->g.y0; J.dx = g->g.dx; J.dy = g->g.dy; J.ds = g->g.ds;
    J.t0 = t0; J.dt = d.dt; J.swl = d.swl; J.depth = d.water_depth;
    J.tan_a = d.tan_a; J.inv_b = 1. / d.tan_b; J.tanab = std::pow(std::tan(d.tan_a), d.tan_b);
    J.eta = g->field(0); J.u = g->field(4); J.v = g->field(5); J.w = g->field(6);
    if (device) {
        J.t = t; J.x = x; J.y = y; J.z = z; J.out = out4;
    } else {
        if (qin_.alloc((size_t)n * 4 * sizeof(double)) || qout_.alloc((size_t)n * 4 * sizeof(double))) return -3;
        double* in = (double*)qin_.get();
        const double* src[4] = {t, x, y, z};
        for (int k = 0; k < 4; ++k)
            CFDWM_HIP(hipMemcpyAsync(in + k * n, src[k], n * sizeof(double), hipMemcpyHostToDevice, st));
        J.t = in; J.x = in + n; J.y = in + 2 * n; J.z = in + 3 * n; J.out = (double*)qout_.get();
    }
    if (n >= kLsOrderMin && n <= INT32_MAX) {
        // cell order: the lanes of a wavefront share a few cells, so their
        // grid loads coalesce; inputs are gathered and outputs written at
        // each query's own index
        int bits = 1;
        while (((int64_t)1 << bits) < (int64_t)g->g.nx * g->g.ny) ++bits;
        size_t tmp = 0;
        CFDWM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                                     (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)n, 0, bits));
        if (okey_.alloc((size_t)n * 2 * sizeof(uint32_t)) || oidx_.alloc((size_t)n * 2 * sizeof(uint32_t)) ||
            otmp_.alloc(std::max<size_t>(tmp, 16)))
            return -3;
        uint32_t* keys = (uint32_t*)okey_.get();
        uint32_t* idx = (uint32_t*)oidx_.get();
        PROF("k_lsquery_sort", launch_ls_order(J, keys, idx, otmp_.get(), tmp, bits, st));
        J.sidx = idx + n;
    }
    if (g->precision() == kFP32)
        PROF("k_lsquery", launch_lsquery<float>(J, st));
    else
        PROF("k_lsquery", launch_lsquery<double>(J, st));
    if (!device) {
        CFDWM_HIP(hipMemcpyAsync(out4, J.out, (size_t)n * 4 * sizeof(double), hipMemcpyDeviceToHost, st));
        CFDWM_HIP(hipStreamSynchronize(st));
    }
    return 0;
}

int LsGrid::write_vtu(const char* path, bool last, const char* time_label) {
    CFDWM_HIP(hipStreamSynchronize(Engine::get().stream()));
    return g->write_vtu_slot(path, last ? 1 : 0, last ? t0 + d.dt : t0, time_label, true);
}

// write_vtk (lsgrid.cpp:1329-1351): directory (created if missing) + prefix
// + update_count as %05d + ".vtu"
int LsGrid::write_vtk_dump(bool last) {
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
    std::snprintf(buf, sizeof buf, "%05d", update_count);
    const std::string path = vtk.dir + vtk.prefix + buf + ".vtu";
    return write_vtu(path.c_str(), last, vtk.label.c_str());
}

}  // namespace cfdwm
