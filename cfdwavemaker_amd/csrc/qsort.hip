// Cell order for large spline-query batches.
//
// The query kernel reads 16 node records and 4 column records around each
// query's cell.  Queries in the caller's order (a random point cloud) make
// every lane of a wavefront read different cells of the packed grid, so each
// record access is a separate last-level-cache line.  Sorted by cell column,
// the lanes of a wavefront share a few columns: their record loads coalesce
// and the grid is read about once per batch.  The sort carries each query's
// coordinates and index as the value, so the query kernel reads its inputs
// contiguously and only its outputs are scattered (to the caller's order).
#include <hip/hip_runtime.h>

#include <hipcub/device/device_radix_sort.hpp>

#include "engine.hpp"
#include "kernels.hpp"

namespace cfdwm {

namespace {

// key = column index i ny + j of the cell holding (x, y), with the edge clamps
// of the query kernel (the order only affects locality, not results)
template <typename TQ>
__global__ __launch_bounds__(256) void k_query_keys(int64_t n, GridGeom g, const TQ* __restrict__ x,
                                                    const TQ* __restrict__ y, const TQ* __restrict__ z,
                                                    uint32_t* __restrict__ key, QRec<TQ>* __restrict__ rec) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const TQ xv = x[p], yv = y[p], zv = z[p];
    double cx = ((double)xv - g.x0) / g.dx;
    cx = cx < (g.nx - 1.) ? cx : (g.nx - 1.);
    cx = cx > 0. ? cx : 0.;
    double cy = ((double)yv - g.y0) / g.dy;
    cy = cy < (g.ny - 1.) ? cy : (g.ny - 1.);
    cy = cy > 0. ? cy : 0.;
    key[p] = (uint32_t)((int64_t)cx * g.ny + (int64_t)cy);
    QRec<TQ> r;
    r.x = xv; r.y = yv; r.z = zv;
    r.p = (decltype(r.p))p;
    rec[p] = r;
}

}  // namespace

template <typename TQ>
hipError_t query_sort_tmp_bytes(int64_t n, int key_bits, size_t* tmp_bytes) {
    if (n <= 0 || n > INT32_MAX || key_bits < 1 || key_bits > 32) return hipErrorInvalidValue;
    *tmp_bytes = 0;
    return hipcub::DeviceRadixSort::SortPairs(nullptr, *tmp_bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                              (const QRec<TQ>*)nullptr, (QRec<TQ>*)nullptr, (int)n, 0, key_bits);
}

template <typename TQ>
hipError_t launch_query_sort(int64_t n, const GridGeom& g, const TQ* x, const TQ* y, const TQ* z, int key_bits,
                             uint32_t* keys, QRec<TQ>* recs, void* tmp, size_t tmp_bytes, hipStream_t st) {
    if (n <= 0 || n > INT32_MAX || key_bits < 1 || key_bits > 32) return hipErrorInvalidValue;
    if ((int64_t)g.nx * g.ny > ((int64_t)1 << key_bits)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_query_keys<TQ>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, g, x, y, z, keys,
                       recs);
    if (hipError_t e = hipGetLastError()) return e;
    size_t tb = tmp_bytes;
    return hipcub::DeviceRadixSort::SortPairs(tmp, tb, keys, keys + n, recs, recs + n, (int)n, 0, key_bits, st);
}

template hipError_t query_sort_tmp_bytes<float>(int64_t, int, size_t*);
template hipError_t query_sort_tmp_bytes<double>(int64_t, int, size_t*);
template hipError_t launch_query_sort<float>(int64_t, const GridGeom&, const float*, const float*, const float*, int,
                                             uint32_t*, QRec<float>*, void*, size_t, hipStream_t);
template hipError_t launch_query_sort<double>(int64_t, const GridGeom&, const double*, const double*, const double*,
                                              int, uint32_t*, QRec<double>*, void*, size_t, hipStream_t);

}  // namespace cfdwm
