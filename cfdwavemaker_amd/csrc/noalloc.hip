// Allocation-free device spline query: the variant of
// cfdwm_grid_query_device_f32 for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  Batches of 65,536 queries or more are evaluated in cell order;
// the keys, query records and radix-sort scratch of that ordering live here in
// a file-scope device array instead of the grid's own buffers.  Each GPU holds
// its own copy of the array until the process ends.  Calls are enqueued on
// the library stream, one after another, so they never overlap; a caller that
// enqueued a query on another stream while one of these runs would race for
// the array.  Nothing in the array needs clearing: the key pass writes every
// key and record, the sort writes its own scratch.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "engine.hpp"

namespace cfdwm {

namespace {
constexpr size_t kScratchBytes = size_t(64) << 20;     // 64 MiB per GPU
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of this GPU's scratch array, looked up once per device with
// that device current.  In host code the array's name is only a stand-in:
// kernels and copies take this address, never the name.
int scratch_address(void** out) {
    *out = nullptr;
    int dev = 0;
    CFDWM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("allocation-free query: device index out of range");
        return -1;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        CFDWM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            set_error("allocation-free query: no device address for the scratch array");
            return -3;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}
}  // namespace

int query_device_f32_noalloc(Grid& grid, int64_t n, double t, const float* x, const float* y, const float* z,
                             float* eta, float* u, float* v, float* w) {
    if (n <= 0) return 0;
    if (int e = Engine::get().init()) return e;
    const int ncol = grid.g.nx * grid.g.ny;
    size_t need = 0;
    if (int e = query_ws_bytes<float>(n, ncol, &need)) return e;
    if (need > kScratchBytes) {
        // the largest batch whose ordering buffers fit the array
        int64_t lo = 0, hi = n;
        while (hi - lo > 1) {
            const int64_t mid = lo + (hi - lo) / 2;
            size_t b = 0;
            if (query_ws_bytes<float>(mid, ncol, &b)) return -3;
            (b <= kScratchBytes ? lo : hi) = mid;
        }
        set_error("allocation-free query: " + std::to_string(n) + " queries exceed the maximum of " +
                  std::to_string(lo) + " for this grid (" + std::to_string(kScratchBytes >> 20) +
                  " MiB of device scratch)");
        return -1;
    }
    void* ws = nullptr;
    if (need > 0)
        if (int e = scratch_address(&ws)) return e;
    return grid.query_device_f32(n, t, x, y, z, eta, u, v, w, nullptr, ws, kScratchBytes);
}

}  // namespace cfdwm
