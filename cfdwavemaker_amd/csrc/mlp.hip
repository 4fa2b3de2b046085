// Per-column MLP of the learned second-order surrogate: hand-written HIP
// kernels for inference and training (no counterpart in the reference; the
// surrogate predicts the second-order correction that src/Irregular.cpp's
// pair sums produce, and the exact kernels stay the fallback).
//
// Network: L linear layers, GELU (erf form) after every layer but the last.
// Master parameters are FP32 (Adam state too); every product runs on the
// matrix cores with both operands rounded to bf16 (round to nearest even)
// and FP32 accumulation, i.e. bf16 weights and activations, FP32 sums.
// Inputs and outputs are standardised inside (per-feature mean / std).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "engine.hpp"
#include "../../include/cfdwm.h"

namespace cfdwm {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint16_t to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// C[M x N] = A[M x K] . B[K x N] with A(m, k) = A[m sam + k sak],
// B(k, n) = B[k sbk + n sbn]; bf16 operands, FP32 accumulation.
// epi 0: C = acc;  1: Z = acc + bias, C = gelu(Z);  2: C = acc + bias.
struct MmArgs {
    int M, N, K;
    const float* A;
    int64_t sam, sak;
    const float* B;
    int64_t sbk, sbn;
    float* C;
    int64_t scm;
    const float* bias;
    float* Z;
    int epi;
};

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float x) {
    return 0.5f * (1.f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

// epilogue of a 32 x 32 accumulator tile (rows m0.., column cn of this lane)
__device__ __forceinline__ void mm_store(const MmArgs& a, const f32x16& acc, int m0, int cn, int hl) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int m = m0 + 8 * (r >> 2) + 4 * hl + (r & 3);
        if (m >= a.M || cn >= a.N) continue;
        float v = acc[r];
        const int64_t o = (int64_t)m * a.scm + cn;
        if (a.epi == 1) {
            v += a.bias[cn];
            a.Z[o] = v;
            a.C[o] = gelu(v);
        } else if (a.epi == 2) {
            a.C[o] = v + a.bias[cn];
        } else {
            a.C[o] = v;
        }
    }
}

// Block 256 threads = 4 waves, 64 x 64 outputs; wave (w & 1, w >> 1) takes a
// 32 x 32 tile (one v_mfma_f32_32x32x16_bf16 accumulator).  Operand element j
// of lane l is row (col) l % 32 at k = k0 + 8 (l / 32) + j; out-of-range
// elements load as zero.
__global__ __launch_bounds__(256) void k_mm(MmArgs a) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int m0 = blockIdx.y * 64 + (w & 1) * 32;
    const int n0 = blockIdx.x * 64 + (w >> 1) * 32;
    const int rm = m0 + (lane & 31), cn = n0 + (lane & 31);
    const int kh = (lane >> 5) * 8;
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.f;
    for (int k0 = 0; k0 < a.K; k0 += 16) {
        uint16_t av[8], bv[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int k = k0 + kh + j;
            const bool kin = k < a.K;
            av[j] = to_bf16(kin && rm < a.M ? a.A[(int64_t)rm * a.sam + (int64_t)k * a.sak] : 0.f);
            bv[j] = to_bf16(kin && cn < a.N ? a.B[(int64_t)k * a.sbk + (int64_t)cn * a.sbn] : 0.f);
        }
        bf16x8 A8, B8;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            A8[j] = __builtin_bit_cast(__bf16, av[j]);
            B8[j] = __builtin_bit_cast(__bf16, bv[j]);
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A8, B8, acc, 0, 0, 0);
    }
    mm_store(a, acc, m0, cn, lane >> 5);
}

// The same product when both operands are contiguous in k (sak = sbk = 1),
// K is a multiple of 16 and rows start 16-byte aligned: each lane's 8 k
// values of A and of B come as two 16-byte loads, and the next k step's are
// requested before this step's product.
__device__ __forceinline__ bf16x8 pack8(const float4& u, const float4& v) {
    bf16x8 r;
    const float f[8] = {u.x, u.y, u.z, u.w, v.x, v.y, v.z, v.w};
#pragma unroll
    for (int j = 0; j < 8; ++j) r[j] = __builtin_bit_cast(__bf16, to_bf16(f[j]));
    return r;
}

__global__ __launch_bounds__(256) void k_mm_vec(MmArgs a) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int m0 = blockIdx.y * 64 + (w & 1) * 32;
    const int n0 = blockIdx.x * 64 + (w >> 1) * 32;
    const int rm = m0 + (lane & 31), cn = n0 + (lane & 31);
    const int kh = (lane >> 5) * 8;
    // rows past M / columns past N read row 0 / column 0 (in bounds) and are
    // not stored
    const float4* pa = reinterpret_cast<const float4*>(a.A + (int64_t)(rm < a.M ? rm : 0) * a.sam + kh);
    const float4* pb = reinterpret_cast<const float4*>(a.B + (int64_t)(cn < a.N ? cn : 0) * a.sbn + kh);
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.f;
    float4 a0 = pa[0], a1 = pa[1], b0 = pb[0], b1 = pb[1];
    for (int k0 = 16; k0 < a.K; k0 += 16) {
        const float4 na0 = pa[k0 / 4], na1 = pa[k0 / 4 + 1], nb0 = pb[k0 / 4], nb1 = pb[k0 / 4 + 1];
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pack8(a0, a1), pack8(b0, b1), acc, 0, 0, 0);
        a0 = na0; a1 = na1; b0 = nb0; b1 = nb1;
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pack8(a0, a1), pack8(b0, b1), acc, 0, 0, 0);
    mm_store(a, acc, m0, cn, lane >> 5);
}

// ------------------------------------------------------------ fused inference
// The whole network for 64 rows per block: the standardised input and each
// hidden layer's output stay in LDS as bf16 (the operand rounding the
// layer-wise kernels apply when they load them), weights come from the FP32
// master parameters, and only the unstandardised outputs reach HBM.  Each
// wave takes 32-column blocks of a layer in turn, both 32-row halves at once
// (k order, products and epilogue as k_mm / k_mm_vec), so the outputs equal
// the layer-wise path's bit for bit.  Limits: at most kFiL layers, widths up to kFiW.
constexpr int kFiR = 64, kFiW = 256, kFiS = kFiW + 8, kFiL = 8;   // rows, max width, LDS row stride

struct MlpFused {
    int64_t n = 0;
    int L = 0, nin = 0, nout = 0;
    int din[kFiL] = {}, dout[kFiL] = {};
    const float* W[kFiL] = {};            // W_l[j][k], dout x din
    const float* b[kFiL] = {};
    const float *mu_in = nullptr, *sd_in = nullptr, *mu_out = nullptr, *sd_out = nullptr;
    const float* X = nullptr;             // n x nin
    float* Y = nullptr;                   // n x nout
};

__global__ __launch_bounds__(256) void k_mlp_fused(MlpFused a) {
    __shared__ __attribute__((aligned(16))) uint16_t H[2][kFiR * kFiS];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t r0 = (int64_t)blockIdx.x * kFiR;
    // layer 0 input, standardised (as k_standardise), zero past nin up to a multiple of 16
    const int k0p = (a.nin + 15) & ~15;
    for (int idx = threadIdx.x; idx < kFiR * k0p; idx += 256) {
        const int r = idx / k0p, c = idx - r * k0p;
        float v = 0.f;
        if (c < a.nin && r0 + r < a.n) v = (a.X[(r0 + r) * a.nin + c] - a.mu_in[c]) / a.sd_in[c];
        H[0][r * kFiS + c] = to_bf16(v);
    }
    __syncthreads();
    const int kh = (lane >> 5) * 8, hl = lane >> 5;
    int cur = 0;
    for (int l = 0; l < a.L; ++l) {
        const int K = a.din[l], N = a.dout[l];
        const int Kp = (K + 15) & ~15;
        const bool last = l == a.L - 1;
        const float* __restrict__ W = a.W[l];
        const float* __restrict__ bias = a.b[l];
        // 16-byte weight loads when every k step lies inside the row (K a multiple of 16)
        const bool wvec = (K & 15) == 0 && (reinterpret_cast<uintptr_t>(W) & 15) == 0;
        const int nct = (N + 31) / 32;
        // column block ct of both 32-row halves: two accumulators sharing each B operand
        for (int ct = w; ct < nct; ct += 4) {
            const int col = ct * 32 + (lane & 31);
            const float* __restrict__ wc = W + (int64_t)(col < N ? col : 0) * K;
            const uint16_t* h0 = &H[cur][(lane & 31) * kFiS + kh];
            const uint16_t* h1 = h0 + 32 * kFiS;
            f32x16 acc0, acc1;
#pragma unroll
            for (int r = 0; r < 16; ++r) acc0[r] = acc1[r] = 0.f;
            if (wvec) {
                // 16-byte weight loads, the next k step's requested before this step's products
                const float4* q = reinterpret_cast<const float4*>(wc + kh);
                float4 u = q[0], v = q[1];
                for (int k0 = 0; k0 < Kp; k0 += 16) {
                    const int kn = k0 + 16 < Kp ? k0 + 16 : k0;
                    const float4 un = q[kn / 4], vn = q[kn / 4 + 1];
                    const float f[8] = {u.x, u.y, u.z, u.w, v.x, v.y, v.z, v.w};
                    bf16x8 B8;
#pragma unroll
                    for (int j = 0; j < 8; ++j) B8[j] = __builtin_bit_cast(__bf16, to_bf16(col < N ? f[j] : 0.f));
                    const bf16x8 A0 = *reinterpret_cast<const bf16x8*>(h0 + k0);
                    const bf16x8 A1 = *reinterpret_cast<const bf16x8*>(h1 + k0);
                    acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A0, B8, acc0, 0, 0, 0);
                    acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A1, B8, acc1, 0, 0, 0);
                    u = un;
                    v = vn;
                }
            } else {
                for (int k0 = 0; k0 < Kp; k0 += 16) {
                    bf16x8 B8;
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        const int k = k0 + kh + j;
                        B8[j] = __builtin_bit_cast(__bf16, to_bf16(k < K && col < N ? wc[k] : 0.f));
                    }
                    const bf16x8 A0 = *reinterpret_cast<const bf16x8*>(h0 + k0);
                    const bf16x8 A1 = *reinterpret_cast<const bf16x8*>(h1 + k0);
                    acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A0, B8, acc0, 0, 0, 0);
                    acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A1, B8, acc1, 0, 0, 0);
                }
            }
            if (col >= N) continue;
            const float bc = bias[col];
#pragma unroll
            for (int rt = 0; rt < 2; ++rt) {
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int row = rt * 32 + 8 * (r >> 2) + 4 * hl + (r & 3);
                    const float v = (rt ? acc1[r] : acc0[r]) + bc;
                    if (!last) {
                        H[cur ^ 1][row * kFiS + col] = to_bf16(gelu(v));
                    } else if (r0 + row < a.n) {
                        a.Y[(r0 + row) * a.nout + col] = v * a.sd_out[col] + a.mu_out[col];
                    }
                }
            }
        }
        if (!last) {
            // columns past N read as zeros by the next layer's last k step
            const int Np = (N + 15) & ~15;
            for (int idx = threadIdx.x; idx < kFiR * (Np - N); idx += 256) {
                const int r = idx / (Np - N), c = N + idx - r * (Np - N);
                H[cur ^ 1][r * kFiS + c] = 0;
            }
        }
        __syncthreads();
        cur ^= 1;
    }
}

// X_hat = (X - mean) / std over n rows of d features (row-major)
__global__ void k_standardise(int64_t n, int d, const float* __restrict__ X, const float* __restrict__ mu,
                              const float* __restrict__ sd, float* __restrict__ Y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * d) return;
    const int j = (int)(i % d);
    Y[i] = (X[i] - mu[j]) / sd[j];
}
// Y = Y_hat * std + mean
__global__ void k_unstandardise(int64_t n, int d, float* __restrict__ Y, const float* __restrict__ mu,
                                const float* __restrict__ sd) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * d) return;
    const int j = (int)(i % d);
    Y[i] = Y[i] * sd[j] + mu[j];
}
// rows of X (and Y) picked by idx into Xb (Yb), standardised
__global__ void k_gather_std(int64_t nb, int d, const int64_t* __restrict__ idx, const float* __restrict__ X,
                             const float* __restrict__ mu, const float* __restrict__ sd, float* __restrict__ Xb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb * d) return;
    const int64_t r = i / d;
    const int j = (int)(i - r * d);
    Xb[i] = (X[idx[r] * d + j] - mu[j]) / sd[j];
}
// dY = 2 (Yp - Y) / (n d) and per-block partial sums of (Yp - Y)^2
__global__ __launch_bounds__(256) void k_mse(int64_t n, int d, const float* __restrict__ Yp, const float* __restrict__ Y,
                                             float* __restrict__ dY, float* __restrict__ part) {
    __shared__ float red[256];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float e2 = 0.f;
    if (i < n * d) {
        const float e = Yp[i] - Y[i];
        dY[i] = 2.f * e / (float)(n * d);
        e2 = e * e;
    }
    red[threadIdx.x] = e2;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
// dZ = dH * gelu'(Z)
__global__ void k_gelu_bwd(int64_t n, const float* __restrict__ dH, const float* __restrict__ Z, float* __restrict__ dZ) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dZ[i] = dH[i] * gelu_d(Z[i]);
}
// db[j] = sum_r dZ[r][j]  (one thread per column, rows in order: deterministic)
__global__ void k_colsum(int64_t n, int d, const float* __restrict__ dZ, float* __restrict__ db) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= d) return;
    float s = 0.f;
    for (int64_t r = 0; r < n; ++r) s += dZ[r * d + j];
    db[j] = s;
}
__global__ void k_adam(int64_t n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

inline unsigned nb256(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

struct Mlp {
    int nin = 0, width = 0, nhid = 0, nout = 0;
    std::vector<int> din, dout;           // per layer
    std::vector<int64_t> woff, boff;      // parameter offsets
    int64_t np = 0;
    DevBuf par, grad, am, av;             // FP32 master parameters, gradients, Adam moments
    DevBuf norm;                          // in_mean, in_std, out_mean, out_std
    DevBuf act;                           // H_0..H_L and Z_0..Z_{L-1} for a batch
    DevBuf tmp, part, idx;
    int64_t cap = 0;                      // rows the activation buffers hold
    int64_t adam_t = 0;
    int layers() const { return (int)din.size(); }
    float* P() const { return (float*)par.get(); }
    float* G() const { return (float*)grad.get(); }
    float* nin_mu() const { return (float*)norm.get(); }
    float* nin_sd() const { return nin_mu() + nin; }
    float* nout_mu() const { return nin_sd() + nin; }
    float* nout_sd() const { return nout_mu() + nout; }
    // activation layout for a batch of cap rows
    int64_t hoff(int l) const {          // H_l, l = 0..L
        int64_t o = 0;
        for (int q = 0; q < l; ++q) o += cap * din[q];
        return o;
    }
    int64_t zoff(int l) const {          // Z_l after all H
        int64_t o = hoff(layers()) + cap * nout;
        for (int q = 0; q < l; ++q) o += cap * dout[q];
        return o;
    }
    int reserve(int64_t rows);
    int forward(int64_t n, hipStream_t st);   // H_0 filled; runs all layers
    int backward(int64_t n, const float* dY, hipStream_t st);
};

int Mlp::reserve(int64_t rows) {
    if (rows <= cap) return 0;
    cap = rows;
    int64_t tot = 0;
    for (int l = 0; l < layers(); ++l) tot += cap * (din[l] + dout[l]);
    tot += cap * nout;
    if (act.alloc((size_t)tot * sizeof(float))) return -3;
    int64_t wmax = 0;
    for (int l = 0; l < layers(); ++l) wmax = std::max<int64_t>(wmax, std::max(din[l], dout[l]));
    if (tmp.alloc((size_t)cap * wmax * 2 * sizeof(float))) return -3;
    if (part.alloc((size_t)(nb256(cap * nout) + 1) * sizeof(float))) return -3;
    if (idx.alloc((size_t)cap * sizeof(int64_t))) return -3;
    return 0;
}

static hipError_t mm(const MmArgs& a, hipStream_t st) {
    if (a.M <= 0 || a.N <= 0) return hipSuccess;
    const dim3 grid((unsigned)((a.N + 63) / 64), (unsigned)((a.M + 63) / 64));
    auto al16 = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
    const bool vec = a.sak == 1 && a.sbk == 1 && a.K >= 16 && a.K % 16 == 0 && a.sam % 4 == 0 && a.sbn % 4 == 0 &&
                     al16(a.A) && al16(a.B);
    if (vec) hipLaunchKernelGGL(k_mm_vec, grid, dim3(256), 0, st, a);
    else hipLaunchKernelGGL(k_mm, grid, dim3(256), 0, st, a);
    return hipGetLastError();
}

int Mlp::forward(int64_t n, hipStream_t st) {
    float* A = (float*)act.get();
    for (int l = 0; l < layers(); ++l) {
        MmArgs a{};
        a.M = (int)n; a.N = dout[l]; a.K = din[l];
        a.A = A + hoff(l); a.sam = din[l]; a.sak = 1;
        a.B = P() + woff[l]; a.sbk = 1; a.sbn = din[l];          // B(k, j) = W[j][k]
        a.bias = P() + boff[l];
        a.scm = dout[l];
        if (l + 1 < layers()) {
            a.C = A + hoff(l + 1);
            a.Z = A + zoff(l);
            a.epi = 1;
        } else {
            a.C = A + hoff(l + 1);
            a.epi = 2;
        }
        CFDWM_HIP(mm(a, st));
    }
    return 0;
}

// dY must sit at the start of tmp (the first of its two halves); dZ and dH
// alternate between the halves.
int Mlp::backward(int64_t n, const float* dY, hipStream_t st) {
    float* A = (float*)act.get();
    int64_t wmax = 0;
    for (int l = 0; l < layers(); ++l) wmax = std::max<int64_t>(wmax, std::max(din[l], dout[l]));
    float* bufA = (float*)tmp.get();
    float* bufB = bufA + cap * wmax;
    if (dY != bufA) return -1;
    float* dZ = bufA;
    float* dH = bufB;
    for (int l = layers() - 1; l >= 0; --l) {
        // dW_l = dZ^T H_l : (dout x n) . (n x din)
        MmArgs g{};
        g.M = dout[l]; g.N = din[l]; g.K = (int)n;
        g.A = dZ; g.sam = 1; g.sak = dout[l];
        g.B = A + hoff(l); g.sbk = din[l]; g.sbn = 1;
        g.C = G() + woff[l]; g.scm = din[l]; g.epi = 0;
        CFDWM_HIP(mm(g, st));
        hipLaunchKernelGGL(k_colsum, dim3(nb256(dout[l])), dim3(256), 0, st, n, dout[l], (const float*)dZ, G() + boff[l]);
        CFDWM_HIP(hipGetLastError());
        if (l == 0) break;
        // dH_l = dZ W_l : (n x dout) . (dout x din);  dZ_{l-1} = dH_l * gelu'(Z_{l-1})
        dH = dZ == bufA ? bufB : bufA;
        MmArgs h{};
        h.M = (int)n; h.N = din[l]; h.K = dout[l];
        h.A = dZ; h.sam = dout[l]; h.sak = 1;
        h.B = P() + woff[l]; h.sbk = din[l]; h.sbn = 1;
        h.C = dH; h.scm = din[l]; h.epi = 0;
        CFDWM_HIP(mm(h, st));
        hipLaunchKernelGGL(k_gelu_bwd, dim3(nb256(n * din[l])), dim3(256), 0, st, n * din[l], (const float*)dH,
                           (const float*)(A + zoff(l - 1)), dH);
        CFDWM_HIP(hipGetLastError());
        dZ = dH;
    }
    return 0;
}

}  // namespace cfdwm

using namespace cfdwm;

struct cfdwm_mlp {
    Mlp m;
};

extern "C" {

int cfdwm_mlp_create(cfdwm_mlp** out, int nin, int width, int nhidden, int nout, uint64_t seed) {
    if (!out || nin < 1 || width < 1 || nhidden < 0 || nout < 1) return -1;
    if (int e = Engine::get().init()) return e;
    auto* h = new cfdwm_mlp();
    Mlp& M = h->m;
    M.nin = nin; M.width = width; M.nhid = nhidden; M.nout = nout;
    int d = nin;
    for (int l = 0; l <= nhidden; ++l) {
        const int o = l < nhidden ? width : nout;
        M.din.push_back(d);
        M.dout.push_back(o);
        d = o;
    }
    for (int l = 0; l < M.layers(); ++l) {
        M.woff.push_back(M.np);
        M.np += (int64_t)M.din[l] * M.dout[l];
        M.boff.push_back(M.np);
        M.np += M.dout[l];
    }
    // He-style uniform initialisation (limit sqrt(6 / fan_in)), zero biases
    std::vector<float> p(M.np, 0.f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l < M.layers(); ++l) {
        const double lim = std::sqrt(6.0 / M.din[l]);
        std::uniform_real_distribution<double> U(-lim, lim);
        for (int64_t q = 0; q < (int64_t)M.din[l] * M.dout[l]; ++q) p[M.woff[l] + q] = (float)U(rng);
    }
    std::vector<float> nrm(2 * nin + 2 * nout, 0.f);
    for (int j = 0; j < nin; ++j) nrm[nin + j] = 1.f;
    for (int j = 0; j < nout; ++j) nrm[2 * nin + nout + j] = 1.f;
    hipStream_t st = Engine::get().stream();
    std::vector<float> z(M.np, 0.f);
    if (M.par.upload(p.data(), p.size() * 4, st) || M.grad.upload(z.data(), z.size() * 4, st) ||
        M.am.upload(z.data(), z.size() * 4, st) || M.av.upload(z.data(), z.size() * 4, st) ||
        M.norm.upload(nrm.data(), nrm.size() * 4, st) || hipStreamSynchronize(st) != hipSuccess) {
        delete h;
        return -3;
    }
    *out = h;
    return 0;
}

void cfdwm_mlp_destroy(cfdwm_mlp* h) {
    if (h) {
        (void)hipStreamSynchronize(Engine::get().stream());
        delete h;
    }
}

int64_t cfdwm_mlp_nparams(cfdwm_mlp* h) { return h ? h->m.np : -1; }

int cfdwm_mlp_get_params(cfdwm_mlp* h, float* out) {
    if (!h || !out) return -1;
    hipStream_t st = Engine::get().stream();
    CFDWM_HIP(hipMemcpyAsync(out, h->m.P(), h->m.np * 4, hipMemcpyDeviceToHost, st));
    CFDWM_HIP(hipStreamSynchronize(st));
    return 0;
}

int cfdwm_mlp_set_params(cfdwm_mlp* h, const float* in) {
    if (!h || !in) return -1;
    hipStream_t st = Engine::get().stream();
    CFDWM_HIP(hipMemcpyAsync(h->m.P(), in, h->m.np * 4, hipMemcpyHostToDevice, st));
    CFDWM_HIP(hipStreamSynchronize(st));
    h->m.adam_t = 0;
    return 0;
}

int cfdwm_mlp_set_norm(cfdwm_mlp* h, const float* in_mean, const float* in_std, const float* out_mean,
                       const float* out_std) {
    if (!h || !in_mean || !in_std || !out_mean || !out_std) return -1;
    const Mlp& M = h->m;
    for (int j = 0; j < M.nin; ++j) if (!(in_std[j] > 0)) return -1;
    for (int j = 0; j < M.nout; ++j) if (!(out_std[j] > 0)) return -1;
    std::vector<float> n(2 * M.nin + 2 * M.nout);
    std::memcpy(n.data(), in_mean, M.nin * 4);
    std::memcpy(n.data() + M.nin, in_std, M.nin * 4);
    std::memcpy(n.data() + 2 * M.nin, out_mean, M.nout * 4);
    std::memcpy(n.data() + 2 * M.nin + M.nout, out_std, M.nout * 4);
    hipStream_t st = Engine::get().stream();
    if (h->m.norm.upload(n.data(), n.size() * 4, st)) return -3;
    CFDWM_HIP(hipStreamSynchronize(st));
    return 0;
}

int cfdwm_mlp_get_norm(cfdwm_mlp* h, float* out) {
    if (!h || !out) return -1;
    hipStream_t st = Engine::get().stream();
    CFDWM_HIP(hipMemcpyAsync(out, h->m.norm.get(), (2 * h->m.nin + 2 * h->m.nout) * 4, hipMemcpyDeviceToHost, st));
    CFDWM_HIP(hipStreamSynchronize(st));
    return 0;
}

// Y = net(X) for n rows (device pointers, raw units), in chunks of at most 65536 rows.
int cfdwm_mlp_forward_device(cfdwm_mlp* h, int64_t n, const float* X, float* Y) {
    if (!h || n < 0 || (n && (!X || !Y))) return -1;
    Mlp& M = h->m;
    hipStream_t st = Engine::get().stream();
    // large batches: the fused kernel (256 blocks of 64 rows fill the GPU);
    // smaller ones run layer by layer, whose 64 x 64 output tiles give more blocks
    bool fits = M.layers() <= kFiL && M.nin <= kFiW;
    for (int l = 0; l < M.layers(); ++l) fits = fits && M.dout[l] <= kFiW;
    if (fits && n >= 256 * kFiR) {
        MlpFused F;
        F.n = n; F.L = M.layers(); F.nin = M.nin; F.nout = M.nout;
        for (int l = 0; l < F.L; ++l) {
            F.din[l] = M.din[l]; F.dout[l] = M.dout[l];
            F.W[l] = M.P() + M.woff[l]; F.b[l] = M.P() + M.boff[l];
        }
        F.mu_in = M.nin_mu(); F.sd_in = M.nin_sd(); F.mu_out = M.nout_mu(); F.sd_out = M.nout_sd();
        F.X = X; F.Y = Y;
        hipLaunchKernelGGL(k_mlp_fused, dim3((unsigned)((n + kFiR - 1) / kFiR)), dim3(256), 0, st, F);
        CFDWM_HIP(hipGetLastError());
        return 0;
    }
    const int64_t chunk = 65536;
    if (M.reserve(std::min(n, chunk))) return -3;
    float* A = (float*)M.act.get();
    for (int64_t r0 = 0; r0 < n; r0 += chunk) {
        const int64_t nr = std::min(chunk, n - r0);
        hipLaunchKernelGGL(k_standardise, dim3(nb256(nr * M.nin)), dim3(256), 0, st, nr, M.nin, X + r0 * M.nin,
                           (const float*)M.nin_mu(), (const float*)M.nin_sd(), A + M.hoff(0));
        CFDWM_HIP(hipGetLastError());
        if (int e = M.forward(nr, st)) return e;
        float* out = A + M.hoff(M.layers());
        hipLaunchKernelGGL(k_unstandardise, dim3(nb256(nr * M.nout)), dim3(256), 0, st, nr, M.nout, out,
                           (const float*)M.nout_mu(), (const float*)M.nout_sd());
        CFDWM_HIP(hipGetLastError());
        CFDWM_HIP(hipMemcpyAsync(Y + r0 * M.nout, out, nr * M.nout * 4, hipMemcpyDeviceToDevice, st));
    }
    return 0;
}

// Loss (mean squared error of the standardised outputs) and its gradient over
// the first n rows (n <= 65536), gradient copied to grad_host when given.
int cfdwm_mlp_loss_grad_device(cfdwm_mlp* h, int64_t n, const float* X, const float* Y, float* loss,
                               float* grad_host) {
    if (!h || n <= 0 || n > 65536 || !X || !Y) return -1;
    Mlp& M = h->m;
    hipStream_t st = Engine::get().stream();
    if (M.reserve(n)) return -3;
    std::vector<int64_t> id(n);
    for (int64_t r = 0; r < n; ++r) id[r] = r;
    if (M.idx.upload(id.data(), n * sizeof(int64_t), st)) return -3;
    float* A = (float*)M.act.get();
    hipLaunchKernelGGL(k_gather_std, dim3(nb256(n * M.nin)), dim3(256), 0, st, n, M.nin, (const int64_t*)M.idx.get(), X,
                       (const float*)M.nin_mu(), (const float*)M.nin_sd(), A + M.hoff(0));
    CFDWM_HIP(hipGetLastError());
    if (int e = M.forward(n, st)) return e;
    // targets standardised into the dY buffer's slot, then MSE in place
    float* Yb = (float*)M.tmp.get() + (M.tmp.size() / sizeof(float)) / 2;
    hipLaunchKernelGGL(k_gather_std, dim3(nb256(n * M.nout)), dim3(256), 0, st, n, M.nout, (const int64_t*)M.idx.get(), Y,
                       (const float*)M.nout_mu(), (const float*)M.nout_sd(), Yb);
    CFDWM_HIP(hipGetLastError());
    const unsigned nbk = nb256(n * M.nout);
    float* dY = (float*)M.tmp.get();
    hipLaunchKernelGGL(k_mse, dim3(nbk), dim3(256), 0, st, n, M.nout, (const float*)(A + M.hoff(M.layers())),
                       (const float*)Yb, dY, (float*)M.part.get());
    CFDWM_HIP(hipGetLastError());
    if (int e = M.backward(n, dY, st)) return e;
    std::vector<float> pt(nbk);
    CFDWM_HIP(hipMemcpyAsync(pt.data(), M.part.get(), nbk * 4, hipMemcpyDeviceToHost, st));
    if (grad_host) CFDWM_HIP(hipMemcpyAsync(grad_host, M.G(), M.np * 4, hipMemcpyDeviceToHost, st));
    CFDWM_HIP(hipStreamSynchronize(st));
    double s = 0;
    for (float v : pt) s += v;
    if (loss) *loss = (float)(s / (double)(n * M.nout));
    return 0;
}

// Adam on minibatches of `batch` rows drawn (with a seeded generator) from the
// n rows of X / Y (device pointers, raw units); learning rate lr with cosine
// decay to lr / 20 over `steps`.  loss_curve[s] (optional, host) receives the
// minibatch loss every `report` steps (others 0).
int cfdwm_mlp_train_device(cfdwm_mlp* h, int64_t n, const float* X, const float* Y, int steps, int batch, float lr,
                           uint64_t seed, int report, float* loss_curve) {
    if (!h || n <= 0 || !X || !Y || steps < 0 || batch <= 0 || batch > 65536) return -1;
    Mlp& M = h->m;
    hipStream_t st = Engine::get().stream();
    const int64_t nbt = std::min<int64_t>(batch, n);
    if (M.reserve(nbt)) return -3;
    std::mt19937_64 rng(seed);
    std::uniform_int_distribution<int64_t> U(0, n - 1);
    std::vector<int64_t> id(nbt);
    float* A = (float*)M.act.get();
    float* dY = (float*)M.tmp.get();
    float* Yb = dY + (M.tmp.size() / sizeof(float)) / 2;
    const unsigned nbk = nb256(nbt * M.nout);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int s = 0; s < steps; ++s) {
        for (auto& v : id) v = U(rng);
        if (M.idx.upload(id.data(), nbt * sizeof(int64_t), st)) return -3;
        hipLaunchKernelGGL(k_gather_std, dim3(nb256(nbt * M.nin)), dim3(256), 0, st, nbt, M.nin,
                           (const int64_t*)M.idx.get(), X, (const float*)M.nin_mu(), (const float*)M.nin_sd(),
                           A + M.hoff(0));
        hipLaunchKernelGGL(k_gather_std, dim3(nb256(nbt * M.nout)), dim3(256), 0, st, nbt, M.nout,
                           (const int64_t*)M.idx.get(), Y, (const float*)M.nout_mu(), (const float*)M.nout_sd(), Yb);
        CFDWM_HIP(hipGetLastError());
        if (int e = M.forward(nbt, st)) return e;
        hipLaunchKernelGGL(k_mse, dim3(nbk), dim3(256), 0, st, nbt, M.nout, (const float*)(A + M.hoff(M.layers())),
                           (const float*)Yb, dY, (float*)M.part.get());
        CFDWM_HIP(hipGetLastError());
        if (int e = M.backward(nbt, dY, st)) return e;
        ++M.adam_t;
        const double frac = steps > 1 ? (double)s / (steps - 1) : 1.0;
        const float lrs = (float)(lr * (0.05 + 0.95 * 0.5 * (1.0 + std::cos(3.14159265358979323846 * frac))));
        const float c1 = (float)(1.0 - std::pow(b1, (double)M.adam_t)), c2 = (float)(1.0 - std::pow(b2, (double)M.adam_t));
        hipLaunchKernelGGL(k_adam, dim3(nb256(M.np)), dim3(256), 0, st, M.np, M.P(), (const float*)M.G(),
                           (float*)M.am.get(), (float*)M.av.get(), lrs, b1, b2, eps, c1, c2);
        CFDWM_HIP(hipGetLastError());
        if (loss_curve) {
            loss_curve[s] = 0.f;
            if (report > 0 && (s % report == 0 || s == steps - 1)) {
                std::vector<float> pt(nbk);
                CFDWM_HIP(hipMemcpyAsync(pt.data(), M.part.get(), nbk * 4, hipMemcpyDeviceToHost, st));
                CFDWM_HIP(hipStreamSynchronize(st));
                double q = 0;
                for (float v : pt) q += v;
                loss_curve[s] = (float)(q / (double)(nbt * M.nout));
            }
        }
    }
    CFDWM_HIP(hipStreamSynchronize(st));
    return 0;
}

}  // extern "C"
