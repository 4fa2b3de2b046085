// Filtered second-order elevation eta2 as a quadratic form on the matrix cores.
//
// The pair sum of Irregular::eta2 (src/Irregular.cpp:1125-1192, pair-cache
// semantics of lsGridSpline's column update) is, per point p,
//
//   eta2_p = sum_i r_i [ c_pi sum_{m>i} Bcc_im c_pm + s_pi sum_{m>i} Bss_im s_pm ]
//
// with c_pi = cos(theta_i(p)), s_pi = sin(theta_i(p)), r_i the frequency ramp
// of component i at the point's time, Bcc = xP + xM and Bss = xM - xP built
// from the filtered pair coefficients (the real form of Re(Z^T B+ Z + Z^H B- Z)
// with Z_i = e^{i theta_i}).  The inner sums are GEMMs (points x components
// times components x components) on v_mfma_f32_32x32x16_bf16; every operand is
// split into bf16 hi + lo parts and the three products hi*hi + hi*lo + lo*hi
// are accumulated in FP32, which keeps about 16 significant bits per product.
// The outer sum over i is a fused epilogue (accumulator times c or s) with a
// cross-lane reduction at the end; the column kernel folds the result into
// its FP64 total.  The same kernel evaluates the z-independent forms of the
// velocity pair sums under the separable depth basis (depth_basis.hpp): per
// skeleton wavenumber mu_q one cos-type form for u and v and one sin-type
// form for w, with B entries cuP c_q(kP) +- cuM c_q(kM) (and cv, cw).
//
// The per-row ramp r_i must be the same for every i at a given time (the
// usual case: all ramps finished, or all components ramped alike); the host
// checks this and passes one factor per time slot.
#include <hip/hip_runtime.h>

#include "engine.hpp"
#include "kernels.hpp"

namespace cfdwm {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr double kInv2PiQ = 0.15915494309189533577;

// float -> bf16 bits, round to nearest even (finite inputs)
__device__ __forceinline__ uint32_t bf16_bits(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return u >> 16;
}
__device__ __forceinline__ float bf16_val(uint32_t b) { return __uint_as_float(b << 16); }

// tiled operand offset of element (row r, k): 32 x 16 tiles, two 32 x 8 halves
__device__ __forceinline__ int64_t qf_off(int64_t r, int k, int nkb) {
    return ((r >> 5) * nkb + (k >> 4)) * 512 + ((k >> 3) & 1) * 256 + (r & 31) * 8 + (k & 7);
}

// ------------------------------------------------------------------ Z operands
// One thread per (point, 8 consecutive components): FP64 phase and range
// reduction, hardware sin/cos, bf16 hi/lo split, 16-byte stores.
__global__ __launch_bounds__(256) void k_qf_zgen(QfJob J) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int nkb = J.npad >> 4;
    const int64_t ntile = (J.ppad >> 5) * nkb;
    if ((tid >> 6) >= ntile) return;
    const int64_t tile = tid >> 6;
    const int h = (int)((tid >> 5) & 1), r = (int)(tid & 31);
    const int64_t rb = tile / nkb;
    const int kb = (int)(tile - rb * nkb);
    const int64_t p = rb * 32 + r;              // point within the chunk
    const int m0 = kb * 16 + h * 8;
    uint32_t ch[8], cl[8], sh[8], sl[8];
    const bool live = p < J.P;
    double x = 0, y = 0, t = 0;
    if (live) {
        const int64_t q = J.p0 + p;
        if (J.grid) {
            const int slot = (int)(q / J.ncol);
            const int64_t cq = q - (int64_t)slot * J.ncol;
            const int64_t col = J.cols ? J.cols[cq] : cq;
            x = J.g.x0 + J.g.dx * (double)(col / J.g.ny);
            y = J.g.y0 + J.g.dy * (double)(col % J.g.ny);
            t = J.slot_t[slot];
        } else {
            x = J.x[q]; y = J.y[q]; t = J.t[q];
        }
    }
    const double dx = x - J.x0, dy = y - J.y0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int m = m0 + j;
        float c = 0.f, s = 0.f;
        if (live && m < J.n) {
            const PhaseRec ph = J.ph[m];
            double a = ph.k * (ph.cth * dx + ph.sth * dy) - ph.omega * t + ph.ph0;
            a *= kInv2PiQ;
            a -= rint(a);
            const float f = (float)a;
            s = __builtin_amdgcn_sinf(f);
            c = __builtin_amdgcn_cosf(f);
        }
        ch[j] = bf16_bits(c); cl[j] = bf16_bits(c - bf16_val(ch[j]));
        sh[j] = bf16_bits(s); sl[j] = bf16_bits(s - bf16_val(sh[j]));
    }
    auto pack = [](const uint32_t* v) {
        return make_uint4(v[0] | (v[1] << 16), v[2] | (v[3] << 16), v[4] | (v[5] << 16), v[6] | (v[7] << 16));
    };
    const int64_t o = tile * 512 + h * 256 + r * 8;
    *reinterpret_cast<uint4*>(J.zc_hi + o) = pack(ch);
    *reinterpret_cast<uint4*>(J.zc_lo + o) = pack(cl);
    *reinterpret_cast<uint4*>(J.zs_hi + o) = pack(sh);
    *reinterpret_cast<uint4*>(J.zs_lo + o) = pack(sl);
}

__device__ __forceinline__ bf16x8 ld8(const uint16_t* p) {
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(p));
}

// butterfly reduce-scatter over the 32 lanes of each half-wave: on return
// v[0] of lane l holds the half-wave sum of the lane values v[l & 31]
template <int O>
__device__ __forceinline__ void bfly(float* v, int lane) {
    const bool up = (lane & O) != 0;
#pragma unroll
    for (int k = 0; k < O; ++k) {
        const float send = up ? v[k] : v[k + O];
        const float keep = up ? v[k + O] : v[k];
        v[k] = keep + __shfl_xor(send, O);
    }
}

// ----------------------------------------------------------- quadratic forms
// Form f of a launch:  Q_f(p) = sum_i e1_pi sum_m B1_f[i][m] c_pm + e2_pi sum_m B2_f[i][m] s_pm
// with (e1, e2) = (c, s) for cos-type forms and (s, c) for sin-type forms.
// Block (blockIdx.x, blockIdx.y = f): 4 waves, 128 points; the block walks
// the i range in tiles of 128, wave (w & 1, w >> 1) taking 64 points x 64 i as
// 2 x 2 MFMA 32 x 32 blocks for each of the two products.  The k (= m) range
// of a 64-wide i tile is [i0, kend): pairs have m > i and the filtered rows
// stop at mhi[i].  Per point the i sums stay in lane registers across tiles;
// one cross-lane reduction and one LDS exchange between the two waves of a
// point range at the end write Q_f(p).
__global__ __launch_bounds__(256) void k_qf_forms(QfJob J) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int nkb = J.npad >> 4;
    const int form = blockIdx.y;
    const size_t nn = (size_t)J.npad * J.npad;
    const uint16_t* __restrict__ B1h = J.b + (size_t)form * 4 * nn;
    const uint16_t* __restrict__ B1l = B1h + nn;
    const uint16_t* __restrict__ B2h = B1h + 2 * nn;
    const uint16_t* __restrict__ B2l = B1h + 3 * nn;
    const bool sint = J.ftype[form] != 0;
    const int64_t P0 = (int64_t)blockIdx.x * 128 + (w & 1) * 64;
    const int64_t rb0 = P0 >> 5;
    const int loff = (lane >> 5) * 256 + (lane & 31) * 8;
    const int hl = lane >> 5, cl32 = lane & 31;
    float v[32];
#pragma unroll
    for (int k = 0; k < 32; ++k) v[k] = 0.f;
    for (int it = 0; it < J.npad; it += 128) {
        const int I0 = it + (w >> 1) * 64;
        const int kb0 = I0 >> 4, kb1 = J.kend[I0 >> 6] >> 4;
        if (kb1 <= kb0) continue;
        const int ib0 = I0 >> 5;
        f32x16 accC[2][2], accS[2][2];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int r = 0; r < 16; ++r) { accC[a][b][r] = 0.f; accS[a][b][r] = 0.f; }
        for (int kb = kb0; kb < kb1; ++kb) {
            bf16x8 zch[2], zcl[2], zsh[2], zsl[2], bch[2], bcl[2], bsh[2], bsl[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const int64_t o = ((rb0 + a) * nkb + kb) * 512 + loff;
                zch[a] = ld8(J.zc_hi + o); zcl[a] = ld8(J.zc_lo + o);
                zsh[a] = ld8(J.zs_hi + o); zsl[a] = ld8(J.zs_lo + o);
            }
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const int64_t o = ((int64_t)(ib0 + b) * nkb + kb) * 512 + loff;
                bch[b] = ld8(B1h + o); bcl[b] = ld8(B1l + o);
                bsh[b] = ld8(B2h + o); bsl[b] = ld8(B2l + o);
            }
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    accC[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(zch[a], bch[b], accC[a][b], 0, 0, 0);
                    accC[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(zch[a], bcl[b], accC[a][b], 0, 0, 0);
                    accC[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(zcl[a], bch[b], accC[a][b], 0, 0, 0);
                    accS[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(zsh[a], bsh[b], accS[a][b], 0, 0, 0);
                    accS[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(zsh[a], bsl[b], accS[a][b], 0, 0, 0);
                    accS[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(zsl[a], bsh[b], accS[a][b], 0, 0, 0);
                }
        }
        // epilogue: accumulator element r of lane l is (row 8(r/4) + 4(l/32) + r%4,
        // column l%32) of its 32 x 32 block; times e1_pi (e2_pi), summed over i
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int64_t p = P0 + a * 32 + 8 * (r >> 2) + 4 * hl + (r & 3);
                float sum = 0.f;
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    const int i = I0 + b * 32 + cl32;
                    const int64_t o = qf_off(p, i, nkb);
                    const float c = bf16_val(J.zc_hi[o]) + bf16_val(J.zc_lo[o]);
                    const float s = bf16_val(J.zs_hi[o]) + bf16_val(J.zs_lo[o]);
                    sum += accC[a][b][r] * (sint ? s : c) + accS[a][b][r] * (sint ? c : s);
                }
                v[a * 16 + r] += sum;
            }
    }
    bfly<16>(v, lane);
    bfly<8>(v, lane);
    bfly<4>(v, lane);
    bfly<2>(v, lane);
    bfly<1>(v, lane);
    // waves (w & 1, 0) and (w & 1, 1) hold the two i halves of the same points
    __shared__ float red[2][64];
    if (w >> 1) red[w & 1][lane] = v[0];
    __syncthreads();
    if (!(w >> 1)) {
        const int k = cl32;
        const int64_t p = P0 + (k >> 4) * 32 + 8 * ((k & 15) >> 2) + 4 * hl + (k & 3);
        if (p < J.P) J.out[(int64_t)form * J.ostride + J.p0 + p] = v[0] + red[w & 1][lane];
    }
}

}  // namespace

static bool qf_shape_ok(const QfJob& J) {
    return !((J.ppad & 127) || (J.npad & 127) || J.ppad < J.P || J.npad < J.n);
}

hipError_t launch_qf_zgen(const QfJob& J, hipStream_t st) {
    if (J.P <= 0) return hipSuccess;
    if (!qf_shape_ok(J)) return hipErrorInvalidValue;
    const int64_t nthr = (J.ppad >> 5) * (J.npad >> 4) * 64;
    hipLaunchKernelGGL(k_qf_zgen, dim3((unsigned)((nthr + 255) / 256)), dim3(256), 0, st, J);
    return hipGetLastError();
}

hipError_t launch_qf_forms(const QfJob& J, hipStream_t st) {
    if (J.P <= 0 || J.nform <= 0) return hipSuccess;
    if (!qf_shape_ok(J)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_qf_forms, dim3((unsigned)(J.ppad / 128), (unsigned)J.nform), dim3(256), 0, st, J);
    return hipGetLastError();
}

}  // namespace cfdwm
