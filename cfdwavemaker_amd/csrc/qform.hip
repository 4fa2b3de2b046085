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

#include <cstdlib>

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

// LDS-staged variant.  Block: 4 waves (one per SIMD), 128 points x 128 i
// per i tile; wave (w & 1, w >> 1) takes 64 points x 64 i.  Per k step (16
// values of m) the block stages the A operands of its 128 points and the B
// operands of its 128 i (4 arrays x 4 row blocks each, 32 KiB) in one of two
// LDS buffers while the other feeds the matrix cores; each operand tile is
// read from L2 once per block instead of once per wave.  Blocks are swizzled
// in groups of 8 forms x 32 point tiles so that the Z and B tables a group
// reads stay in the last-level cache.
constexpr int kQfGF = 8, kQfGP = 32;

__device__ __forceinline__ bf16x8 lds8(const char* base, int off) {
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(base + off));
}

// One (point tile, form) of a job: the whole i range for 128 points.
__device__ __forceinline__ void qf_tile(const QfJob J, int ptile, int form) {
    __shared__ __attribute__((aligned(16))) char lds[2][32768];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wp = w & 1, wi = w >> 1;
    const int nkb = J.npad >> 4;
    const size_t nn = (size_t)J.npad * J.npad;
    const uint16_t* __restrict__ Bf = J.b + (size_t)form * 4 * nn;
    const bool sint = J.ftype[form] != 0;
    const int64_t PB = (int64_t)ptile * 128;          // first point of the block
    const int64_t rbB = PB >> 5;                      // first point row block
    const int64_t P0 = PB + wp * 64;
    const int hl = lane >> 5, cl32 = lane & 31;
    const int loff = ((lane >> 5) * 256 + (lane & 31) * 8) * 2;   // lane's bytes in a 1 KiB operand tile
    float tot = 0.f;                                  // this lane's point after the reductions
    // stage piece q = r * 256 + thread (r < 8): chunk q >> 6 (16 A chunks:
    // array x 4 row blocks, then 16 B chunks), 16 bytes at (q & 63) * 16
#define QF_GL1(R, KB, IB0)                                                                         \
    {                                                                                              \
        const int q = (R) * 256 + threadIdx.x;                                                     \
        const int ch = q >> 6, o = (q & 63) * 8;                                                   \
        const uint16_t* src = ch < 16                                                              \
            ? J.zc_hi + (int64_t)(ch >> 2) * J.zn + ((rbB + (ch & 3)) * nkb + (KB)) * 512 + o       \
            : Bf + (size_t)((ch - 16) >> 2) * nn + ((int64_t)((IB0) + (ch & 3)) * nkb + (KB)) * 512 + o; \
        st##R = *reinterpret_cast<const uint4*>(src);                                              \
    }
#define QF_GLOAD(KB, IB0) \
    QF_GL1(0, KB, IB0) QF_GL1(1, KB, IB0) QF_GL1(2, KB, IB0) QF_GL1(3, KB, IB0) \
    QF_GL1(4, KB, IB0) QF_GL1(5, KB, IB0) QF_GL1(6, KB, IB0) QF_GL1(7, KB, IB0)
#define QF_SW1(R, BUF)                                                                             \
    {                                                                                              \
        const int q = (R) * 256 + threadIdx.x;                                                     \
        *reinterpret_cast<uint4*>((BUF) + (q >> 6) * 1024 + (q & 63) * 16) = st##R;                 \
    }
#define QF_SWRITE(BUF) \
    QF_SW1(0, BUF) QF_SW1(1, BUF) QF_SW1(2, BUF) QF_SW1(3, BUF) \
    QF_SW1(4, BUF) QF_SW1(5, BUF) QF_SW1(6, BUF) QF_SW1(7, BUF)
    uint4 st0, st1, st2, st3, st4, st5, st6, st7;
    for (int it = 0; it < J.npad; it += 128) {
        const int t0 = it >> 6;
        const int kb_lo = it >> 4;
        const int kb_hi = max(J.kend[t0], J.kend[t0 + 1]) >> 4;
        if (kb_hi <= kb_lo) continue;                 // block-uniform
        const int I0 = it + wi * 64;
        const int wk0 = I0 >> 4, wk1 = J.kend[t0 + wi] >> 4;
        const int ib0 = it >> 5;
        f32x16 accC[2][2], accS[2][2];
#pragma unroll
        for (int a2 = 0; a2 < 2; ++a2)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int r = 0; r < 16; ++r) { accC[a2][b][r] = 0.f; accS[a2][b][r] = 0.f; }
        QF_GLOAD(kb_lo, ib0)
        QF_SWRITE(lds[0])
        __syncthreads();
        for (int kb = kb_lo; kb < kb_hi; ++kb) {
            const int cur = (kb - kb_lo) & 1;
            const bool more = kb + 1 < kb_hi;
            if (more) { QF_GLOAD(kb + 1, ib0) }
            if (kb >= wk0 && kb < wk1) {
                const char* L = lds[cur];
#pragma unroll
                for (int a2 = 0; a2 < 2; ++a2) {
                    const int rb = wp * 2 + a2;
                    const bf16x8 zch = lds8(L, (0 * 4 + rb) * 1024 + loff);
                    const bf16x8 zcl = lds8(L, (1 * 4 + rb) * 1024 + loff);
                    const bf16x8 zsh = lds8(L, (2 * 4 + rb) * 1024 + loff);
                    const bf16x8 zsl = lds8(L, (3 * 4 + rb) * 1024 + loff);
#pragma unroll
                    for (int b = 0; b < 2; ++b) {
                        const int ib = wi * 2 + b;
                        const bf16x8 bch = lds8(L, 16384 + (0 * 4 + ib) * 1024 + loff);
                        const bf16x8 bcl = lds8(L, 16384 + (1 * 4 + ib) * 1024 + loff);
                        const bf16x8 bsh = lds8(L, 16384 + (2 * 4 + ib) * 1024 + loff);
                        const bf16x8 bsl = lds8(L, 16384 + (3 * 4 + ib) * 1024 + loff);
                        accC[a2][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(zch, bch, accC[a2][b], 0, 0, 0);
                        accC[a2][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(zch, bcl, accC[a2][b], 0, 0, 0);
                        accC[a2][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(zcl, bch, accC[a2][b], 0, 0, 0);
                        accS[a2][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(zsh, bsh, accS[a2][b], 0, 0, 0);
                        accS[a2][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(zsh, bsl, accS[a2][b], 0, 0, 0);
                        accS[a2][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(zsl, bsh, accS[a2][b], 0, 0, 0);
                    }
                }
            }
            if (more) { QF_SWRITE(lds[cur ^ 1]) }
            __syncthreads();
        }
        if (wk1 > wk0) {                              // wave-uniform
            // element (p, i) of a Z array: ((p >> 5) nkb + (i >> 4)) 512 + ((i >> 3) & 1) 256
            // + (p & 31) 8 + (i & 7); the lane's rows differ by compile-time offsets
            float v[32];
#pragma unroll
            for (int a2 = 0; a2 < 2; ++a2) {
                float sum[16];
#pragma unroll
                for (int r = 0; r < 16; ++r) sum[r] = 0.f;
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    const int i = I0 + b * 32 + cl32;
                    const uint16_t* zc = J.zc_hi + (((P0 >> 5) + a2) * nkb + (i >> 4)) * 512 + ((i >> 3) & 1) * 256
                                       + 4 * hl * 8 + (i & 7);
                    const uint16_t* zcl = zc + J.zn;
                    const uint16_t* zs = zc + 2 * J.zn;
                    const uint16_t* zsl = zc + 3 * J.zn;
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const int ro = (8 * (r >> 2) + (r & 3)) * 8;
                        const float c = bf16_val(zc[ro]) + bf16_val(zcl[ro]);
                        const float sv = bf16_val(zs[ro]) + bf16_val(zsl[ro]);
                        sum[r] += accC[a2][b][r] * (sint ? sv : c) + accS[a2][b][r] * (sint ? c : sv);
                        if ((r & 3) == 3) __asm__ volatile("" ::: "memory");
                    }
                }
#pragma unroll
                for (int r = 0; r < 16; ++r) v[a2 * 16 + r] = sum[r];
            }
            bfly<16>(v, lane);
            bfly<8>(v, lane);
            bfly<4>(v, lane);
            bfly<2>(v, lane);
            bfly<1>(v, lane);
            tot += v[0];
        }
    }
#undef QF_GL1
#undef QF_GLOAD
#undef QF_SW1
#undef QF_SWRITE
    // waves (wp, 0) and (wp, 1) hold the two i halves of the same points
    float* red = reinterpret_cast<float*>(lds[0]);
    __syncthreads();
    if (wi) red[wp * 64 + lane] = tot;
    __syncthreads();
    if (!wi) {
        const int k = cl32;
        const int64_t p = P0 + (k >> 4) * 32 + 8 * ((k & 15) >> 2) + 4 * hl + (k & 3);
        if (p < J.P) J.out[(int64_t)form * J.ostride + J.p0 + p] = tot + red[wp * 64 + lane];
    }
}

// Single-job launch.  The body repeats qf_tile's: compiled as this one
// kernel the schedule runs the N = 1024 fill's 55 forms in 123 ms, through the
// shared inline function in 145 ms (same instructions, different scheduling).
__global__ __launch_bounds__(256) void k_qf_forms_lds(QfJob J, int npt) {
    __shared__ __attribute__((aligned(16))) char lds[2][32768];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wp = w & 1, wi = w >> 1;
    // block swizzle: groups of kQfGF forms x kQfGP point tiles, point tile fastest
    const int id = blockIdx.x;
    const int per = kQfGF * kQfGP;
    const int grp = id / per, in = id - grp * per;
    const int nfg = (J.nform + kQfGF - 1) / kQfGF;
    const int pg = grp / nfg, fg = grp - pg * nfg;
    const int ptile = pg * kQfGP + in % kQfGP;
    const int form = fg * kQfGF + in / kQfGP;
    if (ptile >= npt || form >= J.nform) return;      // whole block, before any barrier
    const int nkb = J.npad >> 4;
    const size_t nn = (size_t)J.npad * J.npad;
    const uint16_t* __restrict__ Bf = J.b + (size_t)form * 4 * nn;
    const bool sint = J.ftype[form] != 0;
    const int64_t PB = (int64_t)ptile * 128;          // first point of the block
    const int64_t rbB = PB >> 5;                      // first point row block
    const int64_t P0 = PB + wp * 64;
    const int hl = lane >> 5, cl32 = lane & 31;
    const int loff = ((lane >> 5) * 256 + (lane & 31) * 8) * 2;   // lane's bytes in a 1 KiB operand tile
    float tot = 0.f;                                  // this lane's point after the reductions
    // stage piece q = r * 256 + thread (r < 8): chunk q >> 6 (16 A chunks:
    // array x 4 row blocks, then 16 B chunks), 16 bytes at (q & 63) * 16
#define QF_GL1(R, KB, IB0)                                                                         \
    {                                                                                              \
        const int q = (R) * 256 + threadIdx.x;                                                     \
        const int ch = q >> 6, o = (q & 63) * 8;                                                   \
        const uint16_t* src = ch < 16                                                              \
            ? J.zc_hi + (int64_t)(ch >> 2) * J.zn + ((rbB + (ch & 3)) * nkb + (KB)) * 512 + o       \
            : Bf + (size_t)((ch - 16) >> 2) * nn + ((int64_t)((IB0) + (ch & 3)) * nkb + (KB)) * 512 + o; \
        st##R = *reinterpret_cast<const uint4*>(src);                                              \
    }
#define QF_GLOAD(KB, IB0) \
    QF_GL1(0, KB, IB0) QF_GL1(1, KB, IB0) QF_GL1(2, KB, IB0) QF_GL1(3, KB, IB0) \
    QF_GL1(4, KB, IB0) QF_GL1(5, KB, IB0) QF_GL1(6, KB, IB0) QF_GL1(7, KB, IB0)
#define QF_SW1(R, BUF)                                                                             \
    {                                                                                              \
        const int q = (R) * 256 + threadIdx.x;                                                     \
        *reinterpret_cast<uint4*>((BUF) + (q >> 6) * 1024 + (q & 63) * 16) = st##R;                 \
    }
#define QF_SWRITE(BUF) \
    QF_SW1(0, BUF) QF_SW1(1, BUF) QF_SW1(2, BUF) QF_SW1(3, BUF) \
    QF_SW1(4, BUF) QF_SW1(5, BUF) QF_SW1(6, BUF) QF_SW1(7, BUF)
    uint4 st0, st1, st2, st3, st4, st5, st6, st7;
    for (int it = 0; it < J.npad; it += 128) {
        const int t0 = it >> 6;
        const int kb_lo = it >> 4;
        const int kb_hi = max(J.kend[t0], J.kend[t0 + 1]) >> 4;
        if (kb_hi <= kb_lo) continue;                 // block-uniform
        const int I0 = it + wi * 64;
        const int wk0 = I0 >> 4, wk1 = J.kend[t0 + wi] >> 4;
        const int ib0 = it >> 5;
        f32x16 accC[2][2], accS[2][2];
#pragma unroll
        for (int a2 = 0; a2 < 2; ++a2)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int r = 0; r < 16; ++r) { accC[a2][b][r] = 0.f; accS[a2][b][r] = 0.f; }
        QF_GLOAD(kb_lo, ib0)
        QF_SWRITE(lds[0])
        __syncthreads();
        for (int kb = kb_lo; kb < kb_hi; ++kb) {
            const int cur = (kb - kb_lo) & 1;
            const bool more = kb + 1 < kb_hi;
            if (more) { QF_GLOAD(kb + 1, ib0) }
            if (kb >= wk0 && kb < wk1) {
                const char* L = lds[cur];
#pragma unroll
                for (int a2 = 0; a2 < 2; ++a2) {
                    const int rb = wp * 2 + a2;
                    const bf16x8 zch = lds8(L, (0 * 4 + rb) * 1024 + loff);
                    const bf16x8 zcl = lds8(L, (1 * 4 + rb) * 1024 + loff);
                    const bf16x8 zsh = lds8(L, (2 * 4 + rb) * 1024 + loff);
                    const bf16x8 zsl = lds8(L, (3 * 4 + rb) * 1024 + loff);
#pragma unroll
                    for (int b = 0; b < 2; ++b) {
                        const int ib = wi * 2 + b;
                        const bf16x8 bch = lds8(L, 16384 + (0 * 4 + ib) * 1024 + loff);
                        const bf16x8 bcl = lds8(L, 16384 + (1 * 4 + ib) * 1024 + loff);
                        const bf16x8 bsh = lds8(L, 16384 + (2 * 4 + ib) * 1024 + loff);
                        const bf16x8 bsl = lds8(L, 16384 + (3 * 4 + ib) * 1024 + loff);
                        accC[a2][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(zch, bch, accC[a2][b], 0, 0, 0);
                        accC[a2][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(zch, bcl, accC[a2][b], 0, 0, 0);
                        accC[a2][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(zcl, bch, accC[a2][b], 0, 0, 0);
                        accS[a2][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(zsh, bsh, accS[a2][b], 0, 0, 0);
                        accS[a2][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(zsh, bsl, accS[a2][b], 0, 0, 0);
                        accS[a2][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(zsl, bsh, accS[a2][b], 0, 0, 0);
                    }
                }
            }
            if (more) { QF_SWRITE(lds[cur ^ 1]) }
            __syncthreads();
        }
        if (wk1 > wk0) {                              // wave-uniform
            // element (p, i) of a Z array: ((p >> 5) nkb + (i >> 4)) 512 + ((i >> 3) & 1) 256
            // + (p & 31) 8 + (i & 7); the lane's rows differ by compile-time offsets
            float v[32];
#pragma unroll
            for (int a2 = 0; a2 < 2; ++a2) {
                float sum[16];
#pragma unroll
                for (int r = 0; r < 16; ++r) sum[r] = 0.f;
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    const int i = I0 + b * 32 + cl32;
                    const uint16_t* zc = J.zc_hi + (((P0 >> 5) + a2) * nkb + (i >> 4)) * 512 + ((i >> 3) & 1) * 256
                                       + 4 * hl * 8 + (i & 7);
                    const uint16_t* zcl = zc + J.zn;
                    const uint16_t* zs = zc + 2 * J.zn;
                    const uint16_t* zsl = zc + 3 * J.zn;
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const int ro = (8 * (r >> 2) + (r & 3)) * 8;
                        const float c = bf16_val(zc[ro]) + bf16_val(zcl[ro]);
                        const float sv = bf16_val(zs[ro]) + bf16_val(zsl[ro]);
                        sum[r] += accC[a2][b][r] * (sint ? sv : c) + accS[a2][b][r] * (sint ? c : sv);
                        if ((r & 3) == 3) __asm__ volatile("" ::: "memory");
                    }
                }
#pragma unroll
                for (int r = 0; r < 16; ++r) v[a2 * 16 + r] = sum[r];
            }
            bfly<16>(v, lane);
            bfly<8>(v, lane);
            bfly<4>(v, lane);
            bfly<2>(v, lane);
            bfly<1>(v, lane);
            tot += v[0];
        }
    }
#undef QF_GL1
#undef QF_GLOAD
#undef QF_SW1
#undef QF_SWRITE
    // waves (wp, 0) and (wp, 1) hold the two i halves of the same points
    float* red = reinterpret_cast<float*>(lds[0]);
    __syncthreads();
    if (wi) red[wp * 64 + lane] = tot;
    __syncthreads();
    if (!wi) {
        const int k = cl32;
        const int64_t p = P0 + (k >> 4) * 32 + 8 * ((k & 15) >> 2) + 4 * hl + (k & 3);
        if (p < J.P) J.out[(int64_t)form * J.ostride + J.p0 + p] = tot + red[wp * 64 + lane];
    }
}

// Persistent launch over the tiles of many jobs (sea states): block b takes
// tiles b, b + gridDim.x, ... of the ordered tile list; every block exits
// after its last tile.
__global__ __launch_bounds__(256) void k_qf_forms_persist(const QfJob* __restrict__ jobs,
                                                          const QfTile* __restrict__ tiles, int ntiles) {
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const QfTile tl = tiles[t];
        __syncthreads();                              // the previous tile's LDS readers are done
        qf_tile(jobs[tl.job], tl.ptile, tl.form);
    }
}

// Form tables: one thread per (i, m) of the filtered pair set.  Depth-basis
// coefficients c_q(kappa) = sum_s einv[q][s] e^{-kappa tau_s} in FP64 (einv is
// ill-conditioned: its entries reach 1e8 while |c| <= 1.4), entries written
// as bf16 hi + lo into the tiled layout.
__global__ __launch_bounds__(256) void k_qf_build(QfBuild B) {
    const int i = blockIdx.y;
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B.n || m <= i || m >= B.fmhi[i]) return;
    const int64_t e = B.frow[i] - (i + 1) + m;
    const UvwRec<float> r = B.uvw[e];
    const XRec<float> xr = B.x[e];
    constexpr double kL2eInv = 0.69314718055994530942;   // ln 2: kappa = pk / log2 e
    const double kp = (double)r.pkP * kL2eInv, km = (double)r.pkM * kL2eInv;
    double eP[32], eM[32];
#pragma unroll
    for (int q = 0; q < 32; ++q) {
        eP[q] = q < B.R ? exp(-kp * B.tau[q]) : 0.0;
        eM[q] = q < B.R ? exp(-km * B.tau[q]) : 0.0;
    }
    const int nkb = B.npad >> 4;
    const size_t nn = (size_t)B.npad * B.npad;
    const size_t o = ((size_t)(i >> 5) * nkb + (m >> 4)) * 512 + ((m >> 3) & 1) * 256 + (i & 31) * 8 + (m & 7);
    auto put = [&](int f, int k, double v) {
        uint16_t* d = B.b + ((size_t)f * 4 + 2 * k) * nn;
        const float x = (float)v;
        const uint32_t hi = bf16_bits(x);
        d[o] = (uint16_t)hi;
        d[o + nn] = (uint16_t)bf16_bits(x - bf16_val(hi));
    };
    put(0, 0, (double)xr.xP + (double)xr.xM);
    put(0, 1, (double)xr.xM - (double)xr.xP);
    const int R = B.R;
    for (int q = 0; q < R; ++q) {
        double cp = 0, cm = 0;
#pragma unroll
        for (int s2 = 0; s2 < 32; ++s2) {
            if (s2 < R) {
                const double w = B.einv[q * R + s2];
                cp += w * eP[s2];
                cm += w * eM[s2];
            }
        }
        put(1 + q, 0, r.cuP * cp + r.cuM * cm);
        put(1 + q, 1, r.cuM * cm - r.cuP * cp);
        put(1 + R + q, 0, r.cvP * cp + r.cvM * cm);
        put(1 + R + q, 1, r.cvM * cm - r.cvP * cp);
        put(1 + 2 * R + q, 0, r.cwP * cp + r.cwM * cm);
        put(1 + 2 * R + q, 1, r.cwP * cp - r.cwM * cm);
    }
}

}  // namespace

hipError_t launch_qf_build(const QfBuild& B, hipStream_t st) {
    if (B.n <= 0 || B.R < 0 || B.R > 32 || (B.npad & 127) || B.npad < B.n) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_qf_build, dim3((unsigned)((B.npad + 255) / 256), (unsigned)B.n), dim3(256), 0, st, B);
    return hipGetLastError();
}

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

hipError_t launch_qf_persist(const QfJob* jobs, const QfTile* tiles, int ntiles, int blocks, hipStream_t st) {
    if (ntiles <= 0) return hipSuccess;
    if (blocks <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_qf_forms_persist, dim3((unsigned)blocks), dim3(256), 0, st, jobs, tiles, ntiles);
    return hipGetLastError();
}

hipError_t launch_qf_forms(const QfJob& J, hipStream_t st) {
    if (J.P <= 0 || J.nform <= 0) return hipSuccess;
    if (!qf_shape_ok(J)) return hipErrorInvalidValue;
    const char* v = std::getenv("CFDWM_QF_KERNEL");
    if (v && std::atoi(v) == 0) {
        hipLaunchKernelGGL(k_qf_forms, dim3((unsigned)(J.ppad / 128), (unsigned)J.nform), dim3(256), 0, st, J);
    } else {
        const int npt = (int)(J.ppad / 128);
        const int nfg = (J.nform + kQfGF - 1) / kQfGF, npg = (npt + kQfGP - 1) / kQfGP;
        const unsigned nblk = (unsigned)((int64_t)nfg * npg * kQfGF * kQfGP);
        hipLaunchKernelGGL(k_qf_forms_lds, dim3(nblk), dim3(256), 0, st, J, npt);
    }
    return hipGetLastError();
}

}  // namespace cfdwm
