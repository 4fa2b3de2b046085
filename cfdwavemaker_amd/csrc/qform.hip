// Filtered second-order elevation eta2 as a quadratic form on the matrix cores.
//
// The pair sum of Irregular::eta2 (src/Irregular.cpp:223-293, pair-cache
// semantics of lsGridSpline's column update) is, per point p,
//
//   eta2_p = sum_i r_i [ c_pi sum_{m>i} Bcc_im c_pm + s_pi sum_{m>i} Bss_im s_pm ]
//
// with c_pi = cos(theta_i(p)), s_pi = sin(theta_i(p)), r_i the frequency ramp
// of component i at the point's time, Bcc = xP + xM and Bss = xM - xP built
// from the filtered pair coefficients (the real form of Re(Z^T B+ Z + Z^H B- Z)
// with Z_i = e^{i theta_i}).  The inner sums are GEMMs (points x components
// times components x components) on v_mfma_f32_32x32x16_bf16.  In the
// three-product mode every operand is split into bf16 hi + lo parts and the
// products hi*hi + hi*lo + lo*hi are accumulated in FP32, which keeps about 16
// significant bits per product; the one-product mode takes hi*hi alone (about
// 8 bits per product, a third of the matrix work and half the staged bytes).
// The outer sum over i is a fused epilogue (accumulator times c or s, summed
// in the lane, two exchanges at the end); the column kernel folds the result into
// its FP64 total.  The same kernel evaluates the z-independent forms of the
// velocity pair sums under the two profile depth bases (depth_basis.hpp): per
// skeleton wavenumber mu_q of the cosh-type basis one cos-type form each for
// u and v, per skeleton wavenumber of the sinh-type basis one sin-type form
// for w, with B entries cuP c_q(kP) +- cuM c_q(kM) (and cv, cw).
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
// bf16 element e (< 4) of four packed in a uint2, as float
__device__ __forceinline__ float bf16_half(uint2 v, int e) {
    const uint32_t w = e < 2 ? v.x : v.y;
    return __uint_as_float((e & 1) ? (w & 0xFFFF0000u) : (w << 16));
}

// ------------------------------------------------------------------ Z operands
// One thread per (point, 8 consecutive components): FP64 phase and range
// reduction, hardware sin/cos, bf16 hi/lo split, 16-byte stores.  A chunk
// holds at most 2^28 (point, component) entries (its Z arrays stay under
// 2 GiB), so thread, tile and s-grid column indices fit 32-bit arithmetic.
__global__ __launch_bounds__(256) void k_qf_zgen(QfJob J) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nkb = (uint32_t)J.npad >> 4;
    const uint32_t ntile = (uint32_t)(J.ppad >> 5) * nkb;
    const uint32_t tile = tid >> 6;
    if (tile >= ntile) return;
    const int h = (int)((tid >> 5) & 1), r = (int)(tid & 31);
    const uint32_t rb = tile / nkb;
    const int kb = (int)(tile - rb * nkb);
    const int64_t p = (int64_t)rb * 32 + r;     // point within the chunk
    const int m0 = kb * 16 + h * 8;
    uint32_t ch[8], cl[8], sh[8], sl[8];
    const bool live = p < J.P;
    double x = 0, y = 0, t = 0;
    if (live) {
        const int64_t q = J.p0 + p;
        if (J.grid) {
            // q < 4 ncol and col < nx ny: 32-bit
            const uint32_t ncol = (uint32_t)J.ncol, ny = (uint32_t)J.g.ny;
            const uint32_t slot = (uint32_t)q / ncol;
            const uint32_t cq = (uint32_t)q - slot * ncol;
            const uint32_t col = J.cols ? (uint32_t)J.cols[cq] : cq;
            const uint32_t ci = col / ny;
            x = J.g.x0 + J.g.dx * (double)ci;
            y = J.g.y0 + J.g.dy * (double)(col - ci * ny);
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
    const int64_t o = (int64_t)tile * 512 + h * 256 + r * 8;
    *reinterpret_cast<uint4*>(J.zc_hi + o) = pack(ch);
    *reinterpret_cast<uint4*>(J.zc_lo + o) = pack(cl);
    *reinterpret_cast<uint4*>(J.zs_hi + o) = pack(sh);
    *reinterpret_cast<uint4*>(J.zs_lo + o) = pack(sl);
}

// ----------------------------------------------------------- quadratic forms
// Form f of a launch:  Q_f(p) = sum_i e1_pi sum_m B1_f[i][m] c_pm + e2_pi sum_m B2_f[i][m] s_pm
// with (e1, e2) = (c, s) for cos-type forms and (s, c) for sin-type forms.
// One tile of work is (128 points, form): 4 waves, one per SIMD, walk the i
// range in tiles of 128; wave (w & 1, w >> 1) takes 64 points x 64 i as
// 2 x 2 MFMA 32 x 32 blocks for each of the two products, computed with
// the table operand first, so that each lane's accumulator holds one point
// and 16 values of i.  The k (= m)
// range of an i tile runs from its first row (pairs have m > i) to the end
// of its filtered rows.  Per k step (16 values of m) the block stages the A
// operands of its 128 points and the B operands of its 128 i (4 arrays x 4
// row blocks each, 32 KiB) in LDS, each operand tile read from L2 once per
// block.  The epilogue of an i tile multiplies the accumulators by c or s
// of the same (point, i), read as runs of four consecutive i (8 bytes), and
// sums over i in the lane; one exchange between the two half-waves and one
// between the two waves of a point range at the end write Q_f(p).  Single-job launches swizzle blocks in
// groups of 8 forms x 32 point tiles so that the Z and B tables a group
// reads stay in the last-level cache.
constexpr int kQfGF = 8, kQfGP = 32;

__device__ __forceinline__ bf16x8 lds8(const char* base, int off) {
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(base + off));
}

// A wave-uniform global pointer as scalar registers: the address arithmetic
// before it stays scalar and is not re-associated with per-lane offsets
// added after it, so loads take the scalar-base + vector-offset form.
typedef const uint16_t __attribute__((address_space(1)))* gu16p;
__device__ __forceinline__ gu16p uni_ptr(const uint16_t* p) {
    const uint64_t a = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
    return reinterpret_cast<gu16p>(((uint64_t)hi << 32) | lo);
}

// Operands move global -> LDS directly (global_load_lds_dwordx4, no
// registers) through four LDS stages, three steps ahead of the products.
// NLV is the number of bf16 products per term: 3 (hi*hi + lo*hi + hi*lo of
// table and Z operands, about 16 significant bits) or 1 (hi*hi, about 8
// significant bits; only the hi arrays are staged).  A stage holds NA = 4
// (NLV 3) or 2 (NLV 1) operand arrays for the block's 128 points and 128 i,
// 8 NA KiB.  Step j of an i tile:
//   1. s_waitcnt vmcnt(2 NA) lgkmcnt(0) + s_barrier: the requests of step
//      j + 1 (issued during step j - 2) have landed for all waves, and all
//      waves' LDS reads of step j - 1 are done;
//   2. the 8 NLV products of step j (operand set X) are issued interleaved
//      with the 4 NA LDS reads of step j + 1's operands (set Y, stage
//      (j + 1) % 4) and the 2 NA requests of step j + 3 (stage (j + 3) % 4 =
//      (j - 1) % 4, last read during step j - 2, before step j - 1's
//      barrier), so the matrix pipe stays busy while they issue.
// The barrier is an asm statement with a memory clobber after an explicit
// s_waitcnt: __syncthreads() would also wait for the requests of steps
// j + 2 and j + 3.  Every wave runs every step of the block's k range: table
// entries below the pair triangle (m <= i) and past a row tile's filtered end
// are zero, so the extra products add exact zeros, and the step body has no
// branch.
template <int NLV>
__device__ __forceinline__ void qf_tile_il(const QfJob& J, int ptile, int form) {
    static_assert(NLV == 1 || NLV == 3, "one or three bf16 products per term");
    constexpr int NB = 4;                         // stages
    constexpr int NA = NLV == 1 ? 2 : 4;          // operand arrays per side in a stage
    constexpr int SB = 8192 * NA;                 // stage bytes
    constexpr int NF = 4 * NA;                    // operand fragments per set
    constexpr int NG = 2 * NA;                    // request chunks per step = issue groups
    constexpr int PG = 8 * NLV / NG;              // products per issue group (3 or 2)
    __shared__ __attribute__((aligned(16))) char lds[NB][SB];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wp = w & 1, wi = w >> 1;
    const int nkb = J.npad >> 4;
    const size_t nn = (size_t)J.npad * J.npad;
    // tables: 4 arrays per form (B1 hi, B1 lo, B2 hi, B2 lo) or 2 (B1 hi, B2 hi)
    const uint16_t* __restrict__ Bf = J.b + (size_t)form * J.nbarr * nn;
    const int bhs = J.nbarr == 4 ? 2 : 1;         // arrays between B1 hi and B2 hi
    const bool sint = J.ftype[form] != 0;
    const int64_t PB = (int64_t)ptile * 128;
    const int64_t rbB = PB >> 5;
    const int64_t P0 = PB + wp * 64;
    const int hl = lane >> 5, cl32 = lane & 31;
    const int loff = ((lane >> 5) * 256 + (lane & 31) * 8) * 2;
    float tp[2] = {0.f, 0.f};                     // per lane: points P0 + cl32 and P0 + 32 + cl32
    const int wv = __builtin_amdgcn_readfirstlane(w);
    const int lo8 = lane * 8;
    typedef __attribute__((address_space(3))) void* lds_vp;
    typedef const __attribute__((address_space(1))) void* glb_vp;
#define QF4_SYNC(VM)                                                                 \
    __builtin_amdgcn_s_waitcnt(((VM) & 15) | (7 << 4) | (((VM) >> 4) << 14));        \
    __asm__ volatile("s_barrier" ::: "memory")
    // chunk r (< 2 NA) of a stage: A (array r, row block rbB + wv) for r < NA,
    // B (array r - NA, i row block ib0 + wv) for r >= NA; gp[r] is its address
    // at k step 0 of the current i tile (wave-uniform)
    const uint16_t* gp[NG];
    auto dma = [&](int r, int kb, int st) {
        __builtin_amdgcn_global_load_lds((glb_vp)(uni_ptr(gp[r] + kb * 512) + lo8),
                                         (lds_vp)(lds[0] + st * SB + (r * 4 + wv) * 1024), 16, 0, 0);
    };
    // operand fragment f of a stage: f < 2 NA A (a2 = f / NA, array f % NA),
    // else B (b = (f - 2 NA) / NA, array f % NA)
    auto frag = [&](int st, int f) {
        const char* L = lds[0] + st * SB;
        return f < 2 * NA ? lds8(L, ((f % NA) * 4 + wp * 2 + f / NA) * 1024 + loff)
                          : lds8(L, NA * 4096 + ((f % NA) * 4 + wi * 2 + (f - 2 * NA) / NA) * 1024 + loff);
    };
    bf16x8 opA[NF], opB[NF];
    for (int it = 0; it < J.npad; it += 128) {
        const int t0 = it >> 6;
        const int kb_lo = it >> 4;
        const int kb_hi = __builtin_amdgcn_readfirstlane(max(J.kend[t0], J.kend[t0 + 1]) >> 4);
        if (kb_hi <= kb_lo) continue;                 // block-uniform
        const int I0 = it + wi * 64;
        const int wk0 = I0 >> 4, wk1 = J.kend[t0 + wi] >> 4;
        const int ib0 = it >> 5;
        f32x16 acc[8];                                // [(a2 * 2 + b) * 2 + (0 cos part, 1 sin part)]
#pragma unroll
        for (int q = 0; q < 8; ++q)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[q][r] = 0.f;
        // product s (< 8 NLV) of a step: level s / 8 (hi x hi, lo x hi, hi x
        // lo of table x Z), accumulator s % 8; an accumulator's products are
        // 8 issues apart
        auto mma = [&](const bf16x8 (&op)[NF], int s) {
            const int lv = s >> 3, q = s & 7, cs = q & 1, a2 = q >> 2, b = (q >> 1) & 1;
            const int fa = NLV == 1 ? a2 * NA + cs : a2 * 4 + 2 * cs + (lv == 2 ? 1 : 0);
            const int fb = NLV == 1 ? 2 * NA + b * NA + cs : 8 + b * 4 + 2 * cs + (lv == 1 ? 1 : 0);
            acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(op[fb], op[fa], acc[q], 0, 0, 0);
        };
        // one step: products of operand set x, reads of step j + 1 into set y
        // from stage s1, requests of step kbd into stage s3; NG chunks of PG
        // products, 2 reads and 1 request in a fixed issue order
        auto step = [&](const bf16x8 (&x)[NF], bf16x8 (&y)[NF], int s1, int s3, int kbd) {
#pragma unroll
            for (int g = 0; g < NG; ++g) {
                mma(x, PG * g);
                y[2 * g] = frag(s1, 2 * g);
                mma(x, PG * g + 1);
                y[2 * g + 1] = frag(s1, 2 * g + 1);
                dma(g, kbd, s3);
                if (PG > 2) mma(x, PG * g + 2);
                __builtin_amdgcn_sched_barrier(0);
            }
        };
        // global array of chunk r: A arrays c_hi, c_lo, s_hi, s_lo (NLV 3) or
        // c_hi, s_hi (NLV 1); B arrays likewise
#pragma unroll
        for (int r = 0; r < NG; ++r) {
            const int a = r % NA;
            const int ga = NLV == 1 ? 2 * a : a;                       // Z array index
            const int gb = NLV == 1 ? a * bhs : a;                     // table array index
            gp[r] = r < NA ? J.zc_hi + (int64_t)ga * J.zn + (rbB + wv) * nkb * 512
                           : Bf + (size_t)gb * nn + (int64_t)(ib0 + wv) * nkb * 512;
        }
        // requests past the last step re-read step kb_hi - 1 (in bounds, unused)
        const int kl = kb_hi - 1;
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int r = 0; r < NG; ++r) dma(r, min(kb_lo + d, kl), d);
        QF4_SYNC(2 * NG);
#pragma unroll
        for (int f = 0; f < NF; ++f) opA[f] = frag(0, f);
        int s3 = 3, s1 = 1;                           // stages of steps j + 3 and j + 1
        for (int kb = kb_lo; kb < kb_hi; kb += 2) {
            QF4_SYNC(NG);
            step(opA, opB, s1, s3, min(kb + 3, kl));
            if (kb + 1 >= kb_hi) break;               // block-uniform
            s3 = (s3 + 1) & 3;
            s1 = (s1 + 1) & 3;
            QF4_SYNC(NG);
            step(opB, opA, s1, s3, min(kb + 4, kl));
            s3 = (s3 + 1) & 3;
            s1 = (s1 + 1) & 3;
        }
        // every request has landed and every wave is past its reads before
        // the next i tile's requests reuse stages 0 .. 2
        QF4_SYNC(0);
        if (wk1 > wk0) {                              // wave-uniform
            // accumulator element r of block (a2, b): point P0 + a2 * 32 + cl32,
            // i = I0 + b * 32 + 8 j + 4 hl + e (j = r / 4, e = r % 4); the four
            // i of one j are consecutive k of one half-tile row (8 bytes)
#pragma unroll
            for (int a2 = 0; a2 < 2; ++a2) {
                const uint16_t* zrow = J.zc_hi + ((P0 >> 5) + a2) * nkb * 512 + cl32 * 8 + 4 * hl;
                float s = 0.f;
#pragma unroll
                for (int b = 0; b < 2; ++b) {
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const uint16_t* zp = zrow + ((I0 >> 4) + 2 * b + (j >> 1)) * 512 + (j & 1) * 256;
                        const uint2 ch = *reinterpret_cast<const uint2*>(zp);
                        const uint2 cl = *reinterpret_cast<const uint2*>(zp + J.zn);
                        const uint2 sh = *reinterpret_cast<const uint2*>(zp + 2 * J.zn);
                        const uint2 sl = *reinterpret_cast<const uint2*>(zp + 3 * J.zn);
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            const int r = 4 * j + e;
                            const float c = bf16_half(ch, e) + bf16_half(cl, e);
                            const float sv = bf16_half(sh, e) + bf16_half(sl, e);
                            s += acc[(a2 * 2 + b) * 2][r] * (sint ? sv : c) + acc[(a2 * 2 + b) * 2 + 1][r] * (sint ? c : sv);
                        }
                    }
                }
                tp[a2] += s;
            }
        }
    }
#undef QF4_SYNC
    // lane l ends with point P0 + l: its own half's i sums plus the other half's
    const float t0 = tp[0] + __shfl_xor(tp[0], 32);
    const float t1 = tp[1] + __shfl_xor(tp[1], 32);
    const float tot = hl ? t1 : t0;
    float* red = reinterpret_cast<float*>(lds[0]);
    __syncthreads();
    if (wi) red[wp * 64 + lane] = tot;
    __syncthreads();
    if (!wi) {
        const int64_t p = P0 + lane;
        if (p < J.P) J.out[(int64_t)form * J.ostride + J.p0 + p] = tot + red[wp * 64 + lane];
    }
}

// Single-job launch: block -> (point tile, form) through the swizzle.
// One product per term stages 64 KiB of LDS per block: with the accumulators
// in VGPRs (no AGPRs) two blocks share a CU, one computing while the other
// waits at its barrier.  Three products stage 128 KiB: one block per CU.
template <int NLV>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NLV == 1 ? 2 : 1, 2)))
void k_qf_forms_il(QfJob J, int npt) {
    const int id = blockIdx.x;
    const int per = kQfGF * kQfGP;
    const int grp = id / per, in = id - grp * per;
    const int nfg = (J.nform + kQfGF - 1) / kQfGF;
    const int pg = grp / nfg, fg = grp - pg * nfg;
    const int ptile = pg * kQfGP + in % kQfGP;
    const int form = fg * kQfGF + in / kQfGP;
    if (ptile >= npt || form >= J.nform) return;      // whole block, before any barrier
    if (!J.clk) {
        qf_tile_il<NLV>(J, ptile, form);
        return;
    }
    const unsigned long long c0 = clock64(), w0 = wall_clock64();
    qf_tile_il<NLV>(J, ptile, form);
    if (threadIdx.x == 0) {
        atomicAdd(J.clk, clock64() - c0);
        atomicAdd(J.clk + 1, wall_clock64() - w0);
    }
}

// Persistent launch over the tiles of many jobs (sea states): block b takes
// tiles b, b + gridDim.x, ... of the ordered tile list; every block exits
// after its last tile.
template <int NLV>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NLV == 1 ? 2 : 1, 2)))
void k_qf_forms_persist(const QfJob* __restrict__ jobs,
                                                          const QfTile* __restrict__ tiles, int ntiles) {
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const QfTile tl = tiles[t];
        const QfJob J = jobs[tl.job];
        __syncthreads();                              // the previous tile's LDS readers are done
        qf_tile_il<NLV>(J, tl.ptile, tl.form);
    }
}

// Form tables: one thread per (i, m) of the filtered pair set.  Profile-basis
// coefficients c_q(kappa) = sum_s finv[q][s] P_kappa(z_s) in FP64 (finv is
// ill-conditioned: its entries are large while |c| stays O(1)), entries
// written as bf16 hi + lo into the tiled layout.
__global__ __launch_bounds__(256) void k_qf_build(QfBuild B) {
    const int i = blockIdx.y;
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B.n || m <= i || m >= B.fmhi[i]) return;
    const int64_t e = B.frow[i] - (i + 1) + m;
    const UvwRec<float> r = B.uvw[e];
    const XRec<float> xr = B.x[e];
    constexpr double kL2eInv = 0.69314718055994530942;   // ln 2: kappa = pk / log2 e
    const double kp = (double)r.pkP * kL2eInv, km = (double)r.pkM * kL2eInv;
    const int nkb = B.npad >> 4;
    const size_t nn = (size_t)B.npad * B.npad;
    const size_t o = ((size_t)(i >> 5) * nkb + (m >> 4)) * 512 + ((m >> 3) & 1) * 256 + (i & 31) * 8 + (m & 7);
    auto put = [&](int f, int k, double v) {
        const bool sep = B.beta && f == 0;            // the eta2 form in its own hi/lo table
        const bool lo = sep || B.nbarr == 4;
        const int fi = B.beta ? f - 1 : f;
        uint16_t* d = sep ? B.beta + (size_t)(2 * k) * nn : B.b + ((size_t)fi * B.nbarr + (lo ? 2 * k : k)) * nn;
        const float x = (float)v;
        const uint32_t hi = bf16_bits(x);
        d[o] = (uint16_t)hi;
        if (lo) d[o + nn] = (uint16_t)bf16_bits(x - bf16_val(hi));
    };
    put(0, 0, (double)xr.xP + (double)xr.xM);
    put(0, 1, (double)xr.xM - (double)xr.xP);
    // profile P_kappa(z) = e^{kappa z} + sg e^{-kappa (z + 2h)} at the
    // skeleton depths, then its basis coefficients c = finv P (one family at a time)
    double eP[32], eM[32];
    auto family = [&](int R, const double* zs, const double* finv, double sg, int f0, int nf) {
#pragma unroll
        for (int q = 0; q < 32; ++q) {
            const double z = q < R ? zs[q] : 0.0;
            eP[q] = q < R ? exp(kp * z) + sg * exp(-kp * (z + 2.0 * B.h)) : 0.0;
            eM[q] = q < R ? exp(km * z) + sg * exp(-km * (z + 2.0 * B.h)) : 0.0;
        }
        for (int q = 0; q < R; ++q) {
            double cp = 0, cm = 0;
#pragma unroll
            for (int s2 = 0; s2 < 32; ++s2) {
                if (s2 < R) {
                    const double w = finv[q * R + s2];
                    cp += w * eP[s2];
                    cm += w * eM[s2];
                }
            }
            if (nf == 2) {                            // u, v (cos-type)
                put(f0 + q, 0, r.cuP * cp + r.cuM * cm);
                put(f0 + q, 1, r.cuM * cm - r.cuP * cp);
                put(f0 + R + q, 0, r.cvP * cp + r.cvM * cm);
                put(f0 + R + q, 1, r.cvM * cm - r.cvP * cp);
            } else {                                  // w (sin-type)
                put(f0 + q, 0, r.cwP * cp + r.cwM * cm);
                put(f0 + q, 1, r.cwP * cp - r.cwM * cm);
            }
        }
    };
    family(B.R, B.zs, B.finv, 1.0, 1, 2);
    family(B.Rw, B.zsw, B.finvw, -1.0, 1 + 2 * B.R, 1);
}

}  // namespace

hipError_t launch_qf_build(const QfBuild& B, hipStream_t st) {
    if (B.n <= 0 || B.R < 0 || B.R > 32 || B.Rw < 0 || B.Rw > 32 || B.nform != 1 + 2 * B.R + B.Rw ||
        (B.npad & 127) || B.npad < B.n || (B.nbarr != 2 && B.nbarr != 4))
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_qf_build, dim3((unsigned)((B.npad + 255) / 256), (unsigned)B.n), dim3(256), 0, st, B);
    return hipGetLastError();
}

static bool qf_shape_ok(const QfJob& J) {
    if ((J.nlv != 1 && J.nlv != 3) || (J.nbarr != 2 && J.nbarr != 4) || (J.nlv == 3 && J.nbarr != 4)) return false;
    return !((J.ppad & 127) || (J.npad & 127) || J.ppad < J.P || J.npad < J.n);
}

hipError_t launch_qf_zgen(const QfJob& J, hipStream_t st) {
    if (J.P <= 0) return hipSuccess;
    if (!qf_shape_ok(J)) return hipErrorInvalidValue;
    const int64_t nthr = (J.ppad >> 5) * (J.npad >> 4) * 64;
    // the kernel's 32-bit indices: at most 2^28 Z entries per chunk, and an
    // s-grid chunk's points within 4 slots of ncol columns
    if ((int64_t)J.ppad * J.npad > ((int64_t)1 << 28) || (J.grid && J.p0 + J.P > ((int64_t)1 << 31))) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_qf_zgen, dim3((unsigned)((nthr + 255) / 256)), dim3(256), 0, st, J);
    return hipGetLastError();
}

hipError_t launch_qf_persist(const QfJob* jobs, const QfTile* tiles, int ntiles, int blocks, int nlv, hipStream_t st) {
    if (ntiles <= 0) return hipSuccess;
    if (blocks <= 0 || (nlv != 1 && nlv != 3)) return hipErrorInvalidValue;
    if (nlv == 1)
        hipLaunchKernelGGL(k_qf_forms_persist<1>, dim3((unsigned)blocks), dim3(256), 0, st, jobs, tiles, ntiles);
    else
        hipLaunchKernelGGL(k_qf_forms_persist<3>, dim3((unsigned)blocks), dim3(256), 0, st, jobs, tiles, ntiles);
    return hipGetLastError();
}

hipError_t launch_qf_forms(const QfJob& J, hipStream_t st) {
    if (J.P <= 0 || J.nform <= 0) return hipSuccess;
    if (!qf_shape_ok(J)) return hipErrorInvalidValue;
    const int npt = (int)(J.ppad / 128);
    const int nfg = (J.nform + kQfGF - 1) / kQfGF, npg = (npt + kQfGP - 1) / kQfGP;
    const unsigned nblk = (unsigned)((int64_t)nfg * npg * kQfGF * kQfGP);
    if (J.nlv == 1)
        hipLaunchKernelGGL(k_qf_forms_il<1>, dim3(nblk), dim3(256), 0, st, J, npt);
    else
        hipLaunchKernelGGL(k_qf_forms_il<3>, dim3(nblk), dim3(256), 0, st, J, npt);
    return hipGetLastError();
}

}  // namespace cfdwm
