// gemm_lowp.hip -- the three real products on 16-bit MFMA (v_mfma_f32_32x32x16_{bf16,f16}).
//
// Same block/wave decomposition, packed-fragment LDS ring and epilogue as the FP32 kernel
// (gemm_f32.hip); one stage is one 16-deep k fragment.
//
//  BF16 split (NP = 2, "bf16x3"): each FP32 plane value x = hi + lo + r with hi = bf16(x),
//    lo = bf16(x - hi) (the pack kernel does the split). Per product
//    P = a_hi*b_hi + a_hi*b_lo + a_lo*b_hi, three MFMAs into one FP32 accumulator; a_lo*b_lo
//    and the remainders r are dropped (about 2^-16 relative per operand).
//  BF16 split (NP = 3, "bf16x6"): x = hi + mid + lo; six MFMAs (hh, hm, mh, hl, lh, mm) keep
//    every term down to 2^-24 relative (mid*lo, lo*mid, lo*lo dropped).
//  FP16 block-scaled (F16S): each 128 (mn) x 128 (k) block of every plane was stored as
//    fp16(x * 2^-e) with its own exponent e, so the MFMA products of k block kb carry the
//    factor 2^(eA + eB). The accumulator holds the running sum in units of the current
//    block's factor: on entering a block whose factor differs, it is rescaled by the exact
//    power of two 2^(e_old - e_new) (ldexp), and at the end multiplied by the last factor.
//    With equal factors along k (the common case) no rescale is issued. Blocks of zeros
//    (kScaleNone) keep the running factor. Exact while the running sum stays in FP32 range:
//    a rescale can overflow only if two blocks along k differ by about 2^80 in magnitude,
//    and a non-finite output is flagged by the correction screen.
#include <type_traits>
#include "device_util.h"
#include "launch.h"

namespace c3m {


// FP16 separate-products kernels: register sets in flight (register form) and LDS ring depth.
// Three sets (two stages of loads ahead) measured about 1.5% faster than four, five or six at
// 8192^3, 4096^3 and 8192 x 8192 x 2048 (products kernel 2.85 against 2.88 ms at 8192^3), and
// 5% faster than two (2.90 against 3.05 ms at 8192^3; 0.390 against 0.400 ms at 4096^3).
#ifndef C3M_F16_NSET
#define C3M_F16_NSET 3
#endif
#ifndef C3M_F16_LDS_RING
#define C3M_F16_LDS_RING 6
#endif

// Factor exponent of one FP16 block product: eA + eB, or kScaleNone if either block is zero.
__device__ __forceinline__ int block_exp(int ea, int eb)
{
    return (ea == kScaleNone || eb == kScaleNone) ? kScaleNone : ea + eb;
}

// Entering a block with factor exponent e: move the running sum (units 2^cur) to units 2^e.
template <int TN, int TM>
__device__ __forceinline__ void enter_block(f32x16 (&acc)[TN][TM], int& cur, int e)
{
    if (e == kScaleNone) return;
    if (cur != kScaleNone && e != cur) {
        const int d = cur - e;
#pragma unroll
        for (int a = 0; a < TN; ++a)
#pragma unroll
            for (int b = 0; b < TM; ++b)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[a][b][r] = ldexpf(acc[a][b][r], d);
    }
    cur = e;
}

// End of the k range: running sum times its factor.
template <int TN, int TM>
__device__ __forceinline__ void finish_blocks(f32x16 (&acc)[TN][TM], int cur)
{
    if (cur == kScaleNone || cur == 0) return;
#pragma unroll
    for (int a = 0; a < TN; ++a)
#pragma unroll
        for (int b = 0; b < TM; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[a][b][r] = ldexpf(acc[a][b][r], cur);
}

template <int BM, int BN, int NP, int WM = 2, int WN = 2>
struct LowTile {
    static constexpr int THREADS = 64 * WM * WN;
    static constexpr int TM = BM / WM / 32, TN = BN / WN / 32;
    static constexpr int FA = BM / 32, FB = BN / 32;
    static constexpr int A16 = 3 * NP * FA * 64;  // 16-byte lane vectors per stage, A side
    static constexpr int B16 = 3 * NP * FB * 64;
    static constexpr int S16 = A16 + B16;
    static constexpr int LOADS = (S16 + THREADS - 1) / THREADS;
    static constexpr int TAIL = S16 - THREADS * (LOADS - 1);  // threads live in the last load
};

__device__ __forceinline__ f32x16 mfma16(u32x4 a, u32x4 b, f32x16 c, bool half)
{
    if (half)
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c,
                                                      0, 0, 0);
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                   0, 0, 0);
}

template <int BM, int BN, int NP, bool F16S, bool PARTIAL, int WM, int WN>
__global__ __launch_bounds__(64 * WM * WN, 1) void gemm3m_lowp_kernel(GemmArgs g)
{
    using T = LowTile<BM, BN, NP, WM, WN>;
    constexpr int NT = T::THREADS;
    __shared__ u32x4 lds[2 * T::S16];

    int part, tmI, tnI;
    map_tile(blockIdx.x, gridDim.x, g.tilesM, g.tilesN, part, tmI, tnI);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave % WM, wn = wave / WM;
    const int ks0 = (int)((long long)g.kSteps * part / g.parts);
    const int ks1 = (int)((long long)g.kSteps * (part + 1) / g.parts);

    const u32x4* Ap = reinterpret_cast<const u32x4*>(g.Ap);
    const u32x4* Bp = reinterpret_cast<const u32x4*>(g.Bp);
    const size_t aPlane = g.aPlaneBytes / 16, bPlane = g.bPlaneBytes / 16;
    const size_t aRow = (size_t)(g.mPad / 32) * 64;  // lane vectors per 16-deep k fragment
    const size_t bRow = (size_t)(g.nPad / 32) * 64;

    const u32x4* src[T::LOADS];
    size_t step[T::LOADS];
    // every load but the last is full; the last covers T::TAIL threads
    const bool tailLive = tid < T::TAIL;
#pragma unroll
    for (int u = 0; u < T::LOADS; ++u) {
        const int e = tid + NT * u;
        const bool live = u + 1 < T::LOADS || tailLive;
        if (e < T::A16) {
            const int row = e / (T::FA * 64);  // plane q * NP + piece
            const int w = e % (T::FA * 64);
            src[u] = Ap + row * aPlane + (size_t)tmI * T::FA * 64 + w;
            step[u] = aRow;
        } else {
            const int eb = live ? e - T::A16 : 0;
            const int row = eb / (T::FB * 64);
            const int w = eb % (T::FB * 64);
            src[u] = Bp + row * bPlane + (size_t)tnI * T::FB * 64 + w;
            step[u] = bRow;
        }
    }

    f32x16 acc[3][T::TN][T::TM];
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int a = 0; a < T::TN; ++a)
#pragma unroll
            for (int b = 0; b < T::TM; ++b)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[q][a][b][r] = 0.0f;
    // FP16: running factor exponent per product; the tile lies in one 128 x 128 scale block
    // of A (row block mb) and of B (nb)
    int eCur[3] = {kScaleNone, kScaleNone, kScaleNone};
    const int* scA = nullptr;
    const int* scB = nullptr;
    if (F16S) {
        scA = g.scaleA + (size_t)((tmI * BM) >> 7) * g.kBlocks128;
        scB = g.scaleB + (size_t)((tnI * BN) >> 7) * g.kBlocks128;
    }
    const size_t scStrideA = F16S ? (size_t)(g.mPad >> 7) * g.kBlocks128 : 0;  // per plane
    const size_t scStrideB = F16S ? (size_t)(g.nPad >> 7) * g.kBlocks128 : 0;

    // Same stage pipeline as the FP32 kernel: fragments of stage s to registers, staged
    // stage s+1 into the other LDS buffer, HBM loads of stage s+2, MFMAs, barrier.
    u32x4 gbuf[T::LOADS];
    if (ks0 < ks1) {
#pragma unroll
        for (int u = 0; u < T::LOADS; ++u)
            if (u + 1 < T::LOADS || tailLive) gbuf[u] = src[u][(size_t)ks0 * step[u]];
#pragma unroll
        for (int u = 0; u < T::LOADS; ++u)
            if (u + 1 < T::LOADS || tailLive) lds[tid + NT * u] = gbuf[u];
        if (ks0 + 1 < ks1) {
#pragma unroll
            for (int u = 0; u < T::LOADS; ++u)
                if (u + 1 < T::LOADS || tailLive) gbuf[u] = src[u][(size_t)(ks0 + 1) * step[u]];
        }
    }
    __syncthreads();

    for (int ks = ks0; ks < ks1; ++ks) {
        const int buf = (ks - ks0) & 1;
        const bool more = ks + 1 < ks1;
        const u32x4* L = lds + buf * T::S16;
        // (piece of B, piece of A) per term, small terms first:
        //   NP = 1: (0,0);  NP = 2: (1,0) (0,1) (0,0);
        //   NP = 3: (1,1) (2,0) (0,2) (1,0) (0,1) (0,0)  -- pieces 0 = hi, 1 = mid, 2 = lo
        constexpr int NTERM = NP == 1 ? 1 : (NP == 2 ? 3 : 6);
        constexpr int TB[6] = {NP == 2 ? 1 : (NP == 3 ? 1 : 0), NP == 2 ? 0 : 2, 0, 1, 0, 0};
        constexpr int TA[6] = {NP == 2 ? 0 : (NP == 3 ? 1 : 0), NP == 2 ? 1 : 0, NP == 2 ? 0 : 2, 0, 1, 0};
        // Fragments held at once: the whole stage (term-major MFMAs, 3*TN*TM independent
        // accumulators between two MFMAs into one accumulator), or one plane at a time when
        // the whole stage would not fit beside the accumulators (three pieces, eight waves).
        constexpr bool PER_PLANE = NP * 3 * (T::TM + T::TN) * 4 > 96;
        if (F16S && (ks == ks0 || (ks & 7) == 0)) {  // first stage of a 128-deep k block
            const int kb = ks >> 3;
#pragma unroll
            for (int q = 0; q < 3; ++q)
                enter_block(acc[q], eCur[q], block_exp(scA[q * scStrideA + kb], scB[q * scStrideB + kb]));
        }
        if (more) {
#pragma unroll
            for (int u = 0; u < T::LOADS; ++u)
                if (u + 1 < T::LOADS || tailLive) lds[(buf ^ 1) * T::S16 + tid + NT * u] = gbuf[u];
        }
        if (ks + 2 < ks1) {
#pragma unroll
            for (int u = 0; u < T::LOADS; ++u)
                if (u + 1 < T::LOADS || tailLive) gbuf[u] = src[u][(size_t)(ks + 2) * step[u]];
        }
        if constexpr (PER_PLANE) {
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                u32x4 af[NP][T::TM], bf[NP][T::TN];
#pragma unroll
                for (int p = 0; p < NP; ++p) {
#pragma unroll
                    for (int b = 0; b < T::TM; ++b) af[p][b] = L[((q * NP + p) * T::FA + wm * T::TM + b) * 64 + lane];
#pragma unroll
                    for (int a = 0; a < T::TN; ++a)
                        bf[p][a] = L[T::A16 + ((q * NP + p) * T::FB + wn * T::TN + a) * 64 + lane];
                }
#pragma unroll
                for (int t = 0; t < NTERM; ++t)
#pragma unroll
                    for (int a = 0; a < T::TN; ++a)
#pragma unroll
                        for (int b = 0; b < T::TM; ++b) {
                            acc[q][a][b] = mfma16(bf[TB[t]][a], af[TA[t]][b], acc[q][a][b], F16S);
                        }
            }
        } else {
            u32x4 af[3][NP][T::TM], bf[3][NP][T::TN];
#pragma unroll
            for (int q = 0; q < 3; ++q)
#pragma unroll
                for (int p = 0; p < NP; ++p) {
#pragma unroll
                    for (int b = 0; b < T::TM; ++b)
                        af[q][p][b] = L[((q * NP + p) * T::FA + wm * T::TM + b) * 64 + lane];
#pragma unroll
                    for (int a = 0; a < T::TN; ++a)
                        bf[q][p][a] = L[T::A16 + ((q * NP + p) * T::FB + wn * T::TN + a) * 64 + lane];
                }
#pragma unroll
            for (int t = 0; t < NTERM; ++t)
#pragma unroll
                for (int q = 0; q < 3; ++q)
#pragma unroll
                    for (int a = 0; a < T::TN; ++a)
#pragma unroll
                        for (int b = 0; b < T::TM; ++b) {
                            acc[q][a][b] = mfma16(bf[q][TB[t]][a], af[q][TA[t]][b], acc[q][a][b], F16S);
                        }
        }
        __syncthreads();
    }
    if (F16S)
#pragma unroll
        for (int q = 0; q < 3; ++q) finish_blocks(acc[q], eCur[q]);

    const int iBase = tmI * BM + wm * (BM / WM) + (lane & 31);
    const int jBase = tnI * BN + wn * (BN / WN) + 4 * (lane >> 5);
    store_tile<T::TN, T::TM, PARTIAL>(g, acc, part, iBase, jBase);
}

// FP16 separate products: is any of workgroup (q, tmI, tnI)'s two row blocks of A or two
// column blocks of B mixed (block exponents that differ along k)? Decided per workgroup, so
// all four waves take the same kernel.
__device__ __forceinline__ bool f16_tile_mixed(const GemmArgs& g, int q, int tmI, int tnI)
{
    const int* sa = g.sumA + q * (g.mPad >> 7) + 2 * tmI;
    const int* sb = g.sumB + q * (g.nPad >> 7) + 2 * tnI;
    return sa[0] == kScaleMixed || sa[1] == kScaleMixed || sb[0] == kScaleMixed || sb[1] == kScaleMixed;
}

// Writes one wave's 128 x (32 TN) outputs of product P_q starting at column block jb of the
// wave's 128 columns (FP16: times 2^e, e the final factor).
template <bool F16S, int TN>
__device__ __forceinline__ void store_products(const GemmArgs& g, f32x16 (&acc)[TN][4], int q, int tmI, int tnI,
                                               int wm, int wn, int jb, int lane, int e)
{
    const int sh = (F16S && e != kScaleNone) ? e : 0;  // ldexp by 0 is exact (identity)
    float* P = g.partial + (size_t)q * g.partStride;
    const int iBase = tmI * 256 + wm * 128 + (lane & 31);
    const int jBase = tnI * 256 + wn * 128 + jb * 32 + 4 * (lane >> 5);
#pragma unroll
    for (int a = 0; a < TN; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int i = iBase + b * 32;
                const int j = jBase + a * 32 + 8 * (r >> 2) + (r & 3);
                P[(size_t)j * g.mPad + i] = F16S ? ldexpf(acc[a][b][r], sh) : acc[a][b][r];
            }
}

// Separate-products form for the 16-bit formats. One workgroup computes one 256 x 256 tile of
// ONE real product P_q; its 2 x 2 waves each own 128 x 128 outputs (16 accumulators of 32 x 32,
// all 256 AGPRs). Holding one product per wave instead of three doubles the MFMAs each loaded
// fragment feeds (4 per term instead of 2), so the operand bytes per MFMA halve and the waves
// can load their fragments straight from the packed planes into registers: no LDS, no
// barriers, NSET register sets so the loads of stage s + NSET - 1 are in flight during the
// MFMAs of stage s. Per accumulator the MFMA order is the fused kernel's (stage by stage, the
// same term order), so each P_q is bit-identical to the fused kernel's product; the FP32
// planes P_q (ld = mPad) are combined and alpha/beta applied by merge_kernel.
//
// FP16 (F16S) runs as two launches. RESCALE = false takes the workgroups whose block exponents
// are uniform along k: one factor for the whole k range, applied at the store, and no
// rescale code in the loop (rescaling 256 AGPR-resident accumulators in the loop makes the
// compiler spill in every stage). RESCALE = true takes the mixed workgroups and rescales the
// running sum at block boundaries like the fused kernel; so that the rescale does not make
// the compiler spill, each wave covers its 128 columns in two passes of 64 (128 accumulator
// registers), reading its A fragments twice. Per accumulator the MFMA order is the same in
// every form, and with uniform exponents the rescale never fires, so the split between the
// two launches does not change any output bit.
template <int NP, int NSET, bool F16S, bool RESCALE>
__global__ __launch_bounds__(256, 1) void gemm3m_lowp_products_kernel(GemmArgs g)
{
    constexpr int TM = 4, TN = RESCALE ? 2 : 4, PASSES = 4 / TN;
    constexpr int NTERM = NP == 1 ? 1 : (NP == 2 ? 3 : 6);
    constexpr int TB[6] = {NP == 1 ? 0 : 1, NP == 2 ? 0 : 2, 0, 1, 0, 0};
    constexpr int TA[6] = {NP == 3 ? 1 : 0, NP == 2 ? 1 : 0, NP == 2 ? 0 : 2, 0, 1, 0};
    // plane = 3 sp + q: product q over K partition sp of g.parts (FP16: one partition)
    int plane, tmI, tnI;
    map_tile(blockIdx.x, gridDim.x, g.tilesM, g.tilesN, plane, tmI, tnI);
    const int q = plane % 3, sp = plane / 3;
    const int ks0 = (int)((long long)g.kSteps * sp / g.parts), ks1 = (int)((long long)g.kSteps * (sp + 1) / g.parts);
    if (F16S && f16_tile_mixed(g, q, tmI, tnI) != RESCALE) return;
    for (int pass = 0; pass < PASSES; ++pass) {
        const int lane = threadIdx.x & 63;
        const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        const int wm = wave & 1, wn = wave >> 1;
        const size_t aRowB = (size_t)(g.mPad / 32) * 1024;  // bytes per 16-deep k fragment row
        const size_t bRowB = (size_t)(g.nPad / 32) * 1024;
        const char* aW = reinterpret_cast<const char*>(g.Ap) + (size_t)(q * NP) * g.aPlaneBytes +
                         (size_t)(tmI * 8 + wm * TM) * 1024;
        const char* bW = reinterpret_cast<const char*>(g.Bp) + (size_t)(q * NP) * g.bPlaneBytes +
                         (size_t)(tnI * 8 + wn * 4 + pass * TN) * 1024;
        const unsigned laneB = (unsigned)lane * 16u;

        struct Set {
            u32x4 a[NP][TM], b[NP][TN];
        };
        auto load = [&](Set& S, int ks) {
#pragma unroll
            for (int p = 0; p < NP; ++p) {
                const char* a = aW + (size_t)p * g.aPlaneBytes + (size_t)ks * aRowB;
                const char* b = bW + (size_t)p * g.bPlaneBytes + (size_t)ks * bRowB;
#pragma unroll
                for (int i = 0; i < TM; ++i) S.a[p][i] = *reinterpret_cast<const u32x4*>(a + i * 1024 + laneB);
#pragma unroll
                for (int j = 0; j < TN; ++j) S.b[p][j] = *reinterpret_cast<const u32x4*>(b + j * 1024 + laneB);
            }
        };
        f32x16 acc[TN][TM];
#pragma unroll
        for (int a = 0; a < TN; ++a)
#pragma unroll
            for (int b = 0; b < TM; ++b)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[a][b][r] = 0.0f;
        // FP16: the wave's 128 x 128 outputs lie in row block mb = 2 tmI + wm of A and nb =
        // 2 tnI + wn of B. Uniform: one factor exponent from the summaries. Rescale: the running
        // exponent, with the next block's loaded one block ahead.
        const int mb = 2 * tmI + wm, nb = 2 * tnI + wn;
        const int* scA = nullptr;
        const int* scB = nullptr;
        int eCur = kScaleNone, eNext = kScaleNone;
        if (F16S && !RESCALE) eCur = block_exp(g.sumA[q * (g.mPad >> 7) + mb], g.sumB[q * (g.nPad >> 7) + nb]);
        if (F16S && RESCALE) {
            scA = g.scaleA + ((size_t)q * (g.mPad >> 7) + mb) * g.kBlocks128;
            scB = g.scaleB + ((size_t)q * (g.nPad >> 7) + nb) * g.kBlocks128;
            eNext = block_exp(scA[0], scB[0]);
        }
        // MFMAs of stage ks (after the block entry on the first stage of a 128-deep k block)
        auto mfma = [&](const Set& S, int ks) {
            if (F16S && RESCALE && (ks & 7) == 0) {
                enter_block(acc, eCur, eNext);
                const int kb = (ks >> 3) + 1;
                if (kb < g.kBlocks128) eNext = block_exp(scA[kb], scB[kb]);
            }
#pragma unroll
            for (int t = 0; t < NTERM; ++t)
#pragma unroll
                for (int a = 0; a < TN; ++a)
#pragma unroll
                    for (int b = 0; b < TM; ++b) acc[a][b] = mfma16(S.b[TB[t]][a], S.a[TA[t]][b], acc[a][b], F16S);
        };

        // Ring of NSET register sets: sets 0..NSET-2 preloaded; stage st computes from set
        // st % NSET after issuing the loads of stage st + NSET - 1 into the set freed last.
        // Stages ks0 .. ks1-1 of this K partition (absolute stage indices).
        const int nst = ks1 - ks0;
        Set S[NSET];
#pragma unroll
        for (int u = 0; u + 1 < NSET; ++u) load(S[u], ks0 + min(u, nst - 1));
        int st = 0;
        for (; st + NSET <= nst; st += NSET) {
#pragma unroll
            for (int u = 0; u < NSET; ++u) {
                // Unconditional (clamped to the last stage): a load skipped on one path would make
                // the wait before the next MFMAs cover the loads just issued.
                load(S[(u + NSET - 1) % NSET], ks0 + min(st + u + NSET - 1, nst - 1));
                // Keep the loads ahead of the MFMAs: the scheduler would otherwise sink each load
                // next to its first use and wait on it there.
                __builtin_amdgcn_sched_barrier(0);
                mfma(S[u], ks0 + st + u);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
#pragma unroll
        for (int u = 0; u + 1 < NSET; ++u)
            if (st + u < nst) mfma(S[u], ks0 + st + u);
        store_products<F16S, TN>(g, acc, plane, tmI, tnI, wm, wn, pass * TN, lane, eCur);
    }
}

// LDS-DMA form of the separate-products kernel. The workgroup's four waves load each stage's
// fragments ONCE per workgroup (instead of once per wave that needs them) with direct
// global->LDS loads (global_load_lds_dwordx4: 64 lanes x 16 B = one 1 KiB fragment per
// instruction, no register staging), into a ring of R stage slots, so R-1 stages are in flight
// while a stage computes. Per stage t: wait for this wave's loads of stage t+1 and for its own
// LDS reads, barrier (all waves' stage-t+1 data in LDS, all reads of stage t's slot done),
// issue the loads of stage t+R into stage t's slot, read stage t+1's fragments into the other
// register set, then the MFMAs of stage t. k is padded to a multiple of 16*U (planner), so the
// unrolled loop needs no tail. Same MFMA order per accumulator as the other forms.
// FP16: the uniform-exponent workgroups only (the register form's RESCALE launch takes the
// others; the decision is per workgroup, so no wave leaves a barrier early).
template <int NP, int R, bool F16S>
__global__ __launch_bounds__(256, 1) void gemm3m_lowp_products_lds_kernel(GemmArgs g)
{
    constexpr int TM = 4, TN = 4;
    constexpr int NTERM = NP == 1 ? 1 : (NP == 2 ? 3 : 6);
    constexpr int TB[6] = {NP == 1 ? 0 : 1, NP == 2 ? 0 : 2, 0, 1, 0, 0};
    constexpr int TA[6] = {NP == 3 ? 1 : 0, NP == 2 ? 1 : 0, NP == 2 ? 0 : 2, 0, 1, 0};
    constexpr int SLOT = 16 * NP * 64;                // u32x4 per stage: 8*NP A + 8*NP B fragments
    constexpr int U = (R % 2 == 0) ? R : 2 * R;       // unroll: slot and register set compile-time
    constexpr int PER = 4 * NP;                       // loads per wave per stage
    constexpr int VM = (R - 2) * PER;                 // loads left in flight when stage t+1 is needed
    constexpr int WAIT = (VM & 15) | ((VM >> 4) << 14) | (7 << 4);  // vmcnt(VM), expcnt max, lgkmcnt(0)
    static_assert(VM < 64, "vmcnt holds at most 63");
    __shared__ u32x4 lds[R * SLOT];

    int q, tmI, tnI;
    map_tile(blockIdx.x, gridDim.x, g.tilesM, g.tilesN, q, tmI, tnI);
    if (F16S && f16_tile_mixed(g, q, tmI, tnI)) return;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = wave & 1, wn = wave >> 1;
    const size_t aRowB = (size_t)(g.mPad / 32) * 1024;
    const size_t bRowB = (size_t)(g.nPad / 32) * 1024;
    const char* aBlk = reinterpret_cast<const char*>(g.Ap) + (size_t)(q * NP) * g.aPlaneBytes + (size_t)tmI * 8 * 1024;
    const char* bBlk = reinterpret_cast<const char*>(g.Bp) + (size_t)(q * NP) * g.bPlaneBytes + (size_t)tnI * 8 * 1024;
    const unsigned laneB = (unsigned)lane * 16u;
    const int nst = g.kSteps;

    // wave w loads mn fragments 2w and 2w+1 of A and of B, every piece
    auto issue = [&](int slot, int ks) {
        u32x4* L = lds + slot * SLOT;
#pragma unroll
        for (int p = 0; p < NP; ++p)
#pragma unroll
            for (int f = 0; f < 2; ++f) {
                const int fr = 2 * wave + f;
                __builtin_amdgcn_global_load_lds(
                    aBlk + (size_t)p * g.aPlaneBytes + (size_t)ks * aRowB + fr * 1024 + laneB,
                    (__attribute__((address_space(3))) void*)(L + (p * 8 + fr) * 64), 16, 0, 0);
                __builtin_amdgcn_global_load_lds(
                    bBlk + (size_t)p * g.bPlaneBytes + (size_t)ks * bRowB + fr * 1024 + laneB,
                    (__attribute__((address_space(3))) void*)(L + (8 * NP + p * 8 + fr) * 64), 16, 0, 0);
            }
    };
    struct Set {
        u32x4 a[NP][TM], b[NP][TN];
    };
    auto read = [&](Set& S, int slot) {
        const u32x4* L = lds + slot * SLOT;
#pragma unroll
        for (int p = 0; p < NP; ++p) {
#pragma unroll
            for (int i = 0; i < TM; ++i) S.a[p][i] = L[(p * 8 + wm * TM + i) * 64 + lane];
#pragma unroll
            for (int j = 0; j < TN; ++j) S.b[p][j] = L[(8 * NP + p * 8 + wn * TN + j) * 64 + lane];
        }
    };
    f32x16 acc[TN][TM];
#pragma unroll
    for (int a = 0; a < TN; ++a)
#pragma unroll
        for (int b = 0; b < TM; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[a][b][r] = 0.0f;
    const int e = F16S ? block_exp(g.sumA[q * (g.mPad >> 7) + 2 * tmI + wm], g.sumB[q * (g.nPad >> 7) + 2 * tnI + wn])
                       : kScaleNone;
    auto mfma = [&](const Set& S) {
#pragma unroll
        for (int t = 0; t < NTERM; ++t)
#pragma unroll
            for (int a = 0; a < TN; ++a)
#pragma unroll
                for (int b = 0; b < TM; ++b) acc[a][b] = mfma16(S.b[TB[t]][a], S.a[TA[t]][b], acc[a][b], F16S);
    };

    Set S[2];
#pragma unroll
    for (int u = 0; u < R; ++u) issue(u, min(u, nst - 1));
    __builtin_amdgcn_s_waitcnt((((R - 1) * PER) & 15) | ((((R - 1) * PER) >> 4) << 14) | (7 << 4) | (15 << 8));
    __builtin_amdgcn_s_barrier();
    read(S[0], 0);
    for (int s0 = 0; s0 < nst; s0 += U) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            __builtin_amdgcn_s_waitcnt(WAIT);
            __builtin_amdgcn_s_barrier();
            __builtin_amdgcn_sched_barrier(0);
            issue(u % R, min(s0 + u + R, nst - 1));
            read(S[(u + 1) & 1], (u + 1) % R);
            __builtin_amdgcn_sched_barrier(0);
            mfma(S[u & 1]);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    // no load may still be writing LDS when the workgroup ends
    __builtin_amdgcn_s_waitcnt(0);
    store_products<F16S, TN>(g, acc, q, tmI, tnI, wm, wn, 0, lane, e);
}

// BF16 products read from the stored operands (bf16x3, bf16x6; no pack launch), for problems
// whose outputs are too few to amortise a pack (tall-skinny shapes, small m and n). A
// workgroup of four waves computes a 128 x 128 block of all three real products over one K
// partition (stages kSteps s / parts .. kSteps (s+1) / parts), each wave a 64 x 64 block: it
// loads op(A)(i, k) and op(B)(k, j) as stored (buffer loads: uniform base, fixed 32-bit lane
// offset, small uniform offset), forms the plane values in registers (the pack's values,
// conjugation as a sign folded into an FMA), splits them into BF16 pieces with bf16_split8
// (the pack's split), and issues per accumulator the same MFMAs in the same order as the
// separate-products kernel on packed planes: stage by stage, the terms in TB/TA order. Its
// output planes (3 s + q, ld = mPad) are therefore the same bits, and merge_kernel combines
// them the same way. A lane holds k = 8 (lane / 32) + e (e = 0..7) of each 16-deep stage:
// a k-contiguous operand gives it 64 contiguous bytes (four 16-byte loads), an m- or
// n-contiguous one eight loads of one k, rows adjacent across lanes. Rows past m (columns past
// n) read the last one and only fill the planes' padding; the stage reaching past K is
// peeled after the ring, clamped and zero-masked. The operand layouts are compile-time
// variants of the k loop (a branch joining loads in flight would make the compiler wait for
// all of them). The host sends a problem here only if every lane offset fits in 31 bits.
template <int NP, int ID>
__global__ __launch_bounds__(256, 1) void gemm3m_bf16_raw_products_kernel(GemmArgs g)
{
    constexpr int TM = 2, TN = 2, NSET = 2;
    constexpr int NTERM = NP == 2 ? 3 : 6;
    constexpr int TB[6] = {1, NP == 2 ? 0 : 2, 0, 1, 0, 0};
    constexpr int TA[6] = {NP == 3 ? 1 : 0, NP == 2 ? 1 : 0, NP == 2 ? 0 : 2, 0, 1, 0};
    typedef float rawf2 __attribute__((ext_vector_type(2)));
    typedef float rawf4 __attribute__((ext_vector_type(4), aligned(8)));
    typedef __attribute__((address_space(1))) const rawf2 gf2;

    int sp, tmI, tnI;
    map_tile(blockIdx.x, gridDim.x, g.tilesM, g.tilesN, sp, tmI, tnI);
    const int lane = threadIdx.x & 63, half = lane >> 5;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int i0 = tmI * 128 + (wave & 1) * 64, j0 = tnI * 128 + (wave >> 1) * 64;
    // a 64 x 64 block wholly past m or n holds only padding, which nothing reads (and its row
    // or column base would lie past the operand): the wave has nothing to do
    if (i0 >= g.m || j0 >= g.n) return;
    const int ks0 = (int)((long long)g.kSteps * sp / g.parts), ks1 = (int)((long long)g.kSteps * (sp + 1) / g.parts);
    const int K = g.src.k;
    const bool aKc = g.src.opA != 111, bKc = g.src.opB == 111;
    const long long aStepB = aKc ? 8 : g.src.lda * 8, bStepB = bKc ? 8 : g.src.ldb * 8;
    const char* aTile = reinterpret_cast<const char*>(g.src.A) + (aKc ? (long long)i0 * g.src.lda * 8 : (long long)i0 * 8);
    const char* bTile = reinterpret_cast<const char*>(g.src.B) + (bKc ? (long long)j0 * g.src.ldb * 8 : (long long)j0 * 8);
    const float aSign = g.src.opA == 113 ? -1.0f : 1.0f, bSign = g.src.opB == 113 ? -1.0f : 1.0f;
    unsigned aOff[TM], bOff[TN];
#pragma unroll
    for (int b = 0; b < TM; ++b) {
        const long long d = min(i0 + 32 * b + (lane & 31), g.m - 1) - i0;
        aOff[b] = (unsigned)((aKc ? d * g.src.lda * 8 : d * 8) + 8 * half * aStepB);
    }
#pragma unroll
    for (int a = 0; a < TN; ++a) {
        const long long d = min(j0 + 32 * a + (lane & 31), g.n - 1) - j0;
        bOff[a] = (unsigned)((bKc ? d * g.src.ldb * 8 : d * 8) + 8 * half * bStepB);
    }
    auto rsrc = [](const char* base) {
        return __builtin_amdgcn_make_buffer_rsrc(const_cast<char*>(base), (short)0, (int)0xffffffff, 0x00020000);
    };
    auto bl2 = [](__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
        return __builtin_bit_cast(rawf2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
    };
    auto bl4 = [](__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
        return __builtin_bit_cast(rawf4, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
    };
    struct Set {
        rawf2 a[TM][8], b[TN][8];  // element e: k = 16 stage + 8 (lane / 32) + e
    };
    auto load = [&](Set& S, int ks, auto aTag, auto bTag) {
        constexpr bool AKC = decltype(aTag)::value, BKC = decltype(bTag)::value;
        const __amdgpu_buffer_rsrc_t ar = rsrc(aTile + (long long)ks * 16 * aStepB);
        const __amdgpu_buffer_rsrc_t br = rsrc(bTile + (long long)ks * 16 * bStepB);
#pragma unroll
        for (int b = 0; b < TM; ++b)
#pragma unroll
            for (int e = 0; e < 8; e += (AKC ? 2 : 1)) {
                if constexpr (AKC) {
                    const rawf4 w = bl4(ar, aOff[b], 8 * e);
                    S.a[b][e] = rawf2{w[0], w[1]};
                    S.a[b][e + 1] = rawf2{w[2], w[3]};
                } else {
                    S.a[b][e] = bl2(ar, aOff[b], (unsigned)(e * aStepB));
                }
            }
#pragma unroll
        for (int a = 0; a < TN; ++a)
#pragma unroll
            for (int e = 0; e < 8; e += (BKC ? 2 : 1)) {
                if constexpr (BKC) {
                    const rawf4 w = bl4(br, bOff[a], 8 * e);
                    S.b[a][e] = rawf2{w[0], w[1]};
                    S.b[a][e + 1] = rawf2{w[2], w[3]};
                } else {
                    S.b[a][e] = bl2(br, bOff[a], (unsigned)(e * bStepB));
                }
            }
    };
    // the stage reaching past K: k clamped to K - 1 per element (the lane's k offset 8 (lane /
    // 32) sits in the lane offsets; a clamped k, per lane, takes it out again)
    auto loadClamped = [&](Set& S, int ks) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const long long kr = min(ks * 16 + 8 * half + e, K - 1) - 8 * half;
            const char* ar = aTile + kr * aStepB;
            const char* br = bTile + kr * bStepB;
#pragma unroll
            for (int b = 0; b < TM; ++b) S.a[b][e] = *(gf2*)(ar + aOff[b]);
#pragma unroll
            for (int a = 0; a < TN; ++a) S.b[a][e] = *(gf2*)(br + bOff[a]);
        }
    };

    f32x16 acc[3][TN][TM];
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int a = 0; a < TN; ++a)
#pragma unroll
            for (int b = 0; b < TM; ++b)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[q][a][b][r] = 0.0f;

    // plane values with the conjugation sign folded in (s * im is exact, so fma(s, im, x)
    // rounds as x + s * im does): plane_value(role, ID, q, re, s * im)
    auto planeA = [&](int q, float re, float im) {
        if (ID == ID_STANDARD) return q == 0 ? re : (q == 1 ? aSign * im : fmaf(aSign, im, re));
        return q == 0 ? fmaf(aSign, im, re) : (q == 1 ? re : aSign * im);  // (Ar+Ai, Ar, Ai)
    };
    auto planeB = [&](int q, float re, float im) {
        if (ID == ID_STANDARD) return q == 0 ? re : (q == 1 ? bSign * im : fmaf(bSign, im, re));
        return q == 0 ? re : (q == 1 ? fmaf(bSign, im, -re) : fmaf(bSign, im, re));  // (Br, Bi-Br, Br+Bi)
    };
    auto mfma = [&](const Set& S, int ks, auto maskTag) {
        constexpr bool MASK = decltype(maskTag)::value;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            u32x4 ah[NP][TM], bh[NP][TN];
#pragma unroll
            for (int b = 0; b < TM; ++b) {
                float x[8];
                u32x4 w[NP];
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const bool kin = !MASK || ks * 16 + 8 * half + e < K;
                    x[e] = kin ? planeA(q, S.a[b][e].x, S.a[b][e].y) : 0.0f;
                }
                bf16_split8<NP>(x, w);
#pragma unroll
                for (int p = 0; p < NP; ++p) ah[p][b] = w[p];
            }
#pragma unroll
            for (int a = 0; a < TN; ++a) {
                float x[8];
                u32x4 w[NP];
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const bool kin = !MASK || ks * 16 + 8 * half + e < K;
                    x[e] = kin ? planeB(q, S.b[a][e].x, S.b[a][e].y) : 0.0f;
                }
                bf16_split8<NP>(x, w);
#pragma unroll
                for (int p = 0; p < NP; ++p) bh[p][a] = w[p];
            }
#pragma unroll
            for (int t = 0; t < NTERM; ++t)
#pragma unroll
                for (int a = 0; a < TN; ++a)
#pragma unroll
                    for (int b = 0; b < TM; ++b) acc[q][a][b] = mfma16(bh[TB[t]][a], ah[TA[t]][b], acc[q][a][b], false);
        }
    };
    typedef std::integral_constant<bool, false> Off;
    typedef std::integral_constant<bool, true> On;

    // full stages through the ring (NSET register sets, unconditional clamped look-ahead
    // loads, scheduling barriers keeping each stage's loads ahead of its MFMAs), then the
    // stages reaching past K, clamped and masked
    const int nst = ks1 - ks0;
    const int nFull = min(nst, max(0, K / 16 - ks0));
    auto ring = [&](auto aTag, auto bTag) {
        Set S[NSET];
        if (nFull > 0) {
#pragma unroll
            for (int u = 0; u + 1 < NSET; ++u) load(S[u], ks0 + min(u, nFull - 1), aTag, bTag);
        }
        int st = 0;
        for (; st + NSET <= nFull; st += NSET) {
#pragma unroll
            for (int u = 0; u < NSET; ++u) {
                load(S[(u + NSET - 1) % NSET], ks0 + min(st + u + NSET - 1, nFull - 1), aTag, bTag);
                __builtin_amdgcn_sched_barrier(0);
                mfma(S[u], ks0 + st + u, Off());
                __builtin_amdgcn_sched_barrier(0);
            }
        }
#pragma unroll
        for (int u = 0; u + 1 < NSET; ++u)
            if (st + u < nFull) mfma(S[u], ks0 + st + u, Off());
    };
    if (aKc) {
        if (bKc)
            ring(On(), On());
        else
            ring(On(), Off());
    } else {
        if (bKc)
            ring(Off(), On());
        else
            ring(Off(), Off());
    }
    for (int t = nFull; t < nst; ++t) {
        Set T;
        loadClamped(T, ks0 + t);
        mfma(T, ks0 + t, On());
    }

    // product q of partition sp into plane 3 sp + q (ld = mPad)
    const int iBase = i0 + (lane & 31), jBase = j0 + 4 * half;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        float* P = g.partial + (size_t)(3 * sp + q) * g.partStride;
#pragma unroll
        for (int a = 0; a < TN; ++a)
#pragma unroll
            for (int b = 0; b < TM; ++b)
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int i = iBase + b * 32;
                    const int j = jBase + a * 32 + 8 * (r >> 2) + (r & 3);
                    P[(size_t)j * g.mPad + i] = acc[q][a][b][r];
                }
    }
}

// g.tilesM / g.tilesN count 128 x 128 tiles, g.parts K partitions; g.partial as for the
// separate-products kernel (3 g.parts planes of mPad x nPad).
hipError_t launch_gemm_bf16_raw_products(const GemmArgs& g, int pieces, hipStream_t s)
{
    const int blocks = g.parts * g.tilesM * g.tilesN;
    const bool std3m = g.identity == ID_STANDARD;
    if (pieces == 2) {
        if (std3m)
            hipLaunchKernelGGL((gemm3m_bf16_raw_products_kernel<2, ID_STANDARD>), dim3(blocks), dim3(256), 0, s, g);
        else
            hipLaunchKernelGGL((gemm3m_bf16_raw_products_kernel<2, ID_FUSED>), dim3(blocks), dim3(256), 0, s, g);
    } else if (pieces == 3) {
        if (std3m)
            hipLaunchKernelGGL((gemm3m_bf16_raw_products_kernel<3, ID_STANDARD>), dim3(blocks), dim3(256), 0, s, g);
        else
            hipLaunchKernelGGL((gemm3m_bf16_raw_products_kernel<3, ID_FUSED>), dim3(blocks), dim3(256), 0, s, g);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// g.tilesM / g.tilesN count 256 x 256 tiles (mPad, nPad multiples of 256); g.partial holds
// 3 g.parts FP32 planes of mPad x nPad, g.partStride floats apart: plane 3 s + q is product q
// over the s-th of g.parts K partitions (stages kSteps s / parts .. kSteps (s+1) / parts).
// pieces: 2 bf16x3, 3 bf16x6, 1 FP16 block-scaled. The LDS form wants kSteps in whole groups
// of lowp_products_lds_unroll(pieces) stages.
int lowp_products_lds_unroll(int pieces)
{
    // the kernel unrolls R stages (R even) or 2R; ring depth R = 4 (bf16x3), 3 (bf16x6)
    const int r = pieces == 2 ? 4 : (pieces == 3 ? 3 : C3M_F16_LDS_RING);
    return r % 2 == 0 ? r : 2 * r;
}

hipError_t launch_gemm_lowp_products(const GemmArgs& g, int pieces, bool lds, hipStream_t s)
{
    // K partitions: the register form of the BF16 products only (FP16's block exponents and
    // the LDS form's unrolled stage groups assume the whole k range)
    if (g.parts < 1 || (g.parts > 1 && (lds || pieces == 1))) return hipErrorInvalidValue;
    const int blocks = 3 * g.parts * g.tilesM * g.tilesN;
    if (lds) {
        if (g.kSteps % lowp_products_lds_unroll(pieces)) return hipErrorInvalidValue;
        if (pieces == 2)
            hipLaunchKernelGGL((gemm3m_lowp_products_lds_kernel<2, 4, false>), dim3(blocks), dim3(256), 0, s, g);
        else if (pieces == 3)
            hipLaunchKernelGGL((gemm3m_lowp_products_lds_kernel<3, 3, false>), dim3(blocks), dim3(256), 0, s, g);
        else if (pieces == 1)
            hipLaunchKernelGGL((gemm3m_lowp_products_lds_kernel<1, C3M_F16_LDS_RING, true>), dim3(blocks), dim3(256),
                               0, s, g);
        else
            return hipErrorInvalidValue;
    } else {
        if (pieces == 2)
            hipLaunchKernelGGL((gemm3m_lowp_products_kernel<2, 3, false, false>), dim3(blocks), dim3(256), 0, s, g);
        else if (pieces == 3)
            hipLaunchKernelGGL((gemm3m_lowp_products_kernel<3, 2, false, false>), dim3(blocks), dim3(256), 0, s, g);
        else if (pieces == 1)
            hipLaunchKernelGGL((gemm3m_lowp_products_kernel<1, C3M_F16_NSET, true, false>), dim3(blocks), dim3(256),
                               0, s, g);
        else
            return hipErrorInvalidValue;
    }
    // FP16: the workgroups with mixed block exponents (most launches return at once)
    if (pieces == 1)
        hipLaunchKernelGGL((gemm3m_lowp_products_kernel<1, C3M_F16_NSET, true, true>), dim3(blocks), dim3(256), 0, s,
                           g);
    return hipGetLastError();
}

template <int BM, int BN, int NP, bool F16S, bool PARTIAL, int WM = 2, int WN = 2>
static hipError_t launch_one(const GemmArgs& g, hipStream_t s)
{
    const int blocks = g.tilesM * g.tilesN * g.parts;
    hipLaunchKernelGGL((gemm3m_lowp_kernel<BM, BN, NP, F16S, PARTIAL, WM, WN>), dim3(blocks), dim3(64 * WM * WN), 0,
                       s, g);
    return hipGetLastError();
}

template <int NP>
static hipError_t launch_bf16_np(const GemmArgs& g, int bm, int bn, int variant, hipStream_t s)
{
    const bool partial = g.parts > 1;
    if (bm == 128 && bn == 128) {
        if (variant == 1)
            return partial ? launch_one<128, 128, NP, false, true, 2, 4>(g, s)
                           : launch_one<128, 128, NP, false, false, 2, 4>(g, s);
        return partial ? launch_one<128, 128, NP, false, true>(g, s) : launch_one<128, 128, NP, false, false>(g, s);
    }
    if (bm == 64 && bn == 64)
        return partial ? launch_one<64, 64, NP, false, true>(g, s) : launch_one<64, 64, NP, false, false>(g, s);
    return hipErrorInvalidValue;
}

// variant 0: 4 waves per 128x128 tile; 1: 8 waves (two per SIMD).
hipError_t launch_gemm_bf16(const GemmArgs& g, int bm, int bn, int pieces, int variant, hipStream_t s)
{
    if (pieces == 2) return launch_bf16_np<2>(g, bm, bn, variant, s);
    if (pieces == 3) return launch_bf16_np<3>(g, bm, bn, variant, s);
    return hipErrorInvalidValue;
}

// variant 0: 4 waves per 128x128 tile; 1: 8 waves (two per SIMD).
hipError_t launch_gemm_f16(const GemmArgs& g, int bm, int bn, int variant, hipStream_t s)
{
    const bool partial = g.parts > 1;
    if (bm == 128 && bn == 128) {
        if (variant == 1)
            return partial ? launch_one<128, 128, 1, true, true, 2, 4>(g, s)
                           : launch_one<128, 128, 1, true, false, 2, 4>(g, s);
        return partial ? launch_one<128, 128, 1, true, true>(g, s) : launch_one<128, 128, 1, true, false>(g, s);
    }
    if (bm == 64 && bn == 64)
        return partial ? launch_one<64, 64, 1, true, true>(g, s) : launch_one<64, 64, 1, true, false>(g, s);
    return hipErrorInvalidValue;
}

}  // namespace c3m
