// gemm_f32.hip -- the three real products on FP32 MFMA (v_mfma_f32_32x32x2_f32), gfx950.
//
// One workgroup = WM x WN waves = one BM x BN output tile; each wave owns a (BM/WM) x (BN/WN)
// block and keeps all THREE product accumulators for it in registers (3 x TM x TN 32x32
// FP32 tiles: 192 registers with 4 waves at 128x128, 96 with 8 waves, two per SIMD). Per k stage (BK = 16) the block copies the
// packed fragments of the 3 A planes and 3 B planes (contiguous runs, see kernel_args.h)
// HBM -> VGPR -> LDS with a double-buffered LDS ring, register staging one stage ahead and
// one barrier per stage; each wave reads its fragments with one 16-byte LDS read per lane
// and issues the MFMAs.
//
// The MFMA "A" operand is the B fragment and the "B" operand the A fragment, so the
// 32x32 result tile has output rows i along the lanes: every epilogue store writes 32
// consecutive rows of a column-major C column (256 contiguous bytes).
//
// Reference roles: products kernel K2 (ops-blas/blas/gemm/arch35/cgemm_compute_kernel.cpp:
// 284-366) and fused three-product tile K3 (:375-895). PARTIAL=false is the fused-tile
// path: products combined, alpha/beta applied and C written in registers, with the
// rounding-correction screen. PARTIAL=true is the K-split path: partition p of the k range
// writes its (re, im) partial planes; merge_kernel sums partitions in index order.
#include <type_traits>
#include "device_util.h"
#include "launch.h"

namespace c3m {

template <int BM, int BN, int WM = 2, int WN = 2>
struct F32Tile {
    static constexpr int THREADS = 64 * WM * WN;        // WM x WN waves
    static constexpr int TM = BM / WM / 32, TN = BN / WN / 32;  // 32x32 MFMA tiles per wave
    static constexpr int FA = BM / 32, FB = BN / 32;  // fragments per plane per k fragment
    static constexpr int KF = 2;                      // k fragments (8 deep) per stage
    static constexpr int A4 = 3 * KF * FA * 64;       // float4 per stage, A side
    static constexpr int B4 = 3 * KF * FB * 64;
    static constexpr int S4 = A4 + B4;
    static constexpr int LOADS = S4 / THREADS;        // float4 loads per thread per stage
    static_assert(A4 % THREADS == 0 && B4 % THREADS == 0, "stage must split evenly over the threads");
    static_assert(TM >= 1 && TN >= 1, "wave tile must hold a 32x32 MFMA tile");
};

// One output tile (tmI, tnI) over K partition `part`, start to finish; lds holds
// 2 * F32Tile<BM, BN, WM, WN>::S4 float4. Used by the one-problem LDS kernels.
template <int BM, int BN, bool PARTIAL, int WM = 2, int WN = 2>
__device__ __forceinline__ void f32_tile(const GemmArgs& g, int part, int tmI, int tnI, f32x4* lds)
{
    using T = F32Tile<BM, BN, WM, WN>;
    constexpr int NT = T::THREADS;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave % WM, wn = wave / WM;
    const int ks0 = (int)((long long)g.kSteps * part / g.parts);
    const int ks1 = (int)((long long)g.kSteps * (part + 1) / g.parts);

    const f32x4* A4p = reinterpret_cast<const f32x4*>(g.Ap);
    const f32x4* B4p = reinterpret_cast<const f32x4*>(g.Bp);
    const size_t aPlane4 = g.aPlaneBytes / 16, bPlane4 = g.bPlaneBytes / 16;
    const size_t aRow4 = (size_t)(g.mPad / 32) * 64;  // float4 per k fragment, all of A's mn
    const size_t bRow4 = (size_t)(g.nPad / 32) * 64;

    // Source offset of each of this thread's loads at stage 0, and the per-stage stride.
    const f32x4* src[T::LOADS];
#pragma unroll
    for (int u = 0; u < T::LOADS; ++u) {
        const int e = tid + NT * u;
        if (u < T::A4 / NT) {
            const int row = e / (T::FA * 64);  // (q, kfl) pair
            const int w = e % (T::FA * 64);
            const int q = row / T::KF, kfl = row % T::KF;
            src[u] = A4p + q * aPlane4 + kfl * aRow4 + (size_t)tmI * T::FA * 64 + w;
        } else {
            const int eb = e - T::A4;
            const int row = eb / (T::FB * 64);
            const int w = eb % (T::FB * 64);
            const int q = row / T::KF, kfl = row % T::KF;
            src[u] = B4p + q * bPlane4 + kfl * bRow4 + (size_t)tnI * T::FB * 64 + w;
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

    // Pipeline per stage s (LDS ring of two buffers, register staging one stage ahead):
    //   1. read every fragment of stage s from LDS into registers;
    //   2. write the staged HBM data of stage s+1 into the other LDS buffer (last read in
    //      stage s-1, before the barrier that ended it);
    //   3. issue the HBM loads of stage s+2 into the staging registers;
    //   4. issue the stage's MFMAs, then the barrier, so the barrier wait and the next
    //      stage's LDS reads overlap the MFMAs still in flight.
    f32x4 gbuf[T::LOADS];
    if (ks0 < ks1) {
#pragma unroll
        for (int u = 0; u < T::LOADS; ++u) {
            const size_t step = (u < T::A4 / NT) ? (size_t)T::KF * aRow4 : (size_t)T::KF * bRow4;
            gbuf[u] = src[u][(size_t)ks0 * step];
        }
#pragma unroll
        for (int u = 0; u < T::LOADS; ++u) lds[tid + NT * u] = gbuf[u];
        if (ks0 + 1 < ks1) {
#pragma unroll
            for (int u = 0; u < T::LOADS; ++u) {
                const size_t step = (u < T::A4 / NT) ? (size_t)T::KF * aRow4 : (size_t)T::KF * bRow4;
                gbuf[u] = src[u][(size_t)(ks0 + 1) * step];
            }
        }
    }
    __syncthreads();

    for (int ks = ks0; ks < ks1; ++ks) {
        const int buf = (ks - ks0) & 1;
        const f32x4* L = lds + buf * T::S4;
        f32x4 af[T::KF][3][T::TM], bf[T::KF][3][T::TN];
#pragma unroll
        for (int kfl = 0; kfl < T::KF; ++kfl)
#pragma unroll
            for (int q = 0; q < 3; ++q) {
#pragma unroll
                for (int b = 0; b < T::TM; ++b)
                    af[kfl][q][b] = L[((q * T::KF + kfl) * T::FA + wm * T::TM + b) * 64 + lane];
#pragma unroll
                for (int a = 0; a < T::TN; ++a)
                    bf[kfl][q][a] = L[T::A4 + ((q * T::KF + kfl) * T::FB + wn * T::TN + a) * 64 + lane];
            }
        if (ks + 1 < ks1) {
#pragma unroll
            for (int u = 0; u < T::LOADS; ++u) lds[(buf ^ 1) * T::S4 + tid + NT * u] = gbuf[u];
        }
        if (ks + 2 < ks1) {
#pragma unroll
            for (int u = 0; u < T::LOADS; ++u) {
                const size_t step = (u < T::A4 / NT) ? (size_t)T::KF * aRow4 : (size_t)T::KF * bRow4;
                gbuf[u] = src[u][(size_t)(ks + 2) * step];
            }
        }
#pragma unroll
        for (int kfl = 0; kfl < T::KF; ++kfl)
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int q = 0; q < 3; ++q)
#pragma unroll
                    for (int a = 0; a < T::TN; ++a)
#pragma unroll
                        for (int b = 0; b < T::TM; ++b)
                            acc[q][a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(bf[kfl][q][a][s], af[kfl][q][b][s],
                                                                                acc[q][a][b], 0, 0, 0);
        __syncthreads();
    }

    // Epilogue: lane holds output row i = ... + lane%32; register r holds column
    // j = ... + 8*(r/4) + 4*(lane/32) + r%4 of each 32x32 tile.
    const int iBase = tmI * BM + wm * (BM / WM) + (lane & 31);
    const int jBase = tnI * BN + wn * (BN / WN) + 4 * (lane >> 5);
    store_tile<T::TN, T::TM, PARTIAL>(g, acc, part, iBase, jBase);
}

template <int BM, int BN, bool PARTIAL, int WM, int WN>
__global__ __launch_bounds__(64 * WM * WN, 1) void gemm3m_f32_kernel(GemmArgs g)
{
    __shared__ f32x4 lds[2 * F32Tile<BM, BN, WM, WN>::S4];
    int part, tmI, tnI;
    map_tile(blockIdx.x, gridDim.x, g.tilesM, g.tilesN, part, tmI, tnI);
    f32_tile<BM, BN, PARTIAL, WM, WN>(g, part, tmI, tnI, lds);
}

// LDS-free form of f32_tile. The packed planes are already in MFMA fragment order, so each
// wave loads its own fragments straight into registers (one 1 KiB coalesced load per
// fragment) and no data passes through LDS: no barriers, and the waves of a workgroup run
// independently. Two register sets alternate: the loads of stage s+1 are issued before the
// MFMAs of stage s, so their latency hides behind a whole stage of MFMAs. Fragments that
// several waves of the workgroup share are read by each of them (served by the CU's vector
// cache and L2). Same MFMA order per accumulator as f32_tile, so the results are identical.
//
// FOUR = true is the four-product form (ID_FOUR): A planes (Ar, Ai, -Ai), B planes (Br, Bi),
// two accumulators per 32x32 tile taking re += Br*Ar then Bi*(-Ai), im += Bi*Ar then Br*Ai
// per k slot; the epilogue stores them as (re, im).
//
// KFD: 8-deep k fragments per register set (stage); NSET: register sets in the ring, so the
// loads of stage s + NSET - 1 are issued before the MFMAs of stage s. SOLO: every wave of the
// workgroup computes a tile of its own (WM = WN = 1; the caller gives each wave its tile).
#ifndef C3M_DIRECT_KFD
#define C3M_DIRECT_KFD 2
#endif
#ifndef C3M_DIRECT_NSET
#define C3M_DIRECT_NSET 2
#endif
template <int BM, int BN, int WM, int WN, int NB, int KFD>
struct DirectSet {
    using T = F32Tile<BM, BN, WM, WN>;
    f32x4 a[KFD][3][T::TM], b[KFD][NB][T::TN];
};

template <int BM, int BN, bool PARTIAL, int WM = 2, int WN = 2, bool FOUR = false, int KFD = C3M_DIRECT_KFD,
          int NSET = C3M_DIRECT_NSET, bool SOLO = false>
__device__ __forceinline__ void f32_tile_direct(const GemmArgs& g, int part, int tmI, int tnI)
{
    static_assert(!SOLO || (WM == 1 && WN == 1), "a solo wave owns its whole tile");
    using T = F32Tile<BM, BN, WM, WN>;
    constexpr int NB = FOUR ? 2 : 3;    // B planes
    constexpr int NACC = FOUR ? 2 : 3;  // accumulators per 32x32 output tile
    static_assert(KFD == 1 || KFD == 2, "a planner stage (16 k) is two 8-deep fragments");
    using Set = DirectSet<BM, BN, WM, WN, NB, KFD>;
    const int lane = threadIdx.x & 63;
    const int wave = SOLO ? 0 : __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = wave % WM, wn = wave / WM;
    // this partition's k fragments [kf0, kf0 + KFD * nst)
    const int kf0 = 2 * (int)((long long)g.kSteps * part / g.parts);
    const int kf1 = 2 * (int)((long long)g.kSteps * (part + 1) / g.parts);

    // Wave-uniform byte addresses (scalar registers) plus one 32-bit lane offset, so every
    // load is a scalar base + vector offset form and addressing costs almost no VGPRs.
    const size_t aRowB = (size_t)(g.mPad / 32) * 1024;  // bytes per k fragment, all of A's mn
    const size_t bRowB = (size_t)(g.nPad / 32) * 1024;
    const char* aW = reinterpret_cast<const char*>(g.Ap) + (size_t)(tmI * T::FA + wm * T::TM) * 1024;
    const char* bW = reinterpret_cast<const char*>(g.Bp) + (size_t)(tnI * T::FB + wn * T::TN) * 1024;
    const unsigned laneB = (unsigned)lane * 16u;

    auto load = [&](Set& S, int st) {
#pragma unroll
        for (int kfl = 0; kfl < KFD; ++kfl)
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const size_t kf = (size_t)(kf0 + st * KFD + kfl);
                const char* a = aW + (size_t)q * g.aPlaneBytes + kf * aRowB;
                const char* b = bW + (size_t)q * g.bPlaneBytes + kf * bRowB;
#pragma unroll
                for (int i = 0; i < T::TM; ++i)
                    S.a[kfl][q][i] = *reinterpret_cast<const f32x4*>(a + i * 1024 + laneB);
                if (q < NB) {
#pragma unroll
                    for (int j = 0; j < T::TN; ++j)
                        S.b[kfl][q < NB ? q : 0][j] = *reinterpret_cast<const f32x4*>(b + j * 1024 + laneB);
                }
            }
    };

    f32x16 acc[NACC][T::TN][T::TM];
#pragma unroll
    for (int q = 0; q < NACC; ++q)
#pragma unroll
        for (int a = 0; a < T::TN; ++a)
#pragma unroll
            for (int b = 0; b < T::TM; ++b)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[q][a][b][r] = 0.0f;

    // term t: accumulator, B plane, A plane
    constexpr int TACC[4] = {0, 1, 0, 1}, TBP[4] = {0, 1, 1, 0}, TAP[4] = {0, 0, 2, 1};
    auto mfma = [&](const Set& S) {
#pragma unroll
        for (int kfl = 0; kfl < KFD; ++kfl)
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                if constexpr (FOUR) {
#pragma unroll
                    for (int t = 0; t < 4; ++t)
#pragma unroll
                        for (int a = 0; a < T::TN; ++a)
#pragma unroll
                            for (int b = 0; b < T::TM; ++b)
                                acc[TACC[t]][a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(
                                    S.b[kfl][TBP[t]][a][s], S.a[kfl][TAP[t]][b][s], acc[TACC[t]][a][b], 0, 0, 0);
                } else {
#pragma unroll
                    for (int q = 0; q < 3; ++q)
#pragma unroll
                        for (int a = 0; a < T::TN; ++a)
#pragma unroll
                            for (int b = 0; b < T::TM; ++b)
                                acc[q][a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(
                                    S.b[kfl][q < NB ? q : 0][a][s], S.a[kfl][q][b][s], acc[q][a][b], 0, 0, 0);
                }
            }
    };

    // Ring of NSET register sets, stages in groups of NSET with one exit, so the accumulators
    // keep their registers; the last NSET - 1 stages are already loaded when the loop ends.
    // Loads are unconditional (look-ahead past the end is clamped to the last stage): a load
    // skipped on one path would make the wait before the next MFMAs cover the loads just
    // issued. Scheduling barriers keep each stage's loads ahead of the MFMAs.
    const int nst = (kf1 - kf0) / KFD;
    Set S[NSET];
    if (nst > 0) {
#pragma unroll
        for (int u = 0; u + 1 < NSET; ++u) load(S[u], min(u, nst - 1));
    }
    int st = 0;
    for (; st + NSET <= nst; st += NSET) {
#pragma unroll
        for (int u = 0; u < NSET; ++u) {
            load(S[(u + NSET - 1) % NSET], min(st + u + NSET - 1, nst - 1));
            __builtin_amdgcn_sched_barrier(0);
            mfma(S[u]);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
#pragma unroll
    for (int u = 0; u + 1 < NSET; ++u)
        if (st + u < nst) mfma(S[u]);

    const int iBase = tmI * BM + wm * (BM / WM) + (lane & 31);
    const int jBase = tnI * BN + wn * (BN / WN) + 4 * (lane >> 5);
    store_tile<T::TN, T::TM, PARTIAL, NACC>(g, acc, part, iBase, jBase);
}

template <int BM, int BN, bool PARTIAL, int WM, int WN, int OCC, bool FOUR = false>
__global__ __launch_bounds__(64 * WM * WN, OCC) void gemm3m_f32_direct_kernel(GemmArgs g)
{
    int part, tmI, tnI;
    map_tile(blockIdx.x, gridDim.x, g.tilesM, g.tilesN, part, tmI, tnI);
    f32_tile_direct<BM, BN, PARTIAL, WM, WN, FOUR>(g, part, tmI, tnI);
}

// Raw-operand form of the 64 x 64 wave tile, for grouped members: the wave reads its operands
// straight from the caller's interleaved complex64 A and B and forms the plane values in
// registers, so the batch needs no pack launch and no plane workspace. Per 16-deep stage a
// lane holds, for each of its two 32-row fragments of op(A) (and two 32-column fragments of
// op(B)) and each 8-deep k fragment, the four consecutive k = 4 * (lane / 32) + s that the
// packed fragment's slots hold: four 8-byte loads (32 lanes cover 32 consecutive rows when
// rows are adjacent in memory, one row's 32 contiguous bytes per lane otherwise).
// Out-of-range rows and k (up to the packed extent kPad, the same stage count) read a clamped
// in-range address and become zeros; conjugation negates the imaginary part; the plane values
// use the pack kernel's formulas (plane_value). So every MFMA operand and the MFMA order per
// accumulator equal the packed 64 x 64 tile's: the results are byte-identical to it.
typedef float rawf2 __attribute__((ext_vector_type(2)));  // one complex64 element (re, im)

template <int ID, int TM, int TN>
struct RawSet {
    rawf2 a[2][TM][4], b[2][TN][4];  // [k fragment][32-row / 32-column fragment][k slot]
};

// TM (TN): 32-row (32-column) fragments of the block that hold outputs. A block at a member's
// last row (column) with at most 32 outputs left takes 1 and skips the MFMAs of the fragment
// that would be padding; the other fragment's accumulators get the same MFMAs in the same
// order, so its outputs are unchanged.
template <int ID, int TM, int TN>
__device__ __forceinline__ void f32_tile_raw64(const GemmArgs& g, int part, int tmI, int tnI)
{
    constexpr int KFD = 2, NSET = 2;
    const int lane = threadIdx.x & 63;
    const int half = lane >> 5;
    const int kf0 = 2 * (int)((long long)g.kSteps * part / g.parts);
    const int kf1 = 2 * (int)((long long)g.kSteps * (part + 1) / g.parts);
    const int K = g.src.k;
    // op(A)(i, k) = A[i + k*lda] (N) or A[k + i*lda] (T, C); op(B)(k, j) = B[k + j*ldb] (N) or
    // B[j + k*ldb] (T, C). Byte address = tile base (the tile's first row or column) + k * step
    // (uniform per stage and slot) + a 32-bit lane offset fixed for the tile (row or column
    // within the tile, and the lane's k offset); the host sends a batch here only if every such
    // offset fits in 31 bits.
    // Rows (columns) past the matrix read the last one: they feed only outputs that are never
    // stored, and the fused store writes in-range outputs only (a K-split partial plane holds
    // them in its padding, which the merge does not read).
    const bool aKc = g.src.opA != 111, bKc = g.src.opB == 111;
    const long long aStepB = aKc ? 8 : g.src.lda * 8, bStepB = bKc ? 8 : g.src.ldb * 8;
    const int i0 = tmI * 64, j0 = tnI * 64;
    const char* aTile = reinterpret_cast<const char*>(g.src.A) + (aKc ? (long long)i0 * g.src.lda * 8 : (long long)i0 * 8);
    const char* bTile = reinterpret_cast<const char*>(g.src.B) + (bKc ? (long long)j0 * g.src.ldb * 8 : (long long)j0 * 8);
    const float aSign = g.src.opA == 113 ? -1.0f : 1.0f, bSign = g.src.opB == 113 ? -1.0f : 1.0f;
    unsigned aOff[TM], bOff[TN];
#pragma unroll
    for (int b = 0; b < TM; ++b) {
        const long long d = min(i0 + 32 * b + (lane & 31), g.m - 1) - i0;
        aOff[b] = (unsigned)((aKc ? d * g.src.lda * 8 : d * 8) + 4 * half * aStepB);
    }
#pragma unroll
    for (int a = 0; a < TN; ++a) {
        const long long d = min(j0 + 32 * a + (lane & 31), g.n - 1) - j0;
        bOff[a] = (unsigned)((bKc ? d * g.src.ldb * 8 : d * 8) + 4 * half * bStepB);
    }
    typedef __attribute__((address_space(1))) const rawf2 gf2;
    typedef float rawf4 __attribute__((ext_vector_type(4), aligned(8)));
    auto at = [](const char* base, unsigned off) { return *(gf2*)(base + off); };
    // Buffer loads for the ring: resource = the stage's uniform base (64-bit scalar math once
    // per stage and slot), 32-bit lane offset in a VGPR, small uniform offset in an SGPR, so a
    // load costs no address arithmetic. No range check (num_records = 2^32 - 1): every offset
    // is in bounds by construction (clamped rows, stages wholly below K).
    auto rsrc = [](const char* base) {
        return __builtin_amdgcn_make_buffer_rsrc(const_cast<char*>(base), (short)0, (int)0xffffffff, 0x00020000);
    };
    auto bl2 = [](__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
        return __builtin_bit_cast(rawf2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
    };
    auto bl4 = [](__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
        return __builtin_bit_cast(rawf4, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
    };
    // One stage: 16 k from k = (kf0 + st * KFD) * 8, the lane's k at 4 * (lane / 32) + u within
    // each 8. Loads only fill registers (no branch around or between loads, no use of a loaded
    // value here): anything consuming a look-ahead load in the load block, or a branch joining
    // loads, would make the wait for it come at once, so the operand layouts are compile-time
    // (AKC, BKC: op(A), op(B) k-contiguous, whose lane's four k are 32 contiguous bytes, read
    // as two 16-byte loads; the other layout reads one k per load, rows adjacent across lanes).
    auto load = [&](RawSet<ID, TM, TN>& S, int st, auto aTag, auto bTag) {
        constexpr bool AKC = decltype(aTag)::value, BKC = decltype(bTag)::value;
        const int sk = (kf0 + st * KFD) * 8;
#pragma unroll
        for (int kfl = 0; kfl < KFD; ++kfl) {
            const __amdgpu_buffer_rsrc_t ar = rsrc(aTile + (long long)(sk + kfl * 8) * aStepB);
            const __amdgpu_buffer_rsrc_t br = rsrc(bTile + (long long)(sk + kfl * 8) * bStepB);
#pragma unroll
            for (int b = 0; b < TM; ++b)
#pragma unroll
                for (int u = 0; u < 4; u += (AKC ? 2 : 1)) {
                    if constexpr (AKC) {
                        const rawf4 w = bl4(ar, aOff[b], 8 * u);
                        S.a[kfl][b][u] = rawf2{w[0], w[1]};
                        S.a[kfl][b][u + 1] = rawf2{w[2], w[3]};
                    } else {
                        S.a[kfl][b][u] = bl2(ar, aOff[b], (unsigned)(u * aStepB));
                    }
                }
#pragma unroll
            for (int a = 0; a < TN; ++a)
#pragma unroll
                for (int u = 0; u < 4; u += (BKC ? 2 : 1)) {
                    if constexpr (BKC) {
                        const rawf4 w = bl4(br, bOff[a], 8 * u);
                        S.b[kfl][a][u] = rawf2{w[0], w[1]};
                        S.b[kfl][a][u + 1] = rawf2{w[2], w[3]};
                    } else {
                        S.b[kfl][a][u] = bl2(br, bOff[a], (unsigned)(u * bStepB));
                    }
                }
        }
    };
    // The stage reaching past K: k clamped to K - 1 per element (the lane's k offset sits in the
    // lane offsets; a clamped k, per lane, takes it out of them again), one k per load.
    auto loadClamped = [&](RawSet<ID, TM, TN>& S, int st) {
        const int sk = (kf0 + st * KFD) * 8;
#pragma unroll
        for (int kfl = 0; kfl < KFD; ++kfl)
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const long long kr = min(sk + kfl * 8 + 4 * half + u, K - 1) - 4 * half;
                const char* ar = aTile + kr * aStepB;
                const char* br = bTile + kr * bStepB;
#pragma unroll
                for (int b = 0; b < TM; ++b) S.a[kfl][b][u] = at(ar, aOff[b]);
#pragma unroll
                for (int a = 0; a < TN; ++a) S.b[kfl][a][u] = at(br, bOff[a]);
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

    // Plane values of one element with the conjugation sign folded in: s * im is exact (s = +-1),
    // so fma(s, im, x) rounds exactly as x + (s * im) does, and the values equal
    // plane_value(role, ID, q, re, s * im), the pack kernel's. The fused identity's third A
    // plane is s * Ai; it is used here as Ai, and that product's accumulators are negated after
    // the k loop when s = -1: rounding is symmetric in sign, so every partial sum of Ai * (Br+Bi)
    // is exactly the negation of the one of (-Ai) * (Br+Bi), and the bytes are the pack path's
    // with one multiply per element and fragment fewer in the loop.
    auto planeA = [&](int q, float re, float im) {
        if (ID == ID_STANDARD) return q == 0 ? re : (q == 1 ? aSign * im : fmaf(aSign, im, re));
        return q == 0 ? fmaf(aSign, im, re) : (q == 1 ? re : im);  // (Ar+Ai, Ar, Ai), Ai unsigned
    };
    auto planeB = [&](int q, float re, float im) {
        if (ID == ID_STANDARD) return q == 0 ? re : (q == 1 ? bSign * im : fmaf(bSign, im, re));
        return q == 0 ? re : (q == 1 ? fmaf(bSign, im, -re) : fmaf(bSign, im, re));  // (Br, Bi-Br, Br+Bi)
    };
    // Row and column energies for the rounding-correction screen, formed from the loaded
    // values in the half-sum order of energy_acc: per stage, the lane's half h sums its 8 k of
    // each row (column) in FP32 (H); T (FP64) sums the stages' H of a 64-deep k tile from
    // zero; R sums the tiles' T from zero in tile order; the two halves' R are added after the
    // loop. Only fused tiles of members with the correction on form them: a uniform branch
    // after each stage's MFMAs (VALU work only; a second copy of the k loops instead would
    // make the compiler wait for all loads in flight at every stage of that copy).
    double T[TM + TN], R[TM + TN];
#pragma unroll
    for (int q = 0; q < TM + TN; ++q) T[q] = R[q] = 0.0;
    const bool energy = g.parts == 1 && g.correction && g.energyInline;

    // MASK: k past K contributes zeros (both operands, so an Inf or NaN read from the clamped
    // address adds nothing), to the products and to the energies.
    auto mfma = [&](const RawSet<ID, TM, TN>& S, int st, auto maskTag) {
        constexpr bool MASK = decltype(maskTag)::value;
        const int sk = (kf0 + st * KFD) * 8;
        const int pos = (kf0 / KFD + st) & 3;  // stage within its 64-deep k tile
#pragma unroll
        for (int kfl = 0; kfl < KFD; ++kfl)
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const bool kin = !MASK || sk + kfl * 8 + 4 * half + s < K;
#pragma unroll
                for (int q = 0; q < 3; ++q)
#pragma unroll
                    for (int a = 0; a < TN; ++a)
#pragma unroll
                        for (int b = 0; b < TM; ++b) {
                            const rawf2 av = S.a[kfl][b][s], bv = S.b[kfl][a][s];
                            const float x = kin ? planeB(q, bv.x, bv.y) : 0.0f;
                            const float y = kin ? planeA(q, av.x, av.y) : 0.0f;
                            acc[q][a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(x, y, acc[q][a][b], 0, 0, 0);
                        }
            }
        if (energy) {  // uniform; VALU only, after the stage's MFMAs are issued
            float H[TM + TN];
#pragma unroll
            for (int q = 0; q < TM + TN; ++q) H[q] = 0.0f;
#pragma unroll
            for (int kfl = 0; kfl < KFD; ++kfl)
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const bool kin = !MASK || sk + kfl * 8 + 4 * half + s < K;
#pragma unroll
                    for (int b = 0; b < TM; ++b) {
                        const rawf2 v = S.a[kfl][b][s];
                        H[b] = energy_acc(H[b], kin ? v.x : 0.0f, kin ? v.y : 0.0f);
                    }
#pragma unroll
                    for (int a = 0; a < TN; ++a) {
                        const rawf2 v = S.b[kfl][a][s];
                        H[TM + a] = energy_acc(H[TM + a], kin ? v.x : 0.0f, kin ? v.y : 0.0f);
                    }
                }
#pragma unroll
            for (int q = 0; q < TM + TN; ++q) {
                T[q] = (pos == 0 ? 0.0 : T[q]) + (double)H[q];
                R[q] = pos == 3 ? R[q] + T[q] : R[q];
            }
        }
    };
    typedef std::integral_constant<bool, false> Off;
    typedef std::integral_constant<bool, true> On;

    // Stages wholly below K run through the ring: NSET register sets, unconditional look-ahead
    // loads (the last full stage again past the end), scheduling barriers keeping each stage's
    // loads ahead of its MFMAs. The stage (at most the partition's last) reaching past K
    // follows, clamped and masked.
    const int nst = (kf1 - kf0) / KFD;
    const int nFull = min(nst, max(0, (K - kf0 * 8) / (8 * KFD)));
    auto ring = [&](auto aTag, auto bTag) {
        RawSet<ID, TM, TN> S[NSET];
        if (nFull > 0) {
#pragma unroll
            for (int u = 0; u + 1 < NSET; ++u) load(S[u], min(u, nFull - 1), aTag, bTag);
        }
        int st = 0;
        for (; st + NSET <= nFull; st += NSET) {
#pragma unroll
            for (int u = 0; u < NSET; ++u) {
                load(S[(u + NSET - 1) % NSET], min(st + u + NSET - 1, nFull - 1), aTag, bTag);
                __builtin_amdgcn_sched_barrier(0);
                mfma(S[u], st + u, Off());
                __builtin_amdgcn_sched_barrier(0);
            }
        }
#pragma unroll
        for (int u = 0; u + 1 < NSET; ++u)
            if (st + u < nFull) mfma(S[u], st + u, Off());
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
        RawSet<ID, TM, TN> Tl;
        loadClamped(Tl, t);
        mfma(Tl, t, On());
    }

    if (ID != ID_STANDARD && g.src.opA == 113) {  // the third product was accumulated with Ai for s * Ai
#pragma unroll
        for (int a = 0; a < TN; ++a)
#pragma unroll
            for (int b = 0; b < TM; ++b)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[2][a][b][r] = -acc[2][a][b][r];
    }
    const int iBase = i0 + (lane & 31), jBase = j0 + 4 * half;
    if (g.parts > 1) {  // K split: the merge screens, from energies computed before this kernel
        store_tile<TN, TM, true, 3>(g, acc, part, iBase, jBase);
    } else if (energy) {
        __shared__ double ebShared[4][64];  // per wave: its 64 columns' energies
        const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        if ((nst & 3) != 0) {  // the last k tile is partial: add its partial sum
#pragma unroll
            for (int q = 0; q < TM + TN; ++q) R[q] += T[q];
        }
        double ea[TM];
#pragma unroll
        for (int b = 0; b < TM; ++b) ea[b] = R[b] + __shfl_xor(R[b], 32);  // half 0 + half 1 (commutes)
#pragma unroll
        for (int a = 0; a < TN; ++a) {
            const double e = R[TM + a] + __shfl_xor(R[TM + a], 32);
            if (half == 0) ebShared[wave][32 * a + lane] = e;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        store_tile<TN, TM, false, 3, true>(g, acc, part, iBase, jBase, ea,
                                           (const __attribute__((address_space(3))) double*)&ebShared[wave][4 * half]);
        // the next tile of this wave rewrites ebShared only after these reads (in order per wave)
        __builtin_amdgcn_wave_barrier();
    } else {
        store_tile<TN, TM, false, 3>(g, acc, part, iBase, jBase);
    }
}

// Grouped batch on 64 x 64 wave tiles: a persistent grid of four-wave workgroups pulls quads
// (four wave tiles, entries 4q .. 4q + 3 of the list) from one queue; wave w computes entry
// 4q + w with the per-wave code of the 128 x 128 LDS-free kernel (a 64 x 64 output block,
// same MFMA order per accumulator, so the bytes equal the one-problem kernels'). The quad of
// a whole 128 x 128 region lists its blocks in that kernel's wave order, so the four waves
// share operand fragments in the CU's cache as they do there; the blocks at the edges of the
// members (where a 128 x 128 tile would be mostly padding) are grouped four at a time by
// cost. An entry with problem < 0 is an empty slot. Quads are sorted by decreasing cost.
template <bool RAW>
__global__ __launch_bounds__(256, 1) void gemm3m_f32_grouped_quad_kernel(const GroupProblem* __restrict__ probs,
                                                                        const GroupTile* __restrict__ tiles, int nquads,
                                                                        unsigned int* queue)
{
    __shared__ int next;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (;;) {
        if (threadIdx.x == 0) next = (int)atomicAdd(queue, 1u);
        __syncthreads();
        const int q = __builtin_amdgcn_readfirstlane(next);
        __syncthreads();
        if (q >= nquads) break;
        const GroupTile gt = tiles[(size_t)q * 4 + wave];
        const int p = __builtin_amdgcn_readfirstlane(gt.problem);
        if (p < 0) continue;
        const GemmArgs g = probs[p].g;
        const int part = __builtin_amdgcn_readfirstlane(gt.part);
        const int tm = __builtin_amdgcn_readfirstlane(gt.tm);
        const int tn = __builtin_amdgcn_readfirstlane(gt.tn);
        if constexpr (RAW) {  // the host sends fused-identity batches only
            const bool m2 = g.m - 64 * tm > 32, n2 = g.n - 64 * tn > 32;
            if (m2 && n2)
                f32_tile_raw64<ID_FUSED, 2, 2>(g, part, tm, tn);
            else if (m2)
                f32_tile_raw64<ID_FUSED, 2, 1>(g, part, tm, tn);
            else if (n2)
                f32_tile_raw64<ID_FUSED, 1, 2>(g, part, tm, tn);
            else
                f32_tile_raw64<ID_FUSED, 1, 1>(g, part, tm, tn);
        } else if (g.parts > 1) {
            f32_tile_direct<64, 64, true, 1, 1, false, C3M_DIRECT_KFD, C3M_DIRECT_NSET, true>(g, part, tm, tn);
        } else {
            f32_tile_direct<64, 64, false, 1, 1, false, C3M_DIRECT_KFD, C3M_DIRECT_NSET, true>(g, part, tm, tn);
        }
    }
}

hipError_t launch_gemm_f32_grouped_quads(const GroupProblem* probs, const GroupTile* tiles, int nquads,
                                         unsigned int* queue, int blocks, bool raw, hipStream_t s)
{
    if (raw)
        hipLaunchKernelGGL(gemm3m_f32_grouped_quad_kernel<true>, dim3(blocks), dim3(256), 0, s, probs, tiles, nquads,
                           queue);
    else
        hipLaunchKernelGGL(gemm3m_f32_grouped_quad_kernel<false>, dim3(blocks), dim3(256), 0, s, probs, tiles, nquads,
                           queue);
    return hipGetLastError();
}

template <int BM, int BN, bool PARTIAL, int WM, int WN>
static hipError_t launch_one(const GemmArgs& g, hipStream_t s)
{
    const int blocks = g.tilesM * g.tilesN * g.parts;
    hipLaunchKernelGGL((gemm3m_f32_kernel<BM, BN, PARTIAL, WM, WN>), dim3(blocks), dim3(64 * WM * WN), 0, s, g);
    return hipGetLastError();
}

template <int BM, int BN, bool PARTIAL, int WM, int WN, int OCC, bool FOUR = false>
static hipError_t launch_direct(const GemmArgs& g, hipStream_t s)
{
    const int blocks = g.tilesM * g.tilesN * g.parts;
    hipLaunchKernelGGL((gemm3m_f32_direct_kernel<BM, BN, PARTIAL, WM, WN, OCC, FOUR>), dim3(blocks),
                       dim3(64 * WM * WN), 0, s, g);
    return hipGetLastError();
}

// 128x128 tiles: variant 0: 4 waves through LDS (64x64 per wave); 1: 8 waves through LDS
// (64x32 per wave, two per SIMD); 2: LDS-free, 4 waves (the default). Measured at 8192^3 on
// MI355X, products kernel alone, complex TFLOP/s: 151.7, 162.5, 185.6. The LDS-free form
// with 8 waves (92) or 64x64 tiles (85, against 150 through LDS) loses: its wave tiles are
// too small for the operands it re-reads from the caches, so it is not built.
hipError_t launch_gemm_f32(const GemmArgs& g, int bm, int bn, int variant, hipStream_t s)
{
    const bool partial = g.parts > 1;
    if (g.identity == ID_FOUR) {  // four products: LDS-free kernels only
        if (bm == 128 && bn == 128)
            return partial ? launch_direct<128, 128, true, 2, 2, 1, true>(g, s)
                           : launch_direct<128, 128, false, 2, 2, 1, true>(g, s);
        if (bm == 64 && bn == 64)
            return partial ? launch_direct<64, 64, true, 2, 2, 2, true>(g, s)
                           : launch_direct<64, 64, false, 2, 2, 2, true>(g, s);
        return hipErrorInvalidValue;
    }
    if (bm == 128 && bn == 128 && variant == 2)
        return partial ? launch_direct<128, 128, true, 2, 2, 1>(g, s) : launch_direct<128, 128, false, 2, 2, 1>(g, s);
    if (bm == 128 && bn == 128) {
        if (variant == 1)
            return partial ? launch_one<128, 128, true, 2, 4>(g, s) : launch_one<128, 128, false, 2, 4>(g, s);
        return partial ? launch_one<128, 128, true, 2, 2>(g, s) : launch_one<128, 128, false, 2, 2>(g, s);
    }
    if (bm == 64 && bn == 64)
        return partial ? launch_one<64, 64, true, 2, 2>(g, s) : launch_one<64, 64, false, 2, 2>(g, s);
    return hipErrorInvalidValue;
}

}  // namespace c3m
