// gemm_f32.hip -- the three real products on FP32 MFMA (v_mfma_f32_32x32x2_f32), gfx950.
//
// One workgroup = 4 waves = one BM x BN output tile; each wave owns a (BM/2) x (BN/2)
// quarter and keeps all THREE product accumulators for it in registers (3 x TM x TN
// 32x32 FP32 tiles, 192 registers at 128x128). Per k stage (BK = 16) the block copies the
// packed fragments of the 3 A planes and 3 B planes (contiguous runs, see kernel_args.h)
// HBM -> VGPR -> LDS with a double-buffered LDS ring and one barrier per stage; each wave
// then reads its fragments with one 16-byte LDS read per lane and issues the MFMAs.
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
#include "device_util.h"
#include "launch.h"

namespace c3m {

template <int BM, int BN>
struct F32Tile {
    static constexpr int TM = BM / 64, TN = BN / 64;  // 32x32 MFMA tiles per wave (2x2 waves)
    static constexpr int FA = BM / 32, FB = BN / 32;  // fragments per plane per k fragment
    static constexpr int KF = 2;                      // k fragments (8 deep) per stage
    static constexpr int A4 = 3 * KF * FA * 64;       // float4 per stage, A side
    static constexpr int B4 = 3 * KF * FB * 64;
    static constexpr int S4 = A4 + B4;
    static constexpr int LOADS = S4 / 256;            // float4 loads per thread per stage
    static_assert(A4 % 256 == 0 && B4 % 256 == 0, "stage must split evenly over 256 threads");
};

// One output tile (tmI, tnI) over K partition `part`, start to finish; lds holds
// 2 * F32Tile<BM, BN>::S4 float4. Used by the one-problem kernel and the grouped kernel.
template <int BM, int BN, bool PARTIAL>
__device__ __forceinline__ void f32_tile(const GemmArgs& g, int part, int tmI, int tnI, f32x4* lds)
{
    using T = F32Tile<BM, BN>;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
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
        const int e = tid + 256 * u;
        if (u < T::A4 / 256) {
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

    f32x4 gbuf[T::LOADS];
    if (ks0 < ks1) {
#pragma unroll
        for (int u = 0; u < T::LOADS; ++u) {
            const size_t step = (u < T::A4 / 256) ? (size_t)T::KF * aRow4 : (size_t)T::KF * bRow4;
            gbuf[u] = src[u][(size_t)ks0 * step];
        }
#pragma unroll
        for (int u = 0; u < T::LOADS; ++u) lds[tid + 256 * u] = gbuf[u];
    }
    __syncthreads();

    for (int ks = ks0; ks < ks1; ++ks) {
        const int buf = (ks - ks0) & 1;
        const bool more = ks + 1 < ks1;
        if (more) {
#pragma unroll
            for (int u = 0; u < T::LOADS; ++u) {
                const size_t step = (u < T::A4 / 256) ? (size_t)T::KF * aRow4 : (size_t)T::KF * bRow4;
                gbuf[u] = src[u][(size_t)(ks + 1) * step];
            }
        }
        const f32x4* L = lds + buf * T::S4;
#pragma unroll
        for (int kfl = 0; kfl < T::KF; ++kfl) {
            f32x4 af[3][T::TM], bf[3][T::TN];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
#pragma unroll
                for (int b = 0; b < T::TM; ++b)
                    af[q][b] = L[((q * T::KF + kfl) * T::FA + wm * T::TM + b) * 64 + lane];
#pragma unroll
                for (int a = 0; a < T::TN; ++a)
                    bf[q][a] = L[T::A4 + ((q * T::KF + kfl) * T::FB + wn * T::TN + a) * 64 + lane];
            }
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int q = 0; q < 3; ++q)
#pragma unroll
                    for (int a = 0; a < T::TN; ++a)
#pragma unroll
                        for (int b = 0; b < T::TM; ++b)
                            acc[q][a][b] =
                                __builtin_amdgcn_mfma_f32_32x32x2f32(bf[q][a][s], af[q][b][s], acc[q][a][b], 0, 0, 0);
        }
        if (more) {
#pragma unroll
            for (int u = 0; u < T::LOADS; ++u) lds[(buf ^ 1) * T::S4 + tid + 256 * u] = gbuf[u];
        }
        __syncthreads();
    }

    // Epilogue: lane holds output row i = ... + lane%32; register r holds column
    // j = ... + 8*(r/4) + 4*(lane/32) + r%4 of each 32x32 tile.
    const int iBase = tmI * BM + wm * (BM / 2) + (lane & 31);
    const int jBase = tnI * BN + wn * (BN / 2) + 4 * (lane >> 5);
    store_tile<T::TN, T::TM, PARTIAL>(g, acc, part, iBase, jBase);
}

template <int BM, int BN, bool PARTIAL>
__global__ __launch_bounds__(256, 1) void gemm3m_f32_kernel(GemmArgs g)
{
    __shared__ f32x4 lds[2 * F32Tile<BM, BN>::S4];
    int part, tmI, tnI;
    map_tile(blockIdx.x, gridDim.x, g.tilesM, g.tilesN, part, tmI, tnI);
    f32_tile<BM, BN, PARTIAL>(g, part, tmI, tnI, lds);
}

// Grouped batch: a persistent grid pulls tiles of many independent problems from one
// queue (tiles pre-sorted by decreasing cost, so the longest start first). Each tile is
// computed start to finish by one workgroup with the same arithmetic as the one-problem
// kernel, so the output does not depend on which workgroup took it. One launch per tile
// size, so each size runs at its own occupancy.
template <int BM>
__global__ __launch_bounds__(256, BM == 128 ? 1 : 2) void gemm3m_f32_grouped_kernel(const GroupProblem* probs,
                                                                                  const GroupTile* tiles,
                                                                                  int ntiles, unsigned int* queue)
{
    __shared__ f32x4 lds[2 * F32Tile<BM, BM>::S4];
    __shared__ int next;
    for (;;) {
        if (threadIdx.x == 0) next = (int)atomicAdd(queue, 1u);
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(next);
        __syncthreads();
        if (t >= ntiles) break;
        const GroupTile gt = tiles[t];
        const int p = __builtin_amdgcn_readfirstlane(gt.problem);
        const GemmArgs g = probs[p].g;
        const int part = __builtin_amdgcn_readfirstlane(gt.part);
        const int tm = __builtin_amdgcn_readfirstlane(gt.tm);
        const int tn = __builtin_amdgcn_readfirstlane(gt.tn);
        if (g.parts > 1)
            f32_tile<BM, BM, true>(g, part, tm, tn, lds);
        else
            f32_tile<BM, BM, false>(g, part, tm, tn, lds);
        __syncthreads();
    }
}

hipError_t launch_gemm_f32_grouped(const GroupProblem* probs, const GroupTile* tiles, int ntiles,
                                   unsigned int* queue, int blocks, int tile, hipStream_t s)
{
    if (tile == 128)
        hipLaunchKernelGGL(gemm3m_f32_grouped_kernel<128>, dim3(blocks), dim3(256), 0, s, probs, tiles, ntiles, queue);
    else
        hipLaunchKernelGGL(gemm3m_f32_grouped_kernel<64>, dim3(blocks), dim3(256), 0, s, probs, tiles, ntiles, queue);
    return hipGetLastError();
}

template <int BM, int BN, bool PARTIAL>
static hipError_t launch_one(const GemmArgs& g, hipStream_t s)
{
    const int blocks = g.tilesM * g.tilesN * g.parts;
    hipLaunchKernelGGL((gemm3m_f32_kernel<BM, BN, PARTIAL>), dim3(blocks), dim3(256), 0, s, g);
    return hipGetLastError();
}

hipError_t launch_gemm_f32(const GemmArgs& g, int bm, int bn, hipStream_t s)
{
    const bool partial = g.parts > 1;
    if (bm == 128 && bn == 128) return partial ? launch_one<128, 128, true>(g, s) : launch_one<128, 128, false>(g, s);
    if (bm == 64 && bn == 64) return partial ? launch_one<64, 64, true>(g, s) : launch_one<64, 64, false>(g, s);
    return hipErrorInvalidValue;
}

}  // namespace c3m
