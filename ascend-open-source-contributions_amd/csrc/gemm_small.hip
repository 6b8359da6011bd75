// gemm_small.hip -- one-launch path for small problems (v_mfma_f32_16x16x4_f32), gfx950.
//
// For small outputs the pack + product launches and a long per-wave MFMA chain set the time,
// not throughput. This kernel reads the caller's interleaved complex64 A and B directly (no
// workspace, no pack launch) and forms the three 3M plane values in registers. One workgroup
// computes one 16x16 output tile with W waves: with W = 4 they split k (wave w takes the
// 16-deep k chunks w, w+4, w+8, ...), so a wave's MFMA chain is a quarter of the k range, and
// the four partial products are added in LDS in wave order (deterministic) before the
// alpha/beta epilogue; with W = 1 (short k and many output tiles) the one wave takes every
// chunk and the workgroup is a quarter the size.
// Reference role: the on-chip short-K path (ops-blas/blas/gemm/arch35/cgemm_onchip_kernel.h:
// 84-353, which deinterleaves into on-chip buffers and needs no workspace), re-designed for
// MFMA fragments.
//
// MFMA operands: the "A" operand comes from op(B) so that output rows i run along the lanes.
// Within a 16-deep chunk, lane l (t = l/16) holds for MFMA step u the k index
// 16c + 4t + u, so each lane reads 4 CONSECUTIVE k of its row i of op(A) and its column j of
// op(B): 32 contiguous bytes when that operand is stored k-contiguous (A as T/C, B as N),
// and 16-lane runs of 128 contiguous bytes otherwise. The 16x16 result holds output
// (i0 + l%16, j0 + 4*(l/16) + r) in register r. Out-of-range elements read a clamped
// in-range address and are replaced by zero, so no load is conditional.
//
// K split in the same launch (slices > 1): each tile gets `slices` workgroups, slice s taking
// the k chunks [chunks*s/slices, chunks*(s+1)/slices); each writes its partial products (and
// energies) to the workspace, then counts itself in on the tile's arrival counter. The one
// that arrives last adds the slices' partials in slice order (whichever arrived last, so the
// result is deterministic), resets the counter to zero and runs the epilogue below.
//
// Rounding correction in the same launch: the loads that feed the MFMAs also give each row's
// |op(A)_i|^2 and each column's |op(B)_j|^2 (FP64, summed in a fixed order); an output that
// fails the screen (as in the other paths) is recomputed by the whole workgroup in FP64 from
// the inputs (thread t takes k = t, t+256, ...; fixed-order reduction), alpha/beta applied in
// FP64, rounded once. Flagged outputs are recomputed one after another in whatever order the
// workgroup listed them; each one's arithmetic is fixed, so the result is deterministic. An
// output that comes out non-finite is listed too, whether or not the correction is on, and
// recomputed by one thread in the reference BLAS's sequential FP32 order (finite exactly when
// that order's result is).
#include "device_util.h"
#include "launch.h"

namespace c3m {

namespace {

// Marks a listed output (tile-local index 0..255) as non-finite rather than screen-flagged.
constexpr int kNonFinite = 0x100;

// 16-deep k chunks per register set. One: 58 VGPRs (six waves per SIMD) against 93 for two
// (four waves); measured faster on most small shapes (581 x 782 x 175: 18.9 -> 16.1 us,
// 512^3: 21.4 -> 19.3 us); four was no faster than two.
constexpr int kSmallG = 1;

struct SmallSet {
    float2 a[kSmallG][4], b[kSmallG][4];
};

}  // namespace

// Workgroup `block` of a launch over `tiles` output tiles x g.slices: tile block % tiles,
// k slice block / tiles.
template <int W>
__device__ __forceinline__ void small_tile(const SmallArgs& g, int tilesM, int tiles, int block)
{
    constexpr int NT = 64 * W;
    __shared__ float red[W][3][256];
    __shared__ double ered[W][2][16];  // per wave: row energies of op(A), column energies of op(B)
    __shared__ double eTot[2][16];     // the tile's row and column energies (all waves, all slices)
    __shared__ double dred[W][2];
    __shared__ int flagged;
    __shared__ int flagList[256];
    __shared__ float2 flagC0[256];
    __shared__ int lastIn;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int tile = block % tiles, slice = block / tiles;
    const int tm = tile % tilesM, tn = tile / tilesM;
    const int i0 = tm * 16, j0 = tn * 16;
    const int i = i0 + (lane & 15), j = j0 + (lane & 15), t = lane >> 4;
    const bool iIn = i < g.m, jIn = j < g.n;
    const int ic = min(i, g.m - 1), jc = min(j, g.n - 1);
    const float2* A = reinterpret_cast<const float2*>(g.A);
    const float2* B = reinterpret_cast<const float2*>(g.B);
    // op(A)(i, k) = A[i + k*lda] (N) or A[k + i*lda] (T, C); op(B)(k, j) = B[k + j*ldb] (N) or
    // B[j + k*ldb] (T, C): element offset = fixed part + k * step.
    const bool aN = g.opA == 111, bN = g.opB == 111;
    const long long aFix = aN ? ic : (long long)ic * g.lda, aStep = aN ? g.lda : 1;
    const long long bFix = bN ? (long long)jc * g.ldb : jc, bStep = bN ? 1 : g.ldb;
    const float aSign = g.opA == 113 ? -1.0f : 1.0f, bSign = g.opB == 113 ? -1.0f : 1.0f;
    const int K = g.k;
    const int chunks = (K + 15) / 16;
    // this slice's chunks [cs0, cs1) (at least one: the planner keeps slices <= chunks)
    const int cs0 = (int)((long long)chunks * slice / g.slices), cs1 = (int)((long long)chunks * (slice + 1) / g.slices);
    const int kEnd = min(K, cs1 * 16);
    const int iters = (cs1 - cs0 + W - 1) / W;  // chunks of this wave: cs0 + wave, cs0 + wave + W, ...

    double ea = 0.0, eb = 0.0;  // this lane's share of |op(A)_i|^2 and |op(B)_j|^2
    // An operand stored k-contiguous (A as T/C, B as N) gives each lane its four consecutive k
    // as 32 contiguous bytes: two 16-byte loads instead of four 8-byte ones (half the load
    // instructions and cache-line lookups), when all four k are in range; otherwise (and for
    // the mn-contiguous orientation) four clamped 8-byte loads.
    typedef float f4 __attribute__((ext_vector_type(4)));
    auto load4 = [&](const float2* X, long long fix, long long step, bool kContig, int k0, float2 (&v)[4]) {
        if (kContig && k0 + 3 < kEnd) {
            f4 lo, hi;
            __builtin_memcpy(&lo, X + fix + k0, 16);
            __builtin_memcpy(&hi, X + fix + k0 + 2, 16);
            v[0] = make_float2(lo[0], lo[1]);
            v[1] = make_float2(lo[2], lo[3]);
            v[2] = make_float2(hi[0], hi[1]);
            v[3] = make_float2(hi[2], hi[3]);
        } else {
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = X[fix + (long long)min(k0 + u, kEnd - 1) * step];
        }
    };
    auto load = [&](SmallSet& S, int it) {
#pragma unroll
        for (int c = 0; c < kSmallG; ++c) {
            const int k0 = (cs0 + (it * kSmallG + c) * W + wave) * 16 + 4 * t;
            float2 a[4], b[4];
            load4(A, aFix, aStep, !aN, k0, a);
            load4(B, bFix, bStep, bN, k0, b);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const bool kIn = k0 + u < kEnd;
                a[u].x = (kIn && iIn) ? a[u].x : 0.0f;
                a[u].y = (kIn && iIn) ? a[u].y * aSign : 0.0f;
                b[u].x = (kIn && jIn) ? b[u].x : 0.0f;
                b[u].y = (kIn && jIn) ? b[u].y * bSign : 0.0f;
                S.a[c][u] = a[u];
                S.b[c][u] = b[u];
            }
        }
    };
    f32x4 acc[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) acc[q] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    auto mfma = [&](const SmallSet& S) {
        if (g.correction) {
#pragma unroll
            for (int c = 0; c < kSmallG; ++c)
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    ea += (double)S.a[c][u].x * S.a[c][u].x + (double)S.a[c][u].y * S.a[c][u].y;
                    eb += (double)S.b[c][u].x * S.b[c][u].x + (double)S.b[c][u].y * S.b[c][u].y;
                }
        }
#pragma unroll
        for (int c = 0; c < kSmallG; ++c)
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int q = 0; q < 3; ++q)
                    acc[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                        plane_value(1, g.identity, q, S.b[c][u].x, S.b[c][u].y),
                        plane_value(0, g.identity, q, S.a[c][u].x, S.a[c][u].y), acc[q], 0, 0, 0);
    };

    // Two register sets of kSmallG chunks; look-ahead loads are unconditional (past the end
    // they read clamped addresses and yield zeros) and kept ahead of the MFMAs.
    const int sets = (iters + kSmallG - 1) / kSmallG;
    SmallSet s0, s1;
    load(s0, 0);
    int it = 0;
    for (; it + 1 < sets; it += 2) {
        load(s1, it + 1);
        __builtin_amdgcn_sched_barrier(0);
        mfma(s0);
        __builtin_amdgcn_sched_barrier(0);
        load(s0, it + 2);
        __builtin_amdgcn_sched_barrier(0);
        mfma(s1);
        __builtin_amdgcn_sched_barrier(0);
    }
    if (it < sets) mfma(s0);

    // Partial products of the W waves (added in wave order below).
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int r = 0; r < 4; ++r) red[wave][q][(lane & 15) + 16 * (4 * t + r)] = acc[q][r];
    const int tid = threadIdx.x;
    if (g.correction) {
        // lanes l, l+16, l+32, l+48 hold the same row i (column j): add them in that order
        const double ea2 = ea + __shfl_down(ea, 16, 64), eb2 = eb + __shfl_down(eb, 16, 64);
        const double ea4 = ea2 + __shfl_down(ea2, 32, 64), eb4 = eb2 + __shfl_down(eb2, 32, 64);
        if (lane < 16) {
            ered[wave][0][lane] = ea4;
            ered[wave][1][lane] = eb4;
        }
    }
    if (tid == 0) flagged = 0;
    __syncthreads();
    // thread t < 32: the tile's energy of row t (t < 16) or column t - 16, waves in order
    if (g.correction && tid < 32) {
        const int side = tid >> 4, r = tid & 15;
        double v = ered[0][side][r];
#pragma unroll
        for (int w = 1; w < W; ++w) v += ered[w][side][r];
        eTot[side][r] = v;
    }
    auto wave_sum = [&](int q, int o) {
        float v = red[0][q][o];
#pragma unroll
        for (int w = 1; w < W; ++w) v += red[w][q][o];
        return v;
    };
    if (g.slices > 1) {
        // Publish this slice and count in; only the last to arrive goes on. The tiles' slices
        // run on different XCDs, whose L2 caches are not coherent with one another: the
        // partials are written and read with agent-scope atomic stores and loads (coherent
        // across the L2s, no cache flush), and the stores have completed (vmcnt 0) before
        // the workgroup's barrier and its arrival. A device-scope fence would instead write
        // back and invalidate this XCD's whole L2 in every workgroup.
        float* mine = g.part + ((size_t)tile * g.slices + slice) * 768;
        for (int o = tid; o < 256; o += NT)
#pragma unroll
            for (int q = 0; q < 3; ++q)
                __hip_atomic_store(mine + q * 256 + o, wave_sum(q, o), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (g.correction && tid < 32)
            __hip_atomic_store(g.epart + ((size_t)tile * g.slices + slice) * 32 + tid, eTot[tid >> 4][tid & 15],
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (tid == 0) {
            const unsigned int prev =
                __hip_atomic_fetch_add(&g.arrive[tile], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            lastIn = prev == (unsigned int)(g.slices - 1);
            if (lastIn)  // ready for the next launch
                __hip_atomic_store(&g.arrive[tile], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
        if (!lastIn) return;
        // every slice's partials, added in slice order, into red[0] (its wave sums were
        // published above and are not read again)
        const float* all = g.part + (size_t)tile * g.slices * 768;
        for (int o = tid; o < 256; o += NT)
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                float v = __hip_atomic_load(all + q * 256 + o, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                for (int sl = 1; sl < g.slices; ++sl)
                    v += __hip_atomic_load(all + (size_t)sl * 768 + q * 256 + o, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                red[0][q][o] = v;
            }
        if (g.correction && tid < 32) {
            const double* ep = g.epart + (size_t)tile * g.slices * 32;
            double v = __hip_atomic_load(ep + tid, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            for (int sl = 1; sl < g.slices; ++sl)
                v += __hip_atomic_load(ep + (size_t)sl * 32 + tid, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            eTot[tid >> 4][tid & 15] = v;
        }
    }
    __syncthreads();
    float2* C = reinterpret_cast<float2*>(g.C);
    const bool bz = g.betaZero != 0;
    for (int o = tid; o < 256; o += NT) {
        const int io = i0 + (o & 15), jo = j0 + (o >> 4);
        if (io >= g.m || jo >= g.n) continue;
        float p[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) p[q] = g.slices > 1 ? red[0][q][o] : wave_sum(q, o);
        float re, im;
        combine3(g.identity, p[0], p[1], p[2], re, im);
        const size_t co = (size_t)jo * g.ldc + io;
        float2 c0 = make_float2(0.0f, 0.0f);
        if (!bz) c0 = C[co];
        const float2 c = apply_scalars(re, im, g.alphaR, g.alphaI, g.betaR, g.betaI, bz, c0);
        // non-finite outputs (listed with kNonFinite) and outputs the screen flags are
        // recomputed below; the rest are stored now
        const bool bad = !finite2(c);
        const bool flag =
            bad || (g.correction && needs_recompute(c, eTot[0][o & 15], eTot[1][o >> 4], g.corrScale, g.corrTau));
        if (flag) {
            const int slot = atomicAdd(&flagged, 1);
            flagList[slot] = o | (bad ? kNonFinite : 0);
            flagC0[slot] = c0;
        } else {
            C[co] = c;
        }
    }
    __syncthreads();
    const int count = flagged;
    if (count == 0) {
        if (tid == 0 && g.flagCount != nullptr) g.flagCount[tile] = 0u;
        return;
    }
    // the reference BLAS's sequential FP32 order, one output per thread: every non-finite
    // output, and every flagged one under correction 2
    for (int e = tid; e < count; e += NT) {
        const int ft = flagList[e];
        if ((ft & kNonFinite) == 0 && g.correction != 2) continue;
        const long long fi = i0 + (ft & 255 & 15), fj = j0 + ((ft & 255) >> 4);
        C[(size_t)fj * g.ldc + fi] = netlib_output(A, g.lda, g.opA, B, g.ldb, g.opB, fi, fj, K,
                                                   make_float2(g.alphaR, g.alphaI), make_float2(g.betaR, g.betaI), bz,
                                                   flagC0[e]);
    }
    if (tid == 0 && g.flagCount != nullptr) g.flagCount[tile] = (unsigned int)count;
    if (g.correction != 1) return;
    // FP64, the whole workgroup on one flagged output at a time
    for (int e = 0; e < count; ++e) {
        const int ft = flagList[e];
        if (ft & kNonFinite) continue;  // the same for every thread
        const long long fi = i0 + (ft & 15), fj = j0 + (ft >> 4);
        double sr = 0.0, si = 0.0;
        for (long long l = tid; l < K; l += NT) {
            const float2 x = aN ? A[(size_t)l * g.lda + fi] : A[(size_t)fi * g.lda + l];  // op(A)(fi, l)
            const float2 y = bN ? B[(size_t)fj * g.ldb + l] : B[(size_t)l * g.ldb + fj];  // op(B)(l, fj)
            const double xr = x.x, xi = aSign * (double)x.y, yr = y.x, yi = bSign * (double)y.y;
            sr = fma(xr, yr, sr);
            sr = fma(-xi, yi, sr);
            si = fma(xr, yi, si);
            si = fma(xi, yr, si);
        }
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            sr += __shfl_xor(sr, d, 64);
            si += __shfl_xor(si, d, 64);
        }
        if (lane == 0) {
            dred[wave][0] = sr;
            dred[wave][1] = si;
        }
        __syncthreads();
        if (tid == 0) {
            double tr = dred[0][0], ti = dred[0][1];
#pragma unroll
            for (int w = 1; w < W; ++w) {
                tr += dred[w][0];
                ti += dred[w][1];
            }
            double cr = g.alphaR * tr - g.alphaI * ti;
            double ci = g.alphaR * ti + g.alphaI * tr;
            if (!bz) {
                const float2 c0 = flagC0[e];
                cr += g.betaR * c0.x - g.betaI * c0.y;
                ci += g.betaR * c0.y + g.betaI * c0.x;
            }
            C[(size_t)fj * g.ldc + fi] = make_float2((float)cr, (float)ci);
        }
        __syncthreads();
    }
}

template <int W>
__global__ __launch_bounds__(64 * W, W == 4 ? 2 : 8) void gemm3m_small_kernel(SmallArgs g, int tilesM, int tiles)
{
    small_tile<W>(g, tilesM, tiles, blockIdx.x);
}

// Grouped batch: one launch runs the small members of a batch; workgroup b finds its member
// by binary search over the members' first-workgroup prefix (members are unsliced).
template <int W>
__global__ __launch_bounds__(64 * W, W == 4 ? 2 : 8) void gemm3m_small_grouped_kernel(const GroupSmall* items,
                                                                                    int nitems)
{
    const long long b = blockIdx.x;
    int lo = 0, hi = nitems - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].firstBlock <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    const int idx = __builtin_amdgcn_readfirstlane(lo);
    const SmallArgs g = items[idx].a;
    small_tile<W>(g, items[idx].tilesM, items[idx].tiles, (int)(b - items[idx].firstBlock));
}

hipError_t launch_gemm_small_grouped(const GroupSmall* items, int nitems, long long blocks, int waves, hipStream_t s)
{
    if (nitems <= 0 || blocks <= 0 || blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    if (waves == 1)
        hipLaunchKernelGGL(gemm3m_small_grouped_kernel<1>, dim3((unsigned)blocks), dim3(64), 0, s, items, nitems);
    else if (waves == 4)
        hipLaunchKernelGGL(gemm3m_small_grouped_kernel<4>, dim3((unsigned)blocks), dim3(256), 0, s, items, nitems);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

// g.arrive must hold zeros when g.slices > 1 (launch_fill_u32 before the first such launch on
// a buffer; every launch leaves them zero).
hipError_t launch_gemm_small(const SmallArgs& g, hipStream_t s)
{
    const int tilesM = (g.m + 15) / 16, tilesN = (g.n + 15) / 16;
    const long long tiles = (long long)tilesM * tilesN;
    const int chunks = (g.k + 15) / 16;
    if (g.slices < 1 || (g.slices > 1 && (g.slices > chunks || g.part == nullptr || g.arrive == nullptr ||
                                          (g.correction && g.epart == nullptr))) ||
        tiles * g.slices > 0x7fffffffLL)
        return hipErrorInvalidValue;
    if (g.waves == 1)
        hipLaunchKernelGGL(gemm3m_small_kernel<1>, dim3((unsigned)(tiles * g.slices)), dim3(64), 0, s, g, tilesM,
                           (int)tiles);
    else if (g.waves == 4)
        hipLaunchKernelGGL(gemm3m_small_kernel<4>, dim3((unsigned)(tiles * g.slices)), dim3(256), 0, s, g, tilesM,
                           (int)tiles);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace c3m
