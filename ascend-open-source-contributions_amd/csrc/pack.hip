// pack.hip -- deinterleave/op kernel (HBM-bound): interleaved complex64 -> packed planes.
//
// Reference role: K1 gemm_cgemm_deinterleave_kernel (ops-blas/blas/gemm/arch35/
// cgemm_kernel.cpp:137-157, plane formulas cgemm_deinterleave_common.h:17-51): read each
// stored element once, apply op (transpose is a layout choice, conjugation negates the
// imaginary part) and write the three real planes of the 3M identity.
//
// gfx950 design: one 256-thread workgroup per 64 (mn) x 64 (k) tile. The load walks the
// stored matrix along its contiguous dimension (8-byte loads, 512 contiguous bytes per
// wave), so both op = N and op = T/C read coalesced; the tile is turned in LDS; the store
// writes whole 1 KiB MFMA fragments (16 bytes per lane, contiguous across the wave) in the
// format the GEMM kernel copies verbatim. Rows >= mn and k >= kk are written as zeros, so
// the GEMM never bounds-checks its inner loop. Optional outputs: per-row partial energies
// sum |x|^2 over the tile's 64 k (FP64, fixed order) for the rounding-correction screen.
#include <climits>

#include "device_util.h"
#include "launch.h"

namespace c3m {

__device__ __forceinline__ void pack_tile(const PackArgs& a, int rt, int ct, float2 (*tile)[65])
{
    const int tid = threadIdx.x;
    const int r0 = rt * 64, c0 = ct * 64;
    const float2* src = reinterpret_cast<const float2*>(a.src);
    // All 16 loads of this thread are issued before any is used: out-of-range elements read
    // a clamped in-range address and are replaced by zero, so there is no branch around a
    // load (pack only runs with mn, kk >= 1).
    float2 v[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) {
        const int e = tid + 256 * t;
        const int fast = e & 63, slow = e >> 6;
        const int r = r0 + (a.mnContig ? fast : slow);
        const int c = c0 + (a.mnContig ? slow : fast);
        const int rr = min(r, a.mn - 1), cc = min(c, a.kk - 1);
        v[t] = a.mnContig ? src[(size_t)cc * a.ld + rr] : src[(size_t)rr * a.ld + cc];
        const bool in = r < a.mn && c < a.kk;
        v[t].x = in ? v[t].x : 0.0f;
        v[t].y = in ? (a.conj ? -v[t].y : v[t].y) : 0.0f;
    }
#pragma unroll
    for (int t = 0; t < 16; ++t) {
        const int e = tid + 256 * t;
        const int fast = e & 63, slow = e >> 6;
        tile[a.mnContig ? fast : slow][a.mnContig ? slow : fast] = v[t];
    }
    __syncthreads();

    if (a.energyPartial != nullptr) {
        // Energy partials in the half-sum order (energy_acc): thread (row, h) sums each stage's
        // k = 16 s + 8 f + 4 h + u in (f, u) order in FP32, the stages in FP64; partial h of k
        // tile ct goes to energyPartial[h][ct][row].
        if (tid < 128) {
            const int row = tid & 63, h = tid >> 6;
            const int kTiles = (a.kPad + 63) / 64;
            double t = 0.0;
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                float hs = 0.0f;
#pragma unroll
                for (int f = 0; f < 2; ++f)
#pragma unroll
                    for (int u = 0; u < 4; ++u) hs = energy_acc(hs, tile[row][16 * s + 8 * f + 4 * h + u]);
                t += (double)hs;
            }
            a.energyPartial[((size_t)h * kTiles + ct) * a.mnPad + r0 + row] = t;
        }
    }

    if (a.planes == nullptr) return;  // energy partials only (grouped members read raw)
    const size_t mnFrags = (size_t)(a.mnPad / 32);
    if (a.format == FMT_F32) {
        // 8 k fragments x 2 mn fragments x 64 lanes; lane holds 4 consecutive k.
        for (int o = tid; o < 1024; o += 256) {
            const int kfl = o >> 7, mfl = (o >> 6) & 1, lane = o & 63;
            const int kf = (c0 >> 3) + kfl;
            if (kf * 8 >= a.kPad) continue;
            const int rl = mfl * 32 + (lane & 31);
            const int cl = kfl * 8 + 4 * (lane >> 5);
            const size_t idx = ((size_t)kf * mnFrags + (size_t)(r0 >> 5) + mfl) * 64 + lane;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                if (q >= a.nplanes) break;
                f32x4 out;
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const float2 w = tile[rl][cl + s];
                    out[s] = plane_value(a.role, a.identity, q, w.x, w.y);
                }
                reinterpret_cast<f32x4*>(static_cast<char*>(a.planes) + q * a.planeBytes)[idx] = out;
            }
        }
    } else {
        // 4 k fragments x 2 mn fragments x 64 lanes; lane holds 8 consecutive k (16 bytes).
        for (int o = tid; o < 512; o += 256) {
            const int kfl = o >> 7, mfl = (o >> 6) & 1, lane = o & 63;
            const int kf = (c0 >> 4) + kfl;
            if (kf * 16 >= a.kPad) continue;
            const int rl = mfl * 32 + (lane & 31);
            const int cl = kfl * 16 + 8 * (lane >> 5);
            const size_t idx = ((size_t)kf * mnFrags + (size_t)(r0 >> 5) + mfl) * 64 + lane;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                if (q >= a.nplanes) break;
                float x[8];
#pragma unroll
                for (int s = 0; s < 8; ++s) {
                    const float2 w = tile[rl][cl + s];
                    x[s] = plane_value(a.role, a.identity, q, w.x, w.y);
                }
                if (a.format == FMT_BF16) {
                    u32x4 w[3];
                    if (a.pieces == 3)
                        bf16_split8<3>(x, w);
                    else
                        bf16_split8<2>(x, *reinterpret_cast<u32x4(*)[2]>(w));
                    for (int p = 0; p < a.pieces; ++p)
                        reinterpret_cast<u32x4*>(static_cast<char*>(a.planes) +
                                                 (size_t)(q * a.pieces + p) * a.planeBytes)[idx] = w[p];
                } else {
                    const int kb = (c0 + cl) >> 7;
                    const int e = a.scaleExp[((size_t)q * (a.mnPad >> 7) + (r0 >> 7)) * ((a.kPad + 127) >> 7) + kb];
                    f16x8 hv;
#pragma unroll
                    for (int s = 0; s < 8; ++s) hv[s] = (_Float16)ldexpf(x[s], -e);
                    reinterpret_cast<f16x8*>(static_cast<char*>(a.planes) + (size_t)q * a.planeBytes)[idx] = hv;
                }
            }
        }
    }
}

__global__ __launch_bounds__(256) void pack_kernel(PackArgs a, PackArgs b, int blocksA)
{
    __shared__ float2 tile[64][65];
    const bool isA = (int)blockIdx.x < blocksA;
    const PackArgs& p = isA ? a : b;
    const int local = isA ? blockIdx.x : blockIdx.x - blocksA;
    const int rowTiles = p.mnPad / 64;
    pack_tile(p, local % rowTiles, local / rowTiles, tile);
}

static int pack_blocks(const PackArgs& a) { return (a.mnPad / 64) * ((a.kPad + 63) / 64); }

hipError_t launch_pack_pair(const PackArgs& a, const PackArgs& b, hipStream_t s)
{
    const int na = pack_blocks(a), nb = pack_blocks(b);
    hipLaunchKernelGGL(pack_kernel, dim3(na + nb), dim3(256), 0, s, a, b, na);
    return hipGetLastError();
}

hipError_t launch_pack(const PackArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(pack_kernel, dim3(pack_blocks(a)), dim3(256), 0, s, a, a, pack_blocks(a));
    return hipGetLastError();
}

// Grouped batch: one launch packs both operands of every problem (or, for members the quad
// kernel reads raw, only computes their energy partials). Workgroup b finds its
// (problem, operand) item by binary search over the items' first-workgroup prefix; the
// first workgroup also resets the grouped GEMM's two tile queues.
__global__ __launch_bounds__(256) void pack_grouped_kernel(const GroupPack* items, int nitems, unsigned int* queue)
{
    __shared__ float2 tile[64][65];
    const long long b = blockIdx.x;
    if (b == 0 && threadIdx.x < 2) queue[threadIdx.x] = 0u;
    int lo = 0, hi = nitems - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].firstBlock <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    const int idx = __builtin_amdgcn_readfirstlane(lo);
    const PackArgs p = items[idx].a;
    const int local = (int)(b - items[idx].firstBlock);
    const int rowTiles = p.mnPad / 64;
    pack_tile(p, local % rowTiles, local / rowTiles, tile);
}

// Energy partials straight from the stored operands, for plans that read them as stored (the
// K-split members of a grouped batch, whose screen runs in the merge; the BF16 products read
// as stored): the pack kernel's partials (thread (row, h): 32 elements of the 64 x 64 tile in
// the half-sum order of energy_acc), computed without the LDS tile or any plane writes. A
// thread issues its 32 loads at once: rows adjacent in memory make each load a 512-byte run
// across the 64 threads of a half; k-contiguous rows give each thread four runs of 32
// contiguous bytes per 16 k. Out-of-range rows and k count as zero (clamped addresses, no
// branch around a load). 128 threads per tile.
__device__ __forceinline__ void energy_partial_tile(const PackArgs& a, int rt, int ct)
{
    const int tid = threadIdx.x, row = tid & 63, h = tid >> 6;
    const int r = rt * 64 + row, c0 = ct * 64;
    const int rr = min(r, a.mn - 1);
    const float2* src = reinterpret_cast<const float2*>(a.src);
    float2 v[32];  // element (s, f, u) at 8 s + 4 f + u: k = c0 + 16 s + 8 f + 4 h + u
    if (a.mnContig) {
#pragma unroll
        for (int e = 0; e < 32; ++e) {
            const int k = c0 + 16 * (e >> 3) + 8 * ((e >> 2) & 1) + 4 * h + (e & 3);
            v[e] = src[(size_t)min(k, a.kk - 1) * a.ld + rr];
        }
    } else if (c0 + 63 < a.kk) {
        typedef float f4 __attribute__((ext_vector_type(4)));
#pragma unroll
        for (int e = 0; e < 32; e += 2) {
            const int k = c0 + 16 * (e >> 3) + 8 * ((e >> 2) & 1) + 4 * h + (e & 3);
            f4 w;
            __builtin_memcpy(&w, src + (size_t)rr * a.ld + k, 16);
            v[e] = make_float2(w[0], w[1]);
            v[e + 1] = make_float2(w[2], w[3]);
        }
    } else {
#pragma unroll
        for (int e = 0; e < 32; ++e) {
            const int k = c0 + 16 * (e >> 3) + 8 * ((e >> 2) & 1) + 4 * h + (e & 3);
            v[e] = src[(size_t)rr * a.ld + min(k, a.kk - 1)];
        }
    }
    double t = 0.0;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        float hs = 0.0f;
#pragma unroll
        for (int e = 8 * s; e < 8 * s + 8; ++e) {
            const int k = c0 + 16 * (e >> 3) + 8 * ((e >> 2) & 1) + 4 * h + (e & 3);
            const bool in = r < a.mn && k < a.kk;
            hs = energy_acc(hs, in ? v[e] : make_float2(0.0f, 0.0f));
        }
        t += (double)hs;
    }
    const int kTiles = (a.kPad + 63) / 64;
    a.energyPartial[((size_t)h * kTiles + ct) * a.mnPad + r] = t;
}

// Grouped: workgroup b finds its (problem, operand) item by binary search over the items'
// first-workgroup prefix.
__global__ __launch_bounds__(128) void energy_partial_grouped_kernel(const GroupPack* items, int nitems)
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
    const PackArgs a = items[idx].a;
    const int local = (int)(b - items[idx].firstBlock);
    const int rowTiles = a.mnPad / 64;
    energy_partial_tile(a, local % rowTiles, local / rowTiles);
}

// One problem: operands a (blocks 0 .. blocksA - 1) and b.
__global__ __launch_bounds__(128) void energy_partial_kernel(PackArgs a, PackArgs b, int blocksA)
{
    const bool isA = (int)blockIdx.x < blocksA;
    const PackArgs& p = isA ? a : b;
    const int local = isA ? blockIdx.x : blockIdx.x - blocksA;
    const int rowTiles = p.mnPad / 64;
    energy_partial_tile(p, local % rowTiles, local / rowTiles);
}

hipError_t launch_energy_partial_pair(const PackArgs& a, const PackArgs& b, hipStream_t s)
{
    const int na = pack_blocks(a), nb = pack_blocks(b);
    hipLaunchKernelGGL(energy_partial_kernel, dim3(na + nb), dim3(128), 0, s, a, b, na);
    return hipGetLastError();
}

hipError_t launch_energy_partial_grouped(const GroupPack* items, int nitems, long long blocks, hipStream_t s)
{
    hipLaunchKernelGGL(energy_partial_grouped_kernel, dim3((unsigned)blocks), dim3(128), 0, s, items, nitems);
    return hipGetLastError();
}

hipError_t launch_pack_grouped(const GroupPack* items, int nitems, long long blocks, unsigned int* queue,
                               hipStream_t s)
{
    hipLaunchKernelGGL(pack_grouped_kernel, dim3((unsigned)blocks), dim3(256), 0, s, items, nitems, queue);
    return hipGetLastError();
}

// FP16 block scale: exponent e per (plane, 128 mn, 128 k) block with max|x| * 2^-e in
// [2^14, 2^15) over the block's finite values, so its largest finite value keeps full FP16
// precision without overflow; kScaleNone for a block with no finite non-zero value.
// One 1024-thread workgroup per (128 mn, 128 k) block: 16 elements per thread, loaded four at
// a time (clamped in-range addresses, out-of-range elements count as zero, no branch around a
// load), then the maxima reduced within each wave and over the 16 waves (max is order-free).
__global__ __launch_bounds__(1024) void scale_exp_kernel(ScaleArgs a0, ScaleArgs a1)
{
    __shared__ float red[3][16];
    const ScaleArgs& a = blockIdx.z == 0 ? a0 : a1;  // grid z: operand A, then B
    const int mb = blockIdx.x, kb = blockIdx.y;
    if (mb >= (a.mnPad >> 7)) return;  // the grid covers the larger operand's row blocks
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const float2* src = reinterpret_cast<const float2*>(a.src);
    float mx[3] = {0.0f, 0.0f, 0.0f};
    for (int grp = 0; grp < 4; ++grp) {
        float2 v[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int e = threadIdx.x + 1024 * (grp * 4 + t);
            const int fast = e & 127, slow = e >> 7;
            const int r = mb * 128 + (a.mnContig ? fast : slow);
            const int c = kb * 128 + (a.mnContig ? slow : fast);
            const int rr = min(r, a.mn - 1), cc = min(c, a.kk - 1);
            v[t] = a.mnContig ? src[(size_t)cc * a.ld + rr] : src[(size_t)rr * a.ld + cc];
            const bool in = r < a.mn && c < a.kk;
            v[t].x = in ? v[t].x : 0.0f;
            v[t].y = in ? (a.conj ? -v[t].y : v[t].y) : 0.0f;
        }
        // The finite maximum only: an Inf or NaN (an input, or a plane sum that overflows)
        // passes through the scaling as itself and must not set the block's scale, which would
        // leave its finite values unscaled (overflowing FP16 or losing precision).
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const float x = fabsf(plane_value(a.role, a.identity, q, v[t].x, v[t].y));
                mx[q] = fmaxf(mx[q], x <= 3.40282347e38f ? x : 0.0f);
            }
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) mx[q] = fmaxf(mx[q], __shfl_xor(mx[q], off));
        if (lane == 0) red[q][wave] = mx[q];
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        float m = red[threadIdx.x][0];
        for (int w = 1; w < 16; ++w) m = fmaxf(m, red[threadIdx.x][w]);
        int e = m == 0.0f ? kScaleNone : 0;
        if (m > 0.0f && isfinite(m)) {
            int ex;
            frexpf(m, &ex);  // m = f * 2^ex, f in [0.5, 1)
            e = ex - 15;
        }
        const int kBlocks = (a.kPad + 127) >> 7;
        a.scaleExp[((size_t)threadIdx.x * (a.mnPad >> 7) + mb) * kBlocks + kb] = e;
    }
}

hipError_t launch_scale_exponents(const ScaleArgs& a, const ScaleArgs& b, hipStream_t s)
{
    if (a.kPad != b.kPad) return hipErrorInvalidValue;  // one grid over both operands' k blocks
    const int rows = (a.mnPad > b.mnPad ? a.mnPad : b.mnPad) / 128;
    hipLaunchKernelGGL(scale_exp_kernel, dim3(rows, (a.kPad + 127) / 128, 2), dim3(1024), 0, s, a, b);
    return hipGetLastError();
}

__device__ __forceinline__ int wave_max(int v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off));
    return v;
}

__device__ __forceinline__ int wave_min(int v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off));
    return v;
}

// One wave per (plane, 128-row block) of A then of B, after the scale pass and before the pack.
// Snap: a block whose exponent is within kSnapOctaves of the row block's largest takes the
// largest, so block maxima that merely straddle a power of two (most data) give one exponent
// along k; its values stay FP16 normals down to 2^(29 - kSnapOctaves) below its own maximum.
// Summary: that common exponent, kScaleNone if every block is zero, kScaleMixed if blocks more
// than kSnapOctaves apart remain.
constexpr int kSnapOctaves = 8;

__global__ __launch_bounds__(256) void scale_summary_kernel(ScaleSummaryArgs a)
{
    const int lane = threadIdx.x & 63;
    const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
    const bool isA = t < a.rowsA;
    const int row = isA ? t : t - a.rowsA;
    if (!isA && row >= a.rowsB) return;  // whole wave
    int* e = (isA ? a.scaleA : a.scaleB) + (size_t)row * a.kBlocks;
    int top = INT_MIN;
    for (int kb = lane; kb < a.kBlocks; kb += 64)
        if (e[kb] != kScaleNone) top = max(top, e[kb]);
    top = wave_max(top);
    int lo = INT_MAX, hi = INT_MIN;
    for (int kb = lane; kb < a.kBlocks; kb += 64) {
        int x = e[kb];
        if (x == kScaleNone) continue;
        if (x != top && x >= top - kSnapOctaves) {
            x = top;
            e[kb] = x;
        }
        lo = min(lo, x);
        hi = max(hi, x);
    }
    lo = wave_min(lo);
    hi = wave_max(hi);
    if (lane == 0) (isA ? a.sumA : a.sumB)[row] = hi == INT_MIN ? kScaleNone : (lo == hi ? hi : kScaleMixed);
}

hipError_t launch_scale_summary(const ScaleSummaryArgs& a, hipStream_t s)
{
    const int rows = a.rowsA + a.rowsB;  // one wave each, four per workgroup
    hipLaunchKernelGGL(scale_summary_kernel, dim3((rows + 3) / 4), dim3(256), 0, s, a);
    return hipGetLastError();
}

// Row energies e[r] = (sum over k tiles, in index order, of the pack kernel's half-0 partials)
// + (the same for half 1), the half-sum order of energy_acc; also clears the correction
// bitmap and the flagged-output counter for this call. Each half is one sequential chain of
// kTiles additions (the order is fixed, so it cannot be split into a tree). Up to 256 tiles a
// thread runs a row's two chains side by side, 16 loads in flight per batch; longer chains
// (tall-skinny problems: 1024 tiles at k = 65536) would be latency-bound that way, so a
// wave takes the row: every lane loads up to 16 tiles of a block of 1024 at once, then the
// wave adds them in tile order, lane by lane (readlane), the running sum the same in every
// lane. Tiles past kTiles load as zero, which adds nothing (the sums are >= 0).
constexpr int kLongChain = 256;

__device__ __forceinline__ double ordered_sum2(const double* p0, const double* p1, size_t ld, int n)
{
    double s0 = 0.0, s1 = 0.0;
    int kt = 0;
    for (; kt + 8 <= n; kt += 8) {
        double v[8], w[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            v[u] = p0[(size_t)(kt + u) * ld];
            w[u] = p1[(size_t)(kt + u) * ld];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            s0 += v[u];
            s1 += w[u];
        }
    }
    for (; kt < n; ++kt) {
        s0 += p0[(size_t)kt * ld];
        s1 += p1[(size_t)kt * ld];
    }
    return s0 + s1;
}

__device__ __forceinline__ double wave_ordered_sum(const double* p, size_t ld, int n, int lane)
{
    double s = 0.0;
    for (int t0 = 0; t0 < n; t0 += 1024) {
        double v[16];
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            const int t = t0 + lane * 16 + c;
            v[c] = t < n ? p[(size_t)t * ld] : 0.0;
        }
        const int live = min(64, (n - t0 + 15) / 16);  // lanes holding tiles of this block
        for (int l = 0; l < live; ++l) {
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                const int lo = __builtin_amdgcn_readlane(__double2loint(v[c]), l);
                const int hi = __builtin_amdgcn_readlane(__double2hiint(v[c]), l);
                s += __hiloint2double(hi, lo);
            }
        }
    }
    return s;
}

__device__ __forceinline__ void energy_body(const EnergyArgs& a)
{
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long stride = (long long)gridDim.x * 256;
    // the two half partials' tile sums, each in tile order, then half 0 + half 1
    const size_t hA = (size_t)a.kTiles * a.mnPadA, hB = (size_t)a.kTiles * a.mnPadB;
    if (a.kTiles <= kLongChain) {
        for (long long r = t; r < (long long)a.mnPadA + a.mnPadB; r += stride) {
            if (r < a.mnPadA)
                a.eA[r] = ordered_sum2(a.partialA + r, a.partialA + hA + r, a.mnPadA, a.kTiles);
            else
                a.eB[r - a.mnPadA] = ordered_sum2(a.partialB + (r - a.mnPadA), a.partialB + hB + (r - a.mnPadA),
                                                  a.mnPadB, a.kTiles);
        }
    } else {
        const int lane = threadIdx.x & 63;
        const long long rows = (long long)a.mnPadA + a.mnPadB;
        for (long long r = t >> 6; r < rows; r += stride >> 6) {
            const bool isA = r < a.mnPadA;
            const long long rr = isA ? r : r - a.mnPadA;
            const size_t ld = isA ? a.mnPadA : a.mnPadB;
            const double* p = (isA ? a.partialA : a.partialB) + rr;
            const double e = wave_ordered_sum(p, ld, a.kTiles, lane) +
                             wave_ordered_sum(p + (isA ? hA : hB), ld, a.kTiles, lane);
            if (lane == 0) (isA ? a.eA : a.eB)[rr] = e;
        }
    }
    for (long long w = t; w < a.bitmapWords; w += stride) a.bitmap[w] = 0ull;
    if (t == 0 && a.counter != nullptr) *a.counter = 0u;
}

__global__ __launch_bounds__(256) void energy_kernel(EnergyArgs a) { energy_body(a); }

// Grouped batch: blockIdx.y picks the member (one with the rounding correction on).
__global__ __launch_bounds__(256) void energy_grouped_kernel(const EnergyArgs* items)
{
    const EnergyArgs a = items[blockIdx.y];
    energy_body(a);
}

hipError_t launch_energy_grouped(const EnergyArgs* items, int nitems, long long maxWork, hipStream_t s)
{
    long long bx = (maxWork + 255) / 256;
    if (bx > 1024) bx = 1024;
    if (bx < 1) bx = 1;
    if (nitems <= 0) return hipErrorInvalidValue;
    // the member is on grid y (at most 65535): larger batches take one launch per 65535
    for (int base = 0; base < nitems; base += 65535) {
        const int ny = nitems - base < 65535 ? nitems - base : 65535;
        hipLaunchKernelGGL(energy_grouped_kernel, dim3((unsigned)bx, (unsigned)ny), dim3(256), 0, s, items + base);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

long long energy_work(int mnPadA, int mnPadB, int kTiles, long long bitmapWords)
{
    long long work = (long long)mnPadA + mnPadB;  // a thread per row
    if (kTiles > kLongChain) work = 64LL * ((long long)mnPadA + mnPadB);  // a wave per row
    return bitmapWords > work ? bitmapWords : work;
}

hipError_t launch_energy(const EnergyArgs& a, hipStream_t s)
{
    const long long work = energy_work(a.mnPadA, a.mnPadB, a.kTiles, a.bitmapWords);
    long long blocks = (work + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(energy_kernel, dim3((unsigned)blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void fill_u64_kernel(unsigned long long* p, long long words)
{
    for (long long w = (long long)blockIdx.x * 256 + threadIdx.x; w < words; w += (long long)gridDim.x * 256)
        p[w] = 0ull;
}

hipError_t launch_fill_u64(unsigned long long* p, long long words, hipStream_t s)
{
    long long blocks = (words + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(fill_u64_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, words);
    return hipGetLastError();
}

}  // namespace c3m
