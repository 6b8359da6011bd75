// epilogue.hip -- HBM-bound output kernels and the FP64 passes.
//
//   merge_kernel     K-split epilogue (reference K9, ops-blas/blas/gemm/arch35/
//                    cgemm_epilogue_kernel.cpp:34-125, :231-332): partitions summed in index
//                    order p = 0..P-1 (no atomics, so reruns are byte-identical), then
//                    alpha*(re, im) + beta*C; C is not read when beta == 0; only rows < m
//                    are written, so ldc padding is never touched.
//   scale_c_kernel   alpha == 0 or k == 0 (reference HandleCgemmAlphaZero,
//                    cgemm_host.cpp:636-663): C = beta*C, or zeros when beta == 0.
//   recompute_kernel rounding correction (reference K5-K8, cgemm3m_rounding_kernel.cpp:
//                    132-497, re-designed): every output flagged by the GEMM epilogue (which
//                    left C's input in place there) is recomputed from the original
//                    complex64 inputs in FP64 by one workgroup, each thread taking a fixed
//                    residue class of k mod 256 in increasing order, each wave's 64 sums
//                    combined by a fixed butterfly and the four wave sums in wave order, then
//                    alpha/beta applied in FP64 and the result rounded once to FP32.
//   residual_kernel  FP64 complex GEMM of the same inputs compared with the FP32 result
//                    (error report for tests and benches only).
#include <algorithm>

#include "device_util.h"
#include "launch.h"

namespace c3m {

// One output (i, j) of a K-split product: partitions summed in index order, then alpha/beta.
// With a.products set, a.partial holds the three real product planes instead (a.partStride
// floats apart), combined exactly as the fused tile combines its accumulators.
__device__ __forceinline__ void merge_one(const MergeArgs& a, int i, int j)
{
    const size_t o = (size_t)j * a.ldp + i;
    const size_t plane = a.partStride;  // floats per partition: re plane then im plane
    const size_t imOff = plane / 2;
    float re = 0.0f, im = 0.0f;
    if (a.products) {
        // separate-products path: each product's K partitions summed in index order (starting
        // from the first, so one partition is the plane itself), then the fused tile's combine
        float pq[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            pq[q] = a.partial[(size_t)q * plane + o];
            for (int sp = 1; sp < a.splits; ++sp) pq[q] += a.partial[(size_t)(3 * sp + q) * plane + o];
        }
        combine3(a.identity, pq[0], pq[1], pq[2], re, im);
    } else {
        for (int p = 0; p < a.parts; ++p) {
            re += a.partial[(size_t)p * plane + o];
            im += a.partial[(size_t)p * plane + imOff + o];
        }
    }
    float2* C = reinterpret_cast<float2*>(a.C);
    const size_t co = (size_t)j * a.ldc + i;
    const bool bz = a.betaZero != 0;
    float2 c0 = make_float2(0.0f, 0.0f);
    if (!bz) c0 = C[co];
    const float2 c = apply_scalars(re, im, a.alphaR, a.alphaI, a.betaR, a.betaI, bz, c0);
    if (!finite2(c)) {
        C[co] = reference_order_output(a.src, i, j, make_float2(a.alphaR, a.alphaI), make_float2(a.betaR, a.betaI), bz,
                                       c0);
        return;
    }
    if (a.correction && needs_recompute(c, a.eA[i], a.eB[j], a.corrScale, a.corrTau)) {
        flag_output(a.bitmap, (long long)j * a.m + i);  // C keeps its input for the recompute
        return;
    }
    C[co] = c;
}

// Rows i0..i0+3 of column j (i0 % 4 == 0, ldp % 4 == 0): the same arithmetic per output as
// merge_one, with the partial or product planes read as 16-byte vectors.
__device__ __forceinline__ void merge_four(const MergeArgs& a, int i0, int j)
{
    typedef float f4 __attribute__((ext_vector_type(4)));
    const size_t o = (size_t)j * a.ldp + i0;
    const size_t plane = a.partStride;
    const size_t imOff = plane / 2;
    const f4* P = reinterpret_cast<const f4*>(a.partial);
    f4 re = {0.0f, 0.0f, 0.0f, 0.0f}, im = {0.0f, 0.0f, 0.0f, 0.0f};
    if (a.products) {
        f4 pq[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            pq[q] = P[((size_t)q * plane + o) / 4];
            for (int sp = 1; sp < a.splits; ++sp) {
                const f4 v = P[((size_t)(3 * sp + q) * plane + o) / 4];
#pragma unroll
                for (int u = 0; u < 4; ++u) pq[q][u] += v[u];
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            float r, q;
            combine3(a.identity, pq[0][u], pq[1][u], pq[2][u], r, q);
            re[u] = r;
            im[u] = q;
        }
    } else {
        for (int p = 0; p < a.parts; ++p) {
            const f4 pr = P[((size_t)p * plane + o) / 4], pi = P[((size_t)p * plane + imOff + o) / 4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                re[u] += pr[u];
                im[u] += pi[u];
            }
        }
    }
    float2* C = reinterpret_cast<float2*>(a.C);
    const size_t co = (size_t)j * a.ldc + i0;
    const bool bz = a.betaZero != 0;
    const int live = a.m - i0 < 4 ? a.m - i0 : 4;
    // Four whole outputs at a 16-byte aligned address (C 16-byte aligned, ldc even): C moves
    // as two 16-byte loads and stores instead of four 8-byte ones.
    const bool wide = live == 4 && a.cWide;
    typedef float f4 __attribute__((ext_vector_type(4)));
    float2 c0[4];
    if (wide) {
        f4 v[2] = {{0.0f, 0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f, 0.0f}};
        if (!bz) {
            v[0] = reinterpret_cast<const f4*>(C + co)[0];
            v[1] = reinterpret_cast<const f4*>(C + co)[1];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) c0[u] = make_float2(v[u >> 1][2 * (u & 1)], v[u >> 1][2 * (u & 1) + 1]);
    } else {
#pragma unroll
        for (int u = 0; u < 4; ++u) c0[u] = (!bz && u < live) ? C[co + u] : make_float2(0.0f, 0.0f);
    }
    // A flagged output gets its input back (what C held, or anything when beta == 0, where the
    // recompute does not read it); a non-finite one is recomputed here in the reference
    // BLAS's order.
    float2 out[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        if (u >= live) break;
        const float2 c = apply_scalars(re[u], im[u], a.alphaR, a.alphaI, a.betaR, a.betaI, bz, c0[u]);
        out[u] = c;
        if (!finite2(c)) {
            out[u] = reference_order_output(a.src, i0 + u, j, make_float2(a.alphaR, a.alphaI),
                                            make_float2(a.betaR, a.betaI), bz, c0[u]);
        } else if (a.correction && needs_recompute(c, a.eA[i0 + u], a.eB[j], a.corrScale, a.corrTau)) {
            flag_output(a.bitmap, (long long)j * a.m + i0 + u);
            out[u] = c0[u];
        }
    }
    if (wide) {
        reinterpret_cast<f4*>(C + co)[0] = f4{out[0].x, out[0].y, out[1].x, out[1].y};
        reinterpret_cast<f4*>(C + co)[1] = f4{out[2].x, out[2].y, out[3].x, out[3].y};
    } else {
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (u < live) C[co + u] = out[u];
    }
}

__global__ __launch_bounds__(256) void merge_kernel(MergeArgs a)
{
    const int i0 = (blockIdx.x * 256 + threadIdx.x) * 4;
    if (i0 >= a.m) return;
    for (int j = blockIdx.y; j < a.n; j += gridDim.y) merge_four(a, i0, j);
}

// Grouped batch: blockIdx.y picks the K-split problem, x walks its m*n outputs.
__global__ __launch_bounds__(256) void merge_grouped_kernel(const GroupMerge* items)
{
    const MergeArgs a = items[blockIdx.y].a;
    const long long total = (long long)a.m * a.n;
    for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < total; t += (long long)gridDim.x * 256)
        merge_one(a, (int)(t % a.m), (int)(t / a.m));
}

// Grouped launches put the member on grid y, which holds at most 65535: larger batches take
// one launch per 65535 members (the members are independent).
constexpr int kMaxGridY = 65535;

hipError_t launch_merge_grouped(const GroupMerge* items, int nitems, int maxOutputs, hipStream_t s)
{
    int bx = (maxOutputs + 255) / 256;
    if (bx > 1024) bx = 1024;
    if (bx < 1) bx = 1;
    if (nitems <= 0) return hipErrorInvalidValue;
    for (int base = 0; base < nitems; base += kMaxGridY) {
        const int ny = std::min(kMaxGridY, nitems - base);
        hipLaunchKernelGGL(merge_grouped_kernel, dim3(bx, ny), dim3(256), 0, s, items + base);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

static unsigned grid_y(int n) { return (unsigned)(n < 65535 ? n : 65535); }

hipError_t launch_merge(const MergeArgs& a, hipStream_t s)
{
    if (a.ldp % 4) return hipErrorInvalidValue;  // planes are padded to 128 rows
    MergeArgs b = a;
    b.cWide = (reinterpret_cast<uintptr_t>(a.C) % 16 == 0 && a.ldc % 2 == 0) ? 1 : 0;
    hipLaunchKernelGGL(merge_kernel, dim3((a.m + 1023) / 1024, grid_y(a.n)), dim3(256), 0, s, b);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void scale_c_kernel(ScaleCArgs a)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.m) return;
    float2* C = reinterpret_cast<float2*>(a.C);
    for (int j = blockIdx.y; j < a.n; j += gridDim.y) {
        const size_t o = (size_t)j * a.ldc + i;
        if (a.betaZero) {
            C[o] = make_float2(0.0f, 0.0f);
        } else {
            // Rounded like the reference-BLAS complex product (no fused multiply-add), so
            // beta*C is bit-exact against it, as the reference's alpha == 0 checks require.
#pragma clang fp contract(off)
            const float2 v = C[o];
            float2 r;
            r.x = a.betaR * v.x - a.betaI * v.y;
            r.y = a.betaR * v.y + a.betaI * v.x;
            C[o] = r;
        }
    }
}

hipError_t launch_scale_c(const ScaleCArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(scale_c_kernel, dim3((a.m + 255) / 256, grid_y(a.n)), dim3(256), 0, s, a);
    return hipGetLastError();
}

// op(X)(row, col): op N reads X(row, col); op T/C reads X(col, row), C conjugated (the sign
// flip is exact, and so is the widening to FP64 at the point of use).
__device__ __forceinline__ float2 load_op(const float2* X, long long ld, int op, long long row, long long col)
{
    const float2 v = (op == 111) ? X[(size_t)col * ld + row] : X[(size_t)row * ld + col];
    return make_float2(v.x, op == 113 ? -v.y : v.y);
}

// The flagged outputs of bitmap word w (bits = its value), recomputed by the whole workgroup.
__device__ __forceinline__ void recompute_word(const RecomputeArgs& a, long long w, unsigned long long bits,
                                               const float2* A, const float2* B, float2* C, double (*part)[2],
                                               int tid, int lane, int wave)
{
    {
        if (a.netlib) {
            // one flagged output per thread, each in the reference BLAS's sequential FP32 order
            if (tid < 64 && ((bits >> tid) & 1ull)) {
                const long long lin = w * 64 + tid;
                const long long i = lin % a.m, j = lin / a.m;
                const float2 c0 = a.betaZero ? make_float2(0.0f, 0.0f) : C[(size_t)j * a.ldc + i];
                C[(size_t)j * a.ldc + i] = netlib_output(A, a.lda, a.opA, B, a.ldb, a.opB, i, j, a.k,
                                                         make_float2(a.alphaRf, a.alphaIf),
                                                         make_float2(a.betaRf, a.betaIf), a.betaZero != 0, c0);
                if (a.counter != nullptr) atomicAdd(a.counter, 1u);
            }
            return;
        }
        while (bits != 0ull) {
            const int b = __ffsll((long long)bits) - 1;
            bits &= bits - 1ull;
            const long long lin = w * 64 + b;
            const long long i = lin % a.m, j = lin / a.m;
            // thread t sums k = t + 256 u in increasing u (the loads of 32 u in flight together,
            // kept in FP32 until used; the adds in u order); each wave's 64 sums by a fixed
            // butterfly; the four wave sums added in wave order
            double sr = 0.0, si = 0.0;
            auto acc = [&](float2 x, float2 y) {
                sr = fma((double)x.x, (double)y.x, sr);
                sr = fma(-(double)x.y, (double)y.y, sr);
                si = fma((double)x.x, (double)y.y, si);
                si = fma((double)x.y, (double)y.x, si);
            };
            long long l = tid;
            for (; l + 256 * 31 < a.k; l += 256 * 32) {
                float2 x[32], y[32];
#pragma unroll
                for (int u = 0; u < 32; ++u) {
                    x[u] = load_op(A, a.lda, a.opA, i, l + 256 * u);   // op(A)(i, l)
                    y[u] = load_op(B, a.ldb, a.opB, l + 256 * u, j);   // op(B)(l, j)
                }
#pragma unroll
                for (int u = 0; u < 32; ++u) acc(x[u], y[u]);
            }
            for (; l < a.k; l += 256) acc(load_op(A, a.lda, a.opA, i, l), load_op(B, a.ldb, a.opB, l, j));
#pragma unroll
            for (int d = 32; d >= 1; d >>= 1) {
                sr += __shfl_xor(sr, d, 64);
                si += __shfl_xor(si, d, 64);
            }
            if (lane == 0) {
                part[wave][0] = sr;
                part[wave][1] = si;
            }
            __syncthreads();
            if (tid == 0) {
                sr = ((part[0][0] + part[1][0]) + part[2][0]) + part[3][0];
                si = ((part[0][1] + part[1][1]) + part[2][1]) + part[3][1];
                double cr = a.alphaR * sr - a.alphaI * si;
                double ci = a.alphaR * si + a.alphaI * sr;
                if (!a.betaZero) {
                    const float2 c0 = C[(size_t)j * a.ldc + i];  // the epilogue left the input there
                    cr += a.betaR * c0.x - a.betaI * c0.y;
                    ci += a.betaR * c0.y + a.betaI * c0.x;
                }
                C[(size_t)j * a.ldc + i] = make_float2((float)cr, (float)ci);
                if (a.counter != nullptr) atomicAdd(a.counter, 1u);
            }
            __syncthreads();  // part[] is rewritten by the next output
        }
    }
}

// span: bitmap words per workgroup and step, 1..256 (from recompute_span).
__device__ __forceinline__ void recompute_body(const RecomputeArgs& a, int span)
{
    __shared__ double part[4][2];
    __shared__ unsigned long long waveMask[4], wordVal[256];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const float2* A = reinterpret_cast<const float2*>(a.A);
    const float2* B = reinterpret_cast<const float2*>(a.B);
    float2* C = reinterpret_cast<float2*>(a.C);
    // Each workgroup scans `span` consecutive bitmap words at a time (one coalesced load, a
    // ballot per wave) and handles the set bits of the nonzero ones in increasing order, all
    // 256 threads on one output. Flagged outputs are rare: the span is chosen so that they
    // spread over many workgroups, since one workgroup's outputs run one after another.
    for (long long base = (long long)blockIdx.x * span; base < a.words; base += (long long)gridDim.x * span) {
        const long long wi = base + tid;
        const unsigned long long mine = (tid < span && wi < a.words) ? a.bitmap[wi] : 0ull;
        const unsigned long long live = __ballot(mine != 0ull);
        wordVal[tid] = mine;
        if (lane == 0) waveMask[wave] = live;
        __syncthreads();
        for (int wv = 0; wv < 4; ++wv) {
            unsigned long long mk = waveMask[wv];
            while (mk != 0ull) {
                const int src = __ffsll((long long)mk) - 1;
                mk &= mk - 1ull;
                const long long w = base + wv * 64 + src;
                recompute_word(a, w, wordVal[wv * 64 + src], A, B, C, part, tid, lane, wave);
            }
        }
        __syncthreads();  // waveMask and wordVal are rewritten by the next group of words
    }
}


__global__ __launch_bounds__(256) void recompute_kernel(RecomputeArgs a, int span) { recompute_body(a, span); }

// Grouped batch: blockIdx.y picks the member (one with the rounding correction on).
__global__ __launch_bounds__(256) void recompute_grouped_kernel(const RecomputeArgs* items, int span)
{
    const RecomputeArgs a = items[blockIdx.y];
    recompute_body(a, span);
}

// Bitmap words per workgroup and step: about `target` workgroups over `words` words, at most
// 256 words (one per thread) and at least 1. With few words (tall-skinny outputs with long k,
// where each flagged output is a long FP64 dot product) every word gets its own workgroup: at
// 256 x 256 x 65536 the 1024 words took 4 workgroups, each recomputing its flagged outputs one
// after another.
static int recompute_span(long long words, long long target)
{
    long long span = (words + target - 1) / target;
    return (int)std::min<long long>(256, std::max<long long>(1, span));
}

hipError_t launch_recompute_grouped(const RecomputeArgs* items, int nitems, long long maxWords, hipStream_t s)
{
    if (nitems <= 0) return hipErrorInvalidValue;
    // about 4096 workgroups over all members
    const int span = recompute_span(maxWords, std::max<long long>(1, 4096 / nitems));
    long long bx = (maxWords + span - 1) / span;
    if (bx > 4096) bx = 4096;
    if (bx < 1) bx = 1;
    for (int base = 0; base < nitems; base += kMaxGridY) {
        const int ny = std::min(kMaxGridY, nitems - base);
        hipLaunchKernelGGL(recompute_grouped_kernel, dim3((unsigned)bx, (unsigned)ny), dim3(256), 0, s, items + base,
                           span);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_recompute(const RecomputeArgs& a, hipStream_t s)
{
    const int span = recompute_span(a.words, 2048);
    long long blocks = (a.words + span - 1) / span;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(recompute_kernel, dim3((unsigned)blocks), dim3(256), 0, s, a, span);
    return hipGetLastError();
}

// FP64 reference: 64 x 64 outputs per 256-thread block, 4 x 4 per thread, k in chunks of
// 16 staged through LDS. Per block it writes [max rel, sum rel, count, max |c-g|, max |g|].
__global__ __launch_bounds__(256) void residual_kernel(ResidualArgs a)
{
    __shared__ double2 As[16][64 + 1];
    __shared__ double2 Bs[16][64 + 1];
    __shared__ double red[5][256];
    const float2* A = reinterpret_cast<const float2*>(a.A);
    const float2* B = reinterpret_cast<const float2*>(a.B);
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const long long i0 = (long long)blockIdx.x * 64, j0 = (long long)blockIdx.y * 64;
    double2 acc[4][4];
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[x][y] = make_double2(0.0, 0.0);
    for (int l0 = 0; l0 < a.k; l0 += 16) {
        for (int e = tid; e < 16 * 64; e += 256) {
            int ll, rr;
            // walk the stored matrix along its contiguous dimension
            if (a.opA == 111) { rr = e & 63; ll = e >> 6; } else { ll = e & 15; rr = e >> 4; }
            const long long gi = i0 + rr, gl = l0 + ll;
            double2 v = make_double2(0.0, 0.0);
            if (gi < a.m && gl < a.k) {
                const float2 f = load_op(A, a.lda, a.opA, gi, gl);
                v = make_double2((double)f.x, (double)f.y);
            }
            As[ll][rr] = v;
            int cc;
            if (a.opB == 111) { ll = e & 15; cc = e >> 4; } else { cc = e & 63; ll = e >> 6; }
            const long long gj = j0 + cc, gl2 = l0 + ll;
            double2 w = make_double2(0.0, 0.0);
            if (gj < a.n && gl2 < a.k) {
                const float2 f = load_op(B, a.ldb, a.opB, gl2, gj);
                w = make_double2((double)f.x, (double)f.y);
            }
            Bs[ll][cc] = w;
        }
        __syncthreads();
#pragma unroll 4
        for (int ll = 0; ll < 16; ++ll) {
            double2 xa[4], yb[4];
#pragma unroll
            for (int x = 0; x < 4; ++x) xa[x] = As[ll][tx + 16 * x];
#pragma unroll
            for (int y = 0; y < 4; ++y) yb[y] = Bs[ll][ty + 16 * y];
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 4; ++y) {
                    acc[x][y].x = fma(xa[x].x, yb[y].x, acc[x][y].x);
                    acc[x][y].x = fma(-xa[x].y, yb[y].y, acc[x][y].x);
                    acc[x][y].y = fma(xa[x].x, yb[y].y, acc[x][y].y);
                    acc[x][y].y = fma(xa[x].y, yb[y].x, acc[x][y].y);
                }
        }
        __syncthreads();
    }
    double mxRel = 0.0, sumRel = 0.0, cnt = 0.0, mxAbs = 0.0, mxRef = 0.0;
    const float2* C = reinterpret_cast<const float2*>(a.C);
    const float2* C0 = reinterpret_cast<const float2*>(a.Cinit);
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) {
            const long long i = i0 + tx + 16 * x, j = j0 + ty + 16 * y;
            if (i < a.m && j < a.n) {
                double gr = a.alphaR * acc[x][y].x - a.alphaI * acc[x][y].y;
                double gi = a.alphaR * acc[x][y].y + a.alphaI * acc[x][y].x;
                if (!a.betaZero) {
                    const float2 c0 = C0[(size_t)j * a.ldc + i];
                    gr += a.betaR * c0.x - a.betaI * c0.y;
                    gi += a.betaR * c0.y + a.betaI * c0.x;
                }
                const float2 c = C[(size_t)j * a.ldc + i];
                const double dr = (double)c.x - gr, di = (double)c.y - gi;
                const double err = sqrt(dr * dr + di * di);
                const double ref = sqrt(gr * gr + gi * gi);
                if (ref > 0.0) {
                    const double rel = err / ref;
                    mxRel = fmax(mxRel, rel);
                    sumRel += rel;
                    cnt += 1.0;
                }
                mxAbs = fmax(mxAbs, err);
                mxRef = fmax(mxRef, ref);
            }
        }
    red[0][tid] = mxRel;
    red[1][tid] = sumRel;
    red[2][tid] = cnt;
    red[3][tid] = mxAbs;
    red[4][tid] = mxRef;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) {
            red[0][tid] = fmax(red[0][tid], red[0][tid + w]);
            red[1][tid] += red[1][tid + w];
            red[2][tid] += red[2][tid + w];
            red[3][tid] = fmax(red[3][tid], red[3][tid + w]);
            red[4][tid] = fmax(red[4][tid], red[4][tid + w]);
        }
        __syncthreads();
    }
    if (tid < 5) a.out[((size_t)blockIdx.y * gridDim.x + blockIdx.x) * 5 + tid] = red[tid][0];
}

hipError_t launch_residual(const ResidualArgs& a, int blocksM, int blocksN, hipStream_t s)
{
    hipLaunchKernelGGL(residual_kernel, dim3(blocksM, blocksN), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace c3m
