// device_util.h -- small device helpers shared by the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernel_args.h"

namespace c3m {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
// 16-byte lane vector as an ext-vector type: arrays of it stay in VGPRs (an array of the
// HIP struct uint4 can be moved to LDS by the compiler, serialising the prefetch loads).
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

// Eight FP32 values split into NP BF16 pieces, each packed as one MFMA operand (element s in
// bits 16 (s % 2) of word s / 2): piece p = bf16(what pieces 0..p-1 left), rounded to nearest
// even by the hardware conversion (v_cvt_pk_bf16_f32; FP32 denormals are kept, as the
// kernels run with denormals on); the remainder x - float(piece) is exact. The pack kernel
// and the kernels reading the operands as stored both split with this, so their planes agree
// bit for bit (NaN payloads aside: a NaN output is recomputed from the stored operands).
template <int NP>
__device__ __forceinline__ void bf16_split8(float (&x)[8], u32x4 (&out)[NP])
{
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
#pragma unroll
    for (int p = 0; p < NP; ++p) {
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            const bf16x2 h = {(__bf16)x[2 * w], (__bf16)x[2 * w + 1]};
            const unsigned u = __builtin_bit_cast(unsigned, h);
            out[p][w] = u;
            if (p + 1 < NP) {
                x[2 * w] -= __uint_as_float(u << 16);
                x[2 * w + 1] -= __uint_as_float(u & 0xffff0000u);
            }
        }
    }
}

// Logical output tile of hardware block `bid`. Workgroups are dealt round-robin to the 8
// XCDs, each with its own L2; remapping gives every XCD one contiguous range of logical
// tiles, and tiles are walked in groups of 8 tile rows so that the tiles in flight on one
// XCD share A row-panels and B column-panels in that L2.
__device__ __forceinline__ void map_tile(int bid, int nblocks, int tilesM, int tilesN, int& part, int& tm,
                                         int& tn)
{
    int logical = bid;
    if ((nblocks & 7) == 0) {
        const int per = nblocks >> 3;
        logical = (bid & 7) * per + (bid >> 3);
    }
    const int T = tilesM * tilesN;
    part = logical / T;
    const int t = logical - part * T;
    constexpr int G = 8;
    const int groupSize = G * tilesN;
    const int g = t / groupSize;
    const int first = g * G;
    const int rows = min(tilesM - first, G);
    const int w = t - g * groupSize;
    tm = first + w % rows;
    tn = w / rows;
}

// One output recomputed in the reference BLAS's (Netlib cgemm's) FP32 order, every complex
// product the naive (ar*br - ai*bi, ar*bi + ai*br) with one rounding per operation:
//   op(A) = N: c = beta*c0 (0 when beta == 0, c0 when beta == 1); for l: c += (alpha*op(B)(l,j))*A(i,l)
//   op(A) = T/C: t = sum_l op(A)(i,l)*op(B)(l,j) in l order; c = alpha*t (+ beta*c0)
// (Netlib reference BLAS 3.10.0 cgemm loop nests, the oracle the reference checks against,
// ops-blas/test/gemm/cgemm3m/cgemm3m_golden.h:46-48.)
__device__ __forceinline__ float2 netlib_cmul(float2 a, float2 b)
{
#pragma clang fp contract(off)
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ float2 netlib_output(const float2* A, long long lda, int opA, const float2* B, long long ldb,
                                                int opB, long long i, long long j, long long k, float2 alpha,
                                                float2 beta, bool betaZero, float2 c0)
{
#pragma clang fp contract(off)
    auto opb = [&](long long l) {
        float2 y = opB == 111 ? B[(size_t)j * ldb + l] : B[(size_t)l * ldb + j];
        if (opB == 113) y.y = -y.y;
        return y;
    };
    const bool betaOne = beta.x == 1.0f && beta.y == 0.0f;
    if (opA == 111) {
        float2 c = betaZero ? make_float2(0.0f, 0.0f) : (betaOne ? c0 : netlib_cmul(beta, c0));
        for (long long l = 0; l < k; ++l) {
            const float2 t = netlib_cmul(alpha, opb(l));
            const float2 p = netlib_cmul(t, A[(size_t)l * lda + i]);
            c.x = c.x + p.x;
            c.y = c.y + p.y;
        }
        return c;
    }
    float2 t = make_float2(0.0f, 0.0f);
    for (long long l = 0; l < k; ++l) {
        float2 x = A[(size_t)i * lda + l];
        if (opA == 113) x.y = -x.y;
        const float2 p = netlib_cmul(x, opb(l));
        t.x = t.x + p.x;
        t.y = t.y + p.y;
    }
    float2 c = netlib_cmul(alpha, t);
    if (!betaZero) {
        const float2 q = netlib_cmul(beta, c0);
        c.x = c.x + q.x;
        c.y = c.y + q.y;
    }
    return c;
}

// Value of plane q for one complex element (re, im) with the imaginary sign already applied
// (role 0: A, 1: B).
__device__ __forceinline__ float plane_value(int role, int identity, int q, float re, float im)
{
    if (identity == ID_STANDARD) {
        return q == 0 ? re : (q == 1 ? im : re + im);
    }
    if (identity == ID_FOUR) return q == 0 ? re : (q == 1 ? im : -im);  // A: (Ar, Ai, -Ai); B: (Br, Bi)
    if (role == 0) return q == 0 ? re + im : (q == 1 ? re : im);  // A: (Ar+Ai, Ar, Ai)
    return q == 0 ? re : (q == 1 ? im - re : re + im);            // B: (Br, Bi-Br, Br+Bi)
}

// The three products of one output combined into (re, im).
__device__ __forceinline__ void combine3(int identity, float p0, float p1, float p2, float& re, float& im)
{
    if (identity == ID_STANDARD) {
        re = p0 - p1;
        im = p2 - p0 - p1;
    } else {
        re = p0 - p2;
        im = p0 + p1;
    }
}

// alpha*(re, im) + beta*c0 in FP32, C never read when beta == 0. Every product and sum is
// rounded on its own (no fused multiply-add: the reference BLAS's order for alpha*temp +
// beta*c), so each kernel that applies the scalars rounds the same way whatever the compiler
// would otherwise contract.
__device__ __forceinline__ float2 apply_scalars(float re, float im, float ar, float ai, float br, float bi,
                                                bool betaZero, float2 c0)
{
#pragma clang fp contract(off)
    float2 c;
    c.x = ar * re - ai * im;
    c.y = ar * im + ai * re;
    if (!betaZero) {
        c.x += br * c0.x - bi * c0.y;
        c.y += br * c0.y + bi * c0.x;
    }
    return c;
}

__device__ __forceinline__ bool finite2(float2 c) { return isfinite(c.x) && isfinite(c.y); }

// An output that came out non-finite (NaN or Inf in a component), recomputed from the stored
// operands in the reference BLAS's FP32 order. The 3M planes can overflow where that order
// does not (A = (3e38, 3e38) times B = (1, 0): Ar + Ai = Inf, the reference gives
// (3e38, 3e38)), and that order can overflow where FP64 does not (two terms of +-4e38 give
// Inf - Inf = NaN), so this keeps an output finite exactly when the reference BLAS's is.
// Every epilogue applies it, whatever the rounding correction does.
__device__ __forceinline__ float2 reference_order_output(const SourceOperands& s, long long i, long long j,
                                                         float2 alpha, float2 beta, bool betaZero, float2 c0)
{
    return netlib_output(reinterpret_cast<const float2*>(s.A), s.lda, s.opA, reinterpret_cast<const float2*>(s.B),
                         s.ldb, s.opB, i, j, s.k, alpha, beta, betaZero, c0);
}

// Squared row norm of op(A) (column norm of op(B)) for the screen below, in one fixed order
// (the half-sum order) that the pack kernels and the grouped FP32 kernel follow, so the plans
// built on them flag the same outputs. Per 16-deep k stage s and half h = 0, 1: the FP32 sum,
// from zero, of the elements' squares (energy_acc) over k = 16 s + 8 f + 4 h + u in (f, u)
// order (f = 0..1, u = 0..3);
// per 64-deep k tile and half: the FP64 sum, from zero, of its four stages' sums in stage
// order; per half, the tiles' partials summed from zero in tile order; the energy = half 0 +
// half 1. (These are the k a lane of the grouped FP32 kernel holds: k = 4 (lane / 32) + u
// within each 8, so that kernel forms the energy from the operands it loads, in registers,
// at two FP32 FMAs per element.) The FP32 squares overflow to Inf for components past
// about 1.8e19, which flags the outputs involved (recomputed in the slow, exact way); they
// lose precision only below about 1e-19, where the products lose it too.
__device__ __forceinline__ float energy_acc(float h, float x, float y) { return fmaf(y, y, fmaf(x, x, h)); }
__device__ __forceinline__ float energy_acc(float h, float2 v) { return energy_acc(h, v.x, v.y); }

// Rounding-correction screen for a finite output: flag it when its FP32 error estimate
// corrScale * |op(A)_i|_2 * |op(B)_j|_2 exceeds corrTau * |c|. Both sides are non-negative,
// so their squares are compared (ea, eb are the squared norms): no square root per output; a
// NaN or infinite estimate still flags.
__device__ __forceinline__ bool needs_recompute(float2 c, double ea, double eb, double corrScale, double corrTau)
{
    const double mag2 = (double)c.x * c.x + (double)c.y * c.y;
    const double est2 = (corrScale * corrScale) * (ea * eb);
    return !(est2 <= (corrTau * corrTau) * mag2);
}

__device__ __forceinline__ void flag_output(unsigned long long* bitmap, long long lin)
{
    atomicOr(&bitmap[lin >> 6], 1ull << (lin & 63));
}

// Output stage shared by the product kernels. acc[q][a][b] is the 32x32 FP32 tile of
// product q for the wave's a-th column tile and b-th row tile; the lane holds output row
// iBase + 32*b and, in register r, column jBase + 32*a + 8*(r/4) + (r%4).
// PARTIAL: write this K partition's (re, im) planes (ld = mPad, padded extent, no checks).
// Otherwise: alpha/beta in registers, C read only when beta != 0, rows >= m and columns
// >= n never touched. An output flagged by the rounding-correction screen is not written
// (the recompute pass reads C's input there); a non-finite one is recomputed by its lane in
// the reference BLAS's order once the tile is stored.
template <int TN, int TM, bool PARTIAL, int NACC = 3, bool OWN_ENERGY = false>
__device__ __forceinline__ void store_tile(const GemmArgs& g, f32x16 (&acc)[NACC][TN][TM], int part, int iBase,
                                           int jBase, const double* eaIn = nullptr,
                                           const __attribute__((address_space(3))) double* ebIn = nullptr)
{
    // OWN_ENERGY: the grouped kernel's own energies, eaIn[b] for row iBase + 32 b and ebIn[t]
    // (in LDS, read with LDS loads) for column jBase + t; else eA / eB from the workspace. A
    // compile-time choice, not a null test: an LDS address can be 0.
    static_assert(TN * TM * 16 <= 64, "one bit per output of the lane");
    // NACC == 3: the three products, combined by the 3M identity; NACC == 2: (re, im) as
    // accumulated by the four-product form.
    auto combine = [&](int a, int b, int r, float& re, float& im) {
        if constexpr (NACC == 2) {
            re = acc[0][a][b][r];
            im = acc[NACC - 1][a][b][r];
        } else {
            combine3(g.identity, acc[0][a][b][r], acc[1][a][b][r], acc[NACC - 1][a][b][r], re, im);
        }
    };
    if (PARTIAL) {
        float* pr = g.partial + (size_t)part * g.partStride;
        float* pi = pr + (size_t)g.mPad * g.nPad;
#pragma unroll
        for (int a = 0; a < TN; ++a)
#pragma unroll
            for (int b = 0; b < TM; ++b)
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int i = iBase + b * 32;
                    const int j = jBase + a * 32 + 8 * (r >> 2) + (r & 3);
                    float re, im;
                    combine(a, b, r, re, im);
                    const size_t o = (size_t)j * g.mPad + i;
                    pr[o] = re;
                    pi[o] = im;
                }
    } else {
        // One 32-column slice (a) at a time: every load the slice needs (old C when beta != 0,
        // the row and column energies when the screen is on) is issued first, from clamped
        // in-range addresses, and waited for once; a load inside the per-output branch would
        // cost one full memory round trip per output.
        float2* C = reinterpret_cast<float2*>(g.C);
        const bool bz = g.betaZero != 0;
        unsigned long long bad = 0ull;  // this lane's non-finite outputs, by (a, b, r)
        double ea[TM];
        if (g.correction) {
#pragma unroll
            for (int b = 0; b < TM; ++b) ea[b] = OWN_ENERGY ? eaIn[b] : g.eA[iBase + b * 32];  // mPad rows
        }
#pragma unroll
        for (int a = 0; a < TN; ++a) {
            float2 c0[TM][16];
            double eb[16];
            if (!bz) {
#pragma unroll
                for (int b = 0; b < TM; ++b)
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const int ic = min(iBase + b * 32, g.m - 1);
                        const int jc = min(jBase + a * 32 + 8 * (r >> 2) + (r & 3), g.n - 1);
                        c0[b][r] = C[(size_t)jc * g.ldc + ic];
                    }
            } else {
#pragma unroll
                for (int b = 0; b < TM; ++b)
#pragma unroll
                    for (int r = 0; r < 16; ++r) c0[b][r] = make_float2(0.0f, 0.0f);
            }
            if (g.correction) {
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int t = a * 32 + 8 * (r >> 2) + (r & 3);
                    eb[r] = OWN_ENERGY ? ebIn[t] : g.eB[jBase + t];  // nPad columns
                }
            }
#pragma unroll
            for (int b = 0; b < TM; ++b)
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int i = iBase + b * 32;
                    const int j = jBase + a * 32 + 8 * (r >> 2) + (r & 3);
                    if (i < g.m && j < g.n) {
                        float re, im;
                        combine(a, b, r, re, im);
                        const size_t o = (size_t)j * g.ldc + i;
                        const float2 c = apply_scalars(re, im, g.alphaR, g.alphaI, g.betaR, g.betaI, bz, c0[b][r]);
                        if (!finite2(c))
                            bad |= 1ull << ((a * TM + b) * 16 + r);
                        else if (g.correction && needs_recompute(c, ea[b], eb[r], g.corrScale, g.corrTau))
                            flag_output(g.bitmap, (long long)j * g.m + i);
                        else
                            C[o] = c;
                    }
                }
        }
        // Non-finite outputs (rare): C still holds the input there.
        while (bad != 0ull) {
            const int t = __builtin_ctzll(bad);
            bad &= bad - 1ull;
            const int r = t & 15, b = (t >> 4) % TM, a = (t >> 4) / TM;
            const int i = iBase + b * 32;
            const int j = jBase + a * 32 + 8 * (r >> 2) + (r & 3);
            const size_t o = (size_t)j * g.ldc + i;
            const float2 c0 = bz ? make_float2(0.0f, 0.0f) : C[o];
            C[o] = reference_order_output(g.src, i, j, make_float2(g.alphaR, g.alphaI), make_float2(g.betaR, g.betaI),
                                          bz, c0);
        }
    }
}

}  // namespace c3m
