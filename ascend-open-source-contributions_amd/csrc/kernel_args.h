// kernel_args.h -- plain-old-data argument blocks passed by value from the host planner
// to the gfx950 kernels (the role the reference's tiling structs play,
// ops-blas/blas/gemm/arch35/cgemm_tiling_data.h:82-126). Shared by host and device code.
#pragma once
#include <stddef.h>
#include <stdint.h>

namespace c3m {

// ---------------------------------------------------------------------------------------
// Packed plane formats. A "plane" is one real operand of one of the three products, stored
// as a grid of MFMA fragments so that the GEMM's global->LDS copy is a straight contiguous
// copy and each LDS->VGPR read is one conflict-free 16-byte read per lane.
//
//   FP32 (v_mfma_f32_32x32x2_f32): fragment = 32 (mn) x 8 (k) floats = 1 KiB.
//       element (r, c) -> frag (c/8, r/32), lane = r%32 + 32*((c%8)/4), slot = c%4.
//       MFMA step s of a fragment uses slot s: lane half h contributes k = 4h + s for both
//       operands, so the 8 k values of a fragment are summed in 4 MFMA steps.
//   BF16/FP16 (v_mfma_f32_32x32x16_{bf16,f16}): fragment = 32 (mn) x 16 (k) halves = 1 KiB.
//       element (r, c) -> frag (c/16, r/32), lane = r%32 + 32*((c%16)/8), slot = c%8.
// Fragments are laid out k-fragment-major: frag (kf, mf) at index kf*(mnPad/32) + mf, so a
// block's run of consecutive mn fragments for one k fragment is contiguous.
// mnPad is a multiple of 128 and kPad a multiple of 32; padding holds zeros.
// ---------------------------------------------------------------------------------------
enum PlaneFormat : int { FMT_F32 = 0, FMT_BF16 = 1, FMT_F16 = 2 };

// Which three planes an operand contributes (the product q uses A-plane q times B-plane q).
//   fused identity    A: (Ar+Ai, Ar, Ai)   B: (Br, Bi-Br, Br+Bi)   re = P0-P2, im = P0+P1
//   standard identity A: (Ar, Ai, Ar+Ai)   B: (Br, Bi, Br+Bi)      re = P0-P1, im = P2-P0-P1
// (textbook 3M identities; the reference uses both: cgemm_deinterleave_common.h:17-51,
//  cgemm_compute_kernel.cpp:406-428, cgemm_epilogue_kernel.cpp:255-268)
// ID_FOUR: four real products, re = Ar*Br - Ai*Bi, im = Ar*Bi + Ai*Br, accumulated directly
// into (re, im): A planes (Ar, Ai, -Ai), B planes (Br, Bi).
enum Identity : int { ID_FUSED = 1, ID_STANDARD = 2, ID_FOUR = 3 };

struct PackArgs {
    const void* src;      // interleaved complex64, column-major, leading dim ld
    long long ld;
    int mn, kk;           // logical op(X) extent: mn = m for A / n for B, kk = k
    int mnPad, kPad;      // padded extents (mnPad % 128 == 0, kPad % 32 == 0)
    int mnContig;         // 1 if consecutive mn indices are adjacent in src (A:N, B:T/C)
    int conj;             // op == C: negate the imaginary part
    int role;             // 0 = A (left operand), 1 = B (right operand)
    int identity;         // Identity
    int format;           // PlaneFormat
    int pieces;           // BF16: 2 (hi, lo) or 3 (hi, mid, lo); F32/F16: 1
    int nplanes;          // planes written: 3, or 2 for the B operand of ID_FOUR
    void* planes;         // base of plane 0; plane q piece p at planes + (q*pieces+p)*planeBytes
                          // (nullptr: the energy partials only)
    size_t planeBytes;
    double* energyPartial; // optional [kPad/64][mnPad]: sum over 64 k of |x|^2 per row
    const int* scaleExp;   // FP16: per (plane, mn/128, k/128) exponent (from the scale pass)
};

struct ScaleArgs {      // FP16 block-scale pass: power-of-two exponent per 128x128 block
    const void* src;
    long long ld;
    int mn, kk, mnPad, kPad, mnContig, conj, role, identity;
    int* scaleExp;      // [3][mnPad/128][kPad/128]
};

// Block exponent of an all-zero (or all-padding) FP16 block: its products are exactly zero,
// so the product kernels leave the running sum's exponent unchanged across it.
constexpr int kScaleNone = -(1 << 20);
// Summary of one (plane, 128-row block) along k, after snapping: its exponent if every non-zero
// block has the same one, kScaleNone if all are zero, kScaleMixed otherwise.
constexpr int kScaleMixed = -(1 << 20) - 1;

struct ScaleSummaryArgs {  // FP16: snap the block exponents, then summarise each row block
    int* scaleA; int rowsA;         // rows = 3 * mnPad / 128, kBlocks exponents each
    int* scaleB; int rowsB;
    int kBlocks;
    int* sumA; int* sumB;           // one int per row
};

struct EnergyArgs {
    const double* partialA; int mnPadA;
    const double* partialB; int mnPadB;
    int kTiles;           // kPad / 64
    double* eA; double* eB;
    unsigned long long* bitmap; long long bitmapWords;
    unsigned int* counter;
};

// The caller's operands as stored (interleaved complex64), for recomputing single outputs in
// the reference BLAS's order (an output the 3M identity left non-finite).
struct SourceOperands {
    const void* A; long long lda; int opA;  // op: 111 N, 112 T, 113 C
    const void* B; long long ldb; int opB;
    int k;
};

struct GemmArgs {
    const void* Ap; size_t aPlaneBytes; int mPad;   // packed A planes (role 0)
    const void* Bp; size_t bPlaneBytes; int nPad;   // packed B planes (role 1)
    int m, n;
    int kSteps;           // number of k stages (kPad / BK)
    int parts;            // K partitions (1 = fused output)
    int tilesM, tilesN;
    int identity;
    float alphaR, alphaI, betaR, betaI;
    int betaZero;
    void* C; long long ldc;                 // output (fused path)
    float* partial; size_t partStride;      // K-split: partition p at partial + p*partStride,
                                            // re plane then im plane, each mPad*nPad, ld = mPad
    // FP16 block scaling: exponents [3][pad/128][kBlocks128] and their per-row-block summaries
    const int* scaleA; const int* scaleB; int kBlocks128;
    const int* sumA; const int* sumB;
    // rounding correction (fused path only); a flagged output is left unwritten (C keeps its
    // input, which the recompute reads when beta != 0)
    int correction;
    int energyInline;                       // grouped FP32 kernel, reading the operands as stored:
                                            // the energies are formed from its loads (eA, eB unused)
    const double* eA; const double* eB;
    double corrScale;                       // |alpha| * u  (error estimate = corrScale * |a_i||b_j|)
    double corrTau;                         // flag when estimate > corrTau * |c|
    unsigned long long* bitmap;
    SourceOperands src;                     // non-finite outputs are recomputed from these
};

struct SmallArgs {      // one-launch path for small problems: complex inputs read directly
    const void* A; long long lda; int opA;  // op: 111 N, 112 T, 113 C
    const void* B; long long ldb; int opB;
    int m, n, k;
    int identity;
    float alphaR, alphaI, betaR, betaI; int betaZero;
    void* C; long long ldc;
    int correction;                         // screen + in-kernel recompute of flagged outputs:
                                            // 1 FP64, 2 the reference BLAS's FP32 order
    double corrScale, corrTau;
    unsigned int* flagCount;                // optional: flagged outputs per output tile (one slot each)
    // K split inside the launch: slices > 1 workgroups per 16x16 tile, each over its own run of
    // 16-deep k chunks; the last to arrive adds the slices' partials in slice order.
    int slices;
    int waves;                              // waves per tile: 4 (each a quarter of the chunks) or 1
    float* part;                            // [tile][slice][3][256] partial products
    double* epart;                          // [tile][slice][32] partial row / column energies
    unsigned int* arrive;                   // [tile] arrival counters, zero at launch (left zero)
};

struct MergeArgs {      // K-split epilogue: fixed-order partition sum + alpha/beta
    const float* partial; size_t partStride; int parts; int ldp;
    int products;       // 1: partial holds the three real products P0, P1, P2 (one plane each),
                        // per K partition: plane 3 s + q for partition s of `splits`
    int identity;       // Identity used to combine P0..P2 when products == 1
    int splits;         // products == 1: K partitions of each product (>= 1), summed in index order
    int m, n;
    float alphaR, alphaI, betaR, betaI; int betaZero;
    void* C; long long ldc;
    int correction;
    const double* eA; const double* eB; double corrScale, corrTau;
    unsigned long long* bitmap;
    SourceOperands src; // non-finite outputs are recomputed from these
    int cWide;          // set by launch_merge: C is 16-byte aligned and ldc is even
};

struct ScaleCArgs {     // alpha == 0 or k == 0: C = beta*C, or zeros when beta == 0
    int m, n; float betaR, betaI; int betaZero; void* C; long long ldc;
};

struct RecomputeArgs {  // FP64 fixed-order recompute of flagged outputs
    const void* A; long long lda; int opA;
    const void* B; long long ldb; int opB;
    int m, n, k;
    double alphaR, alphaI, betaR, betaI; int betaZero;
    const unsigned long long* bitmap; long long words;
    void* C; long long ldc;             // a flagged output still holds its input (read when beta != 0)
    unsigned int* counter;
    int netlib;             // 1: recompute in the reference BLAS's FP32 order instead of FP64
    float alphaRf, alphaIf, betaRf, betaIf;  // the scalars as given (FP32) for that order
};

struct ResidualArgs {   // FP64 reference GEMM compared with the FP32 result
    const void* A; long long lda; int opA;
    const void* B; long long ldb; int opB;
    int m, n, k;
    double alphaR, alphaI, betaR, betaI; int betaZero;
    const void* Cinit; const void* C; long long ldc;
    double* out;        // per block: [maxRel, sumRel, countRel, maxAbsErr, maxAbsRef]
};

// Grouped batch descriptors (device arrays in the workspace).
struct GroupProblem {
    GemmArgs g;           // complete arguments of this problem's product kernel
    int pad_[4];
};
struct GroupTile {        // one output tile of one problem (and one K partition)
    int problem, part, tm, tn;
};
struct GroupPack {        // one operand of one problem for the grouped pack launch
    PackArgs a;
    long long firstBlock; // exclusive prefix of pack workgroups
};
struct GroupMerge {       // one K-split problem for the grouped merge launch
    MergeArgs a;
};
struct GroupSmall {       // one problem of the grouped one-launch (small-problem) kernel
    SmallArgs a;          // unsliced
    int tilesM, tiles;    // 16x16 output tiles: per column of tiles, in all
    long long firstBlock; // exclusive prefix of workgroups
};

}  // namespace c3m
