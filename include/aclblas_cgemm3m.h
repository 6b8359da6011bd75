/*
 * aclblas_cgemm3m.h -- C ABI of the MI355X (gfx950) complex 3M GEMM library.
 *
 * The drop-in surface is the reference's own:
 *   - types, status codes and operation enums   ops-blas/include/cann_ops_blas.h:22-25,
 *                                               ops-blas/docs/zh/api_list.md:31-42, :77-79
 *   - handle functions                          ops-blas/include/cann_ops_blas.h:32-110
 *   - aclblasCgemm3m                            ops-blas/include/cann_ops_blas.h:497-500,
 *                                               ops-blas/blas/gemm/arch35/cgemm3m_host.cpp:17-23
 *   - aclblasCgemm (four products)              ops-blas/include/cann_ops_blas.h:492-495
 *   - aclblasCgemm3mStridedBatched              ops-blas/include/cann_ops_blas.h:652-655
 * Everything after the "extensions" banner has no counterpart in the reference; each
 * declaration says which caller it is for.
 *
 * Conventions (same as the reference): matrices are column-major complex64 in device
 * memory owned by the caller; alpha/beta are HOST pointers read during the call; work
 * is enqueued on the handle's stream and the call returns before it completes.
 * One handle must not be used by two host threads at once.
 */
#ifndef ACLBLAS_CGEMM3M_H_
#define ACLBLAS_CGEMM3M_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ops-blas/docs/zh/api_list.md:31-42 (values are part of the ABI). */
typedef enum {
    ACLBLAS_STATUS_SUCCESS = 0,
    ACLBLAS_STATUS_NOT_INITIALIZED = 1,
    ACLBLAS_STATUS_ALLOC_FAILED = 2,
    ACLBLAS_STATUS_INVALID_VALUE = 3,
    ACLBLAS_STATUS_MAPPING_ERROR = 4,
    ACLBLAS_STATUS_EXECUTION_FAILED = 5,
    ACLBLAS_STATUS_INTERNAL_ERROR = 6,
    ACLBLAS_STATUS_NOT_SUPPORTED = 7,
    ACLBLAS_STATUS_ARCH_MISMATCH = 8,
    ACLBLAS_STATUS_HANDLE_IS_NULLPTR = 9,
    ACLBLAS_STATUS_INVALID_ENUM = 10,
    ACLBLAS_STATUS_UNKNOWN = 11
} aclblasStatus_t;

/* ops-blas/docs/zh/api_list.md:77-79. */
typedef enum {
    ACLBLAS_OP_N = 111,
    ACLBLAS_OP_T = 112,
    ACLBLAS_OP_C = 113
} aclblasOperation_t;

/* Interleaved complex64: 8 bytes, 4-byte aligned (fields used at
 * ops-blas/blas/gemm/arch35/cgemm_host.cpp:719-720). */
typedef struct {
    float real;
    float imag;
} aclblasComplex;

/* ops-blas/include/cann_ops_blas.h:22-25: opaque handle. */
typedef struct _aclblas_handle* aclblasHandle_t;
typedef struct _aclblas_handle* aclblasHandle;

/* The reference's aclrtStream; here it carries a hipStream_t (nullptr = default stream). */
typedef void* aclrtStream;

/* ops-blas/include/cann_ops_blas.h:23: log sink for error text. */
typedef void (*aclblasLogCallback)(const char*);

/* ops-blas/include/cann_ops_blas.h:32-45. *handle must be nullptr on entry. Reserves the
 * 32 MiB default workspace on the current device (ACLBLAS_STATUS_ALLOC_FAILED if it cannot).
 * On a host without a HIP device the handle is still created (argument validation works)
 * and the workspace is reserved by the first call that needs one. */
aclblasStatus_t aclblasCreate(aclblasHandle_t* handle);

/* ops-blas/include/cann_ops_blas.h:47-58. Synchronises the stream, frees the default
 * workspace; a user workspace is forgotten, not freed. */
aclblasStatus_t aclblasDestroy(aclblasHandle_t handle);

/* ops-blas/include/cann_ops_blas.h:60-71. Switching streams reverts to the default workspace. */
aclblasStatus_t aclblasSetStream(aclblasHandle_t handle, aclrtStream stream);

/* ops-blas/include/cann_ops_blas.h:73-82. */
aclblasStatus_t aclblasGetStream(aclblasHandle_t handle, aclrtStream* stream);

/* ops-blas/include/cann_ops_blas.h:84-98. Borrowed, grow-only. When a call needs more
 * than the user workspace holds it fails with ACLBLAS_STATUS_EXECUTION_FAILED
 * (ops-blas/docs/zh/api_list.md:203); query the need with aclblasCgemm3mGetWorkspaceSize. */
aclblasStatus_t aclblasSetWorkspace(aclblasHandle_t handle, void* workspace, size_t workspaceSize);

/* ops-blas/include/cann_ops_blas.h:100-110. MAJOR*10000 + MINOR*100 + PATCH. */
aclblasStatus_t aclblasGetVersion(aclblasHandle_t handle, int* version);

/* C <- alpha*op(A)*op(B) + beta*C with three real products.
 * ops-blas/include/cann_ops_blas.h:497-500; contract ops-blas/blas/gemm/README.md:909-941;
 * validation order ops-blas/blas/gemm/arch35/cgemm_host.cpp:589-634, :687-730. */
aclblasStatus_t aclblasCgemm3m(aclblasHandle_t handle, aclblasOperation_t transa, aclblasOperation_t transb,
                               int m, int n, int k, const aclblasComplex* alpha, const aclblasComplex* A, int lda,
                               const aclblasComplex* B, int ldb, const aclblasComplex* beta, aclblasComplex* C,
                               int ldc);

/* C <- alpha*op(A)*op(B) + beta*C with four real products (re = Ar*Br - Ai*Bi,
 * im = Ar*Bi + Ai*Br). ops-blas/include/cann_ops_blas.h:492-495; contract
 * ops-blas/blas/gemm/README.md (aclblasCgemm section); entry
 * ops-blas/blas/gemm/arch35/cgemm_host.cpp:732-738. Unlike aclblasCgemm3m: a bad op returns
 * ACLBLAS_STATUS_INVALID_ENUM, and A and B must be non-null whenever k > 0 (alpha == 0
 * included). */
aclblasStatus_t aclblasCgemm(aclblasHandle_t handle, aclblasOperation_t transa, aclblasOperation_t transb, int m,
                             int n, int k, const aclblasComplex* alpha, const aclblasComplex* A, int lda,
                             const aclblasComplex* B, int ldb, const aclblasComplex* beta, aclblasComplex* C,
                             int ldc);

/* C_i <- alpha*op(A_i)*op(B_i) + beta*C_i, i = 0..batchCount-1, with A_i = A + i*strideA,
 * B_i = B + i*strideB, C_i = C + i*strideC (complex elements; zero and negative strides allowed).
 * ops-blas/include/cann_ops_blas.h:652-655; contract ops-blas/blas/gemm/README.md, section
 * aclblasCgemm3mStridedBatched: a bad op returns ACLBLAS_STATUS_INVALID_ENUM (not remapped);
 * m, n or batchCount == 0 returns SUCCESS after the argument checks; overlapping output
 * batches (for example strideC == 0) are updated in batch order. */
aclblasStatus_t aclblasCgemm3mStridedBatched(aclblasHandle_t handle, aclblasOperation_t transa,
                                             aclblasOperation_t transb, int m, int n, int k,
                                             const aclblasComplex* alpha, const aclblasComplex* A, int lda,
                                             int64_t strideA, const aclblasComplex* B, int ldb, int64_t strideB,
                                             const aclblasComplex* beta, aclblasComplex* C, int ldc,
                                             int64_t strideC, int batchCount);

/* ========================= extensions (no reference counterpart) ========================= */

/* Arithmetic used for the three real products. FP32 is what aclblasCgemm3m runs. */
typedef enum {
    ACLBLAS_CGEMM3M_FP32 = 0,        /* FP32 MFMA, FP32 accumulation (+ FP64 correction pass) */
    ACLBLAS_CGEMM3M_BF16X3 = 1,      /* each FP32 plane split hi+lo BF16; hi*hi + hi*lo + lo*hi */
    ACLBLAS_CGEMM3M_FP16_SCALED = 2, /* FP16 planes with a power-of-two scale per 128x128 block. Bounds
                                        the NORMWISE error only (about 5e-4 of max|C| at 8192^3 on
                                        uniform data): an output much smaller than the largest of its
                                        block can carry a large relative error (up to 0.66 seen) */
    ACLBLAS_CGEMM3M_BF16X6 = 3       /* three-piece BF16 split, six products (more accurate) */
} aclblasCgemm3mPrecision_t;

/* Planner overrides; 0 everywhere means "let the planner/autotuned table decide". */
typedef struct {
    int precision;  /* aclblasCgemm3mPrecision_t */
    int tile;       /* 0 auto, 64 or 128: output tile edge */
    int kSplit;     /* 0 auto, else 1..16 K partitions reduced in fixed order (1..128 for the
                       BF16 separate products, variants 4 and 7) */
    int correction; /* 0 auto (on when k >= 1024), 1 on: flagged outputs recomputed in FP64,
                       2 on: flagged outputs recomputed in the reference BLAS's FP32 order
                       (Netlib cgemm, bit-identical to it for those outputs), -1 off */
    int identity;   /* 0 auto (fused identity), 1 fused, 2 standard 3M identity, 3 four real
                       products (FP32 only; what aclblasCgemm uses) */
    int variant;    /* 0 auto; 128x128 tiles: 1 four waves through LDS, 2 eight waves through
                       LDS (two per SIMD), 3 four waves loading fragments straight into
                       registers, no LDS (FP32 only); 4 separate products: each real product
                       on its own 256x256 tiles, then a combine pass (16-bit precisions);
                       5 one-launch small-problem kernel reading A and B directly, no
                       workspace (FP32 3M, no K split); 6 separate products with the
                       fragments loaded once per workgroup into an LDS ring by direct
                       global-to-LDS loads (16-bit precisions; measured slower than 4);
                       7 separate products read from the stored A and B, no pack launch
                       (BF16 precisions, kSplit K partitions; chosen by default for
                       tall-skinny and small outputs; the same bytes as 4 with the same
                       partitions). Variants 4 and 7 take kSplit on BF16 precisions only. */
} aclblasCgemm3mConfig;

/* For callers that choose the arithmetic or pin the plan (bench, tests, autotuner, the
 * torch autograd wrapper). Same argument contract as aclblasCgemm3m; config may be null. */
aclblasStatus_t aclblasCgemm3mEx(aclblasHandle_t handle, aclblasOperation_t transa, aclblasOperation_t transb,
                                 int m, int n, int k, const aclblasComplex* alpha, const aclblasComplex* A,
                                 int lda, const aclblasComplex* B, int ldb, const aclblasComplex* beta,
                                 aclblasComplex* C, int ldc, const aclblasCgemm3mConfig* config);

/* For callers that must size a user workspace up front (hipGraph capture, serving loops):
 * bytes the call with these arguments would need. */
aclblasStatus_t aclblasCgemm3mGetWorkspaceSize(aclblasHandle_t handle, aclblasOperation_t transa,
                                               aclblasOperation_t transb, int m, int n, int k,
                                               const aclblasComplex* alpha, const aclblasComplex* beta,
                                               const aclblasCgemm3mConfig* config, size_t* bytes);

/* For callers with many small independent problems: `count` problems in one persistent
 * launch that reads their operands as stored (plus, if any problem has the rounding
 * correction on, one energy launch before it and one recompute launch after it, and if any
 * is K-split, one merge launch). In a batch large enough to fill the GPU without K splits,
 * members with the correction on take none unless the call pins one; problems the autotuned
 * table gives to the one-launch kernel join the persistent launch when the batch's other
 * problems carry at least as much work, and otherwise run in one more launch per form, on
 * a second stream of the handle joined back before the call returns.
 * All arrays are HOST arrays of length count; A/B/C entries are device pointers.
 * Per-problem validation follows aclblasCgemm3m; the first invalid problem's status is
 * returned and nothing is launched. precision is an aclblasCgemm3mPrecision_t. */
aclblasStatus_t aclblasCgemm3mGroupedBatched(aclblasHandle_t handle, int count, const aclblasOperation_t* transa,
                                             const aclblasOperation_t* transb, const int* m, const int* n,
                                             const int* k, const aclblasComplex* alpha,
                                             const aclblasComplex* const* A, const int* lda,
                                             const aclblasComplex* const* B, const int* ldb,
                                             const aclblasComplex* beta, aclblasComplex* const* C, const int* ldc,
                                             int precision);

/* For the autotuner and callers that pin plans: aclblasCgemm3mGroupedBatched with one
 * config applied to every member (precision; tile and kSplit when nonzero; identity;
 * correction when nonzero, else on for members with k >= 1024 as in aclblasCgemm3m; variant 5
 * puts every member on the one-launch kernel, variant 7 packs the persistent kernel's members
 * into planes with one launch first instead of reading their operands as stored: the same
 * bytes, kept for comparison; the standard identity always takes that packed form). */
aclblasStatus_t aclblasCgemm3mGroupedBatchedEx(aclblasHandle_t handle, int count, const aclblasOperation_t* transa,
                                               const aclblasOperation_t* transb, const int* m, const int* n,
                                               const int* k, const aclblasComplex* alpha,
                                               const aclblasComplex* const* A, const int* lda,
                                               const aclblasComplex* const* B, const int* ldb,
                                               const aclblasComplex* beta, aclblasComplex* const* C, const int* ldc,
                                               const aclblasCgemm3mConfig* config);

/* Workspace bytes aclblasCgemm3mGroupedBatched needs for these problems: the larger of its two
 * forms' needs (members read as stored, or packed into planes first when a leading dimension
 * is too wide for the first form's 32-bit offsets; only the call sees the leading dimensions). */
aclblasStatus_t aclblasCgemm3mGroupedBatchedGetWorkspaceSize(aclblasHandle_t handle, int count,
                                                             const aclblasOperation_t* transa,
                                                             const aclblasOperation_t* transb, const int* m,
                                                             const int* n, const int* k, int precision,
                                                             size_t* bytes);

/* Error report for benches and tests: FP64 complex GEMM of the same inputs on the GPU,
 * compared with C (the FP32 result already computed from C0 = Cinit).
 * out[0] = max |c - g| / |g| over g != 0 (elementwise, complex modulus)
 * out[1] = mean of the same
 * out[2] = max |c - g| / max |g|       (normwise)
 * out[3] = max |g|
 * Cinit is the C before the GEMM (may be null when beta == 0). Synchronises the stream. */
aclblasStatus_t aclblasCgemm3mResidual(aclblasHandle_t handle, aclblasOperation_t transa, aclblasOperation_t transb,
                                       int m, int n, int k, const aclblasComplex* alpha, const aclblasComplex* A,
                                       int lda, const aclblasComplex* B, int ldb, const aclblasComplex* beta,
                                       const aclblasComplex* Cinit, const aclblasComplex* C, int ldc,
                                       double* out);

/* Diagnostics of the last aclblasCgemm3m/Ex call on this handle (host-side plan; the
 * flagged count needs a stream sync and is read only when `sync` is nonzero).
 * info[0] path (0 none, 1 fused-tile, 2 K-split, 3 alpha-zero/beta scale)
 * info[1] tile edge, info[2] K partitions, info[3] correction enabled,
 * info[4] precision, info[5] identity (1 fused, 2 standard), info[6] flagged outputs
 * (-1 if not read), info[7] workspace bytes used. */
aclblasStatus_t aclblasCgemm3mGetLastPlan(aclblasHandle_t handle, int sync, long long* info);

/* For benches and profilers: while enabled, every later call on this handle records a HIP
 * event pair around each kernel stage on the handle's stream. */
aclblasStatus_t aclblasCgemm3mSetTiming(aclblasHandle_t handle, int enable);

/* Milliseconds per stage summed over the calls recorded since the last read, then reset:
 * ms[0] pack, ms[1] energy, ms[2] products (GEMM), ms[3] merge, ms[4] recompute,
 * ms[5] beta scale; *calls = number of calls. Synchronises the recorded events. */
aclblasStatus_t aclblasCgemm3mGetTiming(aclblasHandle_t handle, double* ms, long long* calls);

/* No counterpart in the reference. For callers that cannot allocate device memory during
 * the call (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): aclblasCgemm3m with the same arguments, statuses, plan and results, whose
 * temporary storage is a fixed 64 MiB device array per GPU (held from the first such call
 * until the process ends) instead of the handle's workspace. It calls no allocator; a call
 * that needs more than the array holds returns ACLBLAS_STATUS_NOT_SUPPORTED and launches
 * nothing (aclblasCgemm3mGetWorkspaceSize with a null config gives the need). Calls on one GPU
 * share its array: two calls that overlap on one GPU (two streams, two threads) race. */
aclblasStatus_t aclblasCgemm3mNoAlloc(aclblasHandle_t handle, aclblasOperation_t transa, aclblasOperation_t transb,
                                      int m, int n, int k, const aclblasComplex* alpha, const aclblasComplex* A,
                                      int lda, const aclblasComplex* B, int ldb, const aclblasComplex* beta,
                                      aclblasComplex* C, int ldc);

/* Route error text to a callback (null restores stderr). Process-wide. */
aclblasStatus_t aclblasSetLogCallback(aclblasLogCallback callback);

#ifdef __cplusplus
}
#endif

#endif /* ACLBLAS_CGEMM3M_H_ */
