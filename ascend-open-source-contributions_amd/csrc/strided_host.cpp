// strided_host.cpp -- aclblasCgemm3mStridedBatched: C_i = alpha*op(A_i)*op(B_i) + beta*C_i for
// i = 0..batchCount-1, with A_i = A + i*strideA, B_i = B + i*strideB, C_i = C + i*strideC
// (strides in complex elements; zero and negative strides allowed).
//
// Reference: declaration ops-blas/include/cann_ops_blas.h:652-655; contract in
// ops-blas/blas/gemm/README.md, section aclblasCgemm3mStridedBatched: checks before the quick
// return (handle, then op -> INVALID_ENUM, then dimensions, leading dimensions, scalar
// pointers); m, n or batchCount == 0 returns SUCCESS without touching memory; k == 0 or
// alpha == 0 only scales by beta (A and B may be null when alpha == 0); beta == 0 never reads
// C; when output batches overlap (strideC == 0 or closer than one output's extent) C is
// updated in batch order; reruns are byte-identical.
//
// Outputs that do not overlap run as grouped batches (one pack launch, one persistent product
// launch per tile size, one merge launch, descriptors cached on the device), in chunks whose
// workspace fits the budget. Overlapping outputs run batch after batch on the handle's
// stream with the same member plan (grouped_member_config), so each batch gets the same bytes
// either way.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "host_internal.h"

namespace c3m {

// Largest workspace one grouped chunk may take unless the caller lent a workspace.
constexpr size_t kStridedChunkBudget = size_t(4) << 30;

static bool is_op(aclblasOperation_t t) { return t == ACLBLAS_OP_N || t == ACLBLAS_OP_T || t == ACLBLAS_OP_C; }

}  // namespace c3m

using namespace c3m;

extern "C" aclblasStatus_t aclblasCgemm3mStridedBatched(aclblasHandle_t handle, aclblasOperation_t transa,
                                                        aclblasOperation_t transb, int m, int n, int k,
                                                        const aclblasComplex* alpha, const aclblasComplex* A,
                                                        int lda, int64_t strideA, const aclblasComplex* B, int ldb,
                                                        int64_t strideB, const aclblasComplex* beta,
                                                        aclblasComplex* C, int ldc, int64_t strideC,
                                                        int batchCount)
{
    const char* op = "aclblasCgemm3mStridedBatched";
    if (handle == nullptr) {
        log_error(op, "no handle was given (create one with aclblasCreate)");
        return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    }
    // Unlike aclblasCgemm3m, this entry reports a bad op as INVALID_ENUM.
    if (!is_op(transa) || !is_op(transb)) {
        log_error(op, "invalid transa/transb: %d/%d", (int)transa, (int)transb);
        return ACLBLAS_STATUS_INVALID_ENUM;
    }
    if (batchCount < 0) {
        log_error(op, "batchCount %d is negative", batchCount);
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    aclblasStatus_t st = validate(handle, transa, transb, m, n, k, alpha, lda, ldb, beta, ldc, op);
    if (st != ACLBLAS_STATUS_SUCCESS) return st;
    if (m == 0 || n == 0 || batchCount == 0) return ACLBLAS_STATUS_SUCCESS;
    st = validate_pointers(is_zero(*alpha) ? 0 : k, A, B, C, op);
    if (st != ACLBLAS_STATUS_SUCCESS) return st;

    const aclblasComplex al = *alpha, be = *beta;
    auto at = [](const aclblasComplex* base, int64_t stride, int i) {
        return base == nullptr ? nullptr : base + (int64_t)i * stride;
    };
    const long long extent = (long long)(n - 1) * ldc + m;  // elements one output spans
    const long long gap = strideC < 0 ? -(long long)strideC : (long long)strideC;
    // One plan for every batch of the call, whichever form runs it: the K-split rule of the
    // grouped members is decided from the whole call's batches (and the device's CUs; the
    // launches below report a missing device).
    if (ensure_device(handle) != ACLBLAS_STATUS_SUCCESS) handle->cus = 0;
    const int fills =
        grouped_batch_fills((long long)batchCount * ((m + 63) / 64) * ((n + 63) / 64), handle->cus) ? 1 : 0;
    const aclblasCgemm3mConfig cfg = grouped_member_config(m, n, k, nullptr, fills != 0);
    if (batchCount > 1 && gap < extent) {
        for (int i = 0; i < batchCount; ++i) {
            st = run_problem(handle, transa, transb, m, n, k, &al, at(A, strideA, i), lda, at(B, strideB, i), ldb, &be,
                             C + (int64_t)i * strideC, ldc, &cfg, op);
            if (st != ACLBLAS_STATUS_SUCCESS) return st;
        }
        return ACLBLAS_STATUS_SUCCESS;
    }

    int chunk = batchCount;
    if (!is_zero(al) && k > 0) {
        size_t one = 0;
        if (aclblasCgemm3mGroupedBatchedGetWorkspaceSize(handle, 1, &transa, &transb, &m, &n, &k,
                                                         ACLBLAS_CGEMM3M_FP32, &one) == ACLBLAS_STATUS_SUCCESS &&
            one > 0) {
            const size_t budget = handle->userWs != nullptr ? handle->userWsBytes : kStridedChunkBudget;
            chunk = (int)std::max<size_t>(1, std::min<size_t>((size_t)batchCount, budget / one));
        }
    }
    std::vector<aclblasOperation_t> ta(chunk, transa), tb(chunk, transb);
    std::vector<int> mm(chunk, m), nn(chunk, n), kk(chunk, k), la(chunk, lda), lb(chunk, ldb), lc(chunk, ldc);
    std::vector<aclblasComplex> av(chunk, al), bv(chunk, be);
    std::vector<const aclblasComplex*> ap(chunk), bp(chunk);
    std::vector<aclblasComplex*> cp(chunk);
    for (int i0 = 0; i0 < batchCount; i0 += chunk) {
        const int cnt = std::min(chunk, batchCount - i0);
        for (int j = 0; j < cnt; ++j) {
            ap[j] = at(A, strideA, i0 + j);
            bp[j] = at(B, strideB, i0 + j);
            cp[j] = C + (int64_t)(i0 + j) * strideC;
        }
        st = grouped_run(handle, cnt, ta.data(), tb.data(), mm.data(), nn.data(), kk.data(), av.data(), ap.data(),
                         la.data(), bp.data(), lb.data(), bv.data(), cp.data(), lc.data(), ACLBLAS_CGEMM3M_FP32,
                         nullptr, op, fills);
        if (st != ACLBLAS_STATUS_SUCCESS) return st;
    }
    return ACLBLAS_STATUS_SUCCESS;
}
