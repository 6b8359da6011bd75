// noalloc.hip -- aclblasCgemm3mNoAlloc: aclblasCgemm3m for callers that cannot allocate device
// memory during the call (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc). No counterpart in the reference.
//
// Same arguments, validation, plan, kernels and results as aclblasCgemm3m; the only difference
// is where the call's temporary storage (packed planes, K-split partials, energies, the
// correction bitmap, arrival counters) lives: in `scratch`, a fixed array of kScratchBytes
// that every GPU the library runs on holds from the first such call until the process ends.
// Counters that must start at zero are cleared by fill kernels inside the call (no
// hipMemsetAsync), so the call can be captured into a graph. A plan that needs more than the
// array returns ACLBLAS_STATUS_NOT_SUPPORTED before launching anything. Every call on one GPU
// uses that GPU's one array and leaves its contents behind: two calls that overlap on one GPU
// (two streams or two host threads) race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "host_internal.h"

namespace c3m {

constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr int kScratchDevices = 64;

// The device array. In host code this name is only a stand-in for the symbol: its device
// address comes from hipGetSymbolAddress, once per device, and only that address is used.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_scratchMutex;
void* g_scratchPtr[kScratchDevices] = {};
}  // namespace

// Address of the current device's array (looked up on the first call on that device, with
// that device current).
static aclblasStatus_t device_scratch(char** base, size_t* bytes, const char* op)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) {
        (void)hipGetLastError();
        log_error(op, "no current HIP device");
        return ACLBLAS_STATUS_NOT_INITIALIZED;
    }
    if (dev >= kScratchDevices) {
        log_error(op, "device %d is beyond the %d devices this entry serves", dev, kScratchDevices);
        return ACLBLAS_STATUS_NOT_SUPPORTED;
    }
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (g_scratchPtr[dev] == nullptr) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            (void)hipGetLastError();
            log_error(op, "device address of the fixed array not available: %s", hipGetErrorString(e));
            return ACLBLAS_STATUS_INTERNAL_ERROR;
        }
        g_scratchPtr[dev] = p;
    }
    *base = static_cast<char*>(g_scratchPtr[dev]);
    *bytes = kScratchBytes;
    return ACLBLAS_STATUS_SUCCESS;
}

}  // namespace c3m

using namespace c3m;

extern "C" aclblasStatus_t aclblasCgemm3mNoAlloc(aclblasHandle_t handle, aclblasOperation_t transa,
                                                 aclblasOperation_t transb, int m, int n, int k,
                                                 const aclblasComplex* alpha, const aclblasComplex* A, int lda,
                                                 const aclblasComplex* B, int ldb, const aclblasComplex* beta,
                                                 aclblasComplex* C, int ldc)
{
    return run_problem(handle, transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, nullptr,
                       "aclblasCgemm3mNoAlloc", device_scratch);
}
