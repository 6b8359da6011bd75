// handle.cpp -- aclblas handle, stream and workspace functions.
// Reference contract: ops-blas/include/cann_ops_blas.h:32-110,
// ops-blas/docs/zh/api_list.md:195-215 (workspace semantics).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <new>

#include "host_internal.h"

namespace {
std::mutex g_log_mutex;
aclblasLogCallback g_log_cb = nullptr;
}  // namespace

namespace c3m {

void log_error(const char* op, const char* fmt, ...)
{
    char msg[512];
    int n = std::snprintf(msg, sizeof(msg), "[%s] ", op);
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(msg + n, sizeof(msg) - n, fmt, ap);
    va_end(ap);
    std::lock_guard<std::mutex> lock(g_log_mutex);
    if (g_log_cb != nullptr) {
        g_log_cb(msg);
    } else {
        std::fprintf(stderr, "%s\n", msg);
    }
}

aclblasStatus_t ensure_device(aclblasHandle_t h)
{
    if (h->device >= 0) return ACLBLAS_STATUS_SUCCESS;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        (void)hipGetLastError();
        return ACLBLAS_STATUS_NOT_INITIALIZED;
    }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) {
        (void)hipGetLastError();
        return ACLBLAS_STATUS_NOT_INITIALIZED;
    }
    h->device = dev;
    h->cus = cus;
    return ACLBLAS_STATUS_SUCCESS;
}

bool stream_capturing(hipStream_t s)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

// The call's temporary storage: the borrowed user workspace when one is set (too small ->
// EXECUTION_FAILED, api_list.md:203), else the library default, grown on demand. A default
// workspace that a capture used is never freed before aclblasDestroy (a graph may replay
// launches that read it); outgrowing it retires it instead. Growing during a capture
// retires the old buffer too (no synchronisation is allowed then) and allocates in the
// relaxed capture mode.
aclblasStatus_t acquire_workspace(aclblasHandle_t h, size_t bytes, char** base, const char* op)
{
    const bool capturing = stream_capturing(h->stream);
    if (h->userWs != nullptr) {
        if (bytes > h->userWsBytes) {
            log_error(op, "user workspace holds %zu bytes, this call needs %zu", h->userWsBytes, bytes);
            return ACLBLAS_STATUS_EXECUTION_FAILED;
        }
        *base = static_cast<char*>(h->userWs);
        return ACLBLAS_STATUS_SUCCESS;
    }
    const size_t want = std::max(bytes, kDefaultWorkspace);
    if (h->defWs == nullptr || h->defWsBytes < want) {
        if (h->defWs != nullptr) {
            if (capturing || h->defWsCaptured) {
                h->retiredWs.push_back(h->defWs);
            } else {
                // Work already queued may still read the old buffer.
                if (hipStreamSynchronize(h->stream) != hipSuccess) {
                    log_error(op, "stream synchronisation failed before growing the workspace");
                    return ACLBLAS_STATUS_EXECUTION_FAILED;
                }
                (void)hipFree(h->defWs);
            }
            h->defWs = nullptr;
            h->defWsBytes = 0;
            h->defWsCaptured = false;
        }
        void* p = nullptr;
        hipError_t e;
        {
            RelaxedCapture relaxed;
            e = hipMalloc(&p, want);
        }
        if (e != hipSuccess) {
            (void)hipGetLastError();
            log_error(op, "workspace allocation of %zu bytes failed", want);
            return ACLBLAS_STATUS_ALLOC_FAILED;
        }
        h->defWs = p;
        h->defWsBytes = want;
    }
    if (capturing) h->defWsCaptured = true;
    *base = static_cast<char*>(h->defWs);
    return ACLBLAS_STATUS_SUCCESS;
}

}  // namespace c3m

using namespace c3m;

// cann_ops_blas.h:32-46: the handle comes with its 32 MiB default workspace, reserved here on
// the current device (ALLOC_FAILED if that fails). On a host without a HIP device the handle is
// still created, for argument validation, and the workspace is reserved by the first call that
// needs one.
extern "C" aclblasStatus_t aclblasCreate(aclblasHandle_t* handle)
{
    if (handle == nullptr) return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    if (*handle != nullptr) return ACLBLAS_STATUS_INVALID_VALUE;
    _aclblas_handle* h = new (std::nothrow) _aclblas_handle();
    if (h == nullptr) return ACLBLAS_STATUS_ALLOC_FAILED;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        (void)hipGetLastError();
    } else if (ensure_device(h) == ACLBLAS_STATUS_SUCCESS) {
        void* p = nullptr;
        hipError_t e;
        {
            RelaxedCapture relaxed;
            e = hipMalloc(&p, kDefaultWorkspace);
        }
        if (e != hipSuccess) {
            (void)hipGetLastError();
            log_error("aclblasCreate", "default workspace of %zu bytes could not be reserved", kDefaultWorkspace);
            delete h;
            return ACLBLAS_STATUS_ALLOC_FAILED;
        }
        h->defWs = p;
        h->defWsBytes = kDefaultWorkspace;
    }
    *handle = h;
    return ACLBLAS_STATUS_SUCCESS;
}

extern "C" aclblasStatus_t aclblasDestroy(aclblasHandle_t handle)
{
    if (handle == nullptr) return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    aclblasStatus_t st = ACLBLAS_STATUS_SUCCESS;
    if (handle->device >= 0) {
        if (hipStreamSynchronize(handle->stream) != hipSuccess) st = ACLBLAS_STATUS_EXECUTION_FAILED;
        if (handle->defWs != nullptr) (void)hipFree(handle->defWs);
        for (void* w : handle->retiredWs) (void)hipFree(w);
        if (handle->arrivePool != nullptr) (void)hipFree(handle->arrivePool);
        for (auto& sb : handle->staged) {
            (void)hipHostFree(sb.host);
            (void)hipFree(sb.device);
            if (sb.done != nullptr) (void)hipEventDestroy(sb.done);
        }
        for (auto& ev : handle->timed) {
            (void)hipEventDestroy(ev.start);
            (void)hipEventDestroy(ev.stop);
        }
        if (handle->side != nullptr) {
            if (hipStreamSynchronize(handle->side) != hipSuccess) st = ACLBLAS_STATUS_EXECUTION_FAILED;
            (void)hipStreamDestroy(handle->side);
        }
        if (handle->forkEv != nullptr) (void)hipEventDestroy(handle->forkEv);
        if (handle->joinEv != nullptr) (void)hipEventDestroy(handle->joinEv);
    }
    handle->userWs = nullptr;
    delete handle;
    return st;
}

extern "C" aclblasStatus_t aclblasSetStream(aclblasHandle_t handle, aclrtStream stream)
{
    if (handle == nullptr) return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    handle->stream = static_cast<hipStream_t>(stream);
    handle->userWs = nullptr;  // back to the library default workspace
    handle->userWsBytes = 0;
    return ACLBLAS_STATUS_SUCCESS;
}

extern "C" aclblasStatus_t aclblasGetStream(aclblasHandle_t handle, aclrtStream* stream)
{
    if (handle == nullptr) return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    if (stream == nullptr) return ACLBLAS_STATUS_INVALID_VALUE;
    *stream = static_cast<aclrtStream>(handle->stream);
    return ACLBLAS_STATUS_SUCCESS;
}

extern "C" aclblasStatus_t aclblasSetWorkspace(aclblasHandle_t handle, void* workspace, size_t workspaceSize)
{
    if (handle == nullptr) return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    if (workspace == nullptr || workspaceSize == 0) return ACLBLAS_STATUS_INVALID_VALUE;
    if (workspaceSize > handle->userWsBytes) {  // grow-only
        handle->userWs = workspace;
        handle->userWsBytes = workspaceSize;
    }
    return ACLBLAS_STATUS_SUCCESS;
}

extern "C" aclblasStatus_t aclblasGetVersion(aclblasHandle_t /*handle*/, int* version)
{
    if (version == nullptr) return ACLBLAS_STATUS_INVALID_VALUE;
    *version = kVersion;
    return ACLBLAS_STATUS_SUCCESS;
}

extern "C" aclblasStatus_t aclblasSetLogCallback(aclblasLogCallback callback)
{
    std::lock_guard<std::mutex> lock(g_log_mutex);
    g_log_cb = callback;
    return ACLBLAS_STATUS_SUCCESS;
}
