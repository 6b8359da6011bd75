// host_internal.h -- host-side handle state, plan and helpers shared by the C-ABI files.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "../../include/aclblas_cgemm3m.h"
#include "kernel_args.h"

struct _aclblas_handle {
    hipStream_t stream = nullptr;
    void* defWs = nullptr;      // library-owned default workspace (grow-only)
    size_t defWsBytes = 0;
    bool defWsCaptured = false;          // a stream capture recorded launches that use defWs
    std::vector<void*> retiredWs;        // outgrown default workspaces a graph may still use
                                         // (freed by aclblasDestroy)
    void* userWs = nullptr;     // borrowed via aclblasSetWorkspace
    size_t userWsBytes = 0;
    int device = -1;            // device the default workspace lives on
    int cus = 0;                // compute units of that device
    long long lastPlan[8] = {0, 0, 0, 0, 0, 0, -1, 0};
    unsigned int* lastCounter = nullptr;  // device counters of flagged outputs (last call),
    long long lastCounterSlots = 0;       // summed by aclblasCgemm3mGetLastPlan
    int timing = 0;                       // aclblasCgemm3mSetTiming
    struct Timed {
        int stage;
        hipEvent_t start, stop;
    };
    std::vector<Timed> timed;
    long long timedCalls = 0;
    // Grouped-batch descriptor blocks, keyed by content: a pinned host copy and a device
    // copy that is uploaded once. Later calls with the same problems (and hipGraph captures
    // of them) only read the device copy. At most kMaxStaged slots that no capture used are
    // kept; a call with new content reuses the least recently used slot (and its buffers,
    // when large enough) once the event recorded after that slot's last use has completed,
    // so the host waits at most for that one earlier call and nothing is allocated per call.
    // Slots a capture used live until aclblasDestroy, since a graph may replay them.
    struct Staged {
        uint64_t hash;
        size_t bytes;
        size_t capacity;          // bytes both buffers hold
        void* host;
        void* device;
        bool uploaded;
        bool captured;
        uint64_t lastUse;
        hipEvent_t done;          // recorded after the last kernel that read the slot
        hipStream_t doneStream;   // the stream `done` was recorded on (null: never recorded)
        bool recorded;
    };
    std::vector<Staged> staged;
    uint64_t stageClock = 0;
    // Grouped batches: a second stream (default priority) for the members on the one-launch
    // kernel, forked from and joined back to `stream` with the two events; created on first use.
    hipStream_t side = nullptr;
    hipEvent_t forkEv = nullptr, joinEv = nullptr;
    // Arrival counters of k-sliced one-launch calls (kArrivePool of them), zeroed once when
    // created (on an eager call); every launch leaves the counters it used at zero, so later
    // calls need no clearing launch. Freed by aclblasDestroy.
    unsigned int* arrivePool = nullptr;
};

namespace c3m {

constexpr size_t kDefaultWorkspace = 32u << 20;  // cann_ops_blas.h:34-36
constexpr size_t kAlign = 256;
constexpr int kVersion = 10000;
constexpr size_t kMaxStaged = 32;  // descriptor blocks kept per handle besides captured ones
constexpr long long kArrivePool = 1 << 16;  // arrival counters per handle (256 KiB)

void log_error(const char* op, const char* fmt, ...);

// True when `s` is being captured into a graph.
bool stream_capturing(hipStream_t s);

// Switches this thread to the relaxed stream-capture mode for its lifetime, so that
// allocations and stream or event creation made while a caller captures are allowed (a
// capture in the global mode rejects them and is invalidated).
class RelaxedCapture {
  public:
    RelaxedCapture() { (void)hipThreadExchangeStreamCaptureMode(&mode_); }
    ~RelaxedCapture() { (void)hipThreadExchangeStreamCaptureMode(&mode_); }
    RelaxedCapture(const RelaxedCapture&) = delete;
    RelaxedCapture& operator=(const RelaxedCapture&) = delete;

  private:
    hipStreamCaptureMode mode_ = hipStreamCaptureModeRelaxed;
};

enum Stage { ST_PACK = 0, ST_ENERGY = 1, ST_GEMM = 2, ST_MERGE = 3, ST_RECOMPUTE = 4, ST_SCALE = 5, ST_COUNT = 6 };

// Records a HIP event pair around one kernel stage when the handle's timing is enabled.
class StageTimer {
  public:
    StageTimer(aclblasHandle_t h, int stage);  // h == nullptr: records nothing
    ~StageTimer();

  private:
    aclblasHandle_t h_;
    int stage_;
    hipEvent_t start_ = nullptr;
};

// A full execution plan for one product (pure function of the arguments and CU count).
struct Plan {
    int path = 0;         // 0 nothing, 1 fused tile, 2 K-split, 3 beta scale, 4 separate products
    int precision = 0;    // aclblasCgemm3mPrecision_t
    int format = FMT_F32;
    int pieces = 1;
    long long smallTiles = 0;  // path 5: 16x16 output tiles (one flagged-output slot each)
    int smallWaves = 4;        // path 5: waves per output tile (4, or 1 for the single-wave form)
    int bm = 128, bn = 128;
    int bk = 16;          // k depth of one GEMM stage
    int parts = 1;
    int correction = 0;
    int identity = ID_FUSED;
    int variant = 0;      // product-kernel variant: 0 four waves per tile, 1 eight waves
    bool raw = false;     // path 4 on BF16: products read from the stored operands (no pack)
    int mPad = 0, nPad = 0, kPad = 0, kSteps = 0;
    size_t aPlaneBytes = 0, bPlaneBytes = 0;
    int planesA = 3, planesB = 3;  // planes * pieces per operand
    // workspace regions (byte offsets)
    size_t offA = 0, offB = 0, offPartial = 0, offEPA = 0, offEPB = 0, offEA = 0, offEB = 0;
    size_t offBitmap = 0, offCounter = 0, offScaleA = 0, offScaleB = 0,
           offSumA = 0, offSumB = 0, offArrive = 0;
    size_t partStride = 0;       // floats per partition (re plane + im plane)
    long long bitmapWords = 0;
    int kTiles = 0;
    size_t total = 0;
};

struct Problem {
    aclblasOperation_t ta, tb;
    int m, n, k;
    aclblasComplex alpha, beta;
    const aclblasComplex* A; int lda;
    const aclblasComplex* B; int ldb;
    aclblasComplex* C; int ldc;
};

aclblasStatus_t validate(aclblasHandle_t h, aclblasOperation_t ta, aclblasOperation_t tb, int m, int n, int k,
                         const aclblasComplex* alpha, int lda, int ldb, const aclblasComplex* beta, int ldc,
                         const char* op);
bool make_plan(const Problem& p, const aclblasCgemm3mConfig* cfg, int cus, Plan& plan, const char* op);
aclblasStatus_t ensure_device(aclblasHandle_t h);
aclblasStatus_t acquire_workspace(aclblasHandle_t h, size_t bytes, char** base, const char* op);
aclblasStatus_t enqueue(aclblasHandle_t h, const Problem& p, const Plan& plan, char* ws, const char* op,
                        bool noAlloc = false);
aclblasStatus_t validate_pointers(int k, const void* A, const void* B, const void* C, const char* op);
bool is_zero(const aclblasComplex& z);

// Autotuned plan table (tuned_table.h): tile and K partitions for the (log2 m*n, log2 k) class
// of a one-problem call or of a grouped member.
bool tuned_lookup(long long m, long long n, long long k, bool grouped, int* tile, int* parts);

// Temporary storage a call may use instead of the handle's workspace: a fixed device array
// of the current device (base, capacity in bytes). Used by the no-allocation entry.
typedef aclblasStatus_t (*ScratchSource)(char** base, size_t* bytes, const char* op);

// One aclblasCgemm3m(Ex) call: validation, plan, workspace, launches (cgemm3m_host.cpp). With a
// scratch source the call allocates nothing: its temporary storage comes from that array (a
// plan needing more returns ACLBLAS_STATUS_NOT_SUPPORTED).
aclblasStatus_t run_problem(aclblasHandle_t handle, aclblasOperation_t ta, aclblasOperation_t tb, int m, int n,
                            int k, const aclblasComplex* alpha, const aclblasComplex* A, int lda,
                            const aclblasComplex* B, int ldb, const aclblasComplex* beta, aclblasComplex* C,
                            int ldc, const aclblasCgemm3mConfig* cfg, const char* op,
                            ScratchSource scratch = nullptr);
// One grouped batch of FP32 problems (grouped_host.cpp); over: optional plan override. fills:
// whether the batch is large enough that its corrected members take no K split (see
// grouped_batch_fills); -1 decides from this call's members.
aclblasStatus_t grouped_run(aclblasHandle_t handle, int count, const aclblasOperation_t* transa,
                            const aclblasOperation_t* transb, const int* m, const int* n, const int* k,
                            const aclblasComplex* alpha, const aclblasComplex* const* A, const int* lda,
                            const aclblasComplex* const* B, const int* ldb, const aclblasComplex* beta,
                            aclblasComplex* const* C, const int* ldc, int precision,
                            const aclblasCgemm3mConfig* over, const char* op, int fills = -1);
// True when a batch whose members cover `blocks64` output blocks of 64 x 64 in all gives the
// persistent kernel at least two quads (four blocks each) per CU without any K split.
bool grouped_batch_fills(long long blocks64, int cus);
// The plan a grouped member of this shape gets, as a one-problem config (batchFills: the
// batch's grouped_batch_fills).
aclblasCgemm3mConfig grouped_member_config(int m, int n, int k, const aclblasCgemm3mConfig* over, bool batchFills);

}  // namespace c3m
