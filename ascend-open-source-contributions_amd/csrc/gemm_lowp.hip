// gemm_lowp.hip -- split-precision products (BF16 pieces, FP16 block-scaled). Placeholder
// launchers until the kernels land; the host reports NOT_SUPPORTED through them.
#include "device_util.h"
#include "launch.h"

namespace c3m {
hipError_t launch_gemm_bf16(const GemmArgs&, int, int, int, hipStream_t) { return hipErrorNotSupported; }
hipError_t launch_gemm_f16(const GemmArgs&, int, int, hipStream_t) { return hipErrorNotSupported; }
}  // namespace c3m
