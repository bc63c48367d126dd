// Split-BF16 contraction kernels (placeholder until the bf16 MFMA family lands).
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace tp {
hipError_t launch_contract_split(const ContractArgs&, DType, int, int, hipStream_t) { return hipErrorNotSupported; }
}  // namespace tp
