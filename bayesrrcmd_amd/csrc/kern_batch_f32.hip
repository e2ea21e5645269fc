// kern_batch_f32.hip -- the batched-traits kernels with FP32 residual tiles (see kernels.hpp).
#include "batch_kernel.hpp"
#include "kernels.hpp"

namespace brr {
BRR_BATCH_F32_KERNELS(BRR_BATCH_INST)
} // namespace brr
