// kern_sweep_f32.hip -- the sequential BayesR sweep kernels with an FP32
// residual and FP64 refresh (see kernels.hpp).
#include "sweep_kernel.hpp"
#include "kernels.hpp"

namespace brr {
BRR_SWEEP_F32_KERNELS(BRR_SWEEP_INST)
} // namespace brr
