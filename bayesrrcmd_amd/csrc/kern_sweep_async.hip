// kern_sweep_async.hip -- the BayesR sweep kernels in asynchronous windows (see kernels.hpp).
#include "sweep_kernel.hpp"
#include "kernels.hpp"

namespace brr {
BRR_SWEEP_KERNELS_OF(BRR_SWEEP_INST, true)
} // namespace brr
