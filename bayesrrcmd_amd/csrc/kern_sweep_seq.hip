// kern_sweep_seq.hip -- the sequential BayesR sweep kernels (see kernels.hpp).
#include "sweep_kernel.hpp"
#include "kernels.hpp"

namespace brr {
BRR_SWEEP_KERNELS_OF(BRR_SWEEP_INST, false)
} // namespace brr
