// kern_batch.hip -- the batched-traits BayesR kernels (see kernels.hpp).
#include "batch_kernel.hpp"
#include "kernels.hpp"

namespace brr {
BRR_BATCH_KERNELS(BRR_BATCH_INST)
} // namespace brr
