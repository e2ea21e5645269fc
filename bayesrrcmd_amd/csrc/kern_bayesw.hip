// kern_bayesw.hip -- the BayesW sweep kernels (see kernels.hpp).
#include "bayesw_kernel.hpp"
#include "kernels.hpp"

namespace brr {
BRR_BW_KERNELS(BRR_BW_INST)
} // namespace brr
