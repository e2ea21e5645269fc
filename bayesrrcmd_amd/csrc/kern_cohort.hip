// kern_cohort.hip -- the kernels that run several BayesR contexts (cohorts) in
// one persistent launch (see kernels.hpp).
#include "sweep_kernel.hpp"
#include "kernels.hpp"

namespace brr {
BRR_COHORT_KERNELS(BRR_COHORT_INST)
} // namespace brr
