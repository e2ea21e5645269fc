// kernels.hpp -- every kernel instantiation the library launches.
//
// The kern_*.hip files instantiate them (one translation unit per family, so
// they compile in parallel); api.hip declares them extern and takes their
// addresses for the launches and occupancy queries.
#pragma once

// bayesr_sweep_kernel<R, layout, codes, async, fp32>: FP64 residual
#define BRR_SWEEP_KERNELS_OF(X, A)                                                                      \
    X(1, kLayoutDense, true, A, false) X(1, kLayoutRagged, true, A, false)                              \
    X(1, kLayoutRagged, false, A, false) X(1, kLayoutEigen, true, A, false)                             \
    X(2, kLayoutDense, true, A, false) X(2, kLayoutRagged, true, A, false)                              \
    X(2, kLayoutRagged, false, A, false) X(2, kLayoutEigen, true, A, false)                             \
    X(4, kLayoutDense, true, A, false) X(4, kLayoutRagged, true, A, false)                              \
    X(4, kLayoutRagged, false, A, false) X(4, kLayoutEigen, true, A, false)

// FP32 residual with FP64 refresh: sequential sweeps over 2-bit codes
#define BRR_SWEEP_F32_KERNELS(X)                                                                        \
    X(1, kLayoutDense, true, false, true) X(1, kLayoutRagged, true, false, true)                        \
    X(1, kLayoutEigen, true, false, true) X(2, kLayoutDense, true, false, true)                         \
    X(2, kLayoutRagged, true, false, true) X(2, kLayoutEigen, true, false, true)                        \
    X(4, kLayoutDense, true, false, true) X(4, kLayoutRagged, true, false, true)                        \
    X(4, kLayoutEigen, true, false, true)

// bayesr_cohort_kernel<R, fp32>: one or two lane runs per slice (a four-run
// variant gave chains that differ from run to run and stays out)
#define BRR_COHORT_KERNELS(X) X(1, false) X(2, false) X(1, true) X(2, true)

// bayesr_batch_kernel<NT, dense, fp32>: FP64 residual tiles (up to 16 per slice)
#define BRR_BATCH_KERNELS(X) X(4, true, false) X(4, false, false) X(8, true, false) X(8, false, false) \
    X(13, true, false) X(13, false, false) X(16, true, false) X(16, false, false)
// FP32 residual tiles (up to 32 per slice)
#define BRR_BATCH_F32_KERNELS(X) X(8, true, true) X(8, false, true) X(16, true, true) X(16, false, true) \
    X(25, true, true) X(25, false, true) X(32, true, true) X(32, false, true)

// bayesw_sweep_kernel<R, dense>
#define BRR_BW_KERNELS(X) X(1, true) X(1, false) X(2, true) X(2, false) X(4, true) X(4, false)

#define BRR_SWEEP_INST(R, L, C, A, F) template __global__ void bayesr_sweep_kernel<R, L, C, A, F>(SweepParams);
#define BRR_SWEEP_DECL(R, L, C, A, F) extern template __global__ void bayesr_sweep_kernel<R, L, C, A, F>(SweepParams);
#define BRR_COHORT_INST(R, F) template __global__ void bayesr_cohort_kernel<R, F>(CohortTable);
#define BRR_COHORT_DECL(R, F) extern template __global__ void bayesr_cohort_kernel<R, F>(CohortTable);
#define BRR_BATCH_INST(NT, D, F) template __global__ void bayesr_batch_kernel<NT, D, F>(BatchParams);
#define BRR_BATCH_DECL(NT, D, F) extern template __global__ void bayesr_batch_kernel<NT, D, F>(BatchParams);
#define BRR_BW_INST(R, D) template __global__ void bayesw_sweep_kernel<R, D>(BwParams);
#define BRR_BW_DECL(R, D) extern template __global__ void bayesw_sweep_kernel<R, D>(BwParams);
