// api.hip -- the C ABI of libbrr_hip.so (declared in include/brr_hip.h).
//
// Owns all device memory of a context, uploads marker arenas, builds the
// per-workgroup ragged segment table on the device and launches the
// persistent sweep kernel cooperatively (all workgroups co-resident).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "brr_hip.h"
#include "sweep_kernel.hpp"
#include "batch_kernel.hpp"
#include "bayesw_kernel.hpp"
#include "bed_kernels.hpp"
#include "kernels.hpp"

namespace brr {
// instantiated in kern_sweep_seq.hip, kern_sweep_async.hip, kern_batch.hip, kern_bayesw.hip
BRR_SWEEP_KERNELS_OF(BRR_SWEEP_DECL, false)
BRR_SWEEP_KERNELS_OF(BRR_SWEEP_DECL, true)
BRR_SWEEP_F32_KERNELS(BRR_SWEEP_DECL)
BRR_COHORT_KERNELS(BRR_COHORT_DECL)
BRR_BATCH_KERNELS(BRR_BATCH_DECL)
BRR_BATCH_F32_KERNELS(BRR_BATCH_DECL)
BRR_BW_KERNELS(BRR_BW_DECL)
} // namespace brr

using namespace brr;

// words of brr_hip_sweep_noalloc's pinned status block: 4 of control words,
// 16 of counters, 1 of the order flag (a copy of the scratch words from the
// control words on)
constexpr int kStatusWords = 4 + 16 + 1;

// device counters of a launch: [0] rounds, [1] changes, [2..7] phase ticks,
// batched kernel also [8] decision rounds (summed over waves), [10..15] phase ticks 6..11
constexpr int kCounters = 16;

struct brr_hip_ctx {
    brr_hip_config cfg{};
    int device = 0;
    std::string err;
    uint32_t P = 0, B = 0, S = 0, R = 0;
    uint32_t T = 0;              // traits of a batched context (0: one chain)
    uint32_t Tn = 1;             // columns of the per-individual / per-marker state
    uint32_t nt = 0, NT = 0;     // batched: 16-individual tiles per slice, kernel variant
    size_t x1_words = 0, x2_words = 0;
    int layout = 0;
    bool bw = false;             // BayesW context (brr_hip_bw_create)
    uint32_t nq = 0;             // BayesW Gauss-Hermite points
    bool codes = false;          // columns held as 2-bit codes on the device
    uint32_t async = 0;          // asynchronous windows of B markers (brr_hip_set_async)
    uint32_t B_seq = 0;          // block size of the sequential sweep
    bool f32 = false;            // FP32 residual with FP64 refresh (brr_hip_set_precision)
    uint32_t cus = 0;            // compute units of the device
    size_t x1_cap = 0;           // words allocated for the exchange partials
    double* d_num = nullptr;     // per-marker num of the last sweep (brr_hip_record_num)
    double* d_keps = nullptr;    // N: a host residual / residual change of the per-marker kernel calls
    SweepParams* d_cohort = nullptr;   // brr_hip_run_cohorts: the cohorts' parameter blocks
    uint32_t cohort_cap = 0;
    uint64_t* h_status = nullptr; // pinned: control words, counters and order flag of brr_hip_sweep_noalloc
    bool record_num = false;
    uint32_t refresh_markers = 1024;
    double change_rate = 1.0;    // beta changes per marker visit in the last chain launch
    bool have_markers = false;
    bool have_model = false;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    brr_hip_model model{};
    std::vector<double> cva;
    brr_hip_stats stats{};
    double phase_ms[kBPhases] = {};
    // dense arena
    uint32_t* d_dense = nullptr;
    uint64_t col_words = 0, col_bytes = 0;
    double* d_lut = nullptr;
    // ragged arena
    int32_t* d_ones = nullptr;
    int32_t* d_twos = nullptr;
    int32_t* d_miss = nullptr;
    uint64_t* d_ones_off = nullptr;
    uint64_t* d_twos_off = nullptr;
    uint64_t* d_miss_off = nullptr;
    uint32_t* d_seg = nullptr;
    double* d_stats = nullptr;
    // model / state
    int32_t* d_group = nullptr;
    double* d_cva = nullptr;
    double* d_eps = nullptr;
    double* d_beta = nullptr;
    int32_t* d_comp = nullptr;
    double* d_acum = nullptr;
    double* d_scal = nullptr;
    double* d_hist = nullptr;
    size_t hist_cap = 0;
    uint32_t* d_order = nullptr;
    size_t order_cap = 0;
    double* d_u = nullptr;
    double* d_z = nullptr;
    uint64_t* d_xchg1 = nullptr;
    uint64_t* d_xchg2 = nullptr;
    uint32_t* d_ctl = nullptr;
    unsigned long long* d_counters = nullptr;
    double* d_tmp = nullptr;
    // fixed effects: covariates (column-major N x F floats), {sigmaF, gamma[F]}, order
    uint32_t F = 0;
    float* d_fixedX = nullptr;
    double* d_fixed = nullptr;             // BayesW: theta[F]
    uint32_t* d_forder = nullptr;
    double* d_sumf_fix = nullptr;          // BayesW: sum_k X(k,f) failure_k
    // BayesW
    double* d_fail = nullptr;
    double* d_sumf = nullptr;
    unsigned long long last_counters[kCounters] = {};
};

namespace {

int fail(brr_hip_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    return code;
}

#define HIP_TRY(ctx, expr)                                                                   \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail((ctx), BRR_HIP_EDEVICE,                                              \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                  \
    } while (0)

template <typename T>
int dalloc(brr_hip_ctx* c, T** p, size_t count)
{
    if (*p) { (void)hipFree(*p); *p = nullptr; }
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)p, sizeof(T) * count);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(c, BRR_HIP_ENOMEM, "hipMalloc of " + std::to_string(sizeof(T) * count) +
                                           " bytes failed: " + hipGetErrorString(e));
    }
    return BRR_HIP_OK;
}

template <typename T>
void dfree(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

// seg[(j*3 + list)*(P+1) + w] = lower_bound(list_j, w*S) - list_j start
__global__ void build_seg_kernel(const int32_t* ones, const uint64_t* ones_off,
                                 const int32_t* twos, const uint64_t* twos_off,
                                 const int32_t* miss, const uint64_t* miss_off,
                                 uint32_t M, uint32_t P, uint32_t S, uint32_t* seg)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t total = (uint64_t)M * 3u * (P + 1u);
    if (t >= total) return;
    const uint32_t w = (uint32_t)(t % (P + 1u));
    const uint64_t jl = t / (P + 1u);
    const uint32_t list = (uint32_t)(jl % 3u);
    const uint32_t j = (uint32_t)(jl / 3u);
    const int32_t* base;
    uint64_t n;
    if (list == 0) { base = ones + ones_off[j]; n = ones_off[j + 1] - ones_off[j]; }
    else if (list == 1) { base = twos + twos_off[j]; n = twos_off[j + 1] - twos_off[j]; }
    else { base = miss + miss_off[j]; n = miss_off[j + 1] - miss_off[j]; }
    const int64_t key = (int64_t)w * (int64_t)S;
    uint64_t a = 0, b = n;
    while (a < b) {
        const uint64_t mid = (a + b) >> 1;
        if ((int64_t)base[mid] < key) a = mid + 1; else b = mid;
    }
    seg[t] = (uint32_t)a;
}

// BayesRKernel::computeNum for one marker on the device epsilon (one block).
__global__ void compute_num_kernel(SweepParams p, uint32_t j, double beta_old, const double* eps_sum, double* out)
{
    __shared__ double red[kWaves];
    const int tid = threadIdx.x;
    double s = 0.0;
    if (p.layout == kLayoutDense) {
        const double* lt = p.lut + 4 * (size_t)j;
        const double wt[4] = {lt[2], lt[3], lt[1], lt[0]};
        const uint8_t* col = (const uint8_t*)(p.dense + (size_t)j * p.col_words);
        for (uint32_t i = tid; i < p.N; i += kThreads) {
            const uint32_t v = (col[i >> 2] >> ((i & 3u) * 2u)) & 3u;
            s = fma(wt[v], p.eps[i], s);
        }
    } else if (!p.ones) {
        // ragged formulas on the 2-bit codes (no index lists held): genotype 1
        // adds eps, genotype 2 adds 2 eps
        const uint8_t* col = (const uint8_t*)(p.dense + (size_t)j * p.col_words);
        const double wm = (p.layout == kLayoutEigen) ? p.stats[4 * (size_t)j] : 0.0;   // missing: Eigen's mean
        for (uint32_t i = tid; i < p.N; i += kThreads) {
            const uint32_t v = (col[i >> 2] >> ((i & 3u) * 2u)) & 3u;
            if (v == 2u) s += p.eps[i];
            else if (v == 0u) s += 2.0 * p.eps[i];
            else if (v == 1u) s += wm * p.eps[i];
        }
    } else {
        for (uint64_t q = p.ones_off[j] + tid; q < p.ones_off[j + 1]; q += kThreads) s += p.eps[p.ones[q]];
        for (uint64_t q = p.twos_off[j] + tid; q < p.twos_off[j + 1]; q += kThreads) s += 2.0 * p.eps[p.twos[q]];
    }
    s = wave_sum(s);
    if ((tid & 63) == 0) red[tid >> 6] = s;
    __syncthreads();
    if (tid == 0) {
        const double dot = ((red[0] + red[1]) + red[2]) + red[3];
        const double dN = (double)p.N;
        if (p.layout == kLayoutDense) {
            *out = dot + beta_old * (double)(p.N - 1u);
        } else {
            const double mean = p.stats[4 * (size_t)j], sd = p.stats[4 * (size_t)j + 1];
            const double epsSum = *eps_sum;
            *out = beta_old * (dN - 1.0) - mean * epsSum / sd + dot / sd;
        }
    }
}

// calculateEpsilonChange + eps += delta for one marker (grid over N).
__global__ void epsilon_change_kernel(SweepParams p, uint32_t j, double beta_old, double beta)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const double dBeta = beta_old - beta;
    if (p.layout == kLayoutDense) {
        if (i >= p.N) return;
        const double* lt = p.lut + 4 * (size_t)j;
        const double wt[4] = {lt[2], lt[3], lt[1], lt[0]};
        const uint8_t* col = (const uint8_t*)(p.dense + (size_t)j * p.col_words);
        const uint32_t v = (col[i >> 2] >> ((i & 3u) * 2u)) & 3u;
        p.eps[i] = p.eps[i] + dBeta * wt[v];
        return;
    }
    // ragged: classify i by binary search in the three lists
    if (i >= p.N) return;
    auto in_list = [](const int32_t* base, uint64_t n, int32_t key) {
        uint64_t a = 0, b = n;
        while (a < b) {
            const uint64_t mid = (a + b) >> 1;
            if (base[mid] < key) a = mid + 1; else b = mid;
        }
        return a < n && base[a] == key;
    };
    const double mean = p.stats[4 * (size_t)j], sd = p.stats[4 * (size_t)j + 1];
    const double meanAdjustment = dBeta * mean / sd;
    const double oneAdjustment = dBeta / sd;
    double delta = -meanAdjustment;
    if (!p.ones) {
        // 2-bit codes: value 2 = genotype 1, 0 = genotype 2, 1 = missing
        const uint8_t* col = (const uint8_t*)(p.dense + (size_t)j * p.col_words);
        const uint32_t v = (col[i >> 2] >> ((i & 3u) * 2u)) & 3u;
        if (v == 2u) delta += oneAdjustment;
        else if (v == 0u) delta += 2 * oneAdjustment;
        else if (v == 1u) delta += meanAdjustment;
    } else if (in_list(p.ones + p.ones_off[j], p.ones_off[j + 1] - p.ones_off[j], (int32_t)i)) delta += oneAdjustment;
    else if (in_list(p.twos + p.twos_off[j], p.twos_off[j + 1] - p.twos_off[j], (int32_t)i)) delta += 2 * oneAdjustment;
    else if (in_list(p.miss + p.miss_off[j], p.miss_off[j + 1] - p.miss_off[j], (int32_t)i)) delta += meanAdjustment;
    p.eps[i] = p.eps[i] + delta;
    if (i == 0) {
        const double Zsum = p.stats[4 * (size_t)j + 3];
        p.scal[2] += dBeta * Zsum / sd - dBeta * mean * (double)p.N / sd;
    }
}

SweepParams base_params(brr_hip_ctx* c)
{
    SweepParams p{};
    p.num_out = c->record_num ? c->d_num : nullptr;
    p.N = c->cfg.num_inds;
    p.M = c->cfg.num_snps;
    p.K = c->cfg.num_components;
    p.G = c->cfg.num_groups;
    p.P = c->P;
    p.B = c->B;
    p.S = c->S;
    p.layout = c->layout;
    p.async = c->async;
    p.dense = c->d_dense;
    p.col_words = c->col_words;
    p.col_bytes = c->col_bytes;
    p.lut = c->d_lut;
    p.ones = c->d_ones;
    p.twos = c->d_twos;
    p.miss = c->d_miss;
    p.ones_off = c->d_ones_off;
    p.twos_off = c->d_twos_off;
    p.miss_off = c->d_miss_off;
    p.seg = c->d_seg;
    p.stats = c->d_stats;
    p.group = c->d_group;
    p.cva = c->d_cva;
    p.v0E = c->model.v0E;
    p.s02E = c->model.s02E;
    p.v0G = c->model.v0G;
    p.s02G = c->model.s02G;
    p.eps = c->d_eps;
    p.beta = c->d_beta;
    p.comp = c->d_comp;
    p.acum = c->d_acum;
    p.u = c->d_u;          // read by every block load (M values; the chain draws its own)
    p.z = c->d_z;
    p.scal = c->d_scal;
    p.xchg1 = c->d_xchg1;
    p.xchg2 = c->d_xchg2;
    p.ctl = c->d_ctl;
    p.counters = c->d_counters;
    p.F = c->F;
    p.fixedX = c->d_fixedX;
    p.fixed = c->d_fixed;
    p.forder = c->d_forder;
    return p;
}

template <int R, bool kAsync>
const void* kernel_ra(int layout, bool codes)
{
    if (layout == kLayoutDense) return (const void*)bayesr_sweep_kernel<R, kLayoutDense, true, kAsync, false>;
    if (layout == kLayoutEigen)   // Eigen contexts hold codes
        return codes ? (const void*)bayesr_sweep_kernel<R, kLayoutEigen, true, kAsync, false> : nullptr;
    return codes ? (const void*)bayesr_sweep_kernel<R, kLayoutRagged, true, kAsync, false>
                 : (const void*)bayesr_sweep_kernel<R, kLayoutRagged, false, kAsync, false>;
}

// FP32 residual kernels: sequential sweeps over 2-bit codes only
template <int R>
const void* kernel_f32(int layout, bool codes)
{
    if (!codes) return nullptr;
    if (layout == kLayoutDense) return (const void*)bayesr_sweep_kernel<R, kLayoutDense, true, false, true>;
    if (layout == kLayoutEigen) return (const void*)bayesr_sweep_kernel<R, kLayoutEigen, true, false, true>;
    return (const void*)bayesr_sweep_kernel<R, kLayoutRagged, true, false, true>;
}

template <int R>
const void* kernel_r(int layout, bool codes, bool async, bool f32)
{
    if (f32) return async ? nullptr : kernel_f32<R>(layout, codes);
    return async ? kernel_ra<R, true>(layout, codes) : kernel_ra<R, false>(layout, codes);
}

const void* kernel_for(uint32_t R, int layout, bool codes, bool async = false, bool f32 = false)
{
    switch (R) {
    case 1: return kernel_r<1>(layout, codes, async, f32);
    case 2: return kernel_r<2>(layout, codes, async, f32);
    case 4: return kernel_r<4>(layout, codes, async, f32);
    default: return nullptr;
    }
}

const void* cohort_kernel_for(uint32_t R, bool f32)
{
    switch (R) {
    case 1: return f32 ? (const void*)bayesr_cohort_kernel<1, true> : (const void*)bayesr_cohort_kernel<1, false>;
    case 2: return f32 ? (const void*)bayesr_cohort_kernel<2, true> : (const void*)bayesr_cohort_kernel<2, false>;
    default: return nullptr;
    }
}

// Before a cohort launch: the abort words and launch counters of every cohort
// (one workgroup per cohort; a kernel, not a memset, so that it records into a
// graph as a kernel node).
__global__ void cohort_reset_kernel(const SweepParams* params)
{
    const SweepParams& p = params[blockIdx.x];
    if (threadIdx.x == 0) {
        p.ctl[kCtlAbort] = 0u;
        p.ctl[kCtlAbortWhere] = 0u;
    }
    if (threadIdx.x < 16) p.counters[threadIdx.x] = 0ull;
}

// Batched kernels by the 16-individual tiles of a slice: FP64 residual tiles
// in variants of 4, 8, 13 or 16 tiles (13 fit N = 50,000 on 241 workgroups),
// FP32 ones in variants of 8, 16, 25 or 32 (25 fit N = 100,000 on 250).
const void* batch_kernel_for(uint32_t nt, int layout, bool f32)
{
    const bool d = layout == kLayoutDense;
#define BRR_BK(NT, F) (d ? (const void*)bayesr_batch_kernel<NT, true, F> : (const void*)bayesr_batch_kernel<NT, false, F>)
    if (!f32) {
        if (nt <= 4) return BRR_BK(4, false);
        if (nt <= 8) return BRR_BK(8, false);
        if (nt <= 13) return BRR_BK(13, false);
        if (nt <= 16) return BRR_BK(16, false);
        return nullptr;
    }
    if (nt <= 8) return BRR_BK(8, true);
    if (nt <= 16) return BRR_BK(16, true);
    if (nt <= 25) return BRR_BK(25, true);
    if (nt <= 32) return BRR_BK(32, true);
    return nullptr;
#undef BRR_BK
}

const void* bw_kernel_for(uint32_t R, bool dense)
{
    switch (R) {
    case 1: return dense ? (const void*)bayesw_sweep_kernel<1, true> : (const void*)bayesw_sweep_kernel<1, false>;
    case 2: return dense ? (const void*)bayesw_sweep_kernel<2, true> : (const void*)bayesw_sweep_kernel<2, false>;
    case 4: return dense ? (const void*)bayesw_sweep_kernel<4, true> : (const void*)bayesw_sweep_kernel<4, false>;
    default: return nullptr;
    }
}

BwParams bw_params(brr_hip_ctx* c)
{
    BwParams p{};
    p.N = c->cfg.num_inds;
    p.M = c->cfg.num_snps;
    p.K = c->cfg.num_components;
    p.P = c->P;
    p.S = c->S;
    p.nq = c->nq;
    p.codes = c->d_dense;
    p.col_words = c->col_words;
    p.col_bytes = c->col_bytes;
    p.stats = c->d_stats;
    p.lut = c->d_lut;
    p.sumf = c->d_sumf;
    p.fail = c->d_fail;
    p.cls = c->d_cva;
    p.u = c->d_u;
    p.eps = c->d_eps;
    p.beta = c->d_beta;
    p.comp = c->d_comp;
    p.scal = c->d_scal;
    p.xchg1 = c->d_xchg1;
    p.xchg2 = c->d_xchg2;
    p.ctl = c->d_ctl;
    p.counters = c->d_counters;
    p.F = c->F;
    p.fixedX = c->d_fixedX;
    p.theta = c->d_fixed;
    p.sumf_fix = c->d_sumf_fix;
    return p;
}

BatchParams batch_params(brr_hip_ctx* c)
{
    BatchParams p{};
    p.N = c->cfg.num_inds;
    p.M = c->cfg.num_snps;
    p.K = c->cfg.num_components;
    p.T = c->T;
    p.P = c->P;
    p.S = c->S;
    p.nt = c->nt;
    p.nq = (c->T + 15) / 16;
    p.layout = c->layout;
    p.codes = c->d_dense;
    p.col_words = c->col_words;
    p.col_bytes = c->col_bytes;
    p.lut = c->d_lut;
    p.stats = c->d_stats;
    p.cva = c->d_cva;
    p.v0E = c->model.v0E;
    p.s02E = c->model.s02E;
    p.v0G = c->model.v0G;
    p.s02G = c->model.s02G;
    p.eps = c->d_eps;
    p.beta = c->d_beta;
    p.comp = c->d_comp;
    p.acum = c->d_acum;
    p.scal = c->d_scal;
    p.xchg1 = c->d_xchg1;
    p.xchg2 = c->d_xchg2;
    p.ctl = c->d_ctl;
    p.counters = c->d_counters;
    return p;
}

// Ragged lists -> 2-bit PLINK codes, one workgroup per column: every call
// starts as code 3 (genotype 0), ones become 2, twos 0, missing 1.
__global__ void ragged_to_codes_kernel(const int32_t* ones, const uint64_t* ones_off,
                                       const int32_t* twos, const uint64_t* twos_off,
                                       const int32_t* miss, const uint64_t* miss_off,
                                       uint64_t col_words, uint32_t* codes)
{
    const uint32_t j = blockIdx.x;
    uint32_t* col = codes + (size_t)j * col_words;
    for (uint64_t t = threadIdx.x; t < col_words; t += blockDim.x) col[t] = 0xFFFFFFFFu;
    __syncthreads();
    for (uint64_t q = ones_off[j] + threadIdx.x; q < ones_off[j + 1]; q += blockDim.x) {
        const uint32_t i = (uint32_t)ones[q];
        atomicAnd(col + (i >> 4), ~(1u << (2u * (i & 15u))));
    }
    for (uint64_t q = twos_off[j] + threadIdx.x; q < twos_off[j + 1]; q += blockDim.x) {
        const uint32_t i = (uint32_t)twos[q];
        atomicAnd(col + (i >> 4), ~(3u << (2u * (i & 15u))));
    }
    for (uint64_t q = miss_off[j] + threadIdx.x; q < miss_off[j + 1]; q += blockDim.x) {
        const uint32_t i = (uint32_t)miss[q];
        atomicAnd(col + (i >> 4), ~(2u << (2u * (i & 15u))));
    }
}

// Individuals per workgroup of the single-chain kernels. A lane holds R runs
// of 8 individuals, so a workgroup costs the same for any slice up to 512 R,
// while every extra workgroup adds partials to each exchange round: by default
// the fewest runs per lane (1, 2 or 4) that fit N on min(CUs, kMaxP)
// workgroups, then the fewest workgroups at that R, with balanced slices (a
// multiple of 16). An explicit workgroup count spreads N over that many.
uint32_t lane_run_slice(uint32_t N, uint32_t cus, uint32_t requested)
{
    const uint32_t Pmax = std::min(cus, kMaxP);
    uint32_t P = requested ? std::min(requested, Pmax) : 0u;
    if (!P) {
        uint32_t R = (N + kRun * Pmax - 1) / (kRun * Pmax);
        if (R == 3) R = 4;
        P = (N + kRun * R - 1) / (kRun * R);
    }
    const uint32_t S = (N + P - 1) / P;
    return (S + kSliceAlign - 1) / kSliceAlign * kSliceAlign;
}

// The launch shape in the statistics, from context creation on.
void shape_stats(brr_hip_ctx* c)
{
    c->stats.workgroups = c->P;
    c->stats.block_markers = c->T ? (uint32_t)kBB : c->B;
    c->stats.slice = c->S;
    c->stats.storage = c->codes ? 2u : 1u;
}

int reset_exchange(brr_hip_ctx* c)
{
    HIP_TRY(c, hipMemsetAsync(c->d_xchg1, 0, sizeof(uint64_t) * c->x1_words, c->stream));
    HIP_TRY(c, hipMemsetAsync(c->d_xchg2, 0, sizeof(uint64_t) * c->x2_words, c->stream));
    HIP_TRY(c, hipMemsetAsync(c->d_ctl, 0, sizeof(uint32_t) * 8, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return BRR_HIP_OK;
}

// One cooperative launch of the context's kernel with argument block `param`.
int launch_fn(brr_hip_ctx* c, const void* fn, void* param)
{
    if (!fn) return fail(c, BRR_HIP_EINVAL, "unsupported slice size");
    HIP_TRY(c, hipMemsetAsync(c->d_ctl + kCtlAbort, 0, sizeof(uint32_t) * 2, c->stream));
    HIP_TRY(c, hipMemsetAsync(c->d_counters, 0, sizeof(unsigned long long) * kCounters, c->stream));
    void* args[] = {param};
    HIP_TRY(c, hipEventRecord(c->ev0, c->stream));
    HIP_TRY(c, hipLaunchCooperativeKernel(fn, dim3(c->P), dim3(kThreads), args, 0, c->stream));
    HIP_TRY(c, hipEventRecord(c->ev1, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIP_TRY(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    uint32_t ctl[4] = {0, 0, 0, 0};
    unsigned long long counters[kCounters] = {};
    HIP_TRY(c, hipMemcpy(ctl, c->d_ctl, sizeof(ctl), hipMemcpyDeviceToHost));
    HIP_TRY(c, hipMemcpy(counters, c->d_counters, sizeof(counters), hipMemcpyDeviceToHost));
    std::memcpy(c->last_counters, counters, sizeof(counters));
    c->stats.kernel_ms = ms;
    c->stats.rounds = counters[0];
    c->stats.changes = counters[1];
    c->stats.decision_rounds = counters[8];
    for (int q = 0; q < 6; ++q) c->phase_ms[q] = (double)counters[2 + q] * 1e-5;
    for (int q = 6; q < kBPhases; ++q) c->phase_ms[q] = (double)counters[4 + q] * 1e-5;
    shape_stats(c);
    if (ctl[kCtlAbort]) {
        reset_exchange(c);
        if (c->bw && ctl[kCtlAbortWhere] >= 30 && ctl[kCtlAbortWhere] <= 33)
            return fail(c, BRR_HIP_ESAMPLER, std::string("adaptive rejection sampler failed for ") +
                                                 (ctl[kCtlAbortWhere] == 30 ? "mu" :
                                                  ctl[kCtlAbortWhere] == 31 ? "beta" :
                                                  ctl[kCtlAbortWhere] == 32 ? "alpha" : "a fixed effect") +
                                                 " with error " + std::to_string(counters[9]));
        return fail(c, BRR_HIP_ETIMEOUT, "device exchange wait exceeded its bound (site " +
                                             std::to_string(ctl[kCtlAbortWhere]) + ")");
    }
    return BRR_HIP_OK;
}

int launch(brr_hip_ctx* c, SweepParams& p)
{
    p.refresh_blocks = std::max(1u, c->refresh_markers / c->B);
    if (c->f32 && !c->codes)
        return fail(c, BRR_HIP_ESTATE, "the FP32 residual mode needs the columns held as 2-bit codes "
                                       "(ragged_storage 0 or 2)");
    return launch_fn(c, kernel_for(c->R, c->layout, c->codes, c->async != 0, c->f32), &p);
}

// Batched contexts: the same launch with the sweep-mode / chain-mode fields of p.
int launch_batch(brr_hip_ctx* c, const SweepParams& p)
{
    BatchParams b = batch_params(c);
    b.mode = p.mode;
    b.iterations = p.iterations;
    b.iter_base = p.iter_base;
    b.chain = p.chain;
    b.seed = p.seed;
    b.order = p.order;
    b.u = p.u;
    b.z = p.z;
    b.hist = p.hist;
    const void* fn = batch_kernel_for(c->nt, c->layout, c->f32);
    if (!fn)
        return fail(c, BRR_HIP_ESTATE, "batched slices of more than 256 individuals need the FP32 residual "
                                       "(brr_hip_set_precision, BRR_HIP_PRECISION_FP32)");
    return launch_fn(c, fn, &b);
}

int launch_any(brr_hip_ctx* c, SweepParams& p)
{
    return c->T ? launch_batch(c, p) : launch(c, p);
}

// ------------------------------------------------- allocation-free sweep --
// brr_hip_sweep_noalloc keeps its device memory -- the plain sweep kernel's
// tagged exchange words (partials and totals), its control words and launch
// counters, and a bitmap that checks the caller's order -- in this file-scope
// array. Each GPU holds its own copy for the life of the process. Layout in
// 64-bit words: partials, totals, then from kOffCtl on 4 words of control
// words, 16 of counters and 1 of the order flag (the same order as the
// pinned status block), then the bitmap of up to kScratchMaxM markers.
constexpr size_t kScratchX1 = (size_t)2 * kSlots * kMaxP * 2;
constexpr size_t kScratchX2 = (size_t)kCopies * kCopyWords;
constexpr size_t kScratchMaxM = (size_t)1 << 22;
constexpr size_t kOffX2 = kScratchX1;
constexpr size_t kOffCtl = (kOffX2 + kScratchX2 + 31) / 32 * 32;
constexpr size_t kOffCnt = kOffCtl + 4;
constexpr size_t kOffFlag = kOffCnt + 16;
constexpr size_t kOffMap = (kOffFlag + 1 + 31) / 32 * 32;
constexpr size_t kScratchWords = kOffMap + kScratchMaxM / 64;   // 1.53 MiB
static __device__ __align__(256) uint64_t scratch[kScratchWords];

// The device address of this GPU's copy of `scratch`, looked up once per device
// (with that device current) by hipGetSymbolAddress; the array's own name is a
// host-side stand-in and is never passed to a kernel.
std::mutex g_scratch_mu;
uint64_t* g_scratch_ptr[64] = {};

__global__ void scratch_zero_kernel(uint64_t* p, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// Before a launch: the abort words (the round word is kept: tags only grow), the
// counters, the order flag and the first `map_words` bitmap words.
__global__ void scratch_reset_kernel(uint64_t* base, size_t map_words)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t* ctl = (uint32_t*)(base + kOffCtl);
    if (t == 0) {
        ctl[kCtlAbort] = 0u;
        ctl[kCtlAbortWhere] = 0u;
    }
    if (t < 17) base[kOffCnt + t] = 0;
    for (size_t i = t; i < map_words; i += (size_t)gridDim.x * blockDim.x) base[kOffMap + i] = 0;
}

// Marks every marker of the order in the bitmap; flags an id >= M or a repeat.
__global__ void order_check_kernel(const uint32_t* order, uint32_t M, uint64_t* base)
{
    uint32_t* flag = (uint32_t*)(base + kOffFlag);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x) {
        const uint32_t j = order[i];
        if (j >= M) {
            atomicOr(flag, 1u);
            continue;
        }
        const unsigned long long bit = 1ull << (j & 63u);
        const unsigned long long old = atomicOr((unsigned long long*)(base + kOffMap + (j >> 6)), bit);
        if (old & bit) atomicOr(flag, 1u);
    }
}

int scratch_for(brr_hip_ctx* c, uint64_t** out)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (c->device < 0 || c->device >= 64) return fail(c, BRR_HIP_EINVAL, "device ordinal beyond 63");
    if (!g_scratch_ptr[c->device]) {
        void* ptr = nullptr;
        HIP_TRY(c, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return fail(c, BRR_HIP_EDEVICE, "hipGetSymbolAddress returned a null scratch address");
        scratch_zero_kernel<<<256, 256, 0, c->stream>>>((uint64_t*)ptr, kScratchWords);
        HIP_TRY(c, hipGetLastError());
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        g_scratch_ptr[c->device] = (uint64_t*)ptr;
    }
    *out = g_scratch_ptr[c->device];
    return BRR_HIP_OK;
}

// Launch shape of a BayesR context for `workgroups` requested (0: the default
// rule of brr_hip_create) and `block` markers per speculation block (0: keep;
// single-chain contexts): slice, workgroup count, lane runs or 16-row tiles;
// the exchange partials grow if needed and index-list storage gets its segment
// table rebuilt for the new slices. On error the context keeps its shape.
int configure_shape(brr_hip_ctx* c, uint32_t workgroups, uint32_t block)
{
    const uint32_t N = c->cfg.num_inds, M = c->cfg.num_snps;
    const uint32_t Pmax = std::min(c->cus, kMaxP);
    if (block > (uint32_t)kMaxB) return fail(c, BRR_HIP_EINVAL, "block_markers must be 0..64");
    uint32_t S, R = 0, nt = 0, NT = 0;
    const void* probe;
    if (c->T) {
        const uint32_t Preq = workgroups ? std::min(workgroups, Pmax) : Pmax;
        S = (N + Preq - 1) / Preq;
        S = (S + kSliceAlign - 1) / kSliceAlign * kSliceAlign;
        nt = S / 16;
        if (nt > (uint32_t)kMaxNT) return fail(c, BRR_HIP_EINVAL, "batched slices hold at most 512 individuals");
        NT = nt;
        probe = batch_kernel_for(nt, c->layout, c->f32 || nt > (uint32_t)kMaxNT64);
    } else {
        S = lane_run_slice(N, c->cus, workgroups);
        R = (S + kRun - 1) / kRun;
        if (R == 3) R = 4;
        if (R > (uint32_t)kMaxRuns) return fail(c, BRR_HIP_EINVAL, "slices hold at most 2048 individuals");
        probe = kernel_for(R, c->layout, c->codes || c->layout == kLayoutEigen, c->async != 0, c->f32);
    }
    const uint32_t P = (N + S - 1) / S;
    int occ = 0;
    if (!probe || hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, probe, kThreads, 0) != hipSuccess ||
        (uint64_t)occ * c->cus < P)
        return fail(c, BRR_HIP_EDEVICE, "the launch shape does not fit the device");
    const size_t x1 = c->T ? (size_t)2 * P * kBVmax * 2 : (size_t)2 * kSlots * P * 2;
    int rc;
    if (x1 > c->x1_cap) {
        c->x1_cap = 0;
        if ((rc = dalloc(c, &c->d_xchg1, x1))) return rc;
        c->x1_cap = x1;
    }
    const bool seg = !c->T && !c->bw && !c->codes && c->have_markers && c->d_ones;
    if (seg) {
        if ((rc = dalloc(c, &c->d_seg, (size_t)M * 3 * (P + 1)))) return rc;
        const uint64_t total = (uint64_t)M * 3 * (P + 1);
        build_seg_kernel<<<(uint32_t)((total + 255) / 256), 256, 0, c->stream>>>(
            c->d_ones, c->d_ones_off, c->d_twos, c->d_twos_off, c->d_miss, c->d_miss_off, M, P, S, c->d_seg);
        HIP_TRY(c, hipGetLastError());
    }
    c->S = S;
    c->P = P;
    c->x1_words = x1;
    if (c->T) {
        c->nt = nt;
        c->NT = NT;
    } else {
        c->R = R;
        if (block) {
            c->B_seq = block;
            if (!c->async) c->B = block;
        }
    }
    shape_stats(c);
    return reset_exchange(c);
}

int check_ready(brr_hip_ctx* c)
{
    if (!c) return BRR_HIP_EINVAL;
    if (!c->have_markers) return fail(c, BRR_HIP_ESTATE, "markers not uploaded");
    if (!c->have_model) return fail(c, BRR_HIP_ESTATE, "model not set (brr_hip_set_model / brr_hip_bw_set_model)");
    return BRR_HIP_OK;
}

// BayesR calls on a BayesW context and the reverse are refused.
int need_kind(brr_hip_ctx* c, bool bw)
{
    if (!c) return BRR_HIP_EINVAL;
    if (c->bw != bw)
        return fail(c, BRR_HIP_ESTATE, bw ? "BayesW call on a BayesR context" : "BayesR call on a BayesW context");
    return BRR_HIP_OK;
}

int check_order(brr_hip_ctx* c, const uint32_t* orders, uint32_t count)
{
    const uint32_t M = c->cfg.num_snps;
    std::vector<uint8_t> seen(M);
    for (uint32_t it = 0; it < count; ++it) {
        std::fill(seen.begin(), seen.end(), 0);
        const uint32_t* o = orders + (size_t)it * M;
        for (uint32_t i = 0; i < M; ++i) {
            if (o[i] >= M || seen[o[i]]) return fail(c, BRR_HIP_EINVAL, "order is not a permutation of 0..M-1");
            seen[o[i]] = 1;
        }
    }
    return BRR_HIP_OK;
}

} // namespace

extern "C" {

int brr_hip_abi_version(void) { return BRR_HIP_ABI_VERSION; }

int brr_hip_device_count(int* count)
{
    if (!count) return BRR_HIP_EINVAL;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return BRR_HIP_OK;
}

int brr_hip_create(const brr_hip_config* cfg, brr_hip_ctx** out)
{
    if (!cfg || !out) return BRR_HIP_EINVAL;
    *out = nullptr;
    if (cfg->num_inds < 2 || cfg->num_snps < 1) return BRR_HIP_EINVAL;
    if (cfg->num_components < 2 || cfg->num_components > (uint32_t)kMaxK) return BRR_HIP_EINVAL;
    if (cfg->num_groups < 1 || cfg->num_groups > (uint32_t)kMaxG) return BRR_HIP_EINVAL;
    if (cfg->layout != BRR_HIP_LAYOUT_DENSE_2BIT && cfg->layout != BRR_HIP_LAYOUT_RAGGED &&
        cfg->layout != BRR_HIP_LAYOUT_EIGEN_SPARSE)
        return BRR_HIP_EINVAL;
    if (cfg->block_markers > (uint32_t)kMaxB) return BRR_HIP_EINVAL;
    if (cfg->ragged_storage > 2) return BRR_HIP_EINVAL;
    if (cfg->num_traits > (uint32_t)kMaxT) return BRR_HIP_EINVAL;
    if (cfg->num_traits > 0 && cfg->num_groups != 1) return BRR_HIP_EINVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return BRR_HIP_EDEVICE;
    if ((int)cfg->device >= ndev) return BRR_HIP_EINVAL;

    brr_hip_ctx* c = new brr_hip_ctx();
    c->cfg = *cfg;
    c->device = (int)cfg->device;
    c->layout = (int)cfg->layout;
    c->codes = c->layout == kLayoutDense || c->layout == kLayoutEigen;
    if (hipSetDevice(c->device) != hipSuccess) { delete c; return BRR_HIP_EDEVICE; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->device) != hipSuccess) { delete c; return BRR_HIP_EDEVICE; }
    int coop = 0;
    (void)hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, c->device);
    if (!coop) { delete c; return BRR_HIP_EDEVICE; }

    // Slice partition: S multiple of 16 individuals, P = ceil(N / S) <= CUs.
    const uint32_t N = cfg->num_inds;
    const uint32_t cus = (uint32_t)prop.multiProcessorCount;
    c->cus = cus;
    uint32_t Preq = cfg->workgroups ? cfg->workgroups : cus;
    Preq = std::min(std::min(Preq, cus), kMaxP);
    uint32_t S = (N + Preq - 1) / Preq;
    S = (S + kSliceAlign - 1) / kSliceAlign * kSliceAlign;
    c->T = cfg->num_traits;
    c->Tn = c->T ? c->T : 1u;
    const void* probe = nullptr;
    if (c->T) {
        // batched traits: residuals of a slice in registers as 16-row tiles;
        // ragged columns are held as 2-bit codes
        c->nt = S / 16;
        if (c->nt > (uint32_t)kMaxNT) { delete c; return BRR_HIP_EINVAL; }
        c->NT = c->nt;
        c->codes = true;
        c->B = kBB;
        probe = batch_kernel_for(c->nt, c->layout, c->nt > (uint32_t)kMaxNT64);
    } else {
        S = lane_run_slice(N, cus, cfg->workgroups);
        uint32_t R = (S + kRun - 1) / kRun;
        if (R > (uint32_t)kMaxRuns) { delete c; return BRR_HIP_EINVAL; }
        if (R == 3) R = 4;
        c->R = R;
        c->B = cfg->block_markers ? cfg->block_markers : 64u;
        c->B_seq = c->B;
        probe = kernel_for(R, (int)cfg->layout, c->layout == kLayoutEigen);
    }
    c->S = S;
    c->P = (N + S - 1) / S;
    shape_stats(c);
    c->x1_words = c->T ? (size_t)2 * c->P * kBVmax * 2 : (size_t)2 * kSlots * c->P * 2;
    c->x2_words = c->T ? kCopies * kBCopyWords : kCopies * kCopyWords;

    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, probe, kThreads, 0) != hipSuccess ||
        (uint64_t)occ * cus < c->P) {
        delete c;
        return BRR_HIP_EDEVICE;
    }

    const uint32_t M = cfg->num_snps, G = cfg->num_groups, K = cfg->num_components;
    const size_t Tn = c->Tn;
    const ScalarLayout SL{G, K};
    int rc = BRR_HIP_OK;
    if ((rc = dalloc(c, &c->d_eps, N * Tn)) || (rc = dalloc(c, &c->d_beta, M * Tn)) ||
        (rc = dalloc(c, &c->d_comp, M * Tn)) || (rc = dalloc(c, &c->d_acum, M * Tn)) ||
        (rc = dalloc(c, &c->d_scal, SL.size() * Tn)) || (rc = dalloc(c, &c->d_cva, G * (K - 1))) ||
        (rc = dalloc(c, &c->d_xchg1, c->x1_words)) ||
        (rc = dalloc(c, &c->d_xchg2, c->x2_words)) || (rc = dalloc(c, &c->d_ctl, 8)) ||
        (rc = dalloc(c, &c->d_counters, kCounters)) || (rc = dalloc(c, &c->d_tmp, 4)) ||
        (rc = dalloc(c, &c->d_u, M * Tn)) || (rc = dalloc(c, &c->d_z, M * Tn))) {
        brr_hip_destroy(c);
        return rc;
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipHostMalloc((void**)&c->h_status, sizeof(uint64_t) * kStatusWords) != hipSuccess) {
        brr_hip_destroy(c);
        return BRR_HIP_EDEVICE;
    }
    (void)hipMemset(c->d_eps, 0, sizeof(double) * N * Tn);
    (void)hipMemset(c->d_beta, 0, sizeof(double) * M * Tn);
    (void)hipMemset(c->d_comp, 0, sizeof(int32_t) * M * Tn);
    (void)hipMemset(c->d_acum, 0, sizeof(double) * M * Tn);
    (void)hipMemset(c->d_scal, 0, sizeof(double) * SL.size() * Tn);
    c->x1_cap = c->x1_words;
    if (reset_exchange(c) != BRR_HIP_OK) { brr_hip_destroy(c); return BRR_HIP_EDEVICE; }
    *out = c;
    return BRR_HIP_OK;
}

void brr_hip_destroy(brr_hip_ctx* c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    dfree(c->d_dense); dfree(c->d_lut);
    dfree(c->d_ones); dfree(c->d_twos); dfree(c->d_miss);
    dfree(c->d_ones_off); dfree(c->d_twos_off); dfree(c->d_miss_off);
    dfree(c->d_seg); dfree(c->d_stats);
    dfree(c->d_group); dfree(c->d_cva);
    dfree(c->d_fixedX); dfree(c->d_fixed); dfree(c->d_forder); dfree(c->d_sumf_fix);
    dfree(c->d_eps); dfree(c->d_beta); dfree(c->d_comp); dfree(c->d_acum);
    dfree(c->d_scal); dfree(c->d_hist); dfree(c->d_order); dfree(c->d_u); dfree(c->d_z);
    dfree(c->d_xchg1); dfree(c->d_xchg2); dfree(c->d_ctl); dfree(c->d_counters); dfree(c->d_tmp);
    dfree(c->d_fail); dfree(c->d_sumf); dfree(c->d_num); dfree(c->d_keps); dfree(c->d_cohort);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->h_status) (void)hipHostFree(c->h_status);
    delete c;
}

const char* brr_hip_last_error(const brr_hip_ctx* c)
{
    return c ? c->err.c_str() : "null context";
}

int brr_hip_upload_dense(brr_hip_ctx* c, const uint8_t* bed, uint64_t column_stride, const double* lut)
{
    if (!c || !bed || !lut) return fail(c, BRR_HIP_EINVAL, "null argument");
    if (c->layout != BRR_HIP_LAYOUT_DENSE_2BIT && !c->bw)
        return fail(c, BRR_HIP_EINVAL, "context layout is not dense");
    const uint32_t N = c->cfg.num_inds, M = c->cfg.num_snps;
    const uint64_t col_bytes = ((uint64_t)N + 3) / 4;
    if (column_stride < col_bytes) return fail(c, BRR_HIP_EINVAL, "column_stride < ceil(N/4)");
    HIP_TRY(c, hipSetDevice(c->device));
    c->col_bytes = col_bytes;
    c->col_words = (col_bytes + 15) / 16 * 4;   // 16-byte aligned columns
    int rc = dalloc(c, &c->d_dense, (size_t)c->col_words * M);
    if (rc) return rc;
    if ((rc = dalloc(c, &c->d_lut, (size_t)4 * M))) return rc;
    HIP_TRY(c, hipMemset(c->d_dense, 0, sizeof(uint32_t) * c->col_words * M));
    HIP_TRY(c, hipMemcpy2D(c->d_dense, c->col_words * 4, bed, column_stride, col_bytes, M,
                           hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->d_lut, lut, sizeof(double) * 4 * M, hipMemcpyHostToDevice));
    c->have_markers = true;
    if (c->bw) {
        // a BayesW context takes its layout from the markers it is given
        c->layout = kLayoutDense;
        c->cfg.layout = BRR_HIP_LAYOUT_DENSE_2BIT;
        c->have_model = false;   // sum_failure depends on the columns
    }
    return BRR_HIP_OK;
}

int brr_hip_upload_ragged(brr_hip_ctx* c, const int32_t* ones, const uint64_t* ones_off,
                          const int32_t* twos, const uint64_t* twos_off, const int32_t* miss,
                          const uint64_t* miss_off, const double* stats)
{
    if (!c || !ones_off || !twos_off || !miss_off || !stats) return fail(c, BRR_HIP_EINVAL, "null argument");
    if (c->bw) {
        c->layout = kLayoutRagged;
        c->cfg.layout = BRR_HIP_LAYOUT_RAGGED;
    }
    if (c->layout != BRR_HIP_LAYOUT_RAGGED && c->layout != BRR_HIP_LAYOUT_EIGEN_SPARSE)
        return fail(c, BRR_HIP_EINVAL, "context layout is not ragged or EigenSparse");
    const bool eigen = c->layout == kLayoutEigen;
    const uint32_t N = c->cfg.num_inds, M = c->cfg.num_snps;
    const uint64_t n1 = ones_off[M], n2 = twos_off[M], nm = miss_off[M];
    if ((n1 && !ones) || (n2 && !twos) || (nm && !miss)) return fail(c, BRR_HIP_EINVAL, "null list");
    // validate: offsets non-decreasing, indices ascending and in range
    const int32_t* lists[3] = {ones, twos, miss};
    const uint64_t* offs[3] = {ones_off, twos_off, miss_off};
    for (int l = 0; l < 3; ++l) {
        if (offs[l][0] != 0) return fail(c, BRR_HIP_EINVAL, "offsets must start at 0");
        for (uint32_t j = 0; j < M; ++j) {
            if (offs[l][j + 1] < offs[l][j]) return fail(c, BRR_HIP_EINVAL, "offsets not monotone");
            int64_t prev = -1;
            for (uint64_t q = offs[l][j]; q < offs[l][j + 1]; ++q) {
                const int64_t x = lists[l][q];
                if (x <= prev || x >= (int64_t)N)
                    return fail(c, BRR_HIP_EINVAL, "ragged index out of range or not ascending in column " +
                                                       std::to_string(j));
                prev = x;
            }
        }
    }
    HIP_TRY(c, hipSetDevice(c->device));
    int rc;
    if ((rc = dalloc(c, &c->d_ones, n1)) || (rc = dalloc(c, &c->d_twos, n2)) ||
        (rc = dalloc(c, &c->d_miss, nm)) || (rc = dalloc(c, &c->d_ones_off, M + 1)) ||
        (rc = dalloc(c, &c->d_twos_off, M + 1)) || (rc = dalloc(c, &c->d_miss_off, M + 1)) ||
        (rc = dalloc(c, &c->d_stats, (size_t)4 * M)) ||
        (rc = dalloc(c, &c->d_seg, (c->bw || eigen) ? (size_t)1 : (size_t)M * 3 * (c->P + 1))))
        return rc;
    if (n1) HIP_TRY(c, hipMemcpy(c->d_ones, ones, sizeof(int32_t) * n1, hipMemcpyHostToDevice));
    if (n2) HIP_TRY(c, hipMemcpy(c->d_twos, twos, sizeof(int32_t) * n2, hipMemcpyHostToDevice));
    if (nm) HIP_TRY(c, hipMemcpy(c->d_miss, miss, sizeof(int32_t) * nm, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->d_ones_off, ones_off, sizeof(uint64_t) * (M + 1), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->d_twos_off, twos_off, sizeof(uint64_t) * (M + 1), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->d_miss_off, miss_off, sizeof(uint64_t) * (M + 1), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->d_stats, stats, sizeof(double) * 4 * M, hipMemcpyHostToDevice));
    if (!c->bw && !eigen) {
        const uint64_t total = (uint64_t)M * 3 * (c->P + 1);
        const uint32_t blocks = (uint32_t)((total + 255) / 256);
        build_seg_kernel<<<blocks, 256, 0, c->stream>>>(c->d_ones, c->d_ones_off, c->d_twos, c->d_twos_off,
                                                       c->d_miss, c->d_miss_off, M, c->P, c->S, c->d_seg);
        HIP_TRY(c, hipGetLastError());
    }
    // device copy as 2-bit codes when asked, or when smaller than the lists
    const uint64_t list_bytes = 4ull * (n1 + n2 + nm);
    const uint64_t code_bytes = (uint64_t)M * (((uint64_t)N + 3) / 4);
    const uint32_t st = c->cfg.ragged_storage;
    c->codes = c->T > 0 || c->bw || eigen || st == 2 || (st == 0 && code_bytes < list_bytes);
    if (c->codes) {
        c->col_bytes = ((uint64_t)N + 3) / 4;
        c->col_words = (c->col_bytes + 15) / 16 * 4;
        if ((rc = dalloc(c, &c->d_dense, (size_t)c->col_words * M))) return rc;
        ragged_to_codes_kernel<<<M, 256, 0, c->stream>>>(c->d_ones, c->d_ones_off, c->d_twos, c->d_twos_off,
                                                        c->d_miss, c->d_miss_off, c->col_words, c->d_dense);
        HIP_TRY(c, hipGetLastError());
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (c->bw || eigen) {
        // the BayesW and EigenSparse paths read the codes only (the per-marker
        // BayesR kernels too, once the lists are gone)
        dfree(c->d_ones); dfree(c->d_twos); dfree(c->d_miss);
    }
    c->have_markers = true;
    c->have_model = c->have_model && !c->bw;   // sum_failure depends on the columns
    return BRR_HIP_OK;
}

int brr_hip_load_bed(brr_hip_ctx* c, const uint8_t* bed_file, uint64_t file_bytes, int layout)
{
    if (!c || !bed_file) return fail(c, BRR_HIP_EINVAL, "null argument");
    if (layout != BRR_HIP_LAYOUT_DENSE_2BIT && layout != BRR_HIP_LAYOUT_RAGGED &&
        (layout != BRR_HIP_LAYOUT_EIGEN_SPARSE || c->bw))
        return fail(c, BRR_HIP_EINVAL, c->bw ? "layout must be dense or ragged" : "unknown layout");
    if (!c->bw && layout != c->layout) return fail(c, BRR_HIP_EINVAL, "layout differs from the context's");
    const uint32_t N = c->cfg.num_inds, M = c->cfg.num_snps;
    const uint64_t cb = ((uint64_t)N + 3) / 4;
    if (file_bytes < 3 || bed_file[0] != 0x6c || bed_file[1] != 0x1b || bed_file[2] != 0x01)
        return fail(c, BRR_HIP_EINVAL, "not a SNP-major PLINK .bed file (magic 6c 1b 01)");
    if (file_bytes != 3 + cb * M)
        return fail(c, BRR_HIP_EINVAL, "file size " + std::to_string(file_bytes) + " != 3 + M * ceil(N/4) = " +
                                           std::to_string(3 + cb * M));
    if (N < 2 || M == 0) return fail(c, BRR_HIP_EINVAL, "need N >= 2 and M >= 1");
    HIP_TRY(c, hipSetDevice(c->device));
    const bool dense = layout == BRR_HIP_LAYOUT_DENSE_2BIT;
    c->have_markers = false;
    // the packed payload, with 32 zero bytes past its end for the unpacker's
    // five-word reads
    const uint64_t payload = cb * M;
    const uint64_t stage_words = (payload + 32 + 3) / 4;
    uint32_t* d_stage = nullptr;
    uint32_t* d_cnt = nullptr;
    int rc;
    if ((rc = dalloc(c, &d_stage, stage_words)) || (rc = dalloc(c, &d_cnt, (size_t)4 * M + 1))) {
        dfree(d_stage);
        dfree(d_cnt);
        return rc;
    }
    auto cleanup = [&]() { dfree(d_stage); dfree(d_cnt); };
    auto hip_ok = [&](hipError_t e, const char* what) {
        if (e == hipSuccess) return true;
        fail(c, BRR_HIP_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
        cleanup();
        return false;
    };
    if (!hip_ok(hipMemsetAsync((uint8_t*)d_stage + (payload & ~3ull), 0, stage_words * 4 - (payload & ~3ull),
                               c->stream), "hipMemsetAsync") ||
        !hip_ok(hipMemcpyAsync(d_stage, bed_file + 3, payload, hipMemcpyHostToDevice, c->stream), "hipMemcpyAsync"))
        return BRR_HIP_EDEVICE;
    // the arena: columns as 2-bit codes plus the per-marker table / statistics
    c->col_bytes = cb;
    c->col_words = (cb + 15) / 16 * 4;
    dfree(c->d_ones); dfree(c->d_twos); dfree(c->d_miss);
    dfree(c->d_ones_off); dfree(c->d_twos_off); dfree(c->d_miss_off); dfree(c->d_seg);
    if ((rc = dalloc(c, &c->d_dense, (size_t)c->col_words * M)) ||
        (rc = dalloc(c, dense ? &c->d_lut : &c->d_stats, (size_t)4 * M))) {
        cleanup();
        return rc;
    }
    uint32_t* d_bad = d_cnt + 4 * (size_t)M;
    if (!hip_ok(hipMemsetAsync(d_bad, 0xFF, sizeof(uint32_t), c->stream), "hipMemsetAsync") ||
        !hip_ok(hipEventRecord(c->ev0, c->stream), "hipEventRecord"))
        return BRR_HIP_EDEVICE;
    bed_unpack_kernel<<<M, kBedThreads, 0, c->stream>>>(d_stage, cb, N, c->col_words, c->d_dense, d_cnt,
                                                        dense ? nullptr : c->d_stats, d_bad,
                                                        layout == BRR_HIP_LAYOUT_EIGEN_SPARSE ? 1 : 0);
    if (!hip_ok(hipGetLastError(), "bed_unpack_kernel") || !hip_ok(hipEventRecord(c->ev1, c->stream), "hipEventRecord") ||
        !hip_ok(hipStreamSynchronize(c->stream), "bed_unpack_kernel"))
        return BRR_HIP_EDEVICE;
    float ms_unpack = 0.f, ms_table = 0.f;
    (void)hipEventElapsedTime(&ms_unpack, c->ev0, c->ev1);
    if (dense) {
        if (!hip_ok(hipEventRecord(c->ev0, c->stream), "hipEventRecord")) return BRR_HIP_EDEVICE;
        bed_dense_table_kernel<<<(M + kBedThreads - 1) / kBedThreads, kBedThreads, 0, c->stream>>>(
            c->d_dense, c->col_words, N, M, d_cnt, c->d_lut);
        if (!hip_ok(hipGetLastError(), "bed_dense_table_kernel") ||
            !hip_ok(hipEventRecord(c->ev1, c->stream), "hipEventRecord") ||
            !hip_ok(hipStreamSynchronize(c->stream), "bed_dense_table_kernel"))
            return BRR_HIP_EDEVICE;
        (void)hipEventElapsedTime(&ms_table, c->ev0, c->ev1);
    }
    uint32_t bad = 0xFFFFFFFFu;
    if (!hip_ok(hipMemcpy(&bad, d_bad, sizeof(bad), hipMemcpyDeviceToHost), "hipMemcpy")) return BRR_HIP_EDEVICE;
    cleanup();
    c->stats = brr_hip_stats{};
    shape_stats(c);
    c->stats.kernel_ms = (double)ms_unpack + (double)ms_table;
    c->stats.markers = M;
    std::fill(c->phase_ms, c->phase_ms + kBPhases, 0.0);
    c->phase_ms[0] = ms_unpack;
    c->phase_ms[1] = ms_table;
    if (bad != 0xFFFFFFFFu)
        return fail(c, BRR_HIP_EINVAL, "marker " + std::to_string(bad) +
                                           (layout == BRR_HIP_LAYOUT_EIGEN_SPARSE ? " has no nonzero call"
                                                                                  : " has no genotype 1 or 2 calls"));
    c->codes = true;
    if (c->bw) {
        c->layout = dense ? kLayoutDense : kLayoutRagged;
        c->cfg.layout = (uint32_t)layout;
        c->have_model = false;   // sum_failure depends on the columns
    }
    c->have_markers = true;
    return BRR_HIP_OK;
}

int brr_hip_get_marker_tables(brr_hip_ctx* c, double* out)
{
    if (!c || !out) return fail(c, BRR_HIP_EINVAL, "null argument");
    if (!c->have_markers) return fail(c, BRR_HIP_ESTATE, "markers not uploaded");
    const double* src = c->layout == kLayoutDense ? c->d_lut : c->d_stats;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMemcpy(out, src, sizeof(double) * 4 * c->cfg.num_snps, hipMemcpyDeviceToHost));
    return BRR_HIP_OK;
}

int brr_hip_set_async(brr_hip_ctx* c, uint32_t window)
{
    if (int k = need_kind(c, false)) return k;
    if (c->T) return fail(c, BRR_HIP_EINVAL, "asynchronous windows need a single-chain context");
    if (window > (uint32_t)kMaxB) return fail(c, BRR_HIP_EINVAL, "window must be 0 (sequential) or 1..64");
    if (window && c->f32) return fail(c, BRR_HIP_EINVAL, "the FP32 residual mode runs the sequential sweep only");
    c->async = window ? 1u : 0u;
    c->B = window ? window : c->B_seq;
    return BRR_HIP_OK;
}

int brr_hip_set_precision(brr_hip_ctx* c, int precision, uint32_t refresh_markers)
{
    if (int k = need_kind(c, false)) return k;
    if (precision == BRR_HIP_PRECISION_FP64) {
        c->f32 = false;
        return BRR_HIP_OK;
    }
    if (precision == BRR_HIP_PRECISION_FP32) {
        if (!c->T) return fail(c, BRR_HIP_EINVAL, "BRR_HIP_PRECISION_FP32 is the batched contexts' mode "
                                                  "(single chains: BRR_HIP_PRECISION_FP32_REFRESH)");
        c->f32 = true;
        return BRR_HIP_OK;
    }
    if (precision != BRR_HIP_PRECISION_FP32_REFRESH) return fail(c, BRR_HIP_EINVAL, "unknown precision");
    if (c->T) return fail(c, BRR_HIP_EINVAL, "the FP32-with-refresh mode needs a single-chain context "
                                             "(batched contexts: BRR_HIP_PRECISION_FP32)");
    if (c->async) return fail(c, BRR_HIP_EINVAL, "the FP32 residual mode runs the sequential sweep only");
    if (refresh_markers < 1 || refresh_markers > (uint32_t)kMaxLog)
        return fail(c, BRR_HIP_EINVAL, "refresh interval must be 1..1024 markers");
    c->f32 = true;
    c->refresh_markers = refresh_markers;
    return BRR_HIP_OK;
}

int brr_hip_set_launch(brr_hip_ctx* c, uint32_t workgroups, uint32_t block_markers)
{
    if (int k = need_kind(c, false)) return k;
    HIP_TRY(c, hipSetDevice(c->device));
    return configure_shape(c, workgroups, block_markers);
}

int brr_hip_autotune(brr_hip_ctx* c, uint64_t seed, uint32_t iterations, brr_hip_tune_entry* table,
                     uint32_t table_cap, uint32_t* entries)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if ((rc = need_kind(c, false))) return rc;
    if (iterations < 1 || iterations > 16) return fail(c, BRR_HIP_EINVAL, "tuning runs 1..16 iterations per shape");
    HIP_TRY(c, hipSetDevice(c->device));
    const uint32_t N = c->cfg.num_inds, M = c->cfg.num_snps;
    const size_t Tn = c->Tn;
    const size_t sl = ScalarLayout{c->cfg.num_groups, c->cfg.num_components}.size();
    // candidate shapes: (workgroups, block markers)
    const uint32_t Pmax = std::min(c->cus, kMaxP);
    std::vector<std::pair<uint32_t, uint32_t>> cand;
    auto add = [&](uint32_t P, uint32_t B) {
        if (P < 1 || P > Pmax) return;
        for (auto& q : cand)
            if (q.first == P && q.second == B) return;
        cand.emplace_back(P, B);
    };
    if (c->T) {
        for (uint32_t S : {256u, 224u, 192u, 160u, 128u}) add((N + S - 1) / S, 0);
        add(Pmax, 0);
    } else {
        const std::vector<uint32_t> blocks = c->async ? std::vector<uint32_t>{0u} : std::vector<uint32_t>{32u, 48u, 64u};
        for (uint32_t R : {1u, 2u, 4u}) {
            const uint32_t P = (N + kRun * R - 1) / (kRun * R);
            for (uint32_t B : blocks) add(P, B);
        }
        for (uint32_t B : blocks) add(Pmax, B);
    }
    // snapshot of the chain state, restored before every trial and at the end
    double *s_eps = nullptr, *s_beta = nullptr, *s_acum = nullptr, *s_scal = nullptr, *s_fixed = nullptr;
    int32_t* s_comp = nullptr;
    uint32_t* s_forder = nullptr;
    auto release = [&]() { dfree(s_eps); dfree(s_beta); dfree(s_acum); dfree(s_scal); dfree(s_comp);
                           dfree(s_fixed); dfree(s_forder); };
    if ((rc = dalloc(c, &s_eps, N * Tn)) || (rc = dalloc(c, &s_beta, M * Tn)) || (rc = dalloc(c, &s_acum, M * Tn)) ||
        (rc = dalloc(c, &s_scal, sl * Tn)) || (rc = dalloc(c, &s_comp, M * Tn)) ||
        (c->F && ((rc = dalloc(c, &s_fixed, 1 + c->F)) || (rc = dalloc(c, &s_forder, c->F))))) {
        release();
        return rc;
    }
    auto copy_state = [&](bool save) -> int {
        auto cp = [&](void* a, void* b, size_t bytes) { return hipMemcpyAsync(save ? a : b, save ? b : a, bytes,
                                                                               hipMemcpyDeviceToDevice, c->stream); };
        if (cp(s_eps, c->d_eps, sizeof(double) * N * Tn) != hipSuccess ||
            cp(s_beta, c->d_beta, sizeof(double) * M * Tn) != hipSuccess ||
            cp(s_acum, c->d_acum, sizeof(double) * M * Tn) != hipSuccess ||
            cp(s_scal, c->d_scal, sizeof(double) * sl * Tn) != hipSuccess ||
            cp(s_comp, c->d_comp, sizeof(int32_t) * M * Tn) != hipSuccess ||
            (c->F && (cp(s_fixed, c->d_fixed, sizeof(double) * (1 + c->F)) != hipSuccess ||
                      cp(s_forder, c->d_forder, sizeof(uint32_t) * c->F) != hipSuccess)) ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            return fail(c, BRR_HIP_EDEVICE, "autotune: state copy failed");
        return BRR_HIP_OK;
    };
    if ((rc = copy_state(true))) { release(); return rc; }
    std::vector<uint32_t> orders((size_t)iterations * M);
    std::vector<uint32_t> cur(M);
    for (uint32_t i = 0; i < M; ++i) cur[i] = i;
    for (uint32_t it = 0; it < iterations; ++it) {
        brr_shuffle(seed, it, 0, M, cur.data(), 0);
        std::copy(cur.begin(), cur.end(), orders.begin() + (size_t)it * M);
    }
    const uint32_t P0 = c->P, B0 = c->B_seq;
    const double rate0 = c->change_rate;
    double best_ms = 1e300;
    uint32_t best = 0, n = 0;
    for (size_t q = 0; q < cand.size(); ++q) {
        double ms = -1.0;
        if (configure_shape(c, cand[q].first, cand[q].second) == BRR_HIP_OK && copy_state(false) == BRR_HIP_OK) {
            c->change_rate = rate0;
            if (brr_hip_run_chain(c, seed, 0, 0, iterations, orders.data(), nullptr) == BRR_HIP_OK)
                ms = c->stats.kernel_ms / iterations;
        }
        if (ms >= 0.0 && ms < best_ms) { best_ms = ms; best = (uint32_t)q; }
        if (table && n < table_cap)
            table[n++] = brr_hip_tune_entry{c->P, c->T ? (uint32_t)kBB : c->B, c->S, ms};
    }
    if (entries) *entries = n;
    rc = best_ms < 1e300 ? configure_shape(c, cand[best].first, cand[best].second)
                         : configure_shape(c, P0, B0);
    const int rc2 = copy_state(false);
    c->change_rate = rate0;
    release();
    if (rc) return rc;
    if (rc2) return rc2;
    if (best_ms >= 1e300) return fail(c, BRR_HIP_EDEVICE, "autotune: no candidate shape ran");
    return BRR_HIP_OK;
}

int brr_hip_run_cohorts(brr_hip_ctx* const* ctxs, uint32_t n, uint64_t seed, const uint32_t* chains,
                        uint32_t first_iteration, uint32_t iterations, const uint32_t* const* orders,
                        double* const* hists, uint32_t flags, brr_hip_cohort_report* report)
{
    if (!ctxs || n == 0 || !ctxs[0] || !orders) return BRR_HIP_EINVAL;
    brr_hip_ctx* c0 = ctxs[0];
    if (iterations == 0) return fail(c0, BRR_HIP_EINVAL, "zero iterations");
    int rc;
    for (uint32_t q = 0; q < n; ++q) {
        brr_hip_ctx* c = ctxs[q];
        if (!c) return fail(c0, BRR_HIP_EINVAL, "null context");
        if ((rc = check_ready(c)) || (rc = need_kind(c, false))) return rc;
        if (c->T || c->async || c->F || !c->codes)
            return fail(c0, BRR_HIP_EINVAL, "cohort " + std::to_string(q) + ": cohorts are single-chain sequential "
                                            "contexts without fixed effects, columns held as 2-bit codes");
        if (c->R > 2)
            return fail(c0, BRR_HIP_EINVAL, "cohort " + std::to_string(q) + ": cohort launches hold slices of at most "
                                            "1024 individuals (two lane runs; brr_hip_set_launch)");
        if (c->device != c0->device || c->f32 != c0->f32)
            return fail(c0, BRR_HIP_EINVAL, "cohort " + std::to_string(q) + ": cohorts of one call share the device "
                                            "and the residual precision");
        if (!orders[q]) return fail(c0, BRR_HIP_EINVAL, "null orders");
        if ((rc = check_order(c, orders[q], iterations))) return fail(c0, rc, c->err);
    }
    HIP_TRY(c0, hipSetDevice(c0->device));
    const uint32_t Pmax = std::min(c0->cus, kMaxP);
    // per-cohort parameter blocks of a chain launch (orders and history uploaded)
    std::vector<SweepParams> params(n);
    for (uint32_t q = 0; q < n; ++q) {
        brr_hip_ctx* c = ctxs[q];
        const uint32_t M = c->cfg.num_snps, G = c->cfg.num_groups;
        const size_t total = (size_t)iterations * M;
        if (c->P > Pmax) return fail(c0, BRR_HIP_EINVAL, "a cohort needs more workgroups than the device has");
        if (c->order_cap < total) {
            c->order_cap = 0;
            if ((rc = dalloc(c, &c->d_order, total))) return fail(c0, rc, c->err);
            c->order_cap = total;
        }
        const size_t hsz = (size_t)iterations * (4 + G);
        if (c->hist_cap < hsz) {
            c->hist_cap = 0;
            if ((rc = dalloc(c, &c->d_hist, hsz))) return fail(c0, rc, c->err);
            c->hist_cap = hsz;
        }
        HIP_TRY(c0, hipMemcpy(c->d_order, orders[q], sizeof(uint32_t) * total, hipMemcpyHostToDevice));
        SweepParams p = base_params(c);
        p.mode = kModeChain;
        p.iterations = iterations;
        p.iter_base = first_iteration;
        p.chain = chains ? chains[q] : 0u;
        p.seed = seed;
        p.order = c->d_order;
        p.hist = c->d_hist;
        p.refresh_blocks = std::max(1u, c->refresh_markers / c->B);
        params[q] = p;
    }
    // launches: cohorts by decreasing estimated chain length in exchange rounds, first
    // fit into launches of at most one workgroup per compute unit and kMaxCohorts
    // cohorts (so each launch's first cohort sets its length)
    std::vector<uint32_t> byLen(n);
    for (uint32_t q = 0; q < n; ++q) byLen[q] = q;
    // a chain's exchange rounds per iteration are about M / B + its beta changes: the
    // change rate of the context's last launch estimates the latter
    auto rounds_est = [&](uint32_t q) {
        const brr_hip_ctx* c = ctxs[q];
        return (double)c->cfg.num_snps * (1.0 / (double)c->B + std::min(1.0, c->change_rate));
    };
    std::stable_sort(byLen.begin(), byLen.end(), [&](uint32_t a, uint32_t b) { return rounds_est(a) > rounds_est(b); });
    std::vector<std::vector<uint32_t>> launches;
    std::vector<uint32_t> used;
    for (uint32_t q : byLen) {
        size_t l = 0;
        while (l < launches.size() && (used[l] + ctxs[q]->P > Pmax || launches[l].size() >= (size_t)kMaxCohorts)) ++l;
        if (l == launches.size()) { launches.emplace_back(); used.push_back(0); }
        launches[l].push_back(q);
        used[l] += ctxs[q]->P;
    }
    // parameter blocks in launch order, on the device
    std::vector<SweepParams> packed;
    for (auto& L : launches)
        for (uint32_t q : L) packed.push_back(params[q]);
    if (c0->cohort_cap < n) {
        c0->cohort_cap = 0;
        if ((rc = dalloc(c0, &c0->d_cohort, n))) return rc;
        c0->cohort_cap = n;
    }
    HIP_TRY(c0, hipMemcpy(c0->d_cohort, packed.data(), sizeof(SweepParams) * n, hipMemcpyHostToDevice));
    // per launch, the counters' reset and the cooperative cohort kernel: recorded into
    // one graph (or, with BRR_HIP_COHORTS_DIRECT, launched one after another)
    std::vector<CohortTable> tables(launches.size());
    std::vector<const void*> fns(launches.size());
    uint32_t base = 0;
    for (size_t l = 0; l < launches.size(); ++l) {
        CohortTable& t = tables[l];
        t.params = c0->d_cohort + base;
        t.n = (uint32_t)launches[l].size();
        uint32_t wg = 0;
        for (uint32_t i = 0; i < t.n; ++i) { t.first[i] = wg; wg += ctxs[launches[l][i]]->P; }
        t.first[t.n] = wg;
        // the launch's lane runs: the most any of its cohorts needs (a slice shorter than the
        // kernel's runs is masked at its end, as in any context's last workgroup)
        uint32_t R = 1;
        for (uint32_t q : launches[l]) R = std::max(R, ctxs[q]->R);
        fns[l] = cohort_kernel_for(R, c0->f32);
        int occ = 0;
        if (!fns[l] || hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fns[l], kThreads, 0) != hipSuccess ||
            (uint64_t)occ * c0->cus < wg)
            return fail(c0, BRR_HIP_EDEVICE, "a cohort launch does not fit the device");
        base += t.n;
    }
    auto enqueue = [&]() -> hipError_t {
        for (size_t l = 0; l < launches.size(); ++l) {
            CohortTable& t = tables[l];
            cohort_reset_kernel<<<t.n, 64, 0, c0->stream>>>(t.params);
            hipError_t e2 = hipGetLastError();
            if (e2 != hipSuccess) return e2;
            void* args[] = {&t};
            e2 = hipLaunchCooperativeKernel(fns[l], dim3(t.first[t.n]), dim3(kThreads), args, 0, c0->stream);
            if (e2 != hipSuccess) return e2;
        }
        return hipSuccess;
    };
    hipError_t e = hipSuccess;
    if (flags & BRR_HIP_COHORTS_DIRECT) {
        HIP_TRY(c0, hipEventRecord(c0->ev0, c0->stream));
        e = enqueue();
        if (e == hipSuccess) e = hipEventRecord(c0->ev1, c0->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c0->stream);
        if (e != hipSuccess) return fail(c0, BRR_HIP_EDEVICE, std::string("cohort launch: ") + hipGetErrorString(e));
    } else {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        HIP_TRY(c0, hipStreamBeginCapture(c0->stream, hipStreamCaptureModeThreadLocal));
        const hipError_t cap = enqueue();
        const hipError_t endc = hipStreamEndCapture(c0->stream, &graph);
        if (cap != hipSuccess || endc != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return fail(c0, BRR_HIP_EDEVICE, std::string("graph capture of the cohort launches failed: ") +
                                                 hipGetErrorString(cap != hipSuccess ? cap : endc));
        }
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(graph);
            return fail(c0, BRR_HIP_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
        HIP_TRY(c0, hipEventRecord(c0->ev0, c0->stream));
        e = hipGraphLaunch(exec, c0->stream);
        if (e == hipSuccess) e = hipEventRecord(c0->ev1, c0->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c0->stream);
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess)
            return fail(c0, BRR_HIP_EDEVICE, std::string("cohort graph launch: ") + hipGetErrorString(e));
    }
    float ms = 0.f;
    HIP_TRY(c0, hipEventElapsedTime(&ms, c0->ev0, c0->ev1));
    // results per cohort
    int status = BRR_HIP_OK;
    for (uint32_t q = 0; q < n; ++q) {
        brr_hip_ctx* c = ctxs[q];
        uint32_t ctl[4] = {0, 0, 0, 0};
        unsigned long long counters[kCounters] = {};
        HIP_TRY(c0, hipMemcpy(ctl, c->d_ctl, sizeof(ctl), hipMemcpyDeviceToHost));
        HIP_TRY(c0, hipMemcpy(counters, c->d_counters, sizeof(counters), hipMemcpyDeviceToHost));
        std::memcpy(c->last_counters, counters, sizeof(counters));
        c->stats.kernel_ms = ms;
        c->stats.rounds = counters[0];
        c->stats.changes = counters[1];
        c->stats.markers = (uint64_t)iterations * c->cfg.num_snps;
        for (int k = 0; k < 6; ++k) c->phase_ms[k] = (double)counters[2 + k] * 1e-5;
        shape_stats(c);
        c->change_rate = (double)counters[1] / (double)c->stats.markers;
        if (ctl[kCtlAbort]) {
            reset_exchange(c);
            status = fail(c0, BRR_HIP_ETIMEOUT, "cohort " + std::to_string(q) + ": device exchange wait exceeded "
                                                "its bound (site " + std::to_string(ctl[kCtlAbortWhere]) + ")");
            continue;
        }
        if (hists && hists[q])
            HIP_TRY(c0, hipMemcpy(hists[q], c->d_hist, sizeof(double) * iterations * (4 + c->cfg.num_groups),
                                  hipMemcpyDeviceToHost));
    }
    if (report) {
        report->launches = (uint32_t)launches.size();
        report->device_ms = ms;
        for (size_t l = 0; l < launches.size(); ++l)
            for (uint32_t q : launches[l])
                if (q < 64) report->launch_of[q] = (uint32_t)l;
    }
    return status;
}

int brr_hip_probe_exchange(brr_hip_ctx* c, uint32_t rounds, double* us_per_round)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if ((rc = need_kind(c, false))) return rc;
    if (c->T) return fail(c, BRR_HIP_EINVAL, "the exchange probe runs on single-chain contexts");
    if (!us_per_round || rounds < 1 || rounds > (1u << 20)) return fail(c, BRR_HIP_EINVAL, "1..1048576 rounds");
    HIP_TRY(c, hipSetDevice(c->device));
    SweepParams p = base_params(c);
    p.mode = kModeProbe;
    p.iterations = rounds;
    const brr_hip_stats keep = c->stats;
    rc = launch(c, p);
    const double ms = c->stats.kernel_ms;
    c->stats = keep;
    if (rc) return rc;
    *us_per_round = ms * 1e3 / rounds;
    return BRR_HIP_OK;
}

__global__ void calibration_copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

int brr_hip_calibration_copy(uint32_t device, uint64_t bytes, double* ms)
{
    if (!ms || bytes < 16 || bytes % 16) return BRR_HIP_EINVAL;
    if (hipSetDevice((int)device) != hipSuccess) return BRR_HIP_EDEVICE;
    uint4 *a = nullptr, *b = nullptr;
    if (hipMalloc((void**)&a, bytes) != hipSuccess) return BRR_HIP_ENOMEM;
    if (hipMalloc((void**)&b, bytes) != hipSuccess) { (void)hipFree(a); return BRR_HIP_ENOMEM; }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = BRR_HIP_OK;
    if (hipMemset(a, 1, bytes) != hipSuccess || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        rc = BRR_HIP_EDEVICE;
    float t = 0.f;
    if (!rc) {
        (void)hipEventRecord(e0, nullptr);
        calibration_copy_kernel<<<1024, 256>>>(a, b, bytes / 16);
        (void)hipEventRecord(e1, nullptr);
        if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&t, e0, e1) != hipSuccess)
            rc = BRR_HIP_EDEVICE;
    }
    *ms = t;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}

int brr_hip_record_num(brr_hip_ctx* c, int enable)
{
    if (int k = need_kind(c, false)) return k;
    if (c->T) return fail(c, BRR_HIP_EINVAL, "num records need a single-chain context");
    HIP_TRY(c, hipSetDevice(c->device));
    c->record_num = false;
    if (!enable) return BRR_HIP_OK;
    if (!c->d_num) {
        int rc = dalloc(c, &c->d_num, c->cfg.num_snps);
        if (rc) return rc;
    }
    HIP_TRY(c, hipMemset(c->d_num, 0, sizeof(double) * c->cfg.num_snps));
    c->record_num = true;
    return BRR_HIP_OK;
}

int brr_hip_get_num(brr_hip_ctx* c, double* out)
{
    if (int k = need_kind(c, false)) return k;
    if (!out) return fail(c, BRR_HIP_EINVAL, "null output");
    if (!c->d_num) return fail(c, BRR_HIP_ESTATE, "num records are off (brr_hip_record_num)");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMemcpy(out, c->d_num, sizeof(double) * c->cfg.num_snps, hipMemcpyDeviceToHost));
    return BRR_HIP_OK;
}

int brr_hip_set_fixed(brr_hip_ctx* c, const float* X, uint32_t num_fixed)
{
    if (int k = need_kind(c, false)) return k;
    if (c->T) return fail(c, BRR_HIP_EINVAL, "fixed effects need a single-chain context");
    if (num_fixed > (uint32_t)kMaxF) return fail(c, BRR_HIP_EINVAL, "at most 128 fixed-effect covariates");
    if (num_fixed && !X) return fail(c, BRR_HIP_EINVAL, "null covariates");
    HIP_TRY(c, hipSetDevice(c->device));
    const uint32_t N = c->cfg.num_inds;
    c->F = 0;
    if (num_fixed == 0) return BRR_HIP_OK;
    int rc;
    if ((rc = dalloc(c, &c->d_fixedX, (size_t)N * num_fixed)) || (rc = dalloc(c, &c->d_fixed, 1 + num_fixed)) ||
        (rc = dalloc(c, &c->d_forder, num_fixed)))
        return rc;
    HIP_TRY(c, hipMemcpy(c->d_fixedX, X, sizeof(float) * N * num_fixed, hipMemcpyHostToDevice));
    // BayesRBase::runGibbs: gamma = 0 and xI = 0..F-1 before the first iteration
    // (BayesRBase.cpp:170-173); sigmaF = s02F (its value after every iteration)
    std::vector<double> st(1 + num_fixed, 0.0);
    st[0] = 1.0;
    std::vector<uint32_t> ord(num_fixed);
    for (uint32_t f = 0; f < num_fixed; ++f) ord[f] = f;
    HIP_TRY(c, hipMemcpy(c->d_fixed, st.data(), sizeof(double) * st.size(), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->d_forder, ord.data(), sizeof(uint32_t) * num_fixed, hipMemcpyHostToDevice));
    c->F = num_fixed;
    return BRR_HIP_OK;
}

int brr_hip_set_fixed_state(brr_hip_ctx* c, const double* gamma, const uint32_t* order, double sigma_f)
{
    if (int k = need_kind(c, false)) return k;
    if (c->F == 0) return fail(c, BRR_HIP_ESTATE, "no fixed effects (brr_hip_set_fixed)");
    if (!(sigma_f > 0.0)) return fail(c, BRR_HIP_EINVAL, "sigmaF must be > 0");
    HIP_TRY(c, hipSetDevice(c->device));
    if (order) {
        std::vector<uint8_t> seen(c->F, 0);
        for (uint32_t f = 0; f < c->F; ++f) {
            if (order[f] >= c->F || seen[order[f]]) return fail(c, BRR_HIP_EINVAL, "order is not a permutation");
            seen[order[f]] = 1;
        }
        HIP_TRY(c, hipMemcpy(c->d_forder, order, sizeof(uint32_t) * c->F, hipMemcpyHostToDevice));
    }
    HIP_TRY(c, hipMemcpy(c->d_fixed, &sigma_f, sizeof(double), hipMemcpyHostToDevice));
    if (gamma) HIP_TRY(c, hipMemcpy(c->d_fixed + 1, gamma, sizeof(double) * c->F, hipMemcpyHostToDevice));
    return BRR_HIP_OK;
}

int brr_hip_get_fixed_state(brr_hip_ctx* c, double* gamma, uint32_t* order, double* sigma_f)
{
    if (int k = need_kind(c, false)) return k;
    if (c->F == 0) return fail(c, BRR_HIP_ESTATE, "no fixed effects (brr_hip_set_fixed)");
    HIP_TRY(c, hipSetDevice(c->device));
    if (sigma_f) HIP_TRY(c, hipMemcpy(sigma_f, c->d_fixed, sizeof(double), hipMemcpyDeviceToHost));
    if (gamma) HIP_TRY(c, hipMemcpy(gamma, c->d_fixed + 1, sizeof(double) * c->F, hipMemcpyDeviceToHost));
    if (order) HIP_TRY(c, hipMemcpy(order, c->d_forder, sizeof(uint32_t) * c->F, hipMemcpyDeviceToHost));
    return BRR_HIP_OK;
}

int brr_hip_set_groups(brr_hip_ctx* c, const int32_t* group)
{
    if (int k = need_kind(c, false)) return k;
    if (!c || !group) return fail(c, BRR_HIP_EINVAL, "null argument");
    const uint32_t M = c->cfg.num_snps;
    for (uint32_t j = 0; j < M; ++j)
        if (group[j] < 0 || group[j] >= (int32_t)c->cfg.num_groups)
            return fail(c, BRR_HIP_EINVAL, "group label out of range");
    HIP_TRY(c, hipSetDevice(c->device));
    int rc = dalloc(c, &c->d_group, M);
    if (rc) return rc;
    HIP_TRY(c, hipMemcpy(c->d_group, group, sizeof(int32_t) * M, hipMemcpyHostToDevice));
    return BRR_HIP_OK;
}

int brr_hip_set_model(brr_hip_ctx* c, const brr_hip_model* m)
{
    if (int k = need_kind(c, false)) return k;
    if (!c || !m || !m->cva) return fail(c, BRR_HIP_EINVAL, "null argument");
    const uint32_t G = c->cfg.num_groups, K = c->cfg.num_components;
    for (uint32_t t = 0; t < G * (K - 1); ++t)
        if (!(m->cva[t] > 0.0)) return fail(c, BRR_HIP_EINVAL, "variance components must be > 0");
    HIP_TRY(c, hipSetDevice(c->device));
    c->model = *m;
    c->cva.assign(m->cva, m->cva + G * (K - 1));
    c->model.cva = c->cva.data();
    HIP_TRY(c, hipMemcpy(c->d_cva, m->cva, sizeof(double) * G * (K - 1), hipMemcpyHostToDevice));
    c->have_model = true;
    return BRR_HIP_OK;
}

int brr_hip_set_state(brr_hip_ctx* c, const double* eps, const double* beta, const int32_t* comp)
{
    if (!c) return BRR_HIP_EINVAL;
    HIP_TRY(c, hipSetDevice(c->device));
    const size_t N = c->cfg.num_inds, M = c->cfg.num_snps, Tn = c->Tn;
    if (eps) HIP_TRY(c, hipMemcpy(c->d_eps, eps, sizeof(double) * N * Tn, hipMemcpyHostToDevice));
    if (beta) HIP_TRY(c, hipMemcpy(c->d_beta, beta, sizeof(double) * M * Tn, hipMemcpyHostToDevice));
    if (comp) HIP_TRY(c, hipMemcpy(c->d_comp, comp, sizeof(int32_t) * M * Tn, hipMemcpyHostToDevice));
    return BRR_HIP_OK;
}

int brr_hip_set_scalars(brr_hip_ctx* c, const brr_hip_scalars* s)
{
    return brr_hip_set_trait_scalars(c, 0, s);
}

int brr_hip_set_trait_scalars(brr_hip_ctx* c, uint32_t trait, const brr_hip_scalars* s)
{
    if (int k = need_kind(c, false)) return k;
    if (!c || !s || !s->sigmaG || !s->pi) return fail(c, BRR_HIP_EINVAL, "null argument");
    if (trait >= c->Tn) return fail(c, BRR_HIP_EINVAL, "trait index out of range");
    HIP_TRY(c, hipSetDevice(c->device));
    const uint32_t G = c->cfg.num_groups, K = c->cfg.num_components;
    const ScalarLayout SL{G, K};
    std::vector<double> h(SL.size(), 0.0);
    h[SL.mu()] = s->mu;
    h[SL.sigmaE()] = s->sigmaE;
    h[SL.epsSum()] = s->epsilon_sum;
    for (uint32_t g = 0; g < G; ++g) h[SL.sigmaG() + g] = s->sigmaG[g];
    for (uint32_t t = 0; t < G * K; ++t) h[SL.pi() + t] = s->pi[t];
    HIP_TRY(c, hipMemcpy(c->d_scal + (size_t)trait * SL.size(), h.data(), sizeof(double) * h.size(),
                         hipMemcpyHostToDevice));
    return BRR_HIP_OK;
}

int brr_hip_get_state(brr_hip_ctx* c, double* eps, double* beta, int32_t* comp, double* acum)
{
    if (!c) return BRR_HIP_EINVAL;
    HIP_TRY(c, hipSetDevice(c->device));
    const size_t N = c->cfg.num_inds, M = c->cfg.num_snps, Tn = c->Tn;
    if (acum && c->bw) return fail(c, BRR_HIP_EINVAL, "a BayesW context keeps no acum");
    if (eps) HIP_TRY(c, hipMemcpy(eps, c->d_eps, sizeof(double) * N * Tn, hipMemcpyDeviceToHost));
    if (beta) HIP_TRY(c, hipMemcpy(beta, c->d_beta, sizeof(double) * M * Tn, hipMemcpyDeviceToHost));
    if (comp) HIP_TRY(c, hipMemcpy(comp, c->d_comp, sizeof(int32_t) * M * Tn, hipMemcpyDeviceToHost));
    if (acum) HIP_TRY(c, hipMemcpy(acum, c->d_acum, sizeof(double) * M * Tn, hipMemcpyDeviceToHost));
    return BRR_HIP_OK;
}

int brr_hip_get_scalars(brr_hip_ctx* c, brr_hip_scalars* s)
{
    return brr_hip_get_trait_scalars(c, 0, s);
}

int brr_hip_get_trait_scalars(brr_hip_ctx* c, uint32_t trait, brr_hip_scalars* s)
{
    if (int k = need_kind(c, false)) return k;
    if (!c || !s) return BRR_HIP_EINVAL;
    if (trait >= c->Tn) return fail(c, BRR_HIP_EINVAL, "trait index out of range");
    HIP_TRY(c, hipSetDevice(c->device));
    const uint32_t G = c->cfg.num_groups, K = c->cfg.num_components;
    const ScalarLayout SL{G, K};
    std::vector<double> h(SL.size());
    HIP_TRY(c, hipMemcpy(h.data(), c->d_scal + (size_t)trait * SL.size(), sizeof(double) * h.size(),
                         hipMemcpyDeviceToHost));
    s->mu = h[SL.mu()];
    s->sigmaE = h[SL.sigmaE()];
    s->epsilon_sum = h[SL.epsSum()];
    if (s->sigmaG) for (uint32_t g = 0; g < G; ++g) s->sigmaG[g] = h[SL.sigmaG() + g];
    if (s->pi) for (uint32_t t = 0; t < G * K; ++t) s->pi[t] = h[SL.pi() + t];
    if (s->v) for (uint32_t t = 0; t < G * K; ++t) s->v[t] = h[SL.v() + t];
    if (s->betasqn) for (uint32_t g = 0; g < G; ++g) s->betasqn[g] = h[SL.betasqn() + g];
    return BRR_HIP_OK;
}

int brr_hip_sweep(brr_hip_ctx* c, const uint32_t* order, const double* u, const double* z)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if ((rc = need_kind(c, false))) return rc;
    if (!order || !u || !z) return fail(c, BRR_HIP_EINVAL, "null argument");
    const uint32_t M = c->cfg.num_snps;
    std::vector<uint8_t> seen(M, 0);
    for (uint32_t i = 0; i < M; ++i) {
        if (order[i] >= M || seen[order[i]]) return fail(c, BRR_HIP_EINVAL, "order is not a permutation of 0..M-1");
        seen[order[i]] = 1;
    }
    HIP_TRY(c, hipSetDevice(c->device));
    if (c->order_cap < M) {
        c->order_cap = 0;
        if ((rc = dalloc(c, &c->d_order, M))) return rc;
        c->order_cap = M;
    }
    HIP_TRY(c, hipMemcpy(c->d_order, order, sizeof(uint32_t) * M, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->d_u, u, sizeof(double) * M * c->Tn, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->d_z, z, sizeof(double) * M * c->Tn, hipMemcpyHostToDevice));
    SweepParams p = base_params(c);
    p.mode = kModeSweep;
    p.iterations = 1;
    p.order = c->d_order;
    p.u = c->d_u;
    p.z = c->d_z;
    rc = launch_any(c, p);
    c->stats.markers = (uint64_t)M * c->Tn;
    return rc;
}

int brr_hip_sweep_noalloc(brr_hip_ctx* c, const uint32_t* d_order, const double* d_u, const double* d_z)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if ((rc = need_kind(c, false))) return rc;
    if (c->T) return fail(c, BRR_HIP_EINVAL, "brr_hip_sweep_noalloc: single-chain contexts only");
    if (!d_order || !d_u || !d_z) return fail(c, BRR_HIP_EINVAL, "null argument");
    const uint32_t M = c->cfg.num_snps;
    if ((size_t)M > kScratchMaxM)
        return fail(c, BRR_HIP_EINVAL, "brr_hip_sweep_noalloc: at most 4194304 markers (its scratch bitmap)");
    if (c->x1_words > kScratchX1 || c->x2_words > kScratchX2)
        return fail(c, BRR_HIP_EINVAL, "brr_hip_sweep_noalloc: the exchange does not fit its scratch");
    if (c->f32 && !c->codes)
        return fail(c, BRR_HIP_ESTATE, "the FP32 residual mode needs the columns held as 2-bit codes "
                                       "(ragged_storage 0 or 2)");
    const void* fn = kernel_for(c->R, c->layout, c->codes, c->async != 0, c->f32);
    if (!fn) return fail(c, BRR_HIP_EINVAL, "unsupported slice size");
    HIP_TRY(c, hipSetDevice(c->device));
    uint64_t* base = nullptr;
    if ((rc = scratch_for(c, &base))) return rc;
    const size_t map_words = ((size_t)M + 63) / 64;
    scratch_reset_kernel<<<64, 256, 0, c->stream>>>(base, map_words);
    HIP_TRY(c, hipGetLastError());
    order_check_kernel<<<std::min<uint32_t>((M + 255) / 256, 1024u), 256, 0, c->stream>>>(d_order, M, base);
    HIP_TRY(c, hipGetLastError());
    SweepParams p = base_params(c);
    p.mode = kModeSweep;
    p.iterations = 1;
    p.order = d_order;
    p.u = d_u;
    p.z = d_z;
    p.xchg1 = base;
    p.xchg2 = base + kOffX2;
    p.ctl = (uint32_t*)(base + kOffCtl);
    p.counters = (unsigned long long*)(base + kOffCnt);
    p.gate = (const uint32_t*)(base + kOffFlag);
    p.refresh_blocks = std::max(1u, c->refresh_markers / c->B);
    void* args[] = {&p};
    HIP_TRY(c, hipEventRecord(c->ev0, c->stream));
    HIP_TRY(c, hipLaunchCooperativeKernel(fn, dim3(c->P), dim3(kThreads), args, 0, c->stream));
    HIP_TRY(c, hipEventRecord(c->ev1, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->h_status, base + kOffCtl, sizeof(uint64_t) * kStatusWords, hipMemcpyDeviceToHost,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIP_TRY(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    const uint32_t* ctl = (const uint32_t*)c->h_status;
    const unsigned long long* counters = (const unsigned long long*)(c->h_status + 4);
    const uint32_t flag = (uint32_t)c->h_status[20];
    if (flag) return fail(c, BRR_HIP_EINVAL, "order is not a permutation of 0..M-1");
    c->stats.kernel_ms = ms;
    c->stats.rounds = counters[0];
    c->stats.changes = counters[1];
    c->stats.decision_rounds = counters[8];
    c->stats.markers = M;
    for (int q = 0; q < 6; ++q) c->phase_ms[q] = (double)counters[2 + q] * 1e-5;
    shape_stats(c);
    if (ctl[kCtlAbort]) {
        // the exchange may hold tags of the aborted rounds: clear it and the round word
        scratch_zero_kernel<<<256, 256, 0, c->stream>>>(base, kOffCtl + 4);
        HIP_TRY(c, hipGetLastError());
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        return fail(c, BRR_HIP_ETIMEOUT, "device exchange wait exceeded its bound (site " +
                                             std::to_string(ctl[kCtlAbortWhere]) + ")");
    }
    return BRR_HIP_OK;
}

int brr_hip_fixed_effects(brr_hip_ctx* c, uint64_t seed, uint32_t iteration, uint32_t chain)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if ((rc = need_kind(c, false))) return rc;
    if (c->F == 0) return fail(c, BRR_HIP_ESTATE, "no fixed effects (brr_hip_set_fixed)");
    HIP_TRY(c, hipSetDevice(c->device));
    SweepParams p = base_params(c);
    p.mode = kModeFixed;
    p.iterations = 1;
    p.iter_base = iteration;
    p.chain = chain;
    p.seed = seed;
    rc = launch_any(c, p);
    c->stats.markers = 0;
    return rc;
}

int brr_hip_run_chain(brr_hip_ctx* c, uint64_t seed, uint32_t chain, uint32_t first_iteration,
                      uint32_t iterations, const uint32_t* orders, double* hist)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if ((rc = need_kind(c, false))) return rc;
    if (!orders || iterations == 0) return fail(c, BRR_HIP_EINVAL, "null orders or zero iterations");
    const uint32_t M = c->cfg.num_snps, G = c->cfg.num_groups;
    const size_t total = (size_t)iterations * M;
    std::vector<uint8_t> seen(M);
    for (uint32_t it = 0; it < iterations; ++it) {
        std::fill(seen.begin(), seen.end(), 0);
        const uint32_t* o = orders + (size_t)it * M;
        for (uint32_t i = 0; i < M; ++i) {
            if (o[i] >= M || seen[o[i]]) return fail(c, BRR_HIP_EINVAL, "order is not a permutation of 0..M-1");
            seen[o[i]] = 1;
        }
    }
    HIP_TRY(c, hipSetDevice(c->device));
    if (c->order_cap < total) {
        c->order_cap = 0;
        if ((rc = dalloc(c, &c->d_order, total))) return rc;
        c->order_cap = total;
    }
    const size_t hsz = (size_t)iterations * c->Tn * (4 + G);
    if (c->hist_cap < hsz) {
        c->hist_cap = 0;
        if ((rc = dalloc(c, &c->d_hist, hsz))) return rc;
        c->hist_cap = hsz;
    }
    HIP_TRY(c, hipMemcpy(c->d_order, orders, sizeof(uint32_t) * total, hipMemcpyHostToDevice));
    SweepParams p = base_params(c);
    p.mode = kModeChain;
    p.iterations = iterations;
    p.iter_base = first_iteration;
    p.chain = chain;
    p.seed = seed;
    p.order = c->d_order;
    p.hist = c->d_hist;
    rc = launch_any(c, p);
    c->stats.markers = total * c->Tn;
    if (rc) return rc;
    c->change_rate = (double)c->stats.changes / (double)(total * c->Tn);
    if (hist) HIP_TRY(c, hipMemcpy(hist, c->d_hist, sizeof(double) * hsz, hipMemcpyDeviceToHost));
    return BRR_HIP_OK;
}

int brr_hip_compute_num(brr_hip_ctx* c, uint32_t marker, double beta_old, double* num)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if ((rc = need_kind(c, false))) return rc;
    if (c->T) return fail(c, BRR_HIP_ESTATE, "per-marker calls need a single-chain context");
    if (!num || marker >= c->cfg.num_snps) return fail(c, BRR_HIP_EINVAL, "bad marker or null output");
    HIP_TRY(c, hipSetDevice(c->device));
    SweepParams p = base_params(c);
    compute_num_kernel<<<1, kThreads, 0, c->stream>>>(p, marker, beta_old, c->d_scal + 2, c->d_tmp);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(num, c->d_tmp, sizeof(double), hipMemcpyDeviceToHost));
    return BRR_HIP_OK;
}

int brr_hip_epsilon_change(brr_hip_ctx* c, uint32_t marker, double beta_old, double beta)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if ((rc = need_kind(c, false))) return rc;
    if (c->T) return fail(c, BRR_HIP_ESTATE, "per-marker calls need a single-chain context");
    if (marker >= c->cfg.num_snps) return fail(c, BRR_HIP_EINVAL, "bad marker");
    HIP_TRY(c, hipSetDevice(c->device));
    SweepParams p = base_params(c);
    const uint32_t blocks = (c->cfg.num_inds + kThreads - 1) / kThreads;
    epsilon_change_kernel<<<blocks, kThreads, 0, c->stream>>>(p, marker, beta_old, beta);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return BRR_HIP_OK;
}

int brr_hip_kernel_compute_num(brr_hip_ctx* c, uint32_t marker, const double* epsilon, double epsilon_sum,
                               double beta_old, double* num)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if ((rc = need_kind(c, false))) return rc;
    if (c->T) return fail(c, BRR_HIP_ESTATE, "per-marker calls need a single-chain context");
    if (!epsilon || !num || marker >= c->cfg.num_snps) return fail(c, BRR_HIP_EINVAL, "bad marker or null argument");
    const uint32_t N = c->cfg.num_inds;
    HIP_TRY(c, hipSetDevice(c->device));
    if (!c->d_keps && (rc = dalloc(c, &c->d_keps, N))) return rc;
    HIP_TRY(c, hipMemcpy(c->d_keps, epsilon, sizeof(double) * N, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->d_tmp + 1, &epsilon_sum, sizeof(double), hipMemcpyHostToDevice));
    SweepParams p = base_params(c);
    p.eps = c->d_keps;
    compute_num_kernel<<<1, kThreads, 0, c->stream>>>(p, marker, beta_old, c->d_tmp + 1, c->d_tmp);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(num, c->d_tmp, sizeof(double), hipMemcpyDeviceToHost));
    return BRR_HIP_OK;
}

int brr_hip_kernel_epsilon_change(brr_hip_ctx* c, uint32_t marker, double beta_old, double beta, double* delta,
                                  double* epsilon_sum_delta)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if ((rc = need_kind(c, false))) return rc;
    if (c->T) return fail(c, BRR_HIP_ESTATE, "per-marker calls need a single-chain context");
    if (!delta || marker >= c->cfg.num_snps) return fail(c, BRR_HIP_EINVAL, "bad marker or null argument");
    const uint32_t N = c->cfg.num_inds;
    HIP_TRY(c, hipSetDevice(c->device));
    if (!c->d_keps && (rc = dalloc(c, &c->d_keps, N))) return rc;
    HIP_TRY(c, hipMemsetAsync(c->d_keps, 0, sizeof(double) * N, c->stream));
    SweepParams p = base_params(c);
    p.eps = c->d_keps;
    const uint32_t blocks = (N + kThreads - 1) / kThreads;
    epsilon_change_kernel<<<blocks, kThreads, 0, c->stream>>>(p, marker, beta_old, beta);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(delta, c->d_keps, sizeof(double) * N, hipMemcpyDeviceToHost));
    if (epsilon_sum_delta) {
        // SparseBayesRKernel::computeEpsilonSumUpdate (sparsebayesrkernel.cpp:29-34); 0 for dense markers
        double d = 0.0;
        if (c->layout != kLayoutDense) {
            double st[4];
            HIP_TRY(c, hipMemcpy(st, c->d_stats + 4 * (size_t)marker, sizeof(st), hipMemcpyDeviceToHost));
            const double dBeta = beta_old - beta;
            d = dBeta * st[3] / st[1] - dBeta * st[0] * (double)N / st[1];
        }
        *epsilon_sum_delta = d;
    }
    return BRR_HIP_OK;
}

int brr_hip_last_phase_ms(const brr_hip_ctx* c, double* out, uint32_t n)
{
    if (!c || !out) return BRR_HIP_EINVAL;
    for (uint32_t q = 0; q < n && q < (uint32_t)kBPhases; ++q) out[q] = c->phase_ms[q];
    return BRR_HIP_OK;
}

int brr_hip_last_stats(const brr_hip_ctx* c, brr_hip_stats* s)
{
    if (!c || !s) return BRR_HIP_EINVAL;
    *s = c->stats;
    return BRR_HIP_OK;
}

// ---------------------------------------------------------------- BayesW --

int brr_hip_bw_create(const brr_hip_bw_config* cfg, brr_hip_ctx** out)
{
    if (!cfg || !out) return BRR_HIP_EINVAL;
    *out = nullptr;
    if (cfg->num_inds < 2 || cfg->num_snps < 1) return BRR_HIP_EINVAL;
    if (cfg->num_components < 2 || cfg->num_components > (uint32_t)kMaxK) return BRR_HIP_EINVAL;
    if (cfg->quad_points < 3 || cfg->quad_points > 15 || (cfg->quad_points % 2) == 0) return BRR_HIP_EINVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return BRR_HIP_EDEVICE;
    if ((int)cfg->device >= ndev) return BRR_HIP_EINVAL;

    brr_hip_ctx* c = new brr_hip_ctx();
    c->cfg.num_inds = cfg->num_inds;
    c->cfg.num_snps = cfg->num_snps;
    c->cfg.num_components = cfg->num_components;
    c->cfg.num_groups = 1;
    c->cfg.layout = BRR_HIP_LAYOUT_RAGGED;
    c->cfg.workgroups = cfg->workgroups;
    c->cfg.device = cfg->device;
    c->cfg.ragged_storage = 2;
    c->device = (int)cfg->device;
    c->layout = kLayoutRagged;
    c->bw = true;
    c->codes = true;
    c->nq = cfg->quad_points;
    if (hipSetDevice(c->device) != hipSuccess) { delete c; return BRR_HIP_EDEVICE; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->device) != hipSuccess) { delete c; return BRR_HIP_EDEVICE; }
    int coop = 0;
    (void)hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, c->device);
    if (!coop) { delete c; return BRR_HIP_EDEVICE; }

    const uint32_t N = cfg->num_inds, M = cfg->num_snps, K = cfg->num_components;
    const uint32_t cus = (uint32_t)prop.multiProcessorCount;
    const uint32_t S = lane_run_slice(N, cus, cfg->workgroups);
    uint32_t R = (S + kRun - 1) / kRun;
    if (R > (uint32_t)kMaxRuns) { delete c; return BRR_HIP_EINVAL; }
    if (R == 3) R = 4;
    c->R = R;
    c->S = S;
    c->P = (N + S - 1) / S;
    c->B = kBwB;
    shape_stats(c);
    c->x1_words = (size_t)2 * kBwV * c->P * 2;
    c->x2_words = kCopies * kBwCopyWords;
    for (int dense = 0; dense < 2; ++dense) {
        int occ = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, bw_kernel_for(R, dense != 0), kThreads, 0) !=
                hipSuccess ||
            (uint64_t)occ * cus < c->P) {
            delete c;
            return BRR_HIP_EDEVICE;
        }
    }
    const size_t nscal = 5 + 2 * (size_t)K;
    int rc = BRR_HIP_OK;
    if ((rc = dalloc(c, &c->d_eps, N)) || (rc = dalloc(c, &c->d_beta, M)) || (rc = dalloc(c, &c->d_comp, M)) ||
        (rc = dalloc(c, &c->d_acum, 1)) || (rc = dalloc(c, &c->d_scal, nscal)) ||
        (rc = dalloc(c, &c->d_cva, K - 1)) || (rc = dalloc(c, &c->d_xchg1, c->x1_words)) ||
        (rc = dalloc(c, &c->d_xchg2, c->x2_words)) || (rc = dalloc(c, &c->d_ctl, 8)) ||
        (rc = dalloc(c, &c->d_counters, kCounters)) || (rc = dalloc(c, &c->d_tmp, 4)) ||
        (rc = dalloc(c, &c->d_u, M)) || (rc = dalloc(c, &c->d_z, 1)) || (rc = dalloc(c, &c->d_fail, N)) ||
        (rc = dalloc(c, &c->d_sumf, M))) {
        brr_hip_destroy(c);
        return rc;
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        brr_hip_destroy(c);
        return BRR_HIP_EDEVICE;
    }
    (void)hipMemset(c->d_eps, 0, sizeof(double) * N);
    (void)hipMemset(c->d_beta, 0, sizeof(double) * M);
    (void)hipMemset(c->d_comp, 0, sizeof(int32_t) * M);
    (void)hipMemset(c->d_scal, 0, sizeof(double) * nscal);
    if (reset_exchange(c) != BRR_HIP_OK) { brr_hip_destroy(c); return BRR_HIP_EDEVICE; }
    *out = c;
    return BRR_HIP_OK;
}

int brr_hip_bw_set_model(brr_hip_ctx* c, const double* mixture_C, const double* failure)
{
    int rc;
    if ((rc = need_kind(c, true))) return rc;
    if (!mixture_C || !failure) return fail(c, BRR_HIP_EINVAL, "null argument");
    if (!c->have_markers) return fail(c, BRR_HIP_ESTATE, "markers not uploaded");
    const uint32_t N = c->cfg.num_inds, M = c->cfg.num_snps, K = c->cfg.num_components;
    for (uint32_t k = 0; k + 1 < K; ++k)
        if (!(mixture_C[k] > 0.0)) return fail(c, BRR_HIP_EINVAL, "mixture variances must be > 0");
    for (uint32_t i = 0; i < N; ++i)
        if (failure[i] != 0.0 && failure[i] != 1.0)
            return fail(c, BRR_HIP_EINVAL, "failure indicators must be 0 or 1");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMemcpy(c->d_cva, mixture_C, sizeof(double) * (K - 1), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->d_fail, failure, sizeof(double) * N, hipMemcpyHostToDevice));
    if (c->layout == kLayoutDense)
        bw_dense_sum_failure_kernel<<<M, kThreads, 0, c->stream>>>(c->d_dense, c->col_words, N, c->d_fail, c->d_lut,
                                                                   c->d_sumf);
    else
        bw_sum_failure_kernel<<<M, kThreads, 0, c->stream>>>(c->d_dense, c->col_words, N, c->d_fail, c->d_stats,
                                                             c->d_sumf);
    HIP_TRY(c, hipGetLastError());
    if (c->F) {
        bw_fixed_sumf_kernel<<<c->F, kThreads, 0, c->stream>>>(c->d_fixedX, N, c->d_fail, c->d_sumf_fix);
        HIP_TRY(c, hipGetLastError());
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->have_model = true;
    return BRR_HIP_OK;
}

int brr_hip_bw_set_fixed(brr_hip_ctx* c, const float* X, uint32_t num_fixed)
{
    int rc;
    if ((rc = need_kind(c, true))) return rc;
    if (num_fixed > (uint32_t)kMaxF) return fail(c, BRR_HIP_EINVAL, "at most 128 fixed-effect covariates");
    if (num_fixed && !X) return fail(c, BRR_HIP_EINVAL, "null covariates");
    HIP_TRY(c, hipSetDevice(c->device));
    const uint32_t N = c->cfg.num_inds;
    c->F = 0;
    if (num_fixed == 0) return BRR_HIP_OK;
    if ((rc = dalloc(c, &c->d_fixedX, (size_t)N * num_fixed)) || (rc = dalloc(c, &c->d_fixed, num_fixed)) ||
        (rc = dalloc(c, &c->d_sumf_fix, num_fixed)))
        return rc;
    HIP_TRY(c, hipMemcpy(c->d_fixedX, X, sizeof(float) * N * num_fixed, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemset(c->d_fixed, 0, sizeof(double) * num_fixed));   // m_theta.setZero()
    HIP_TRY(c, hipMemset(c->d_sumf_fix, 0, sizeof(double) * num_fixed));
    c->F = num_fixed;
    if (c->have_model) {
        bw_fixed_sumf_kernel<<<c->F, kThreads, 0, c->stream>>>(c->d_fixedX, N, c->d_fail, c->d_sumf_fix);
        HIP_TRY(c, hipGetLastError());
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    return BRR_HIP_OK;
}

int brr_hip_bw_set_theta(brr_hip_ctx* c, const double* theta)
{
    int rc;
    if ((rc = need_kind(c, true))) return rc;
    if (c->F == 0) return fail(c, BRR_HIP_ESTATE, "no fixed effects (brr_hip_bw_set_fixed)");
    if (!theta) return fail(c, BRR_HIP_EINVAL, "null argument");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMemcpy(c->d_fixed, theta, sizeof(double) * c->F, hipMemcpyHostToDevice));
    return BRR_HIP_OK;
}

int brr_hip_bw_get_theta(brr_hip_ctx* c, double* theta)
{
    int rc;
    if ((rc = need_kind(c, true))) return rc;
    if (c->F == 0) return fail(c, BRR_HIP_ESTATE, "no fixed effects (brr_hip_bw_set_fixed)");
    if (!theta) return fail(c, BRR_HIP_EINVAL, "null argument");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMemcpy(theta, c->d_fixed, sizeof(double) * c->F, hipMemcpyDeviceToHost));
    return BRR_HIP_OK;
}

int brr_hip_bw_set_scalars(brr_hip_ctx* c, const brr_hip_bw_scalars* s)
{
    int rc;
    if ((rc = need_kind(c, true))) return rc;
    if (!s || !s->pi) return fail(c, BRR_HIP_EINVAL, "null argument");
    const uint32_t K = c->cfg.num_components;
    std::vector<double> h(5 + 2 * (size_t)K, 0.0);
    h[kBwMu] = s->mu;
    h[kBwAlpha] = s->alpha;
    h[kBwSigmaB] = s->sigma_b;
    h[kBwD] = s->d;
    for (uint32_t k = 0; k < K; ++k) {
        h[kBwPi + k] = s->pi[k];
        h[kBwPi + K + k] = s->v ? s->v[k] : 1.0;
    }
    if (!(s->alpha > 0.0) || !(s->sigma_b > 0.0)) return fail(c, BRR_HIP_EINVAL, "alpha and sigma_b must be > 0");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMemcpy(c->d_scal, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
    return BRR_HIP_OK;
}

int brr_hip_bw_get_scalars(brr_hip_ctx* c, brr_hip_bw_scalars* s)
{
    int rc;
    if ((rc = need_kind(c, true))) return rc;
    if (!s) return fail(c, BRR_HIP_EINVAL, "null argument");
    const uint32_t K = c->cfg.num_components;
    std::vector<double> h(5 + 2 * (size_t)K);
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMemcpy(h.data(), c->d_scal, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
    s->mu = h[kBwMu];
    s->alpha = h[kBwAlpha];
    s->sigma_b = h[kBwSigmaB];
    s->d = h[kBwD];
    s->betasq = h[kBwBetaSq];
    if (s->pi) for (uint32_t k = 0; k < K; ++k) s->pi[k] = h[kBwPi + k];
    if (s->v) for (uint32_t k = 0; k < K; ++k) s->v[k] = h[kBwPi + K + k];
    return BRR_HIP_OK;
}

int brr_hip_bw_sweep(brr_hip_ctx* c, const uint32_t* order, const double* u, uint64_t seed, uint32_t iteration,
                     uint32_t chain)
{
    int rc;
    if ((rc = need_kind(c, true)) || (rc = check_ready(c))) return rc;
    if (!order || !u) return fail(c, BRR_HIP_EINVAL, "null argument");
    if ((rc = check_order(c, order, 1))) return rc;
    const uint32_t M = c->cfg.num_snps;
    HIP_TRY(c, hipSetDevice(c->device));
    if (c->order_cap < M) {
        c->order_cap = 0;
        if ((rc = dalloc(c, &c->d_order, M))) return rc;
        c->order_cap = M;
    }
    HIP_TRY(c, hipMemcpy(c->d_order, order, sizeof(uint32_t) * M, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->d_u, u, sizeof(double) * M, hipMemcpyHostToDevice));
    BwParams p = bw_params(c);
    p.mode = kModeSweep;
    p.iterations = 1;
    p.iter_base = iteration;
    p.chain = chain;
    p.seed = seed;
    p.order = c->d_order;
    rc = launch_fn(c, bw_kernel_for(c->R, c->layout == kLayoutDense), &p);
    c->stats.markers = M;
    return rc;
}

int brr_hip_bw_fixed_effects(brr_hip_ctx* c, uint64_t seed, uint32_t iteration, uint32_t chain)
{
    int rc;
    if ((rc = need_kind(c, true)) || (rc = check_ready(c))) return rc;
    if (c->F == 0) return fail(c, BRR_HIP_ESTATE, "no fixed effects (brr_hip_bw_set_fixed)");
    HIP_TRY(c, hipSetDevice(c->device));
    BwParams p = bw_params(c);
    p.mode = kModeFixed;
    p.iterations = 1;
    p.iter_base = iteration;
    p.chain = chain;
    p.seed = seed;
    rc = launch_fn(c, bw_kernel_for(c->R, c->layout == kLayoutDense), &p);
    c->stats.markers = 0;
    return rc;
}

int brr_hip_bw_run_chain(brr_hip_ctx* c, uint64_t seed, uint32_t chain, uint32_t first_iteration,
                         uint32_t iterations, const uint32_t* orders, double* hist)
{
    int rc;
    if ((rc = need_kind(c, true)) || (rc = check_ready(c))) return rc;
    if (!orders || iterations == 0) return fail(c, BRR_HIP_EINVAL, "null orders or zero iterations");
    if ((rc = check_order(c, orders, iterations))) return rc;
    const uint32_t M = c->cfg.num_snps;
    const size_t total = (size_t)iterations * M;
    HIP_TRY(c, hipSetDevice(c->device));
    if (c->order_cap < total) {
        c->order_cap = 0;
        if ((rc = dalloc(c, &c->d_order, total))) return rc;
        c->order_cap = total;
    }
    const size_t hsz = (size_t)iterations * 5;
    if (c->hist_cap < hsz) {
        c->hist_cap = 0;
        if ((rc = dalloc(c, &c->d_hist, hsz))) return rc;
        c->hist_cap = hsz;
    }
    HIP_TRY(c, hipMemcpy(c->d_order, orders, sizeof(uint32_t) * total, hipMemcpyHostToDevice));
    BwParams p = bw_params(c);
    p.mode = kModeChain;
    p.iterations = iterations;
    p.iter_base = first_iteration;
    p.chain = chain;
    p.seed = seed;
    p.order = c->d_order;
    p.hist = c->d_hist;
    rc = launch_fn(c, bw_kernel_for(c->R, c->layout == kLayoutDense), &p);
    c->stats.markers = total;
    if (rc) return rc;
    if (hist) HIP_TRY(c, hipMemcpy(hist, c->d_hist, sizeof(double) * hsz, hipMemcpyDeviceToHost));
    return BRR_HIP_OK;
}

int brr_hip_bw_set_vi(brr_hip_ctx* c, uint32_t marker, double alpha, double* out4)
{
    int rc;
    if ((rc = need_kind(c, true))) return rc;
    if (!c->have_markers) return fail(c, BRR_HIP_ESTATE, "markers not uploaded");
    if (!out4 || marker >= c->cfg.num_snps) return fail(c, BRR_HIP_EINVAL, "bad marker or null output");
    HIP_TRY(c, hipSetDevice(c->device));
    bw_set_vi_kernel<<<1, kThreads, 0, c->stream>>>(c->d_dense, c->col_words, c->cfg.num_inds, c->d_eps, alpha,
                                                    marker, c->d_tmp);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(out4, c->d_tmp, sizeof(double) * 4, hipMemcpyDeviceToHost));
    return BRR_HIP_OK;
}

int brr_hip_bw_residual_update(brr_hip_ctx* c, uint32_t marker, double beta, int sign)
{
    int rc;
    if ((rc = need_kind(c, true))) return rc;
    if (!c->have_markers) return fail(c, BRR_HIP_ESTATE, "markers not uploaded");
    if (marker >= c->cfg.num_snps || (sign != 1 && sign != -1)) return fail(c, BRR_HIP_EINVAL, "bad marker or sign");
    HIP_TRY(c, hipSetDevice(c->device));
    const uint32_t blocks = (c->cfg.num_inds + kThreads - 1) / kThreads;
    bw_residual_kernel<<<blocks, kThreads, 0, c->stream>>>(c->d_dense, c->col_words, c->cfg.num_inds, c->d_stats,
                                                           c->layout == kLayoutDense ? c->d_lut : nullptr, marker,
                                                           beta, sign, c->d_eps);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return BRR_HIP_OK;
}

int brr_hip_bw_sum_failure(brr_hip_ctx* c, double* out)
{
    int rc;
    if ((rc = need_kind(c, true))) return rc;
    if (!c->have_model) return fail(c, BRR_HIP_ESTATE, "model not set (brr_hip_bw_set_model)");
    if (!out) return fail(c, BRR_HIP_EINVAL, "null argument");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMemcpy(out, c->d_sumf, sizeof(double) * c->cfg.num_snps, hipMemcpyDeviceToHost));
    return BRR_HIP_OK;
}

} // extern "C"
