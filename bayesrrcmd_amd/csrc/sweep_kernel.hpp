// sweep_kernel.hpp -- persistent gfx950 kernel for the BayesR per-marker Gibbs sweep.
//
// What it computes: BayesRBase::processColumn (BayesRBase.cpp:287-392) over a
// marker order, i.e. for each marker j in sequence
//   num_j   = x_j . eps + beta_j (N-1)              (densebayesrkernel.cpp:10-15,
//             sparsebayesrkernel.cpp:24-27 + raggedbayesrkernel.cpp:32-35)
//   mixture log-likelihoods, acum, component and beta draw  (BayesRBase.cpp:300-366)
//   eps += (beta_old - beta_new) x_j unless both are 0     (BayesRBase.cpp:370-378)
// optionally wrapped in whole runGibbs iterations (updateMu, sigma/pi draws).
//
// How (MI355X-first): one cooperative launch of P workgroups, one per CU. Each
// workgroup owns a fixed slice of individuals and keeps that slice of eps in
// registers for the whole launch, so eps never travels to HBM; only genotype
// codes stream in. Markers are decided in speculation blocks of B: every
// workgroup computes its partial dots for all undecided markers of the block,
// the partials are combined by a deterministic reduce-scatter/all-gather
// through agent-coherent tagged words, and every workgroup then evaluates the
// same FP64 decisions redundantly (identical inputs and instruction order give
// identical bits). Decisions before the first marker whose beta changes are
// exactly those the one-marker-at-a-time sweep makes, so they are accepted;
// the changed marker's update is applied to each slice and the remaining
// markers are re-dotted in the next round. One exchange round per block plus
// one per beta change replaces one grid barrier per marker.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "brr_rng.h"

namespace brr {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kMaxB = 64;      // markers per speculation block (one wave of decisions)
constexpr int kMaxK = 8;       // mixture components incl. zero
constexpr int kMaxG = 32;      // marker groups
constexpr int kMaxRuns = 4;    // slice <= 512 * kMaxRuns individuals
constexpr uint32_t kRun = 512; // individuals per lane-run pass: 64 lanes x 8
constexpr uint32_t kSliceAlign = 16;
constexpr uint64_t kSpinLimit = 1ull << 22;

enum : int { kModeSweep = 0, kModeChain = 1 };
enum : int { kLayoutDense = 1, kLayoutRagged = 2 };

// ctl words (persist across launches)
enum : int { kCtlRound = 0, kCtlAbort = 1, kCtlAbortWhere = 2 };

struct SweepParams {
    uint32_t N, M, K, G;
    uint32_t P, B, S;
    int layout, mode;
    uint32_t iterations, iter_base, chain, pad0;
    uint64_t seed;
    // dense arena: column j at dense + j * col_words (32-bit words)
    const uint32_t* dense;
    uint64_t col_words, col_bytes;
    const double* lut;            // M x 4, by dosage 0,1,2,missing
    // ragged arena
    const int32_t* ones;
    const int32_t* twos;
    const int32_t* miss;
    const uint64_t* ones_off;
    const uint64_t* twos_off;
    const uint64_t* miss_off;
    const uint32_t* seg;          // M x 3 x (P+1) slice starts, relative to the column's list
    const double* stats;          // M x 4 {mean, sd, sqrdZ, Zsum}
    const int32_t* group;         // M or null
    const double* cva;            // G x (K-1)
    double v0E, s02E, v0G, s02G;
    const uint32_t* order;        // M (sweep) or iterations x M (chain)
    const double* u;              // sweep mode, by marker id
    const double* z;
    double* eps;                  // N
    double* beta;                 // M (agent-coherent words)
    int32_t* comp;                // M
    double* acum;                 // M
    double* scal;                 // scalar block, see ScalarLayout
    double* hist;                 // iterations x (4 + G) or null
    uint64_t* xchg1;              // [2][kMaxB][P][2] tagged partials
    uint64_t* xchg2;              // [2][kMaxB][2] tagged totals
    uint32_t* ctl;                // persistent control words
    unsigned long long* counters; // [0] rounds, [1] changes, [2] markers
};

// Scalar block (doubles): mu, sigmaE, epsSum, sigmaG[G], pi[G*K], v[G*K], betasqn[G]
struct ScalarLayout {
    uint32_t G, K;
    __host__ __device__ uint32_t mu() const { return 0; }
    __host__ __device__ uint32_t sigmaE() const { return 1; }
    __host__ __device__ uint32_t epsSum() const { return 2; }
    __host__ __device__ uint32_t sigmaG() const { return 3; }
    __host__ __device__ uint32_t pi() const { return 3 + G; }
    __host__ __device__ uint32_t v() const { return 3 + G + G * K; }
    __host__ __device__ uint32_t betasqn() const { return 3 + G + 2 * G * K; }
    __host__ __device__ uint32_t size() const { return 3 + 2 * G + 2 * G * K; }
};

template <int R>
struct Shared {
    uint32_t codes[kMaxB][R][32];       // 2-bit PLINK codes, lane-run layout
    double dot[kMaxB];
    double beta_old[kMaxB];
    double wt[kMaxB][4];                // column value per 2-bit code
    double mstat[kMaxB][3];             // ragged: mean, sd, Zsum
    double uu[kMaxB];
    double zz[kMaxB];
    double beta_new[kMaxB];
    double acumv[kMaxB];
    uint32_t mk[kMaxB];
    int32_t gr[kMaxB];
    int32_t compv[kMaxB];
    double delta[4];
    double red[kWaves];
    double mu, sigmaE, epsSum;
    double sigmaG[kMaxG];
    double pi[kMaxG * kMaxK];
    double v[kMaxG * kMaxK];
    double betasqn[kMaxG];
    double cva[kMaxG * (kMaxK - 1)];
    double gc[kMaxG * kMaxK];           // log pi_k - 0.5 log(logLScale S_k + 1)
    double gden[kMaxG * kMaxK];         // denom_k (index k >= 1)
    double gbsd[kMaxG * kMaxK];         // sqrt(sigmaE / denom_k)
    double dnext;                       // scalar exchange result
    int change;
    int abort;
};

__device__ __forceinline__ uint64_t ld_agent(const uint64_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void st_agent(uint64_t* p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ double wave_sum(double x)
{
    // xor butterfly: every lane ends with bit-identical sums (fp add commutes)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

__device__ __forceinline__ double wt_of(uint32_t v, double w0, double w1, double w2, double w3)
{
    return (v & 2u) ? ((v & 1u) ? w3 : w2) : ((v & 1u) ? w1 : w0);
}

__device__ __forceinline__ void publish_pair(uint64_t* s, uint32_t rnd, double val)
{
    const uint64_t b = double_as_bits(val);
    const uint64_t tag = (uint64_t)rnd << 32;
    st_agent(s, tag | (b >> 32));
    st_agent(s + 1, tag | (b & 0xFFFFFFFFull));
}

// Returns false on timeout or abort.
__device__ __forceinline__ bool wait_pair(const uint64_t* s, uint32_t rnd, const uint32_t* ctl,
                                          double* out)
{
    uint64_t h = ld_agent(s);
    uint64_t l = ld_agent(s + 1);
    uint64_t spins = 0;
    while ((uint32_t)(h >> 32) != rnd || (uint32_t)(l >> 32) != rnd) {
        if (++spins > kSpinLimit) return false;
        if ((spins & 255u) == 0 &&
            __hip_atomic_load(ctl + kCtlAbort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            return false;
        __builtin_amdgcn_s_sleep(1);
        h = ld_agent(s);
        l = ld_agent(s + 1);
    }
    *out = bits_as_double((h << 32) | (l & 0xFFFFFFFFull));
    return true;
}

__device__ __forceinline__ void raise_abort(const SweepParams& p, uint32_t where)
{
    __hip_atomic_store(p.ctl + kCtlAbort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(p.ctl + kCtlAbortWhere, where, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t* slot1(const SweepParams& p, uint32_t rnd, int k, uint32_t w)
{
    return p.xchg1 + ((((size_t)(rnd & 1u) * kMaxB + (size_t)k) * p.P + w) * 2u);
}

__device__ __forceinline__ uint64_t* slot2(const SweepParams& p, uint32_t rnd, int k)
{
    return p.xchg2 + (((size_t)(rnd & 1u) * kMaxB + (size_t)k) * 2u);
}

// Reduce-scatter / all-gather of the partials of slots [pos, nb) published at
// round rnd. Slot k is reduced by workgroup k % P. Afterwards L.dot[k] holds
// the same total in every workgroup. Returns false (in all threads) on abort.
template <int R>
__device__ bool exchange(const SweepParams& p, Shared<R>& L, uint32_t rnd, int pos, int nb,
                         uint32_t w, int wave, int lane)
{
    int cnt = 0;
    bool ok = true;
    for (int k = pos; k < nb; ++k) {
        if ((uint32_t)k % p.P != w) continue;
        if ((cnt++ % kWaves) != wave) continue;
        double acc = 0.0;
        for (uint32_t w2 = (uint32_t)lane; w2 < p.P; w2 += 64u) {
            double x;
            if (!wait_pair(slot1(p, rnd, k, w2), rnd, p.ctl, &x)) { ok = false; break; }
            acc += x;
        }
        if (__any(!ok)) { ok = false; break; }
        acc = wave_sum(acc);
        if (lane == 0) publish_pair(slot2(p, rnd, k), rnd, acc);
    }
    if (!ok) {
        if (lane == 0) { L.abort = 1; raise_abort(p, 1); }
    }
    if (wave == 0 && pos + lane < nb) {
        double x;
        if (!wait_pair(slot2(p, rnd, pos + lane), rnd, p.ctl, &x)) {
            L.abort = 1;
            raise_abort(p, 2);
        } else {
            L.dot[pos + lane] = x;
        }
    }
    __syncthreads();
    return L.abort == 0;
}

// A scalar all-reduce through slot 0. Every wave holds the whole slice, so the
// slice partial is wave 0's sum alone.
template <int R>
__device__ bool allreduce_scalar(const SweepParams& p, Shared<R>& L, uint32_t& rnd, double part,
                                 uint32_t w, int wave, int lane, double* total)
{
    part = wave_sum(part);
    ++rnd;
    if (threadIdx.x == 0) publish_pair(slot1(p, rnd, 0, w), rnd, part);
    if (!exchange<R>(p, L, rnd, 0, 1, w, wave, lane)) return false;
    *total = L.dot[0];
    return true;
}

template <int R>
__device__ void group_constants(const SweepParams& p, Shared<R>& L)
{
    // Per-iteration constants of processColumn (BayesRBase.cpp:309-333); they
    // depend only on sigmaE, sigmaG[g], pi[g], S[g].
    const uint32_t K = p.K;
    const double NM1 = (double)(p.N - 1u);
    for (uint32_t t = threadIdx.x; t < p.G * K; t += kThreads) {
        const uint32_t g = t / K, k = t % K;
        const double sigmaG = L.sigmaG[g];
        const double sigmaE = L.sigmaE;
        if (k == 0) {
            L.gc[t] = log(L.pi[t]);
            L.gden[t] = 0.0;
            L.gbsd[t] = 0.0;
        } else {
            const double S = L.cva[g * (K - 1) + (k - 1)];
            const double sigmaEOverSigmaG = sigmaE / sigmaG;
            const double den = NM1 + sigmaEOverSigmaG * (1.0 / S);
            const double logLScale = sigmaG / sigmaE * NM1;
            L.gden[t] = den;
            L.gc[t] = log(L.pi[t]) - 0.5 * log(logLScale * S + 1.0);
            L.gbsd[t] = __dsqrt_rn(sigmaE / den);
        }
    }
}

__device__ __forceinline__ bool guard700(const double* logL, int K, int r)
{
    bool g = false;
    for (int i = 1; i < K; ++i) g = g || (fabs(logL[i] - logL[r]) > 700.0);
    return g;
}

__device__ __forceinline__ double inv_sum_exp(const double* logL, int K, int r)
{
    double s = 0.0;
    for (int k = 0; k < K; ++k) s += exp(logL[k] - logL[r]);
    return 1.0 / s;
}

template <int R>
__global__ void __launch_bounds__(kThreads)
bayesr_sweep_kernel(SweepParams p)
{
    __shared__ Shared<R> L;
    const uint32_t w = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const uint32_t lo = w * p.S;
    const uint32_t hi = min(lo + p.S, p.N);
    const int K = (int)p.K;
    const ScalarLayout SL{p.G, p.K};
    const double dN = (double)p.N;
    const double NM1 = (double)(p.N - 1u);

    // eps slice in registers: individual lo + 512 r + 8 lane + e
    double eps[R][8];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const uint32_t idx = lo + (uint32_t)r * kRun + 8u * (uint32_t)lane + (uint32_t)e;
            eps[r][e] = (idx < hi) ? p.eps[idx] : 0.0;
        }

    if (tid == 0) {
        L.abort = 0;
        L.mu = p.scal[SL.mu()];
        L.sigmaE = p.scal[SL.sigmaE()];
        L.epsSum = p.scal[SL.epsSum()];
    }
    for (uint32_t t = tid; t < p.G; t += kThreads) {
        L.sigmaG[t] = p.scal[SL.sigmaG() + t];
        L.betasqn[t] = 0.0;
    }
    for (uint32_t t = tid; t < p.G * p.K; t += kThreads) {
        L.pi[t] = p.scal[SL.pi() + t];
        L.v[t] = 0.0;
    }
    for (uint32_t t = tid; t < p.G * (p.K - 1); t += kThreads) L.cva[t] = p.cva[t];
    uint32_t rnd = __hip_atomic_load(p.ctl + kCtlRound, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned long long n_rounds = 0, n_changes = 0;
    __syncthreads();

    const uint32_t iters = (p.mode == kModeChain) ? p.iterations : 1u;
    for (uint32_t it = 0; it < iters; ++it) {
        const uint32_t iteration = p.iter_base + it;
        const uint32_t* order = p.order + (size_t)((p.mode == kModeChain) ? it : 0u) * p.M;

        if (p.mode == kModeChain) {
            // ---- updateMu (DenseBayesRRmz.cpp:36-42 / SparseBayesRRG.cpp:106-113)
            const double old_mu = L.mu;
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const uint32_t idx = lo + (uint32_t)r * kRun + 8u * (uint32_t)lane + (uint32_t)e;
                    if (idx < hi) eps[r][e] = eps[r][e] + old_mu;
                }
            double mean_for_mu;
            if (p.layout == kLayoutDense) {
                double part = 0.0;
#pragma unroll
                for (int r = 0; r < R; ++r)
#pragma unroll
                    for (int e = 0; e < 8; ++e) part += eps[r][e];
                double total;
                if (!allreduce_scalar<R>(p, L, rnd, part, w, wave, lane, &total)) goto done;
                ++n_rounds;
                mean_for_mu = total / dN;
            } else {
                __syncthreads();
                if (tid == 0) L.epsSum += old_mu * dN;
                __syncthreads();
                mean_for_mu = L.epsSum / dN;
            }
            const double new_mu = norm_rng(p.seed, iteration, kDrawMu, p.chain, mean_for_mu,
                                           L.sigmaE / dN);
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const uint32_t idx = lo + (uint32_t)r * kRun + 8u * (uint32_t)lane + (uint32_t)e;
                    if (idx < hi) eps[r][e] = eps[r][e] - new_mu;
                }
            __syncthreads();
            if (tid == 0) {
                L.mu = new_mu;
                if (p.layout != kLayoutDense) L.epsSum -= new_mu * dN;
            }
            for (uint32_t t = tid; t < p.G; t += kThreads) L.betasqn[t] = 0.0;
            for (uint32_t t = tid; t < p.G * p.K; t += kThreads) L.v[t] = 0.0;
            __syncthreads();
        }
        group_constants<R>(p, L);
        __syncthreads();

        // ---- the sweep (AnalysisGraph::exec over the order)
        for (uint32_t b0 = 0; b0 < p.M; b0 += p.B) {
            const int nb = (int)min(p.B, p.M - b0);

            // block load: per-marker scalars
            if (tid < nb) {
                const uint32_t j = order[b0 + tid];
                L.mk[tid] = j;
                L.gr[tid] = p.group ? p.group[j] : 0;
                L.beta_old[tid] = bits_as_double(ld_agent((const uint64_t*)(p.beta + j)));
                if (p.mode == kModeChain) {
                    double uu, zz;
                    marker_uz(p.seed, iteration, j, p.chain, &uu, &zz);
                    L.uu[tid] = uu;
                    L.zz[tid] = zz;
                } else {
                    L.uu[tid] = p.u[j];
                    L.zz[tid] = p.z[j];
                }
                if (p.layout == kLayoutDense) {
                    const double* lt = p.lut + 4 * (size_t)j;
                    // 2-bit value v -> dosage: 0->2, 1->missing, 2->1, 3->0
                    L.wt[tid][0] = lt[2];
                    L.wt[tid][1] = lt[3];
                    L.wt[tid][2] = lt[1];
                    L.wt[tid][3] = lt[0];
                } else {
                    const double* st = p.stats + 4 * (size_t)j;
                    L.mstat[tid][0] = st[0];
                    L.mstat[tid][1] = st[1];
                    L.mstat[tid][2] = st[3];
                    L.wt[tid][0] = 2.0;   // genotype 2
                    L.wt[tid][1] = 0.0;   // missing (not in the ones/twos lists)
                    L.wt[tid][2] = 1.0;   // genotype 1
                    L.wt[tid][3] = 0.0;   // genotype 0
                }
            }
            // codes
            if (p.layout == kLayoutDense) {
                const uint32_t per = (uint32_t)R * 32u;
                for (uint32_t t = tid; t < (uint32_t)nb * per; t += kThreads) {
                    const uint32_t k = t / per, rem = t % per;
                    const uint32_t j = order[b0 + k];
                    const uint64_t byte_off = (uint64_t)(lo >> 2) + 4ull * rem;
                    uint32_t word = 0u;
                    if (byte_off < p.col_bytes)
                        word = p.dense[(size_t)j * p.col_words + (lo >> 4) + rem];
                    (&L.codes[0][0][0])[k * per + rem] = word;
                }
                __syncthreads();
            } else {
                const uint32_t per = (uint32_t)R * 32u;
                for (uint32_t t = tid; t < (uint32_t)nb * per; t += kThreads)
                    (&L.codes[0][0][0])[t] = 0xFFFFFFFFu;   // all genotype 0 (code 3)
                __syncthreads();
                for (int k = 0; k < nb; ++k) {
                    const uint32_t j = order[b0 + k];
#pragma unroll
                    for (int list = 0; list < 3; ++list) {
                        const uint32_t* sg = p.seg + ((size_t)j * 3u + (uint32_t)list) * (p.P + 1u);
                        const uint32_t s0 = sg[w], s1 = sg[w + 1];
                        const int32_t* base = (list == 0) ? p.ones + p.ones_off[j]
                                            : (list == 1) ? p.twos + p.twos_off[j]
                                                          : p.miss + p.miss_off[j];
                        // ones: code 3 -> 2 (clear bit 0); twos: -> 0 (clear both); missing: -> 1 (clear bit 1)
                        const uint32_t clr = (list == 0) ? 1u : (list == 1) ? 3u : 2u;
                        for (uint32_t q = s0 + (uint32_t)tid; q < s1; q += kThreads) {
                            const uint32_t loc = (uint32_t)base[q] - lo;
                            const uint32_t r = loc >> 9, ln = (loc >> 3) & 63u, e = loc & 7u;
                            const uint32_t bit = 16u * (ln & 1u) + 2u * e;
                            atomicAnd(&L.codes[k][r][ln >> 1], ~(clr << bit));
                        }
                    }
                }
                __syncthreads();
            }

            int pos = 0;
            while (pos < nb) {
                // ---- partial dots of slots [pos, nb) over this slice
                ++rnd;
                ++n_rounds;
                for (int k = pos + wave; k < nb; k += kWaves) {
                    const double w0 = L.wt[k][0], w1 = L.wt[k][1], w2 = L.wt[k][2], w3 = L.wt[k][3];
                    double s = 0.0;
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const uint32_t bits = L.codes[k][r][lane >> 1] >> (16 * (lane & 1));
#pragma unroll
                        for (int e = 0; e < 8; ++e)
                            s = fma(wt_of((bits >> (2 * e)) & 3u, w0, w1, w2, w3), eps[r][e], s);
                    }
                    s = wave_sum(s);
                    if (lane == 0) publish_pair(slot1(p, rnd, k, w), rnd, s);
                }
                if (!exchange<R>(p, L, rnd, pos, nb, w, wave, lane)) goto done;

                // ---- decisions (wave 0, lane = slot), BayesRBase.cpp:300-366
                if (wave == 0) {
                    const int k = pos + lane;
                    bool changed = false;
                    if (k < nb) {
                        const int g = L.gr[k];
                        const double* gc = L.gc + g * K;
                        const double* gden = L.gden + g * K;
                        const double* gbsd = L.gbsd + g * K;
                        const double beta_old = L.beta_old[k];
                        double num;
                        if (p.layout == kLayoutDense) {
                            num = L.dot[k] + beta_old * NM1;
                        } else {
                            const double mean = L.mstat[k][0], sd = L.mstat[k][1];
                            num = beta_old * (dN - 1.0) - mean * L.epsSum / sd + L.dot[k] / sd;
                        }
                        double logL[kMaxK], muk[kMaxK];
                        logL[0] = gc[0];
                        muk[0] = 0.0;
                        for (int c = 1; c < K; ++c) {
                            muk[c] = num / gden[c];
                            logL[c] = gc[c] + 0.5 * (muk[c] * num) / L.sigmaE;
                        }
                        double acum = guard700(logL, K, 0) ? 0.0 : inv_sum_exp(logL, K, 0);
                        const double acum0 = acum;
                        const double uu = L.uu[k];
                        int sel = -1;
                        double bnew = beta_old;
                        for (int c = 0; c < K; ++c) {
                            if (uu <= acum) {
                                sel = c;
                                bnew = (c == 0) ? 0.0 : L.zz[k] * gbsd[c] + muk[c];
                                break;
                            } else if (c + 1 < K) {
                                if (!guard700(logL, K, c + 1)) acum += inv_sum_exp(logL, K, c + 1);
                            }
                        }
                        L.beta_new[k] = bnew;
                        L.acumv[k] = acum0;
                        L.compv[k] = sel;
                        changed = (bnew != beta_old);
                    }
                    const unsigned long long m = __ballot(changed);
                    if (lane == 0) L.change = m ? pos + (int)__ffsll((long long)m) - 1 : nb;
                }
                __syncthreads();
                const int c = L.change;
                const int last = (c < nb) ? c : nb - 1;

                // ---- commit accepted slots [pos, last] in order
                if (tid == 0) {
                    for (int k = pos; k <= last; ++k) {
                        const int sel = L.compv[k];
                        if (sel < 0) continue;
                        const int g = L.gr[k];
                        L.v[g * K + sel] += 1.0;
                        if (sel > 0) L.betasqn[g] += L.beta_new[k] * L.beta_new[k];
                    }
                    if (c < nb) {
                        const double dBeta = L.beta_old[c] - L.beta_new[c];
                        if (p.layout == kLayoutDense) {
#pragma unroll
                            for (int q = 0; q < 4; ++q) L.delta[q] = dBeta * L.wt[c][q];
                        } else {
                            // raggedbayesrkernel.cpp:10-30 and sparsebayesrkernel.cpp:29-34
                            const double mean = L.mstat[c][0], sd = L.mstat[c][1], Zsum = L.mstat[c][2];
                            const double meanAdjustment = dBeta * mean / sd;
                            const double oneAdjustment = dBeta / sd;
                            L.delta[3] = -meanAdjustment;
                            L.delta[2] = -meanAdjustment + oneAdjustment;
                            L.delta[0] = -meanAdjustment + 2 * oneAdjustment;
                            L.delta[1] = -meanAdjustment + meanAdjustment;
                            L.epsSum += dBeta * Zsum / sd - dBeta * mean * dN / sd;
                        }
                    }
                }
                if (w == 0 && tid >= pos && tid <= last) {
                    const uint32_t j = L.mk[tid];
                    st_agent((uint64_t*)(p.beta + j), double_as_bits(L.beta_new[tid]));
                    if (L.compv[tid] >= 0) p.comp[j] = L.compv[tid];
                    p.acum[j] = L.acumv[tid];
                }
                __syncthreads();
                if (c < nb) {
                    ++n_changes;
                    const double d0 = L.delta[0], d1 = L.delta[1], d2 = L.delta[2], d3 = L.delta[3];
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const uint32_t bits = L.codes[c][r][lane >> 1] >> (16 * (lane & 1));
#pragma unroll
                        for (int e = 0; e < 8; ++e) {
                            const uint32_t idx = lo + (uint32_t)r * kRun + 8u * (uint32_t)lane + (uint32_t)e;
                            if (idx < hi) eps[r][e] = eps[r][e] + wt_of((bits >> (2 * e)) & 3u, d0, d1, d2, d3);
                        }
                    }
                }
                pos = last + 1;
            }
            // workgroup 0's beta stores must be complete before later rounds
            // let other workgroups read them
            if (w == 0) __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
        }

        if (p.mode == kModeChain) {
            // ---- sigmaE, sigmaG, pi draws (BayesRBase.cpp:237-247)
            double part = 0.0;
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int e = 0; e < 8; ++e) part += eps[r][e] * eps[r][e];
            double e2;
            if (!allreduce_scalar<R>(p, L, rnd, part, w, wave, lane, &e2)) goto done;
            ++n_rounds;
            if (tid == 0) {
                L.sigmaE = inv_scaled_chisq_rng(p.seed, iteration, kDrawSigmaE, p.chain, p.v0E + p.N,
                                                (e2 + p.v0E * p.s02E) / (p.v0E + p.N));
                double m0_total = 0.0;
                for (uint32_t g = 0; g < p.G; ++g) {
                    double vsum = 0.0;
                    for (int k = 0; k < K; ++k) vsum += L.v[g * K + k];
                    const double m0 = vsum - L.v[g * K];
                    m0_total += m0;
                    L.sigmaG[g] = inv_scaled_chisq_rng(p.seed, iteration, kDrawSigmaG0 + g, p.chain,
                                                       p.v0G + m0,
                                                       (L.betasqn[g] * m0 + p.v0G * p.s02G) / (p.v0G + m0));
                    double gs[kMaxK];
                    double sum = 0.0;
                    for (int k = 0; k < K; ++k)
                        gs[k] = rgamma(p.seed, iteration, 2u + p.G + g * (uint32_t)K + (uint32_t)k,
                                       p.chain, L.v[g * K + k] + 1.0, 1.0);
                    for (int k = 0; k < K; ++k) sum += gs[k];
                    for (int k = 0; k < K; ++k) L.pi[g * K + k] = gs[k] / sum;
                }
                if (w == 0 && p.hist) {
                    double* h = p.hist + (size_t)it * (4u + p.G);
                    h[0] = L.mu;
                    h[1] = L.sigmaE;
                    h[2] = m0_total;
                    h[3] = e2;
                    for (uint32_t g = 0; g < p.G; ++g) h[4 + g] = L.sigmaG[g];
                }
            }
            __syncthreads();
        }
    }

done:
    __syncthreads();
    if (L.abort == 0) {
        if (wave == 0) {
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const uint32_t idx = lo + (uint32_t)r * kRun + 8u * (uint32_t)lane + (uint32_t)e;
                    if (idx < hi) p.eps[idx] = eps[r][e];
                }
        }
        if (w == 0) {
            if (tid == 0) {
                p.scal[SL.mu()] = L.mu;
                p.scal[SL.sigmaE()] = L.sigmaE;
                p.scal[SL.epsSum()] = L.epsSum;
                __hip_atomic_store(p.ctl + kCtlRound, rnd, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                p.counters[0] = n_rounds;
                p.counters[1] = n_changes;
            }
            for (uint32_t t = tid; t < p.G; t += kThreads) {
                p.scal[SL.sigmaG() + t] = L.sigmaG[t];
                p.scal[SL.betasqn() + t] = L.betasqn[t];
            }
            for (uint32_t t = tid; t < p.G * p.K; t += kThreads) {
                p.scal[SL.pi() + t] = L.pi[t];
                p.scal[SL.v() + t] = L.v[t];
            }
        }
    }
}

} // namespace brr
