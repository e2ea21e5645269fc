// batch_kernel.hpp -- persistent gfx950 kernel for T BayesR chains (traits)
// that share one genotype matrix.
//
// What it computes: for every trait t < T an independent BayesRBase::processColumn
// sweep (BayesRBase.cpp:287-392) over one marker order, each trait with its own
// residual, beta, components and chain scalars. That is what T separate runs of
// the reference over T phenotypes compute when they are handed the same order
// (each run is a valid random-scan chain; the order stream is shared).
//
// How (MI355X-first): one cooperative launch of P workgroups, one per CU. A
// workgroup owns a slice of at most 16 * kMaxNT individuals and keeps its
// residuals for all traits in registers (FP64). Markers go in blocks of 32:
//   1. matrix cores (v_mfma_f64_16x16x4f64) form the slice's partial dots
//      C[k][t] = x_k . eps_t for the 32 markers x T traits, and the partial Gram
//      G[i][k] = u_i . w_k (i < k) of the block, where w_k is marker k's dot
//      weight and u_i the residual change per unit beta change of marker i;
//   2. one deterministic reduce-scatter / all-gather (agent-coherent tagged
//      words) makes every workgroup hold the full C and G;
//   3. every workgroup evaluates the same FP64 decisions: all pending
//      (marker, trait) pairs in parallel; per trait the pairs up to its first
//      beta change are accepted, the dots of that trait's later markers are
//      corrected with the change times the Gram row, and the rest is decided
//      again -- the one-marker-at-a-time result, with the dot of a later marker
//      formed as C + sum(delta * G) instead of a fresh x . eps;
//   4. matrix cores apply eps_t += sum_k delta_{k,t} u_k to the slice (only
//      4-marker steps that hold a change are issued).
// So one exchange per block of 32 markers serves all T traits.
#pragma once

#include <type_traits>

#include "sweep_kernel.hpp"

namespace brr {

constexpr int kBB = 32;                      // markers per block (two 16-marker tiles)
constexpr int kMaxT = 64;                    // traits per launch
constexpr int kMaxNT = 32;                   // 16-individual tiles per workgroup slice (FP32 residual)
constexpr int kMaxNT64 = 16;                 // ... with the FP64 residual
constexpr int kBRow = 4 * kMaxNT + 4;        // LDS bytes per code row (padded: rows fall in distinct banks)
constexpr int kBGram = kBB * (kBB - 1) / 2;  // strictly-upper Gram entries of a block
constexpr uint32_t kBVmax = (uint32_t)(kBB * kMaxT + kBGram);   // values exchanged per block
// phase timers: 0 block commit, 1 dot tiles, 2 exchange gather, 3 decision draws,
// 4 residual update, 5 iteration scalars, 6 Gram tiles, 7 u / z draws,
// 8 exchange reduce, 9 accept / correct, 10 block-end barrier, 11 spare
constexpr int kBPhases = 12;

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

struct BatchParams {
    uint32_t N, M, K, T;
    uint32_t P, S, nt, nq;            // workgroups, slice, 16-row tiles per slice, 16-trait tiles
    int layout, mode;
    uint32_t iterations, iter_base, chain, pad0;
    uint64_t seed;
    const uint32_t* codes;            // 2-bit PLINK columns, column j at codes + j * col_words
    uint64_t col_words, col_bytes;
    const double* lut;                // dense: M x 4 by dosage 0, 1, 2, missing
    const double* stats;              // ragged: M x 4 {mean, sd, sqrdZ, Zsum}
    const double* cva;                // K - 1 mixture variances
    double v0E, s02E, v0G, s02G;
    const uint32_t* order;            // M (sweep) or iterations x M (chain)
    const double* u;                  // sweep mode: M x T, by marker id
    const double* z;
    double* eps;                      // N x T (trait fastest)
    double* beta;                     // M x T (agent-coherent words)
    int32_t* comp;                    // M x T
    double* acum;                     // M x T
    double* scal;                     // T scalar blocks of ScalarLayout{1, K}
    double* hist;                     // iterations x T x 5 or null
    uint64_t* xchg1;                  // [2][P][kBVmax][2] tagged partials
    uint64_t* xchg2;                  // [2][kBVmax][2] tagged totals
    uint32_t* ctl;
    unsigned long long* counters;     // [0] rounds, [1] changes, [2..7] phase ticks
};

// Per-block inputs, double-buffered: the current block reads buffer cb while
// the next block's data is written to cb ^ 1.
struct BBlock {
    uint8_t codes[kBB][kBRow];        // byte b of row k: individuals lo + 4b .. 4b + 3 of marker k
    double wt[kBB][4];                // dot weight per 2-bit code (0 for markers past the block)
    double ut[kBB][4];                // residual change per unit beta change, per code
    double mstat[kBB][3];             // ragged: mean, sd, Zsum
    double bo[kBB][kMaxT];            // beta before the block
    uint32_t mk[kBB];                 // marker ids
};

struct BShared {
    BBlock blk[2];
    double cd[kBB][kMaxT];            // dots, corrected dots, then the accepted beta deltas
    double gm[kBB][kBB];              // Gram u_i . w_k, i < k
    double bn[kBB][kMaxT];            // decided beta
    double acm[kBB];                  // acum of the latest draw of this workgroup's output trait
                                      // (at most one per block while T <= P)
    int8_t sl[kBB][kMaxT];            // decided component (-1: none selected)
    double uu[kBB][kMaxT];            // uniform of (marker slot, trait), drawn while the exchange runs
    uint64_t zw[kBB][kMaxT];          // its normal: Philox words (chain) or the given z's bits
    uint16_t plist[kBB * kMaxT];      // pending pairs of a round after the first: k << 8 | t
    double gc[kMaxK][kMaxT];          // log pi_k - 0.5 log(logLScale S_k + 1)
    double gden[kMaxK][kMaxT];        // denom_k
    double gbsd[kMaxK][kMaxT];        // sqrt(sigmaE / denom_k)
    double dg[kMaxK][kMaxT];          // logL_k - logL_0 = dg_k + num^2 hk_k
    double hk[kMaxK][kMaxT];
    double mu[kMaxT], sigmaE[kMaxT], epsSum[kMaxT], sigmaG[kMaxT], betasqn[kMaxT];
    double pi[kMaxK][kMaxT];
    double v[kMaxK][kMaxT];
    double tvec[kMaxT];               // per-trait all-reduce results / new mu
    double cva[kMaxK];                // S_k for k >= 1
    double red[kWaves][8];
    int pos[kMaxT];                   // first undecided marker of the block, per trait
    int fc[kMaxT];                    // first change among this round's pairs, per trait
    int pref[kMaxT + 1];              // exclusive prefix of pending pairs over traits
    uint32_t flag[kBB / 4];           // bit q: 4-marker step holds a change in trait tile q
    uint64_t tk[kBPhases];            // phase timers of workgroup 0 (100 MHz ticks)
    unsigned long long sum_changes, sum_drounds;
    uint64_t tlast;
    int abort;
};

// Phase timer of workgroup 0, thread 0, kept in LDS.
#define BRR_BTICK(slot)                                                          \
    do {                                                                         \
        if (tid == 0) {                                                          \
            const uint64_t t_now_ = __builtin_amdgcn_s_memrealtime();           \
            L.tk[(slot)] += t_now_ - L.tlast;                                    \
            L.tlast = t_now_;                                                    \
        }                                                                        \
    } while (0)

__device__ __forceinline__ uint64_t* bslot1(const BatchParams& p, uint32_t rnd, uint32_t w, uint32_t v)
{
    return p.xchg1 + ((((size_t)(rnd & 1u) * p.P + w) * kBVmax + v) * 2u);
}

// totals in kCopies copies, as the BayesR sweep's (copy w % kCopies is read by workgroup w)
constexpr size_t kBCopyWords = 2 * (size_t)kBVmax * 2 + 32;

__device__ __forceinline__ uint64_t* bslot2(const BatchParams& p, uint32_t rnd, uint32_t v, uint32_t copy)
{
    return p.xchg2 + copy * kBCopyWords + (((size_t)(rnd & 1u) * kBVmax + v) * 2u);
}

// Orders the LDS accesses of one wave across its lanes: the wave's LDS
// operations complete in issue order, so a compiler-level barrier suffices.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ dbl4 mfma16x16x4(double a, double b, dbl4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ flt4 mfma16x16x4(float a, float b, flt4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Row of a 16 x 16 matrix-core accumulator held in entry r of lane (kq, tq)
// (column tq): the f64 instruction interleaves rows over the lane groups
// (4 r + kq), the f32 one blocks them (4 kq + r), as measured on the device.
template <bool kF32>
__device__ __forceinline__ int acc_row(int r, int kq)
{
    return kF32 ? 4 * kq + r : 4 * r + kq;
}

__device__ __forceinline__ dbl4 to_dbl4(flt4 x) { return dbl4{(double)x[0], (double)x[1], (double)x[2], (double)x[3]}; }
__device__ __forceinline__ dbl4 to_dbl4(dbl4 x) { return x; }

// Reduce-scatter of the V values published at round rnd. Chunks of 8
// consecutive values belong to workgroup (chunk % P); thread (g, e) of the
// owner sums partial (w2, 8 chunk + e) for w2 = g, g + 32, ... in order, the
// 8 lane groups of a wave are folded by a fixed butterfly, the 4 waves in
// index order, and the total is published. Every total is formed once, so all
// workgroups read the same bits. Returns false (in all threads) on abort.
__device__ __forceinline__ bool breduce(const BatchParams& p, BShared& L, uint32_t rnd, uint32_t V,
                                        uint32_t w, int tid)
{
    const int lane = tid & 63, wave = tid >> 6;
    const uint32_t nch = (V + 7u) / 8u;
    for (uint32_t c = w; c < nch; c += p.P) {
        const uint32_t v = 8u * c + (uint32_t)(tid & 7);
        // all of a thread's partials are loaded at once and only stale ones polled
        // again; summed in w2 order as before
        double acc = 0.0;
        bool ok = true;
        if (v < V)
            ok = wait_sum_strided<kMaxP / 32>(p.xchg1, [&](uint32_t w2) { return bslot1(p, rnd, w2, v); },
                                              (uint32_t)(tid >> 3), 32u, p.P, rnd, p.ctl, &acc);
        if (!ok) { L.abort = 1; raise_abort(p, 11); }
        acc += __shfl_xor(acc, 8, 64);
        acc += __shfl_xor(acc, 16, 64);
        acc += __shfl_xor(acc, 32, 64);
        if (lane < 8) L.red[wave][lane] = acc;
        __syncthreads();
        if (L.abort) return false;
        if (tid < 8 * kCopies && v < V) {
            const int e = tid & 7;
            publish_pair(p.xchg2, bslot2(p, rnd, v, (uint32_t)(tid >> 3)), rnd,
                         ((L.red[0][e] + L.red[1][e]) + L.red[2][e]) + L.red[3][e]);
        }
        __syncthreads();
    }
    return true;
}

// All-reduce of one value per trait (published at v = trait): totals to L.tvec.
__device__ __forceinline__ bool ballreduce_traits(const BatchParams& p, BShared& L, uint32_t rnd, uint32_t w,
                                                  int tid)
{
    if (!breduce(p, L, rnd, p.T, w, tid)) return false;
    if (tid < (int)p.T) {
        double x;
        if (!wait_pair(p.xchg2, bslot2(p, rnd, (uint32_t)tid, w % kCopies), rnd, p.ctl, &x)) { L.abort = 1; raise_abort(p, 12); }
        else L.tvec[tid] = x;
    }
    __syncthreads();
    return L.abort == 0;
}

// ------------------------------------------------------------ block loading --

// Everything of the next block that can be fetched before it is needed, in
// registers: code words, per-marker tables and beta of (marker, trait) pairs.
constexpr int kBWords = kBB * kMaxNT / kThreads;   // code words of a block per thread

struct BNext {
    uint32_t words[kBWords];    // kBB * kMaxNT words over 256 threads
    double m0, m1, m2, m3;      // dense: table by dosage; ragged: mean, sd, Zsum
    double b[8];                // beta of pair u * 256 + tid = (k = e >> 6, t = e & 63)
};

__device__ __forceinline__ void bissue(const BatchParams& p, const uint32_t* ids, int nb, uint32_t lo, int tid,
                                       BNext& x)
{
    const uint32_t per = p.nt;   // 32-bit words (16 individuals each) per marker row
#pragma unroll
    for (int u = 0; u < kBWords; ++u) {
        const uint32_t t = (uint32_t)u * kThreads + (uint32_t)tid;
        x.words[u] = 0u;
        if (t < (uint32_t)nb * per) {
            const uint32_t k = t / per, rem = t % per;
            const uint64_t byte_off = (uint64_t)(lo >> 2) + 4ull * rem;
            if (byte_off < p.col_bytes) x.words[u] = p.codes[(size_t)ids[k] * p.col_words + (lo >> 4) + rem];
        }
    }
    x.m0 = x.m1 = x.m2 = x.m3 = 0.0;
    if (tid < nb) {
        const uint32_t j = ids[tid];
        if (p.layout == kLayoutDense) {
            const double* lt = p.lut + 4 * (size_t)j;
            x.m0 = lt[0]; x.m1 = lt[1]; x.m2 = lt[2]; x.m3 = lt[3];
        } else {
            const double* st = p.stats + 4 * (size_t)j;
            x.m0 = st[0]; x.m1 = st[1]; x.m2 = st[3];
        }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const int e = u * kThreads + tid;
        const int k = e >> 6, t = e & 63;
        x.b[u] = 0.0;
        if (k < nb && t < (int)p.T)
            x.b[u] = bits_as_double(ld_agent((const uint64_t*)(p.beta + (size_t)ids[k] * p.T + t)));
    }
}

__device__ __forceinline__ void bstore(const BatchParams& p, BBlock& L, const BNext& x, int nb, int tid)
{
    const uint32_t per = p.nt;
#pragma unroll
    for (int u = 0; u < kBWords; ++u) {
        const uint32_t t = (uint32_t)u * kThreads + (uint32_t)tid;
        if (t < (uint32_t)nb * per) *(uint32_t*)&L.codes[t / per][4u * (t % per)] = x.words[u];
    }
    if (tid < kBB) {
        const int k = tid;
        if (k >= nb) {
#pragma unroll
            for (int c = 0; c < 4; ++c) { L.wt[k][c] = 0.0; L.ut[k][c] = 0.0; }
            L.mstat[k][0] = 0.0; L.mstat[k][1] = 1.0; L.mstat[k][2] = 0.0;
        } else if (p.layout == kLayoutDense) {
            // 2-bit code v -> dosage: 0->2, 1->missing, 2->1, 3->0
            L.wt[k][0] = x.m2; L.wt[k][1] = x.m3; L.wt[k][2] = x.m1; L.wt[k][3] = x.m0;
            L.ut[k][0] = x.m2; L.ut[k][1] = x.m3; L.ut[k][2] = x.m1; L.ut[k][3] = x.m0;
        } else {
            // ragged dot over the ones/twos lists (missing counts 0), and the
            // residual change per unit beta change (raggedbayesrkernel.cpp:10-30)
            const double mean = x.m0, sd = x.m1;
            L.mstat[k][0] = mean; L.mstat[k][1] = sd; L.mstat[k][2] = x.m2;
            L.wt[k][0] = 2.0; L.wt[k][2] = 1.0; L.wt[k][3] = 0.0;
            L.wt[k][1] = (p.layout == kLayoutEigen) ? mean : 0.0;   // EigenBayesRKernel: missing at the mean
            const double meanAdjustment = mean / sd;
            const double oneAdjustment = 1.0 / sd;
            L.ut[k][0] = -meanAdjustment + 2 * oneAdjustment;
            L.ut[k][1] = 0.0;
            L.ut[k][2] = -meanAdjustment + oneAdjustment;
            L.ut[k][3] = -meanAdjustment;
        }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const int e = u * kThreads + tid;
        const int k = e >> 6, t = e & 63;
        if (k < nb && t < (int)p.T) L.bo[k][t] = x.b[u];
    }
}

// ------------------------------------------------------------- decisions --

// num of processColumn for slot k of trait t from the current dot in L.cd
// (densebayesrkernel.cpp:10-15; sparsebayesrkernel.cpp:24-27)
template <bool kDenseMath>
__device__ __forceinline__ double bnum(const BShared& L, const BBlock& B, int k, int t, double dN, double NM1)
{
    const double beta_old = B.bo[k][t];
    if (kDenseMath) return L.cd[k][t] + beta_old * NM1;
    const double mean = B.mstat[k][0], sd = B.mstat[k][1];
    return beta_old * (dN - 1.0) - mean * L.epsSum[t] / sd + L.cd[k][t] / sd;
}

// logL and mu of every component (BayesRBase.cpp:309-333)
template <int kK>
__device__ __forceinline__ void blogl(const BShared& L, int t, int K, double num, double (&logL)[kK],
                                      double (&muk)[kK])
{
    const double sigmaE = L.sigmaE[t];
    logL[0] = L.gc[0][t];
    muk[0] = 0.0;
#pragma unroll
    for (int c = 1; c < kK; ++c) {
        muk[c] = 0.0;
        logL[c] = 0.0;
        if (c < K) {
            muk[c] = num / L.gden[c][t];
            logL[c] = L.gc[c][t] + 0.5 * (muk[c] * num) / sigmaE;
        }
    }
}

// The draw of processColumn (BayesRBase.cpp:335-366) from num, u and z. With
// d_k = logL_k - logL_0 = dg_k + num^2 hk_k, the probability terms
// 1 / sum_i exp(logL_i - logL_c) equal exp(d_c) / sum_i exp(d_i); while every
// |d_k| <= 350 no 700 guard can fire and no exponential overflows, so they are
// formed from one set of K - 1 exponentials. Otherwise the reference's own
// formulation, guards included, is evaluated.
struct BDraw {
    double bnew, acum0;
    int sel;
};

template <int kK>
__device__ __forceinline__ BDraw bdraw_k(const BShared& L, int t, int K, double num, double beta_old, double u,
                                       uint64_t zw, bool zwords)
{
    const double n2 = num * num;
    double d[kK];
    d[0] = 0.0;
    bool far = false;
#pragma unroll
    for (int c = 1; c < kK; ++c) {
        d[c] = 0.0;
        if (c < K) {
            d[c] = L.dg[c][t] + n2 * L.hk[c][t];
            far = far || fabs(d[c]) > 350.0;
        }
    }
    BDraw r{beta_old, 0.0, -1};
    if (!far) {
        double e[kK];
        double S = 1.0;
        e[0] = 1.0;
#pragma unroll
        for (int c = 1; c < kK; ++c) {
            e[c] = 0.0;
            if (c < K) {
                e[c] = exp(d[c]);
                S += e[c];
            }
        }
        const double invS = 1.0 / S;
        double acum = invS;
        r.acum0 = acum;
#pragma unroll
        for (int c = 0; c < kK; ++c) {
            if (c < K && r.sel < 0) {
                if (u <= acum) r.sel = c;
                else if (c + 1 < kK && c + 1 < K) acum += e[c + 1 < kK ? c + 1 : 0] * invS;
            }
        }
    } else {
        double logL[kK], muk[kK];
        blogl<kK>(L, t, K, num, logL, muk);
        double acum = guard700(logL, K, logL[0]) ? 0.0 : inv_sum_exp(logL, K, logL[0]);
        r.acum0 = acum;
#pragma unroll 1
        for (int c = 0; c < K; ++c) {
            if (u <= acum) { r.sel = c; break; }
            if (c + 1 < K) {
                double ref = 0.0;
#pragma unroll
                for (int q = 1; q < kK; ++q)
                    if (q == c + 1) ref = logL[q];
                if (!guard700(logL, K, ref)) acum += inv_sum_exp(logL, K, ref);
            }
        }
    }
    if (r.sel == 0) {
        r.bnew = 0.0;
    } else if (r.sel > 0) {
        const double z = zwords ? unit_normal((uint32_t)(zw >> 32), (uint32_t)zw) : bits_as_double(zw);
        r.bnew = z * L.gbsd[r.sel][t] + num / L.gden[r.sel][t];
    }
    return r;
}

// The loops unrolled to 4 components when K <= 4 (the usual mixture), else to kMaxK.
__device__ __forceinline__ BDraw bdraw(const BShared& L, int t, int K, double num, double beta_old, double u,
                                       uint64_t zw, bool zwords)
{
    return K <= 4 ? bdraw_k<4>(L, t, K, num, beta_old, u, zw, zwords)
                  : bdraw_k<kMaxK>(L, t, K, num, beta_old, u, zw, zwords);
}

// -------------------------------------------------------------------- kernel --

// kF32: the residual tiles are FP32 on the f32 matrix-core instruction (up to
// 32 tiles, 512 individuals, per workgroup instead of 16); every 16-individual
// tile's FP32 dot and Gram products are accumulated across tiles in FP64, and
// the decisions, exchange and scalar draws stay FP64. The residual is not
// refreshed: its FP32 rounding accumulates over the chain.
template <int NT, bool kDenseMath, bool kF32>
__global__ void __launch_bounds__(kThreads)
bayesr_batch_kernel(BatchParams p)
{
    using ET = typename std::conditional<kF32, float, double>::type;
    using ET4 = typename std::conditional<kF32, flt4, dbl4>::type;
    __shared__ BShared L;
    const uint32_t w = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int tq = lane & 15;          // matrix-core row / column index of this lane
    const int kq = lane >> 4;          // matrix-core k index of this lane
    const uint32_t lo = w * p.S;
    const uint32_t nloc = min(p.S, p.N - lo);
    const int T = (int)p.T;
    const int K = (int)p.K;
    const int nt = (int)p.nt;
    const int tl = 16 * wave + tq;     // trait of this lane's residual column
    const bool tval = tl < T;
    const ScalarLayout SL{1u, p.K};
    const uint32_t sls = SL.size();
    const double dN = (double)p.N;
    const double NM1 = (double)(p.N - 1u);
    constexpr bool dense = kDenseMath;

    // residuals in the matrix-core accumulator layout (lane = 16 * (row % 4) + column,
    // entry r = row / 4): E[s][r] = eps of individual lo + 16 s + 4 r + kq, trait tl
    // (FP32 tiles: lo + 16 s + 4 kq + r, the f32 instruction's blocked rows).
    // As the B operand of step (s, r) the same registers hold individuals
    // 16 s + 4 r + 0..3 (k index kq), so a tile feeds the dots without shuffles.
    ET4 E[NT];
#pragma unroll
    for (int s = 0; s < NT; ++s)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const uint32_t i = 16u * (uint32_t)s + (uint32_t)acc_row<kF32>(r, kq);
            double e = 0.0;
            if (s < nt && i < nloc && tval) e = p.eps[(size_t)(lo + i) * T + tl];
            E[s][r] = (ET)e;
        }

    if (tid < kMaxT) {
        const int t = tid;
        L.mu[t] = 0.0; L.sigmaE[t] = 1.0; L.epsSum[t] = 0.0; L.sigmaG[t] = 1.0; L.betasqn[t] = 0.0;
        for (int k = 0; k < kMaxK; ++k) { L.pi[k][t] = 0.0; L.v[k][t] = 0.0; }
        if (t < T) {
            const double* sc = p.scal + (size_t)t * sls;
            L.mu[t] = sc[SL.mu()];
            L.sigmaE[t] = sc[SL.sigmaE()];
            L.epsSum[t] = sc[SL.epsSum()];
            L.sigmaG[t] = sc[SL.sigmaG()];
            for (int k = 0; k < K; ++k) L.pi[k][t] = sc[SL.pi() + k];
        }
    }
    if (tid < kMaxK) L.cva[tid] = (tid >= 1 && tid < K) ? p.cva[tid - 1] : 0.0;
    for (int e = tid; e < kBB * kMaxT; e += kThreads) (&L.cd[0][0])[e] = 0.0;
    if (tid == 0) L.abort = 0;
    uint32_t rnd = __hip_atomic_load(p.ctl + kCtlRound, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned long long n_rounds = 0, n_changes = 0, n_drounds = 0;
    if (tid < kBPhases) L.tk[tid] = 0;
    if (tid == 0) { L.sum_changes = 0; L.sum_drounds = 0; }
    if (tid == 0) L.tlast = __builtin_amdgcn_s_memrealtime();
    int cb = 0;     // current block buffer
    __syncthreads();

    const uint32_t iters = (p.mode == kModeChain) ? p.iterations : 1u;
    const uint32_t nblocks = (p.M + kBB - 1) / kBB;
    for (uint32_t it = 0; it < iters; ++it) {
        const uint32_t iteration = p.iter_base + it;
        const uint32_t* order = p.order + (size_t)((p.mode == kModeChain) ? it : 0u) * p.M;

        if (p.mode == kModeChain) {
            // ---- updateMu per trait (DenseBayesRRmz.cpp:36-42 / SparseBayesRRG.cpp:106-113)
            const double om = L.mu[tl];
#pragma unroll
            for (int s = 0; s < NT; ++s)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const uint32_t i = 16u * (uint32_t)s + (uint32_t)acc_row<kF32>(r, kq);
                    if (s < nt && i < nloc && tval) E[s][r] = (ET)((double)E[s][r] + om);
                }
            if (dense) {
                double part = 0.0;
#pragma unroll
                for (int s = 0; s < NT; ++s)
#pragma unroll
                    for (int r = 0; r < 4; ++r) part += (double)E[s][r];
                part += __shfl_xor(part, 16, 64);
                part += __shfl_xor(part, 32, 64);
                ++rnd;
                ++n_rounds;
                if (kq == 0 && tval) publish_pair(p.xchg1, bslot1(p, rnd, w, (uint32_t)tl), rnd, part);
                if (!ballreduce_traits(p, L, rnd, w, tid)) goto done;
            } else {
                __syncthreads();
                if (tid < T) L.epsSum[tid] += L.mu[tid] * dN;
                __syncthreads();
            }
            if (tid < T) {
                const double mean = dense ? L.tvec[tid] / dN : L.epsSum[tid] / dN;
                L.tvec[tid] = norm_rng(p.seed, iteration, kDrawMu, p.chain + (uint32_t)tid, mean,
                                       L.sigmaE[tid] / dN);
            }
            __syncthreads();
            const double nm = L.tvec[tl];
#pragma unroll
            for (int s = 0; s < NT; ++s)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const uint32_t i = 16u * (uint32_t)s + (uint32_t)acc_row<kF32>(r, kq);
                    if (s < nt && i < nloc && tval) E[s][r] = (ET)((double)E[s][r] - nm);
                }
            __syncthreads();
            if (tid < T) {
                L.mu[tid] = L.tvec[tid];
                if (!dense) L.epsSum[tid] -= L.tvec[tid] * dN;
                L.betasqn[tid] = 0.0;
            }
            for (int e = tid; e < kMaxK * kMaxT; e += kThreads) (&L.v[0][0])[e] = 0.0;
            __syncthreads();
        }

        // ---- per-iteration constants of processColumn (BayesRBase.cpp:309-333)
        for (int e = tid; e < T * K; e += kThreads) {
            const int t = e / K, k = e % K;
            const double sigmaG = L.sigmaG[t], sigmaE = L.sigmaE[t];
            if (k == 0) {
                L.gc[0][t] = log(L.pi[0][t]);
                L.gden[0][t] = 0.0;
                L.gbsd[0][t] = 0.0;
            } else {
                const double S = L.cva[k];
                const double sigmaEOverSigmaG = sigmaE / sigmaG;
                const double den = NM1 + sigmaEOverSigmaG * (1.0 / S);
                const double logLScale = sigmaG / sigmaE * NM1;
                L.gden[k][t] = den;
                L.gc[k][t] = log(L.pi[k][t]) - 0.5 * log(logLScale * S + 1.0);
                L.gbsd[k][t] = __dsqrt_rn(sigmaE / den);
                L.hk[k][t] = 0.5 / (den * sigmaE);
            }
        }
        __syncthreads();
        for (int e = tid; e < T * K; e += kThreads) {
            const int t = e / K, k = e % K;
            L.dg[k][t] = (k == 0) ? 0.0 : L.gc[k][t] - L.gc[0][t];
        }
        // ---- block 0 loaded in place; ids of block 1 staged
        {
            const int nb0 = (int)min((uint32_t)kBB, p.M);
            if (tid < nb0) L.blk[cb].mk[tid] = order[tid];
            __syncthreads();
            BNext x;
            bissue(p, L.blk[cb].mk, nb0, lo, tid, x);
            bstore(p, L.blk[cb], x, nb0, tid);
            if (nblocks > 1) {
                const int nb1 = (int)min((uint32_t)kBB, p.M - kBB);
                if (tid < nb1) L.blk[cb ^ 1].mk[tid] = order[kBB + tid];
            }
            __syncthreads();
        }
        BRR_BTICK(5);

        for (uint32_t blk = 0; blk < nblocks; ++blk) {
            const uint32_t b0 = blk * kBB;
            const int nb = (int)min((uint32_t)kBB, p.M - b0);
            const bool has1 = blk + 1 < nblocks;
            const int nb1 = has1 ? (int)min((uint32_t)kBB, p.M - b0 - kBB) : 0;
            const bool has2 = blk + 2 < nblocks;
            const int nb2 = has2 ? (int)min((uint32_t)kBB, p.M - b0 - 2 * kBB) : 0;

            BBlock& B = L.blk[cb];
            // prefetch of block blk+1 into registers (stored to the other buffer
            // once the matrix-core phase has hidden the latency), ids of blk+2
            BNext x;
            if (has1) bissue(p, L.blk[cb ^ 1].mk, nb1, lo, tid, x);
            uint32_t jn2 = 0;
            if (has2 && tid < nb2) jn2 = order[b0 + 2 * kBB + tid];
            ++rnd;
            ++n_rounds;
            BRR_BTICK(0);

            // ---- 1. partial dots C[k][t] over this slice (wave q: traits 16q..16q+15)
            // and the partial Gram tiles (0,0), (0,1), (1,1) on waves 0, 1, 2. The
            // lane's code rows (markers tq and 16 + tq) and their tables sit in
            // registers, so each matrix-core step decodes its operand with a few
            // VALU selects instead of dependent LDS reads.
            {
                // (FP32 tiles: up to 32 of them, so their code words are read from LDS
                // tile by tile instead of being held in registers)
                constexpr int NR = kF32 ? 1 : NT;
                uint32_t rw0[NR], rw1[NR];
                if constexpr (!kF32) {
#pragma unroll
                    for (int s = 0; s < NT; ++s) {
                        rw0[s] = (s < nt) ? *(const uint32_t*)&B.codes[tq][4 * s] : 0u;
                        rw1[s] = (s < nt) ? *(const uint32_t*)&B.codes[16 + tq][4 * s] : 0u;
                    }
                }
                const double w00 = B.wt[tq][0], w01 = B.wt[tq][1], w02 = B.wt[tq][2], w03 = B.wt[tq][3];
                const double w10 = B.wt[16 + tq][0], w11 = B.wt[16 + tq][1], w12 = B.wt[16 + tq][2],
                             w13 = B.wt[16 + tq][3];
                if (wave < (int)p.nq) {
                    dbl4 c0, c1;
                    if constexpr (!kF32) {
                        // two accumulator chains per marker tile (even / odd residual tiles)
                        dbl4 acc0[2] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
                        dbl4 acc1[2] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
                        // straight-line code (no per-step branches, which would force the
                        // accumulators out of the matrix-core registers at each join):
                        // tiles past the slice hold zero residuals, slots past the block
                        // zero weights
#pragma unroll
                        for (int s = 0; s < NT; ++s) {
#pragma unroll
                            for (int r = 0; r < 4; ++r) {
                                const uint32_t sh = 8u * (uint32_t)r + 2u * (uint32_t)kq;
                                acc0[s & 1] = mfma16x16x4(wt_of((rw0[s < NR ? s : 0] >> sh) & 3u, w00, w01, w02, w03), E[s][r],
                                                          acc0[s & 1]);
                                acc1[s & 1] = mfma16x16x4(wt_of((rw1[s < NR ? s : 0] >> sh) & 3u, w10, w11, w12, w13), E[s][r],
                                                          acc1[s & 1]);
                            }
                        }
                        c0 = acc0[0] + acc0[1];
                        c1 = acc1[0] + acc1[1];
                    } else {
                        // per residual tile an FP32 product of 16 individuals, the tiles
                        // summed in FP64; entry r of the B operand is individual 4 kq + r
                        float f00 = (float)w00, f01 = (float)w01, f02 = (float)w02, f03 = (float)w03;
                        float f10 = (float)w10, f11 = (float)w11, f12 = (float)w12, f13 = (float)w13;
                        // converted once here (an empty asm keeps the compiler from selecting
                        // among the FP64 values and converting per step)
                        asm volatile("" : "+v"(f00), "+v"(f01), "+v"(f02), "+v"(f03));
                        asm volatile("" : "+v"(f10), "+v"(f11), "+v"(f12), "+v"(f13));
                        c0 = dbl4{0.0, 0.0, 0.0, 0.0};
                        c1 = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
                        for (int s = 0; s < NT; ++s) {
                            // loaded without a branch (every row holds kMaxNT words), masked after
                            const uint32_t x0 = *(const uint32_t*)&B.codes[tq][4 * s];
                            const uint32_t x1 = *(const uint32_t*)&B.codes[16 + tq][4 * s];
                            const uint32_t a0 = (s < nt) ? x0 : 0u;
                            const uint32_t a1 = (s < nt) ? x1 : 0u;
                            flt4 t0 = {0.f, 0.f, 0.f, 0.f}, t1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                            for (int r = 0; r < 4; ++r) {
                                const uint32_t sh = 8u * (uint32_t)kq + 2u * (uint32_t)r;
                                t0 = mfma16x16x4(wt_of_f((a0 >> sh) & 3u, f00, f01, f02, f03), (float)E[s][r], t0);
                                t1 = mfma16x16x4(wt_of_f((a1 >> sh) & 3u, f10, f11, f12, f13), (float)E[s][r], t1);
                            }
                            c0 += to_dbl4(t0);
                            c1 += to_dbl4(t1);
                        }
                    }
                    if (tval) {
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int k0 = acc_row<kF32>(r, kq);
                            if (k0 < nb) publish_pair(p.xchg1, bslot1(p, rnd, w, (uint32_t)(k0 * T + tl)), rnd, c0[r]);
                            if (16 + k0 < nb)
                                publish_pair(p.xchg1, bslot1(p, rnd, w, (uint32_t)((16 + k0) * T + tl)), rnd, c1[r]);
                        }
                    }
                }
                BRR_BTICK(1);
                if (wave < 3) {
                    const bool mi = wave == 2;        // row tile of u (A operand)
                    const bool mc = wave != 0;        // column tile of w (B operand)
                    if (16 * (int)mc < nb) {
                        const int ra = 16 * (int)mi + tq;
                        const double u0 = B.ut[ra][0], u1 = B.ut[ra][1], u2 = B.ut[ra][2], u3 = B.ut[ra][3];
                        const double b0 = mc ? w10 : w00, b1 = mc ? w11 : w01, b2 = mc ? w12 : w02,
                                     b3 = mc ? w13 : w03;
                        dbl4 g;
                        if constexpr (!kF32) {
                            dbl4 gg[2] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
#pragma unroll
                            for (int s = 0; s < NT; ++s) {
                                const uint32_t rowa = mi ? rw1[s < NR ? s : 0] : rw0[s < NR ? s : 0];
                                const uint32_t rowb = mc ? rw1[s < NR ? s : 0] : rw0[s < NR ? s : 0];
#pragma unroll
                                for (int r = 0; r < 4; ++r) {
                                    const uint32_t sh = 8u * (uint32_t)r + 2u * (uint32_t)kq;
                                    const uint32_t i = 16u * (uint32_t)s + 4u * (uint32_t)r + (uint32_t)kq;
                                    const double av = (i < nloc) ? wt_of((rowa >> sh) & 3u, u0, u1, u2, u3) : 0.0;
                                    gg[r & 1] = mfma16x16x4(av, wt_of((rowb >> sh) & 3u, b0, b1, b2, b3), gg[r & 1]);
                                }
                            }
                            g = gg[0] + gg[1];
                        } else {
                            // per tile an FP32 product, the tiles summed in FP64
                            float fu0 = (float)u0, fu1 = (float)u1, fu2 = (float)u2, fu3 = (float)u3;
                            float fb0 = (float)b0, fb1 = (float)b1, fb2 = (float)b2, fb3 = (float)b3;
                            // converted once (as for the dot tiles)
                            asm volatile("" : "+v"(fu0), "+v"(fu1), "+v"(fu2), "+v"(fu3));
                            asm volatile("" : "+v"(fb0), "+v"(fb1), "+v"(fb2), "+v"(fb3));
                            g = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
                            for (int s = 0; s < NT; ++s) {
                                // loaded without a branch (every row holds kMaxNT words), masked after
                                const uint32_t xa = *(const uint32_t*)&B.codes[16 * (int)mi + tq][4 * s];
                                const uint32_t xb = *(const uint32_t*)&B.codes[16 * (int)mc + tq][4 * s];
                                const uint32_t rowa = (s < nt) ? xa : 0u, rowb = (s < nt) ? xb : 0u;
                                flt4 gt = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                                for (int r = 0; r < 4; ++r) {
                                    const uint32_t sh = 8u * (uint32_t)r + 2u * (uint32_t)kq;
                                    const uint32_t i = 16u * (uint32_t)s + 4u * (uint32_t)r + (uint32_t)kq;
                                    const float av = (i < nloc) ? wt_of_f((rowa >> sh) & 3u, fu0, fu1, fu2, fu3) : 0.f;
                                    gt = mfma16x16x4(av, wt_of_f((rowb >> sh) & 3u, fb0, fb1, fb2, fb3), gt);
                                }
                                g += to_dbl4(gt);
                            }
                        }
                        const int CT = nb * T;
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int i = 16 * (int)mi + acc_row<kF32>(r, kq), k = 16 * (int)mc + tq;
                            if (i < k && k < nb)
                                publish_pair(p.xchg1, bslot1(p, rnd, w, (uint32_t)(CT + k * (k - 1) / 2 + i)), rnd, g[r]);
                        }
                    }
                }
            }
            BRR_BTICK(6);
            // the prefetched block goes to the other buffer (read after this block)
            if (has1) bstore(p, L.blk[cb ^ 1], x, nb1, tid);
            // u and z of every (slot, trait) pair, drawn while the partials travel
            for (int e = tid; e < nb * kMaxT; e += kThreads) {
                const int k = e >> 6, t = e & 63;
                if (t >= T) continue;
                const uint32_t j = B.mk[k];
                if (p.mode == kModeChain) {
                    const U4 xr = draw_block(p.seed, j, iteration, kStreamMarker, p.chain + (uint32_t)t);
                    L.uu[k][t] = unit_u53(xr.x[0], xr.x[1]);
                    L.zw[k][t] = ((uint64_t)xr.x[2] << 32) | xr.x[3];   // z formed when it is needed
                } else {
                    L.uu[k][t] = p.u[(size_t)j * p.T + t];
                    L.zw[k][t] = double_as_bits(p.z[(size_t)j * p.T + t]);
                }
            }
            BRR_BTICK(7);

            // ---- 2. exchange: every workgroup ends with the full C and G
            {
                const uint32_t CT = (uint32_t)nb * (uint32_t)T;
                const uint32_t V = CT + (uint32_t)(nb * (nb - 1) / 2);
                if (!breduce(p, L, rnd, V, w, tid)) goto done;
                BRR_BTICK(8);
                // totals v = tid + 256 i, four loads in flight per thread (one at a time
                // costs a full memory round trip per value)
                constexpr int kG = 4;
                const float invT = 1.0f / (float)T;
                for (uint32_t v0 = (uint32_t)tid; v0 < V; v0 += kG * kThreads) {
                    uint64_t hh[kG], ll[kG];
                    uint32_t live = 0;
#pragma unroll
                    for (int q = 0; q < kG; ++q) {
                        const uint32_t v = v0 + (uint32_t)q * kThreads;
                        hh[q] = 0;
                        ll[q] = 0;
                        if (v < V) {
                            ld_pair(p.xchg2, bslot2(p, rnd, v, w % kCopies), hh[q], ll[q]);
                            live |= 1u << q;
                        }
                    }
                    uint64_t spins = 0;
                    for (;;) {
                        uint32_t stale = 0;
#pragma unroll
                        for (int q = 0; q < kG; ++q)
                            if (((live >> q) & 1u) && ((uint32_t)(hh[q] >> 32) != rnd || (uint32_t)(ll[q] >> 32) != rnd))
                                stale |= 1u << q;
                        if (!stale) break;
                        if (++spins > kSpinLimit ||
                            ((spins & 255u) == 0 && __hip_atomic_load(p.ctl + kCtlAbort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                            L.abort = 1;
                            raise_abort(p, 13);
                            live = 0;
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
#pragma unroll
                        for (int q = 0; q < kG; ++q)
                            if ((stale >> q) & 1u) ld_pair(p.xchg2, bslot2(p, rnd, v0 + (uint32_t)q * kThreads, w % kCopies), hh[q], ll[q]);
                    }
#pragma unroll
                    for (int q = 0; q < kG; ++q) {
                        if (!((live >> q) & 1u)) continue;
                        const uint32_t v = v0 + (uint32_t)q * kThreads;
                        const double xv = bits_as_double((hh[q] << 32) | (ll[q] & 0xFFFFFFFFull));
                        if (v < CT) {
                            // k = v / T without a division: a float estimate, corrected
                            uint32_t k = (uint32_t)((float)v * invT);
                            if (k * (uint32_t)T > v) --k;
                            else if ((k + 1) * (uint32_t)T <= v) ++k;
                            L.cd[k][v - k * (uint32_t)T] = xv;
                        } else {
                            const uint32_t e = v - CT;
                            uint32_t k = (uint32_t)((1.0f + sqrtf(1.0f + 8.0f * (float)e)) * 0.5f);
                            while (k > 1 && k * (k - 1) / 2 > e) --k;
                            while ((k + 1) * k / 2 <= e) ++k;
                            L.gm[e - k * (k - 1) / 2][k] = xv;
                        }
                    }
                    if (!live && L.abort) break;
                }
                if (tid < kMaxT) { L.pos[tid] = (tid < T) ? 0 : nb; L.fc[tid] = nb; }
                if (tid < kBB / 4) L.flag[tid] = 0u;
                __syncthreads();
                if (L.abort) goto done;
            }
            BRR_BTICK(2);

            // ---- 3. decisions, wave q for its traits 16q .. 16q+15 (traits are
            // independent chains, so each wave resolves its tile without
            // workgroup barriers). Rounds: every pending (slot, trait) pair is
            // drawn in parallel; each trait accepts its pairs up to its first beta
            // change, the dots of its later slots are corrected by delta * Gram
            // row (BayesRBase.cpp:370-378) and those slots are drawn again.
            if (wave < (int)p.nq) {
                // the traits whose outputs of this block this workgroup writes: t with
                // (t + 7 blk) % P == w (one division per block, not one per trait)
                unsigned long long own = 0;
                {
                    const uint32_t sh = (blk * 7u) % p.P;
                    for (uint32_t t = w >= sh ? w - sh : w + p.P - sh; t < (uint32_t)T; t += p.P) own |= 1ull << t;
                }
                const int tt = 16 * wave + tq;             // trait of this lane's accept work
                const bool tok = tt < T;
                const bool mine = tok && ((own >> tt) & 1ull);   // output workgroup
                const bool single_out = p.T <= p.P;   // then a workgroup outputs at most one trait per block
                uint16_t* plist = L.plist + wave * (kBB * 16);
                for (bool first = true;; first = false) {
                    const int p0 = tok ? L.pos[tt] : nb;
                    const int cnt = nb - p0;
                    int incl = cnt;
#pragma unroll
                    for (int off = 1; off < 16; off <<= 1) {
                        const int y = __shfl_up(incl, off, 16);
                        if (tq >= off) incl += y;
                    }
                    const int total = __shfl(incl, 15, 64);
                    if (total == 0) break;
                    if (!first && kq == 0)
                        for (int i = 0; i < cnt; ++i) plist[incl - cnt + i] = (uint16_t)(((p0 + i) << 8) | tt);
                    wave_lds_sync();
                    const int span = first ? nb * 16 : total;
                    for (int e = lane; e < span; e += 64) {
                        int k, t;
                        if (first) {
                            k = e >> 4;
                            t = 16 * wave + (e & 15);
                            if (t >= T) continue;
                        } else {
                            const int kt = plist[e];
                            k = kt >> 8;
                            t = kt & 0xFF;
                        }
                        const double beta_old = B.bo[k][t];
                        const BDraw r = bdraw(L, t, K, bnum<kDenseMath>(L, B, k, t, dN, NM1), beta_old,
                                              L.uu[k][t], L.zw[k][t], p.mode == kModeChain);
                        L.bn[k][t] = r.bnew;
                        L.sl[k][t] = (int8_t)r.sel;
                        if ((own >> t) & 1ull) L.acm[k] = r.acum0;
                        if (r.bnew != beta_old) atomicMin(&L.fc[t], k);
                    }
                    ++n_drounds;
                    wave_lds_sync();
                    BRR_BTICK(3);
                    // accept the trait's slots up to its first change; correct later dots
                    if (tok) {
                        const int c = L.fc[tt];
                        const double dl = (c < nb) ? B.bo[c][tt] - L.bn[c][tt] : 0.0;
                        for (int k = p0 + kq; k < nb; k += 4) {
                            if (k <= c) {
                                const int sel = L.sl[k][tt];
                                if (sel >= 0) atomicAdd(&L.v[sel][tt], 1.0);
                                if (mine) {
                                    const size_t o = (size_t)B.mk[k] * p.T + (size_t)tt;
                                    const double ac = single_out ? L.acm[k]
                                                                 : bdraw(L, tt, K, bnum<kDenseMath>(L, B, k, tt, dN, NM1),
                                                                         B.bo[k][tt], 2.0, 0ull, false).acum0;
                                    // agent-scope stores: the same slot may be written by
                                    // another workgroup (another XCD's L2) in a later iteration
                                    st_agent((uint64_t*)(p.beta + o), double_as_bits(L.bn[k][tt]));
                                    if (sel >= 0)
                                        __hip_atomic_store(p.comp + o, sel, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                                    st_agent((uint64_t*)(p.acum + o), double_as_bits(ac));
                                }
                                L.cd[k][tt] = (k == c) ? dl : 0.0;
                                if (k == c) atomicOr(&L.flag[k >> 2], 1u << wave);
                            } else {
                                L.cd[k][tt] += dl * L.gm[c][k];
                            }
                        }
                    }
                    wave_lds_sync();
                    if (kq == 0 && tok) {
                        const int c = L.fc[tt];
                        if (c < nb) {
                            if (!dense) {
                                // epsilonSum (sparsebayesrkernel.cpp:29-34)
                                const double dBeta = B.bo[c][tt] - L.bn[c][tt];
                                const double mean = B.mstat[c][0], sd = B.mstat[c][1], Zsum = B.mstat[c][2];
                                L.epsSum[tt] += dBeta * Zsum / sd - dBeta * mean * dN / sd;
                            }
                            L.pos[tt] = c + 1;
                            ++n_changes;
                        } else {
                            L.pos[tt] = nb;
                        }
                        L.fc[tt] = nb;
                    }
                    wave_lds_sync();
                    BRR_BTICK(9);
                }
                // betasqnG += beta^2 in marker order (BayesRBase.cpp:353)
                if (kq == 0 && tok) {
                    double bs = L.betasqn[tt];
#pragma unroll 8
                    for (int k = 0; k < nb; ++k) {
                        const double b = L.bn[k][tt];
                        if (L.sl[k][tt] > 0) bs += b * b;
                    }
                    L.betasqn[tt] = bs;
                }
                // ---- 4. eps_t += sum_k delta_{k,t} u_k on the matrix cores, for the
                // 4-marker steps holding a change of this wave's traits
                uint32_t fmask = 0u;
#pragma unroll
                for (int p4 = 0; p4 < kBB / 4; ++p4)
                    if ((L.flag[p4] >> wave) & 1u) fmask |= 1u << p4;
                if (fmask) {
                    // two halves of the slice's tiles, so the per-step code rows and
                    // tables are loaded once per 4-marker step into registers
                    constexpr int HS = (NT + 1) / 2;
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        ET4 acc[HS];
#pragma unroll
                        for (int i2 = 0; i2 < HS; ++i2) acc[i2] = ET4{0, 0, 0, 0};
                        for (uint32_t m = fmask; m; m &= m - 1u) {
                            const int km = 4 * __builtin_ctz(m) + kq;   // < kBB: the rows read below exist
                            const bool kv = km < nb;
                            // straight-line loads: a row past the slice's tiles or past the block
                            // reads what the buffer holds and is masked to zero after the load
                            uint32_t row[HS];
#pragma unroll
                            for (int i2 = 0; i2 < HS; ++i2) {
                                const int s2 = h * HS + i2;
                                const uint32_t wrow = *(const uint32_t*)&B.codes[km][4 * (s2 < NT ? s2 : 0)];
                                row[i2] = (s2 < nt && kv) ? wrow : 0u;
                            }
                            ET u0 = (ET)B.ut[km][0], u1 = (ET)B.ut[km][1], u2 = (ET)B.ut[km][2],
                               u3 = (ET)B.ut[km][3];
                            // FP32 tiles: the four values converted once here (an empty asm keeps the
                            // compiler from selecting among the FP64 values and converting per tile)
                            if constexpr (kF32) asm volatile("" : "+v"(u0), "+v"(u1), "+v"(u2), "+v"(u3));
                            const ET cdv = (ET)L.cd[km][16 * wave + tq];
                            const ET bv = kv ? cdv : (ET)0;
                            // rows past the slice are masked to zero by i < nloc (a select, not a branch)
#pragma unroll
                            for (int i2 = 0; i2 < HS; ++i2) {
                                const int s2 = h * HS + i2;
                                if (s2 < NT) {
                                    const uint32_t i = 16u * (uint32_t)s2 + (uint32_t)tq;
                                    const ET wv = wt_t((row[i2] >> (2u * (uint32_t)tq)) & 3u, u0, u1, u2, u3);
                                    const ET av = (i < nloc) ? wv : (ET)0;
                                    acc[i2] = mfma16x16x4(av, bv, acc[i2]);
                                }
                            }
                        }
#pragma unroll
                        for (int i2 = 0; i2 < HS; ++i2) {
                            const int s2 = h * HS + i2;
                            if (s2 < NT) E[s2 < NT ? s2 : 0] += acc[i2];
                        }
                    }
                }
                BRR_BTICK(4);
            }
            __syncthreads();
            BRR_BTICK(10);

            // ---- block blk+1 becomes current; its successor's ids go to the freed buffer
            cb ^= 1;
            if (has2 && tid < nb2) L.blk[cb ^ 1].mk[tid] = jn2;
            __syncthreads();
            BRR_BTICK(0);
        }
        // other workgroups read this workgroup's beta stores only in the block
        // loads of the next iteration, which follow its next publish
        __builtin_amdgcn_s_waitcnt(0);

        if (p.mode == kModeChain) {
            // ---- sigmaE, sigmaG, pi draws per trait (BayesRBase.cpp:237-247)
            double part = 0.0;
#pragma unroll
            for (int s = 0; s < NT; ++s)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const double e = (double)E[s][r];
                    part += e * e;
                }
            part += __shfl_xor(part, 16, 64);
            part += __shfl_xor(part, 32, 64);
            ++rnd;
            ++n_rounds;
            if (kq == 0 && tval) publish_pair(p.xchg1, bslot1(p, rnd, w, (uint32_t)tl), rnd, part);
            if (!ballreduce_traits(p, L, rnd, w, tid)) goto done;
            if (tid < T) {
                const int t = tid;
                const uint32_t ch = p.chain + (uint32_t)t;
                const double e2 = L.tvec[t];
                L.sigmaE[t] = inv_scaled_chisq_rng(p.seed, iteration, kDrawSigmaE, ch, p.v0E + p.N,
                                                   (e2 + p.v0E * p.s02E) / (p.v0E + p.N));
                double vsum = 0.0;
                for (int k = 0; k < K; ++k) vsum += L.v[k][t];
                const double m0 = vsum - L.v[0][t];
                L.sigmaG[t] = inv_scaled_chisq_rng(p.seed, iteration, kDrawSigmaG0, ch, p.v0G + m0,
                                                   (L.betasqn[t] * m0 + p.v0G * p.s02G) / (p.v0G + m0));
                // Dirichlet(v + 1): draw ids 2 + G + k with one group
                double sum = 0.0;
                for (int k = 0; k < K; ++k) {
                    L.pi[k][t] = rgamma(p.seed, iteration, 3u + (uint32_t)k, ch, L.v[k][t] + 1.0, 1.0);
                    sum += L.pi[k][t];
                }
                for (int k = 0; k < K; ++k) L.pi[k][t] = L.pi[k][t] / sum;
                if (w == 0 && p.hist) {
                    double* h = p.hist + ((size_t)it * p.T + (size_t)t) * 5u;
                    h[0] = L.mu[t];
                    h[1] = L.sigmaE[t];
                    h[2] = m0;
                    h[3] = e2;
                    h[4] = L.sigmaG[t];
                }
            }
            __syncthreads();
            BRR_BTICK(5);
        }
    }

done:
    __syncthreads();
    if (w == 0) {
        // per-lane change counts (lanes kq == 0 own a trait) and per-wave round counts
        atomicAdd(&L.sum_changes, n_changes);
        if (lane == 0) atomicAdd(&L.sum_drounds, n_drounds);
    }
    __syncthreads();
    if (L.abort == 0) {
#pragma unroll
        for (int s = 0; s < NT; ++s)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t i = 16u * (uint32_t)s + (uint32_t)acc_row<kF32>(r, kq);
                if (s < nt && i < nloc && tval) p.eps[(size_t)(lo + i) * T + tl] = (double)E[s][r];
            }
        if (w == 0) {
            if (tid < T) {
                const int t = tid;
                double* sc = p.scal + (size_t)t * sls;
                sc[SL.mu()] = L.mu[t];
                sc[SL.sigmaE()] = L.sigmaE[t];
                sc[SL.epsSum()] = L.epsSum[t];
                sc[SL.sigmaG()] = L.sigmaG[t];
                sc[SL.betasqn()] = L.betasqn[t];
                for (int k = 0; k < K; ++k) {
                    sc[SL.pi() + k] = L.pi[k][t];
                    sc[SL.v() + k] = L.v[k][t];
                }
            }
            if (tid == 0) {
                __hip_atomic_store(p.ctl + kCtlRound, rnd, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                p.counters[0] = n_rounds;
                p.counters[1] = L.sum_changes;
                for (int q = 0; q < 6; ++q) p.counters[2 + q] = L.tk[q];
                p.counters[8] = L.sum_drounds;
                for (int q = 6; q < kBPhases; ++q) p.counters[4 + q] = L.tk[q];
            }
        }
    }
}

} // namespace brr
