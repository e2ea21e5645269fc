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
// markers are re-dotted in the next round. A marker already in the model
// changes for certain, so when it is the round's first such marker its Gram
// entries against the later markers ride along with the dots, and after its
// change the later markers are decided again within the round. One exchange
// round per block plus about one per other beta change replaces one grid
// barrier per marker.
//
// The genotype codes and per-marker constants of block b+1 are loaded into
// registers while block b waits on its exchanges and are written to LDS when
// block b ends, so HBM latency stays off the critical path (ragged lists need a
// second pipeline stage: their per-slice segment bounds are fetched one block
// earlier still).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "brr_rng.h"

namespace brr {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kMaxB = 64;      // markers per speculation block (one wave of decisions)
constexpr int kMaxF = 128;   // fixed-effect covariates per context
constexpr int kMaxK = 8;       // mixture components incl. zero
constexpr int kMaxG = 32;      // marker groups
constexpr int kMaxRuns = 4;    // slice <= 512 * kMaxRuns individuals
constexpr uint32_t kRun = 512; // individuals per lane-run pass: 64 lanes x 8
constexpr uint32_t kSliceAlign = 16;
constexpr uint64_t kSpinLimit = 1ull << 22;
constexpr uint32_t kMaxP = 256;   // workgroups of a launch (the exchange reducers read <= 4 per lane)
constexpr int kIdxCap = 32;    // prefetched ragged indices per thread per block

// kModeFixed: only the fixed-effects block of one iteration (brr_hip_fixed_effects).
// kModeProbe: `iterations` exchange rounds of a full block with nothing else
// (brr_hip_probe_exchange: the latency floor of a round at this launch shape).
enum : int { kModeSweep = 0, kModeChain = 1, kModeFixed = 2, kModeProbe = 3 };
// kLayoutEigen: the ragged formulas with EigenBayesRKernel's dot, whose
// missing calls count at the marker mean (eigenbayesrkernel.cpp:17-20).
enum : int { kLayoutDense = 1, kLayoutRagged = 2, kLayoutEigen = 3 };
constexpr int kPhases = 6;  // block load, dots, exchange, decisions, commit/update, iteration scalars

// ctl words (persist across launches)
enum : int { kCtlRound = 0, kCtlAbort = 1, kCtlAbortWhere = 2 };

// Phase timer of workgroup 0, thread 0 (s_memrealtime: 100 MHz constant clock).
#define BRR_TICK(slot)                                                           \
    do {                                                                         \
        if (tid == 0) {                                                          \
            const uint64_t t_now_ = __builtin_amdgcn_s_memrealtime();           \
            tk[(slot)] += t_now_ - t_last;                                       \
            t_last = t_now_;                                                     \
        }                                                                        \
    } while (0)

struct SweepParams {
    uint32_t N, M, K, G;
    uint32_t P, B, S;
    int layout, mode;
    uint32_t iterations, iter_base, chain;
    uint32_t async;               // 1: each block is one window of the asynchronous sweep
    // fixed effects (BayesRBase.cpp:212-235): F covariates, column-major N x F
    // floats (Data::X); state {sigmaF, gamma[F]} and the carried covariate order
    const float* fixedX;
    double* fixed;
    uint32_t* forder;
    uint32_t F;
    uint32_t refresh_blocks;      // FP32 residual: FP64 refresh after every this many blocks (>= 1)
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
    double* num_out;              // M or null: the num of every marker's decision (computeNum)
    const uint32_t* gate;         // null, or a word that must be 0 for the launch to run (else all exit at once)
    double* scal;                 // scalar block, see ScalarLayout
    double* hist;                 // iterations x (4 + G) or null
    uint64_t* xchg1;              // [2][kMaxB][P][2] tagged partials
    uint64_t* xchg2;              // [2][kMaxB][2] tagged totals
    uint32_t* ctl;                // persistent control words
    unsigned long long* counters; // [0] rounds, [1] changes, [2..7] phase ticks (100 MHz)
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
    double gram[kMaxB];                 // Gram entries against the round's first in-model slot
    // current block's per-marker constants
    double beta_old[kMaxB];
    double wt[kMaxB][4];                // column value per 2-bit code
    double mstat[kMaxB][3];             // ragged: mean, sd, Zsum
    double uu[kMaxB];
    uint64_t zw[kMaxB];                // z (bits)
    // chain: the next block's u and z, formed by wave 1 while wave 0 decides
    double uu_next[kMaxB];
    uint64_t zw_next[kMaxB];
    uint32_t mk[kMaxB];
    int32_t gr[kMaxB];
    uint32_t mk_next[kMaxB];            // marker ids of the next block
    uint32_t mk_next2[kMaxB];           // marker ids of the block after it
    // ragged: segments of the next block (bounds fetched a block ahead)
    uint64_t seg_start[3 * kMaxB];
    uint32_t seg_pref[3 * kMaxB + 1];   // exclusive prefix; [3 nb] = total
    double dl[kMaxB][4];                // residual change per 2-bit code of each applied slot
    // column values of the current block by pairs of codes: entry n of slot k is
    // (wt[k][n & 3], wt[k][n >> 2]), the values of two adjacent individuals whose
    // codes form the 4-bit field n; FP64, or FP32 in the FP32-residual mode
    union alignas(16) {
        double d[kMaxB][16][2];
        float f[kMaxB][16][2];
    } pt;
    unsigned long long cmask;           // the slots whose change is applied this round
    double mu, sigmaE, epsSum;
    double sigmaG[kMaxG];
    double pi[kMaxG * kMaxK];
    double v[kMaxG * kMaxK];
    double betasqn[kMaxG];
    double cva[kMaxG * (kMaxK - 1)];
    double gc[kMaxG * kMaxK];           // log pi_k - 0.5 log(logLScale S_k + 1)
    double gden[kMaxG * kMaxK];         // denom_k (index k >= 1)
    double gbsd[kMaxG * kMaxK];         // sqrt(sigmaE / denom_k)
    double gdg[kMaxG * kMaxK];          // logL_k - logL_0 = gdg_k + num^2 ghk_k
    double ghk[kMaxG * kMaxK];
    // fixed effects
    double gamma[kMaxF];
    uint32_t forder[kMaxF];
    double sigmaF;
    int change;                         // first applied slot of the round (nb: none)
    int change2;                        // second applied slot (nb: none)
    int next;                           // first undecided slot after the round
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

// The value held by lane (lane ^ kBit), moved without the LDS crossbar: DPP moves
// within rows of 16 lanes (quad permutations for 1 and 2, row rotations for 4
// and 8) and gfx950's row swaps for 16 and 32. permlane16_swap(x, x) leaves
// rows (0, 0, 2, 2) in its first result and (1, 1, 3, 3) in its second;
// permlane32_swap(x, x) halves (0, 0) and (1, 1); row_ror:n gives lane i the
// value of lane (i - n) mod 16 of its row. Every lane of the wave must be active
// (a DPP read of an inactive lane returns 0).
template <int kBit>
__device__ __forceinline__ uint32_t lane_xor_u32(uint32_t x, int lane)
{
    if constexpr (kBit == 1) {
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0xB1, 0xF, 0xF, true);    // quad_perm 1,0,3,2
    } else if constexpr (kBit == 2) {
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x4E, 0xF, 0xF, true);    // quad_perm 2,3,0,1
    } else if constexpr (kBit == 4) {
        const uint32_t dn = (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x124, 0xF, 0xF, true);   // row_ror:4
        const uint32_t up = (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x12C, 0xF, 0xF, true);   // row_ror:12
        return (lane & 4) ? dn : up;
    } else if constexpr (kBit == 8) {
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x128, 0xF, 0xF, true);   // row_ror:8
    } else if constexpr (kBit == 16) {
        const auto sw = __builtin_amdgcn_permlane16_swap(x, x, false, false);
        return (lane & 16) ? sw[0] : sw[1];
    } else {
        static_assert(kBit == 32, "lane_xor: 1, 2, 4, 8, 16 or 32");
        const auto sw = __builtin_amdgcn_permlane32_swap(x, x, false, false);
        return (lane & 32) ? sw[0] : sw[1];
    }
}
template <int kBit>
__device__ __forceinline__ float lane_xor(float x, int lane)
{
    return __uint_as_float(lane_xor_u32<kBit>(__float_as_uint(x), lane));
}
template <int kBit>
__device__ __forceinline__ double lane_xor(double x, int lane)
{
    const uint64_t b = double_as_bits(x);
    const uint64_t lo = lane_xor_u32<kBit>((uint32_t)b, lane);
    const uint64_t hi = lane_xor_u32<kBit>((uint32_t)(b >> 32), lane);
    return bits_as_double((hi << 32) | lo);
}

// The value of lane `src` (the same index in every lane) read into a scalar
// register: no LDS round trip, and inactive source lanes are read as well.
__device__ __forceinline__ int lane_read(int x, int src)
{
    return __builtin_amdgcn_readlane(x, src);
}
__device__ __forceinline__ double lane_read(double x, int src)
{
    const uint64_t b = double_as_bits(x);
    const uint64_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, src);
    const uint64_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), src);
    return bits_as_double((hi << 32) | lo);
}

// wave_sum's butterfly (the same additions in the same order) with lane_xor's
// moves: every lane of the wave must be active.
__device__ __forceinline__ double wave_sum_full(double x, int lane)
{
    x += lane_xor<32>(x, lane);
    x += lane_xor<16>(x, lane);
    x += lane_xor<8>(x, lane);
    x += lane_xor<4>(x, lane);
    x += lane_xor<2>(x, lane);
    x += lane_xor<1>(x, lane);
    return x;
}

// Sums 16 per-lane values over the 64 lanes of a wave at once: halving
// exchanges (xor 32, 16, 8, 4) leave lanes 4i..4i+3 with value i's partial
// over their quarter, then a butterfly over xor 2, 1 completes it. 17 shuffles
// instead of 16 x 6; the addition tree is fixed, so the result is the same in
// every workgroup.
// One halving stage with compile-time extents, so every array index is static
// (a runtime index would turn each access into a 16-way select chain).
template <int kHalf, int kBit, int kN>
__device__ __forceinline__ void transpose_stage(double (&v)[kN], int lane)
{
    const bool up = (lane & kBit) != 0;
#pragma unroll
    for (int i = 0; i < kHalf; ++i) {
        const double send = up ? v[i] : v[i + kHalf];
        const double recv = lane_xor<kBit>(send, lane);
        v[i] = (up ? v[i + kHalf] : v[i]) + recv;
    }
}

__device__ __forceinline__ double transpose_reduce16(double (&v)[16], int lane)
{
    transpose_stage<8, 32>(v, lane);
    transpose_stage<4, 16>(v, lane);
    transpose_stage<2, 8>(v, lane);
    transpose_stage<1, 4>(v, lane);
    double x = v[0];
    x += lane_xor<2>(x, lane);
    x += lane_xor<1>(x, lane);
    return x;   // total of value ((lane >> 2) & 15)
}

// The same for 8 values: three halving exchanges (xor 32, 16, 8), then a
// butterfly over xor 4, 2, 1; lanes 8i..8i+7 end with value i's total.
__device__ __forceinline__ double transpose_reduce8(double (&v)[8], int lane)
{
    transpose_stage<4, 32>(v, lane);
    transpose_stage<2, 16>(v, lane);
    transpose_stage<1, 8>(v, lane);
    double x = v[0];
    x += lane_xor<4>(x, lane);
    x += lane_xor<2>(x, lane);
    x += lane_xor<1>(x, lane);
    return x;   // total of value ((lane >> 3) & 7)
}

template <bool B>
struct BoolC {
    static constexpr bool value = B;
};
template <int N>
struct IntC {
    static constexpr int value = N;
};

// Value of a 2-bit code: two flat selects on 64-bit halves (no control flow).
__device__ __forceinline__ double wt_of(uint32_t v, double w0, double w1, double w2, double w3)
{
    const bool b0 = (v & 1u) != 0;
    const bool b1 = (v & 2u) != 0;
    const uint64_t lo = b0 ? double_as_bits(w1) : double_as_bits(w0);
    const uint64_t hi = b0 ? double_as_bits(w3) : double_as_bits(w2);
    return bits_as_double(b1 ? hi : lo);
}

// The same for an FP32 column table.
__device__ __forceinline__ float wt_of_f(uint32_t v, float w0, float w1, float w2, float w3)
{
    const bool b0 = (v & 1u) != 0;
    const bool b1 = (v & 2u) != 0;
    const float lo = b0 ? w1 : w0;
    const float hi = b0 ? w3 : w2;
    return b1 ? hi : lo;
}

__device__ __forceinline__ float wt_t(uint32_t v, float w0, float w1, float w2, float w3)
{
    return wt_of_f(v, w0, w1, w2, w3);
}
__device__ __forceinline__ double wt_t(uint32_t v, double w0, double w1, double w2, double w3)
{
    return wt_of(v, w0, w1, w2, w3);
}

__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

typedef double dbl2 __attribute__((ext_vector_type(2)));
typedef float flt2 __attribute__((ext_vector_type(2)));

// Knuth's TwoSum: s + e == a + b exactly (needs IEEE rounding and no
// contraction, which the build's -ffp-contract=off guarantees).
__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e)
{
    s = a + b;
    const float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// transpose_reduce16 on FP32 partials with compensated additions: every add of
// the exchange tree is a TwoSum whose rounding error is carried in a second
// FP32 word and exchanged alongside, so the wave total is hi + lo, accurate to
// about 2^-48 of the summed magnitudes instead of 2^-24 (Ogita, Rump & Oishi,
// "Accurate sum and dot product", SIAM J. Sci. Comput. 26, 2005: Sum2).
template <int kHalf, int kBit>
__device__ __forceinline__ void transpose_stage_c(float (&hi)[16], float (&lo)[16], int lane)
{
    const bool up = (lane & kBit) != 0;
#pragma unroll
    for (int i = 0; i < kHalf; ++i) {
        const float sh = up ? hi[i] : hi[i + kHalf];
        const float sl = up ? lo[i] : lo[i + kHalf];
        const float rh = lane_xor<kBit>(sh, lane);
        const float rl = lane_xor<kBit>(sl, lane);
        const float kh = up ? hi[i + kHalf] : hi[i];
        const float kl = up ? lo[i + kHalf] : lo[i];
        float s, e;
        two_sum(kh, rh, s, e);
        hi[i] = s;
        lo[i] = (kl + rl) + e;
    }
}

__device__ __forceinline__ double transpose_reduce16_c(float (&v)[16], int lane)
{
    float lo[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) lo[i] = 0.f;
    transpose_stage_c<8, 32>(v, lo, lane);
    transpose_stage_c<4, 16>(v, lo, lane);
    transpose_stage_c<2, 8>(v, lo, lane);
    transpose_stage_c<1, 4>(v, lo, lane);
    float h = v[0], l = lo[0];
    auto step = [&](float rh, float rl) {
        float s, e;
        two_sum(h, rh, s, e);
        h = s;
        l = (l + rl) + e;
    };
    step(lane_xor<2>(h, lane), lane_xor<2>(l, lane));
    step(lane_xor<1>(h, lane), lane_xor<1>(l, lane));
    return (double)h + (double)l;   // total of value ((lane >> 2) & 15)
}

// FP32-residual mode: the residual changes applied since the last FP64
// refresh, in the order they were applied (marker id, residual change per
// 2-bit code), and a staging area that hands wave 0's refreshed FP64 slice to
// the other waves.
constexpr int kMaxLog = 1024;
template <bool kF32, int R>
struct RefreshLog {
    uint32_t j[kMaxLog];
    double dl[kMaxLog][4];
    double ed[R * kRun];
    int n;
};
template <int R>
struct RefreshLog<false, R> {
    int n;
};

// A tagged pair is two 64-bit words (round tag << 32 | 32 bits of the value),
// written and read as one 16-byte buffer access with the cache policy of an
// agent-scope atomic (sc1). Each word carries the tag, so a reader that sees
// both tags current has both halves of this round's value.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
constexpr int kCpolDevice = 16;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t xbuf(const uint64_t* base)
{
    return __builtin_amdgcn_make_buffer_rsrc((void*)base, 0, 0x7FFFFFFF, 0x00020000);
}

__device__ __forceinline__ void publish_pair(const uint64_t* base, uint64_t* s, uint32_t rnd, double val)
{
    const uint64_t b = double_as_bits(val);
    const u32x4 v = {(uint32_t)(b >> 32), rnd, (uint32_t)b, rnd};
    __builtin_amdgcn_raw_buffer_store_b128(v, xbuf(base), (uint32_t)((const char*)s - (const char*)base), 0,
                                           kCpolDevice);
}

__device__ __forceinline__ void ld_pair(const uint64_t* base, const uint64_t* s, uint64_t& h, uint64_t& l)
{
    uint32_t off = (uint32_t)((const char*)s - (const char*)base);
    asm volatile("" : "+v"(off));   // opaque offset: every poll issues a fresh load
    const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(xbuf(base), off, 0, kCpolDevice);
    h = ((uint64_t)v.y << 32) | v.x;
    l = ((uint64_t)v.w << 32) | v.z;
}

// Returns false on timeout or abort.
__device__ __forceinline__ bool wait_pair(const uint64_t* base, const uint64_t* s, uint32_t rnd, const uint32_t* ctl,
                                          double* out)
{
    uint64_t h, l;
    ld_pair(base, s, h, l);
    uint64_t spins = 0;
    while ((uint32_t)(h >> 32) != rnd || (uint32_t)(l >> 32) != rnd) {
        if (++spins > kSpinLimit) return false;
        if ((spins & 255u) == 0 &&
            __hip_atomic_load(ctl + kCtlAbort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            return false;
        __builtin_amdgcn_s_sleep(1);
        ld_pair(base, s, h, l);
    }
    *out = bits_as_double((h << 32) | (l & 0xFFFFFFFFull));
    return true;
}

// Waits for the tagged pairs of workgroups w2 = first, first + stride, ... (< P,
// at most kPer of them) and adds them to *acc in that order (acc = ((acc +
// x_first) + x_first+stride) + ...). All of a thread's loads are in flight at
// once; only pairs whose tag is stale are polled again. `slot(w2)` gives the
// pair's address. False on timeout or abort.
template <int kPer, class SlotFn>
__device__ __forceinline__ bool wait_sum_strided(const uint64_t* base, SlotFn slot, uint32_t first, uint32_t stride, uint32_t P,
                                                 uint32_t rnd, const uint32_t* ctl, double* acc_io)
{
    uint64_t h[kPer], l[kPer];
    uint32_t live = 0;
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const uint32_t w2 = first + stride * q;
        h[q] = 0;
        l[q] = 0;
        if (w2 < P) {
            ld_pair(base, slot(w2), h[q], l[q]);
            live |= 1u << q;
        }
    }
    uint64_t spins = 0;
    for (;;) {
        uint32_t stale = 0;
#pragma unroll
        for (int q = 0; q < kPer; ++q)
            if (((live >> q) & 1u) && ((uint32_t)(h[q] >> 32) != rnd || (uint32_t)(l[q] >> 32) != rnd))
                stale |= 1u << q;
        if (!stale) break;
        if (++spins > kSpinLimit) return false;
        if ((spins & 255u) == 0 &&
            __hip_atomic_load(ctl + kCtlAbort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            return false;
        __builtin_amdgcn_s_sleep(1);
#pragma unroll
        for (int q = 0; q < kPer; ++q)
            if ((stale >> q) & 1u) ld_pair(base, slot(first + stride * q), h[q], l[q]);
    }
    double acc = *acc_io;
#pragma unroll
    for (int q = 0; q < kPer; ++q)
        if ((live >> q) & 1u) acc += bits_as_double((h[q] << 32) | (l[q] & 0xFFFFFFFFull));
    *acc_io = acc;
    return true;
}

// One lane's share of a P-workgroup reduction: workgroups lane, lane + 64, ...
template <class SlotFn>
__device__ __forceinline__ bool wait_sum_lanes(const uint64_t* base, SlotFn slot, uint32_t P, uint32_t rnd,
                                               const uint32_t* ctl, int lane, double* out)
{
    *out = 0.0;
    return wait_sum_strided<kMaxP / 64>(base, slot, (uint32_t)lane, 64u, P, rnd, ctl, out);
}

// The slot whose partial dot lane `lane` of wave `wave` publishes in a round
// starting at slot 0 (the transposed reduction leaves value (lane >> 2) & 15).
__device__ __forceinline__ int pos0_slot(int wave, int lane)
{
    return wave + kWaves * ((lane >> 2) & 15);
}

template <typename Params>
__device__ __forceinline__ void raise_abort(const Params& p, uint32_t where)
{
    __hip_atomic_store(p.ctl + kCtlAbort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(p.ctl + kCtlAbortWhere, where, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exchange slots of a round: k < kMaxB carries slot k's dot, kMaxB + k its
// Gram entry against the round's first in-model slot.
constexpr int kSlots = 2 * kMaxB;

__device__ __forceinline__ uint64_t* slot1(const SweepParams& p, uint32_t rnd, int s, uint32_t w)
{
    return p.xchg1 + ((((size_t)(rnd & 1u) * kSlots + (size_t)s) * p.P + w) * 2u);
}

// Each slot total is published in kCopies copies (copy c read by the
// workgroups w with w % kCopies == c, one XCD's worth under the round-robin
// dispatch), so that no single address is polled by every workgroup. Copies
// are kCopyWords apart, 256 bytes past a 4 KB multiple.
constexpr int kCopies = 8;
constexpr size_t kCopyWords = 2 * kSlots * 2 + 32;

__device__ __forceinline__ uint64_t* slot2(const SweepParams& p, uint32_t rnd, int s, uint32_t copy)
{
    return p.xchg2 + copy * kCopyWords + (((size_t)(rnd & 1u) * kSlots + (size_t)s) * 2u);
}

// Reduce-scatter / all-gather of the partials published at round rnd: the dots
// of slots [pos, nb) and the Gram entries of slots [gpos, nb). Slot k (both of
// its values) is reduced by workgroup k % P. A Gram value is published before
// the dot of its slot, so only the dots are polled; the Gram words are read
// once the dot has arrived. Afterwards L.dot[k] and L.gram[k] hold the same
// totals in every workgroup. Returns false (in all threads) on abort.
template <int R>
__device__ __forceinline__ bool exchange(const SweepParams& p, Shared<R>& L, uint32_t rnd, int pos, int nb,
                         uint32_t w, int wave, int lane, int gpos = kMaxB)
{
    bool ok = true;
    // the slots k in [pos, nb) with k % P == w in increasing order, the c-th of
    // them reduced by wave c % kWaves (one division per round: a division per
    // slot, a long scalar sequence, cost more than the rest of the exchange)
    const uint32_t r0 = (uint32_t)pos % p.P;
    int c = 0;
    for (int k = pos + (int)(w >= r0 ? w - r0 : w + p.P - r0); k < nb; k += (int)p.P, ++c) {
        if ((c % kWaves) != wave) continue;
        double acc = 0.0, gacc = 0.0;
        ok = wait_sum_lanes(p.xchg1, [&](uint32_t w2) { return slot1(p, rnd, k, w2); }, p.P, rnd, p.ctl, lane, &acc);
        if (ok && k >= gpos)
            ok = wait_sum_lanes(p.xchg1, [&](uint32_t w2) { return slot1(p, rnd, kMaxB + k, w2); }, p.P, rnd, p.ctl, lane,
                                &gacc);
        if (__any(!ok)) { ok = false; break; }
        acc = wave_sum_full(acc, lane);
        if (k >= gpos) {
            gacc = wave_sum_full(gacc, lane);
            if (lane < kCopies) publish_pair(p.xchg2, slot2(p, rnd, kMaxB + k, (uint32_t)lane), rnd, gacc);
        }
        if (lane < kCopies) publish_pair(p.xchg2, slot2(p, rnd, k, (uint32_t)lane), rnd, acc);
    }
    if (!ok) {
        if (lane == 0) { L.abort = 1; raise_abort(p, 1); }
    }
    const int k = pos + lane;
    if (wave == 0 && k < nb) {
        // the slot's dot and Gram totals are loaded together and only stale ones
        // polled again (one round trip when both have arrived)
        const uint32_t copy = w % kCopies;
        const uint32_t need = 1u | (k >= gpos ? 2u : 0u);
        uint64_t hh[2], ll[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            hh[q] = 0;
            ll[q] = 0;
            if ((need >> q) & 1u) ld_pair(p.xchg2, slot2(p, rnd, q * kMaxB + k, copy), hh[q], ll[q]);
        }
        uint64_t spins = 0;
        bool ok2 = true;
        for (;;) {
            uint32_t stale = 0;
#pragma unroll
            for (int q = 0; q < 2; ++q)
                if (((need >> q) & 1u) && ((uint32_t)(hh[q] >> 32) != rnd || (uint32_t)(ll[q] >> 32) != rnd))
                    stale |= 1u << q;
            if (!stale) break;
            if (++spins > kSpinLimit || ((spins & 255u) == 0 &&
                                         __hip_atomic_load(p.ctl + kCtlAbort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                ok2 = false;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
#pragma unroll
            for (int q = 0; q < 2; ++q)
                if ((stale >> q) & 1u) ld_pair(p.xchg2, slot2(p, rnd, q * kMaxB + k, copy), hh[q], ll[q]);
        }
        if (!ok2) {
            L.abort = 1;
            raise_abort(p, 2);
        } else {
            L.dot[k] = bits_as_double((hh[0] << 32) | (ll[0] & 0xFFFFFFFFull));
            L.gram[k] = (need & 2u) ? bits_as_double((hh[1] << 32) | (ll[1] & 0xFFFFFFFFull)) : 0.0;
        }
    }
    __syncthreads();
    return L.abort == 0;
}

// A scalar all-reduce through slot 0. Every wave holds the whole slice, so the
// slice partial is wave 0's sum alone.
template <int R>
__device__ __forceinline__ bool allreduce_scalar(const SweepParams& p, Shared<R>& L, uint32_t& rnd, double part,
                                 uint32_t w, int wave, int lane, double* total)
{
    part = wave_sum(part);
    ++rnd;
    if (threadIdx.x == 0) publish_pair(p.xchg1, slot1(p, rnd, 0, w), rnd, part);
    if (!exchange<R>(p, L, rnd, 0, 1, w, wave, lane)) return false;
    *total = L.dot[0];
    return true;
}

// The fixed-effects block of BayesRBase::runGibbs (BayesRBase.cpp:212-235) on
// this workgroup's slice; every workgroup draws the same gammas. The covariate
// order (carried across iterations like the reference's xI) takes the
// libstdc++ random_shuffle rule on its own stream; gamma_f is scalar draw
// 2 + G + G K + f. epsilonSum is left as is, as in the reference.
template <int R>
__device__ __forceinline__ bool fixed_effects(const SweepParams& p, Shared<R>& L, uint32_t& rnd,
                                              unsigned long long& n_rounds, double (&eps)[R][8],
                                              uint32_t iteration, uint32_t lo, uint32_t hi, uint32_t w,
                                              int wave, int lane)
{
    const int tid = threadIdx.x;
    if (tid == 0) {
        for (uint32_t i = 1; i < p.F; ++i) {
            const U4 x = draw_block(p.seed, i, iteration, kStreamFixed, p.chain);
            const uint64_t r = ((uint64_t)x.x[0] << 32) | x.x[1];
            const uint32_t j = (uint32_t)(r % ((uint64_t)i + 1));
            if (j != i) {
                const uint32_t t = L.forder[i];
                L.forder[i] = L.forder[j];
                L.forder[j] = t;
            }
        }
    }
    __syncthreads();
    const double dNm1 = (double)(p.N - 1u);
    const double sigE_sigF = L.sigmaE / L.sigmaF;
    const uint32_t draw0 = 2u + p.G + p.G * p.K;
    for (uint32_t i = 0; i < p.F; ++i) {
        const uint32_t f = L.forder[i];
        const float* x = p.fixedX + (size_t)f * p.N;
        const double gamma_old = L.gamma[f];
        double xv[R][8];
        double part = 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const uint32_t idx = lo + (uint32_t)r * kRun + 8u * (uint32_t)lane + (uint32_t)e;
                xv[r][e] = (idx < hi) ? (double)x[idx] : 0.0;
                part += xv[r][e] * (eps[r][e] + gamma_old * xv[r][e]);
            }
        double num;
        if (!allreduce_scalar<R>(p, L, rnd, part, w, wave, lane, &num)) return false;
        ++n_rounds;
        const double denom = dNm1 + sigE_sigF;
        const double g = norm_rng(p.seed, iteration, draw0 + f, p.chain, num / denom, L.sigmaE / denom);
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const uint32_t idx = lo + (uint32_t)r * kRun + 8u * (uint32_t)lane + (uint32_t)e;
                if (idx < hi) eps[r][e] = eps[r][e] + (gamma_old - g) * xv[r][e];
            }
        __syncthreads();
        if (tid == 0) L.gamma[f] = g;
        __syncthreads();
    }
    if (tid == 0) L.sigmaF = 1.0;   // s02F (BayesRBase.hpp:90)
    __syncthreads();
    return true;
}

template <int R>
__device__ __forceinline__ void group_constants(const SweepParams& p, Shared<R>& L)
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
            L.gdg[t] = 0.0;
            L.ghk[t] = 0.0;
        } else {
            const double S = L.cva[g * (K - 1) + (k - 1)];
            const double sigmaEOverSigmaG = sigmaE / sigmaG;
            const double den = NM1 + sigmaEOverSigmaG * (1.0 / S);
            const double logLScale = sigmaG / sigmaE * NM1;
            L.gden[t] = den;
            L.gc[t] = log(L.pi[t]) - 0.5 * log(logLScale * S + 1.0);
            L.gbsd[t] = __dsqrt_rn(sigmaE / den);
            L.gdg[t] = L.gc[t] - log(L.pi[g * K]);
            L.ghk[t] = 0.5 / (den * sigmaE);
        }
    }
}

// any_{i=1..K-1} |logL[i] - logL[r]| > 700 (BayesRBase.cpp:335, 360); loops run
// to the array size N >= K with a predicate so the arrays stay in registers
template <int N>
__device__ __forceinline__ bool guard700(const double (&logL)[N], int K, double ref)
{
    bool g = false;
#pragma unroll
    for (int i = 1; i < N; ++i)
        if (i < K) g = g || (fabs(logL[i] - ref) > 700.0);
    return g;
}

// 1 / sum_k exp(logL[k] - ref) in index order (BayesRBase.cpp:338, 363)
template <int N>
__device__ __forceinline__ double inv_sum_exp(const double (&logL)[N], int K, double ref)
{
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < N; ++k)
        if (k < K) s += exp(logL[k] - ref);
    return 1.0 / s;
}

// ------------------------------------------------------------ block loading --

// Per-marker constants of the next block, held in registers by thread t < nb.
struct MetaNext {
    uint32_t j;
    int32_t g;
    uint64_t beta_bits;
    double w0, w1, w2, w3;   // dense table by 2-bit code / ragged mean, sd, Zsum in w0..w2
    double u;
    uint64_t zw;             // z (chain mode, block loaded in place)
    double u_in, z_in;       // the given u and z (sweep mode)
};

// A marker's u and z in a chain (as marker_uz).
__device__ __forceinline__ void chain_draws(const SweepParams& p, uint32_t j, uint32_t iteration, double& u,
                                            uint64_t& zw)
{
    const U4 x = draw_block(p.seed, j, iteration, kStreamMarker, p.chain);
    u = unit_u53(x.x[0], x.x[1]);
    zw = double_as_bits(unit_normal(x.x[2], x.x[3]));
}

// With `draws` (a block loaded in place) the chain's draws are formed here,
// before any load is issued; a prefetched block's draws come from wave 1
// (L.uu_next / L.zw_next). The given u and z are loaded in every mode (p.u and
// p.z always point at M values), and the column table is chosen at compile
// time: a load in a branch not taken still counts as outstanding for the
// registers it would have written, and a temporary placed in one of them made
// every block start wait for all of its prefetch loads.
template <int kLay>
__device__ __forceinline__ void meta_issue(const SweepParams& p, uint32_t j, uint32_t iteration, bool draws,
                                           MetaNext& m)
{
    m.j = j;
    m.u = 0.0;
    m.zw = 0;
    if (draws && p.mode == kModeChain) chain_draws(p, j, iteration, m.u, m.zw);
    m.u_in = p.u[j];
    m.z_in = p.z[j];
    m.g = p.group ? p.group[j] : 0;
    m.beta_bits = ld_agent((const uint64_t*)(p.beta + j));
    if constexpr (kLay == kLayoutDense) {
        const double* lt = p.lut + 4 * (size_t)j;
        // 2-bit value v -> dosage: 0->2, 1->missing, 2->1, 3->0
        m.w0 = lt[2];
        m.w1 = lt[3];
        m.w2 = lt[1];
        m.w3 = lt[0];
    } else {
        const double* st = p.stats + 4 * (size_t)j;
        m.w0 = st[0];
        m.w1 = st[1];
        m.w2 = st[3];
        m.w3 = 0.0;
    }
}

template <int R, int kLay>
__device__ __forceinline__ void meta_store(const SweepParams& p, Shared<R>& L, int t, const MetaNext& m,
                                           bool draws)
{
    L.mk[t] = m.j;
    L.gr[t] = m.g;
    L.beta_old[t] = bits_as_double(m.beta_bits);
    if (p.mode != kModeChain) {
        L.uu[t] = m.u_in;
        L.zw[t] = double_as_bits(m.z_in);
    } else if (draws) {
        L.uu[t] = m.u;
        L.zw[t] = m.zw;
    } else {
        L.uu[t] = L.uu_next[t];
        L.zw[t] = L.zw_next[t];
    }
    // a run-time branch on purpose: with a compile-time one the compiler sinks
    // the prefetch loads of meta_issue towards this commit, and the block then
    // waits for their full latency
    if (p.layout == kLayoutDense) {
        L.wt[t][0] = m.w0;
        L.wt[t][1] = m.w1;
        L.wt[t][2] = m.w2;
        L.wt[t][3] = m.w3;
    } else {
        L.mstat[t][0] = m.w0;
        L.mstat[t][1] = m.w1;
        L.mstat[t][2] = m.w2;
        L.wt[t][0] = 2.0;   // genotype 2
        // missing: absent from the ragged dot, at the marker mean in EigenBayesRKernel's
        L.wt[t][1] = (kLay == kLayoutEigen) ? m.w0 : 0.0;
        L.wt[t][2] = 1.0;   // genotype 1
        L.wt[t][3] = 0.0;   // genotype 0
    }
}

// The pair table of the current block's slots [0, nb) from L.wt (all threads;
// the caller synchronises before the table is read).
template <bool kF32, int R>
__device__ __forceinline__ void pair_table_store(Shared<R>& L, int nb, int tid)
{
    for (int t = tid; t < nb * 16; t += kThreads) {
        const int k = t >> 4, n = t & 15;
        const double a = L.wt[k][n & 3], b = L.wt[k][n >> 2];
        if constexpr (kF32) {
            L.pt.f[k][n][0] = (float)a;
            L.pt.f[k][n][1] = (float)b;
        } else {
            L.pt.d[k][n][0] = a;
            L.pt.d[k][n][1] = b;
        }
    }
}

// The column values of the two individuals whose codes form field n, slot k.
template <int R>
__device__ __forceinline__ dbl2 pair_at(const Shared<R>& L, int k, uint32_t n, double)
{
    return *reinterpret_cast<const dbl2*>(&L.pt.d[k][n][0]);
}
template <int R>
__device__ __forceinline__ flt2 pair_at(const Shared<R>& L, int k, uint32_t n, float)
{
    return *reinterpret_cast<const flt2*>(&L.pt.f[k][n][0]);
}

// Dense codes of the next block: word t of the flattened [nb][R][32] array.
template <int R>
__device__ __forceinline__ void dense_issue(const SweepParams& p, const Shared<R>& L, int nb, uint32_t lo,
                                            int tid, uint32_t (&words)[8 * R])
{
    // word t = u8 * 256 + tid is word rem = tid % (32 R) of slot k = t / (32 R): the
    // same word of every slot for one thread (256 is a multiple of 32 R). The marker
    // ids are read first and the loads issued without branches (an invalid word
    // reads word 0; dense_store drops it), so the LDS reads and the loads go out
    // together and nothing waits for a load before the block's commit.
    constexpr uint32_t per = (uint32_t)R * 32u;
    constexpr uint32_t kStep = kThreads / per;
    const uint32_t rem = (uint32_t)tid % per;
    const uint32_t k0 = (uint32_t)tid / per;
    const bool inb = (uint64_t)(lo >> 2) + 4ull * rem < p.col_bytes;
    const size_t off = (size_t)(lo >> 4) + rem;
    uint32_t jj[8 * R];
#pragma unroll
    for (int u8 = 0; u8 < 8 * R; ++u8) {
        const uint32_t k = k0 + (uint32_t)u8 * kStep;
        jj[u8] = L.mk_next[k < (uint32_t)nb ? k : 0u];
    }
#pragma unroll
    for (int u8 = 0; u8 < 8 * R; ++u8) {
        const bool ok = k0 + (uint32_t)u8 * kStep < (uint32_t)nb && inb;
        words[u8] = p.dense[ok ? (size_t)jj[u8] * p.col_words + off : 0];
    }
}

template <int R>
__device__ __forceinline__ void dense_store(const SweepParams& p, Shared<R>& L, int nb, uint32_t lo, int tid,
                                            const uint32_t (&words)[8 * R])
{
    const uint32_t total = (uint32_t)nb * (uint32_t)R * 32u;
    const uint32_t rem = (uint32_t)tid % ((uint32_t)R * 32u);
    const bool inb = (uint64_t)(lo >> 2) + 4ull * rem < p.col_bytes;   // past the column: zero
#pragma unroll
    for (int u8 = 0; u8 < 8 * R; ++u8) {
        const uint32_t t = (uint32_t)u8 * kThreads + (uint32_t)tid;
        if (t < total) (&L.codes[0][0][0])[t] = inb ? words[u8] : 0u;
    }
}

// Ragged stage 1: segment bounds of (marker, list) pairs of the next block.
__device__ __forceinline__ void seg_issue(const SweepParams& p, uint32_t j, int list, uint32_t w,
                                          uint64_t& start, uint32_t& len)
{
    const uint32_t* sg = p.seg + ((size_t)j * 3u + (uint32_t)list) * (p.P + 1u);
    const uint32_t s0 = sg[w], s1 = sg[w + 1];
    uint64_t off;
    if (list == 0) off = p.ones_off[j];
    else if (list == 1) off = p.twos_off[j];
    else off = p.miss_off[j];
    start = off + s0;
    len = s1 - s0;
}

// Stores the stage-1 results and forms the exclusive prefix (all threads call).
template <int R>
__device__ __forceinline__ void seg_store_prefix(Shared<R>& L, int nb, int tid, int wave, int lane,
                                                 uint64_t start, uint32_t len)
{
    if (tid < 3 * nb) {
        L.seg_start[tid] = start;
        L.seg_pref[tid] = len;     // lengths for now
    }
    __syncthreads();
    if (wave == 0) {
        uint32_t a = 0, b = 0, c = 0;
        if (lane < nb) { a = L.seg_pref[3 * lane]; b = L.seg_pref[3 * lane + 1]; c = L.seg_pref[3 * lane + 2]; }
        const uint32_t x = a + b + c;
        uint32_t incl = x;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(incl, off, 64);
            if (lane >= off) incl += y;
        }
        const uint32_t excl = incl - x;
        if (lane < nb) {
            L.seg_pref[3 * lane] = excl;
            L.seg_pref[3 * lane + 1] = excl + a;
            L.seg_pref[3 * lane + 2] = excl + a + b;
        }
        if (lane == 63) L.seg_pref[3 * nb] = incl;
    }
    __syncthreads();
}

template <int R>
__device__ __forceinline__ int seg_item_of(const Shared<R>& L, int items, uint32_t q)
{
    int a = 0, b = items - 1;           // last item whose prefix <= q
    while (a < b) {
        const int mid = (a + b + 1) >> 1;
        if (L.seg_pref[mid] <= q) a = mid; else b = mid - 1;
    }
    return a;
}

// One ragged list entry (explicit branches keep the kernel-argument struct in
// scalar registers; a select between its fields becomes a dynamic offset into it)
__device__ __forceinline__ int32_t list_entry(const SweepParams& p, int list, uint64_t at)
{
    if (list == 0) return p.ones[at];
    if (list == 1) return p.twos[at];
    return p.miss[at];
}

// Ragged stage 2: list entries q = u8 * 256 + tid (u8 < kIdxCap) of the next block.
template <int R>
__device__ __forceinline__ void idx_issue(const SweepParams& p, const Shared<R>& L, int nb, int tid,
                                          int32_t (&idx)[kIdxCap], uint32_t (&itm)[kIdxCap / 4])
{
    const uint32_t total = L.seg_pref[3 * nb];
    const int items = 3 * nb;
#pragma unroll
    for (int u4 = 0; u4 < kIdxCap / 4; ++u4) itm[u4] = 0xFFFFFFFFu;
#pragma unroll
    for (int u8 = 0; u8 < kIdxCap; ++u8) {
        const uint32_t q = (uint32_t)u8 * kThreads + (uint32_t)tid;
        idx[u8] = 0;
        if (q < total) {
            const int a = seg_item_of(L, items, q);
            idx[u8] = list_entry(p, a % 3, L.seg_start[a] + (q - L.seg_pref[a]));
            itm[u8 >> 2] = (itm[u8 >> 2] & ~(0xFFu << (8 * (u8 & 3)))) | ((uint32_t)a << (8 * (u8 & 3)));
        }
    }
}

template <int R>
__device__ __forceinline__ void code_clear(Shared<R>& L, int k, uint32_t loc, int list)
{
    // ones: code 3 -> 2 (clear bit 0); twos: -> 0 (clear both); missing: -> 1 (clear bit 1)
    const uint32_t clr = (list == 0) ? 1u : (list == 1) ? 3u : 2u;
    const uint32_t r = loc >> 9, ln = (loc >> 3) & 63u, e = loc & 7u;
    const uint32_t bit = 16u * (ln & 1u) + 2u * e;
    atomicAnd(&L.codes[k][r][ln >> 1], ~(clr << bit));
}

// Builds the code maps of the block whose segments are in L.seg_*: the
// prefetched entries first, then any beyond the per-thread cap.
template <int R>
__device__ __forceinline__ void ragged_store(const SweepParams& p, Shared<R>& L, int nb, uint32_t lo,
                                             int tid, const int32_t (&idx)[kIdxCap],
                                             const uint32_t (&itm)[kIdxCap / 4])
{
    const uint32_t per = (uint32_t)R * 32u;
    for (uint32_t t = tid; t < (uint32_t)nb * per; t += kThreads) (&L.codes[0][0][0])[t] = 0xFFFFFFFFu;
    __syncthreads();
    const uint32_t total = L.seg_pref[3 * nb];
#pragma unroll
    for (int u8 = 0; u8 < kIdxCap; ++u8) {
        const uint32_t a = (itm[u8 >> 2] >> (8 * (u8 & 3))) & 0xFFu;
        if (a == 0xFFu) continue;
        code_clear(L, (int)(a / 3), (uint32_t)idx[u8] - lo, (int)(a % 3));
    }
    const int items = 3 * nb;
    for (uint32_t q = (uint32_t)kIdxCap * kThreads + (uint32_t)tid; q < total; q += kThreads) {
        const int a = seg_item_of(L, items, q);
        const int list = a % 3;
        code_clear(L, a / 3, (uint32_t)list_entry(p, list, L.seg_start[a] + (q - L.seg_pref[a])) - lo, list);
    }
}

// -------------------------------------------------------------------- kernel --

// kLay: kLayoutDense (DenseBayesRRmz formulas), kLayoutRagged or kLayoutEigen
// (SparseBayesRRG formulas with RaggedBayesRKernel's / EigenBayesRKernel's dot).
// kCodes: columns stored as 2-bit codes (else as ragged index lists).
// kAsync: each block is one window of the asynchronous sweep.
// kF32: the residual slice is held in FP32 and the dots run in FP32 with
// compensated wave reductions; an FP64 copy of the residual (p.eps) is brought
// up to date every p.refresh_blocks blocks and at every iteration boundary by
// replaying the logged residual changes in FP64, and the FP32 copy is then
// re-rounded from it. The decisions, the scalar draws, updateMu, the fixed
// effects and |eps|^2 are FP64 throughout.
// The sweep of one context by one of its P workgroups (w), with the LDS of
// the workgroup passed in: the body of bayesr_sweep_kernel and of the cohort
// kernel below, which runs several contexts side by side in one launch.
template <int R, int kLay, bool kCodes, bool kAsync, bool kF32>
__device__ __forceinline__ void sweep_body(const SweepParams& p, Shared<R>& L, RefreshLog<kF32, R>& RL,
                                           const uint32_t w)
{
    static_assert(!kF32 || kCodes, "the FP32 residual mode reads its columns as 2-bit codes");
    using E = typename std::conditional<kF32, float, double>::type;
    if (p.gate && __hip_atomic_load(p.gate, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) return;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const uint32_t lo = w * p.S;
    const uint32_t hi = min(lo + p.S, p.N);
    const int K = (int)p.K;
    const ScalarLayout SL{p.G, p.K};
    const double dN = (double)p.N;
    const double NM1 = (double)(p.N - 1u);
    constexpr bool dense = kLay == kLayoutDense;

    // eps slice in registers: individual lo + 512 r + 8 lane + e
    E eps[R][8];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const uint32_t idx = lo + (uint32_t)r * kRun + 8u * (uint32_t)lane + (uint32_t)e;
            eps[r][e] = (idx < hi) ? (E)p.eps[idx] : (E)0;
        }
    if (tid == 0) RL.n = 0;

    // Runs body(ed) on the FP64 residual slice ed[R][8] held by every wave.
    // FP64 mode: ed is the register slice itself. FP32 mode: wave 0 loads the
    // FP64 residual of the slice, replays the logged changes on it in order
    // (the same additions, in the same order, as the FP64 kernel makes), and
    // hands it to the other waves through LDS; after the body wave 0 stores it
    // back and every wave re-rounds its FP32 slice from it.
    auto fp64_region = [&](auto&& body) -> bool {
        if constexpr (!kF32) {
            return body(eps);
        } else {
            __syncthreads();
            if (RL.n > kMaxLog) {          // cannot happen: refresh_blocks * B <= kMaxLog
                if (tid == 0) { L.abort = 1; raise_abort(p, 40); }
                __syncthreads();
                return false;
            }
            double ed[R][8];
            if (wave == 0) {
#pragma unroll
                for (int r = 0; r < R; ++r)
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        const uint32_t idx = lo + (uint32_t)r * kRun + 8u * (uint32_t)lane + (uint32_t)e;
                        ed[r][e] = (idx < hi) ? p.eps[idx] : 0.0;
                    }
                const int n = RL.n;
                constexpr int kChunk = 8;
                for (int n0 = 0; n0 < n; n0 += kChunk) {
                    uint32_t wd[kChunk][R];
#pragma unroll
                    for (int q = 0; q < kChunk; ++q) {
                        const uint32_t j = (n0 + q < n) ? RL.j[n0 + q] : 0u;
#pragma unroll
                        for (int r = 0; r < R; ++r) {
                            const uint32_t rem = (uint32_t)r * 32u + (uint32_t)(lane >> 1);
                            const uint64_t byte_off = (uint64_t)(lo >> 2) + 4ull * rem;
                            wd[q][r] = (n0 + q < n && byte_off < p.col_bytes)
                                           ? p.dense[(size_t)j * p.col_words + (lo >> 4) + rem] : 0u;
                        }
                    }
#pragma unroll
                    for (int q = 0; q < kChunk; ++q) {
                        if (n0 + q >= n) break;
                        const double d0 = RL.dl[n0 + q][0], d1 = RL.dl[n0 + q][1];
                        const double d2 = RL.dl[n0 + q][2], d3 = RL.dl[n0 + q][3];
#pragma unroll
                        for (int r = 0; r < R; ++r) {
                            const uint32_t bits = wd[q][r] >> (16 * (lane & 1));
#pragma unroll
                            for (int e = 0; e < 8; ++e) {
                                const uint32_t idx = lo + (uint32_t)r * kRun + 8u * (uint32_t)lane + (uint32_t)e;
                                if (idx < hi) ed[r][e] = ed[r][e] + wt_of((bits >> (2 * e)) & 3u, d0, d1, d2, d3);
                            }
                        }
                    }
                }
#pragma unroll
                for (int r = 0; r < R; ++r)
#pragma unroll
                    for (int e = 0; e < 8; ++e) RL.ed[r * kRun + 8 * lane + e] = ed[r][e];
            }
            __syncthreads();
            if (wave != 0) {
#pragma unroll
                for (int r = 0; r < R; ++r)
#pragma unroll
                    for (int e = 0; e < 8; ++e) ed[r][e] = RL.ed[r * kRun + 8 * lane + e];
            }
            __syncthreads();
            if (tid == 0) RL.n = 0;
            const bool ok = body(ed);
            if (ok && wave == 0) {
#pragma unroll
                for (int r = 0; r < R; ++r)
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        const uint32_t idx = lo + (uint32_t)r * kRun + 8u * (uint32_t)lane + (uint32_t)e;
                        if (idx < hi) p.eps[idx] = ed[r][e];
                    }
            }
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int e = 0; e < 8; ++e) eps[r][e] = (float)ed[r][e];
            return ok;
        }
    };

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
    if (tid == 0) L.sigmaF = p.F ? p.fixed[0] : 1.0;
    for (uint32_t t = tid; t < p.F; t += kThreads) {
        L.gamma[t] = p.fixed[1 + t];
        L.forder[t] = p.forder[t];
    }
    uint32_t rnd = __hip_atomic_load(p.ctl + kCtlRound, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned long long n_rounds = 0, n_changes = 0;
    uint64_t tk[kPhases] = {0, 0, 0, 0, 0, 0};
    uint64_t t_last = __builtin_amdgcn_s_memrealtime();
    __syncthreads();

    if (p.mode == kModeProbe) {
        // every slot of a block publishes its (zero) partial dot, then the
        // reduce-scatter / all-gather of a sweep round; no state is touched
        const int nb = (int)min(p.B, p.M);
        for (uint32_t q = 0; q < p.iterations; ++q) {
            ++rnd;
            ++n_rounds;
            const int k = pos0_slot(wave, lane);
            if ((lane & 3) == 0 && k < nb) publish_pair(p.xchg1, slot1(p, rnd, k, w), rnd, 0.0);
            if (!exchange<R>(p, L, rnd, 0, nb, w, wave, lane)) return;   // abort raised: nothing to store
        }
        BRR_TICK(2);
        if (w == 0 && tid == 0) {
            __hip_atomic_store(p.ctl + kCtlRound, rnd, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            p.counters[0] = n_rounds;
            for (int q = 0; q < kPhases; ++q) p.counters[2 + q] = tk[q];
        }
        return;
    }

    const uint32_t iters = (p.mode == kModeChain) ? p.iterations : 1u;
    const uint32_t nblocks = (p.M + p.B - 1) / p.B;
    for (uint32_t it = 0; it < iters; ++it) {
        const uint32_t iteration = p.iter_base + it;
        const uint32_t* order = p.order + (size_t)((p.mode == kModeChain) ? it : 0u) * p.M;
        if (p.mode == kModeFixed) {
            if (!fp64_region([&](double (&ed)[R][8]) {
                    return fixed_effects<R>(p, L, rnd, n_rounds, ed, iteration, lo, hi, w, wave, lane);
                }))
                goto done;
            break;
        }

        if (p.mode == kModeChain) {
            // ---- updateMu (DenseBayesRRmz.cpp:36-42 / SparseBayesRRG.cpp:106-113)
            const double old_mu = L.mu;
            double new_mu = 0.0;
            const bool ok = fp64_region([&](double (&ed)[R][8]) {
#pragma unroll
                for (int r = 0; r < R; ++r)
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        const uint32_t idx = lo + (uint32_t)r * kRun + 8u * (uint32_t)lane + (uint32_t)e;
                        if (idx < hi) ed[r][e] = ed[r][e] + old_mu;
                    }
                double mean_for_mu;
                if (dense) {
                    double part = 0.0;
#pragma unroll
                    for (int r = 0; r < R; ++r)
#pragma unroll
                        for (int e = 0; e < 8; ++e) part += ed[r][e];
                    double total;
                    if (!allreduce_scalar<R>(p, L, rnd, part, w, wave, lane, &total)) return false;
                    ++n_rounds;
                    mean_for_mu = total / dN;
                } else {
                    __syncthreads();
                    if (tid == 0) L.epsSum += old_mu * dN;
                    __syncthreads();
                    mean_for_mu = L.epsSum / dN;
                }
                new_mu = norm_rng(p.seed, iteration, kDrawMu, p.chain, mean_for_mu, L.sigmaE / dN);
#pragma unroll
                for (int r = 0; r < R; ++r)
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        const uint32_t idx = lo + (uint32_t)r * kRun + 8u * (uint32_t)lane + (uint32_t)e;
                        if (idx < hi) ed[r][e] = ed[r][e] - new_mu;
                    }
                return true;
            });
            if (!ok) goto done;
            __syncthreads();
            if (tid == 0) {
                L.mu = new_mu;
                if (!dense) L.epsSum -= new_mu * dN;
            }
            for (uint32_t t = tid; t < p.G; t += kThreads) L.betasqn[t] = 0.0;
            for (uint32_t t = tid; t < p.G * p.K; t += kThreads) L.v[t] = 0.0;
            __syncthreads();
        }
        group_constants<R>(p, L);
        __syncthreads();

        // ---- block 0 of this iteration, loaded in place; the ids (and for
        // ragged the segments) of block 1 staged for the prefetch pipeline
        {
            const int nb0 = (int)min(p.B, p.M);
            if (tid < nb0) L.mk_next[tid] = order[tid];
            __syncthreads();
            MetaNext m{};
            if (tid < nb0) meta_issue<kLay>(p, L.mk_next[tid], iteration, true, m);
            if constexpr (kCodes) {
                uint32_t words[8 * R];
                dense_issue<R>(p, L, nb0, lo, tid, words);
                dense_store<R>(p, L, nb0, lo, tid, words);
            } else {
                uint64_t st = 0;
                uint32_t ln = 0;
                if (tid < 3 * nb0) seg_issue(p, L.mk_next[tid / 3], tid % 3, w, st, ln);
                seg_store_prefix<R>(L, nb0, tid, wave, lane, st, ln);
                int32_t idx[kIdxCap];
                uint32_t itm[kIdxCap / 4];
                idx_issue<R>(p, L, nb0, tid, idx, itm);
                ragged_store<R>(p, L, nb0, lo, tid, idx, itm);
            }
            if (tid < nb0) meta_store<R, kLay>(p, L, tid, m, true);
            __syncthreads();
            pair_table_store<kF32, R>(L, nb0, tid);
            __syncthreads();
            if (nblocks > 1) {
                const int nb1 = (int)min(p.B, p.M - p.B);
                if (tid < nb1) L.mk_next[tid] = order[p.B + tid];
                __syncthreads();
                if constexpr (!kCodes) {
                    uint64_t st = 0;
                    uint32_t ln = 0;
                    if (tid < 3 * nb1) seg_issue(p, L.mk_next[tid / 3], tid % 3, w, st, ln);
                    seg_store_prefix<R>(L, nb1, tid, wave, lane, st, ln);
                }
            }
        }
        BRR_TICK(5);

        // ---- the sweep (AnalysisGraph::exec over the order)
        for (uint32_t blk = 0; blk < nblocks; ++blk) {
            const uint32_t b0 = blk * p.B;
            const int nb = (int)min(p.B, p.M - b0);
            const bool has1 = blk + 1 < nblocks;
            const int nb1 = has1 ? (int)min(p.B, p.M - b0 - p.B) : 0;
            const bool has2 = blk + 2 < nblocks;
            const int nb2 = has2 ? (int)min(p.B, p.M - b0 - 2 * p.B) : 0;

            // prefetch issue: everything of block blk+1 whose address is known
            // (ids in L.mk_next, ragged segments in L.seg_*), and the ids of blk+2
            // The ids of blk+2 are loaded first and without a branch (a zeroed register
            // that a load may still be writing would make the block wait for every
            // load issued before it, the prefetch included); only tid < nb2 use them.
            const uint32_t jnext2 = order[has2 && tid < nb2 ? b0 + 2 * p.B + tid : 0u];
            MetaNext mnext{};
            uint64_t st2 = 0;
            uint32_t ln2 = 0;
            uint32_t words[kCodes ? 8 * R : 1];
            int32_t idx[kCodes ? 1 : kIdxCap];
            uint32_t itm[kCodes ? 1 : kIdxCap / 4];
            if (has1) {
                if (tid < nb1) meta_issue<kLay>(p, L.mk_next[tid], iteration, false, mnext);
                if constexpr (kCodes) dense_issue<R>(p, L, nb1, lo, tid, words);
                else idx_issue<R>(p, L, nb1, tid, idx, itm);
            }
            BRR_TICK(0);

            int pos = 0;
            bool first_round = true;
            while (pos < nb) {
                // ---- partial dots of slots [pos, nb) over this slice: wave q
                // handles slots pos + q + 4 i, reduced together
                ++rnd;
                ++n_rounds;
                // the round's first slot already in the model (beta_old != 0) changes
                // for certain, so the sequential sweep also sums its Gram entries
                // G[k] = sum_i u_h(x_hi) w_k(x_ki) against the later slots k, with
                // u_h the residual change per unit dBeta: when slot h is the round's
                // first change, dot_k + dBeta_h G[k] is slot k's dot after that change
                int h = nb;
                if constexpr (!kAsync) {
                    const unsigned long long im =
                        __ballot(pos + lane < nb && L.beta_old[pos + lane < nb ? pos + lane : 0] != 0.0);
                    if (im) h = pos + (int)__ffsll((long long)im) - 1;
                }
                const int gpos = h + 1 < nb ? h + 1 : kMaxB;
                auto dots = [&](auto gram) {
                    constexpr bool kG = decltype(gram)::value;
                    E uh[kG ? R : 1][8];
                    if constexpr (kG) {
                        double u0, u1, u2, u3;
                        if (dense) {
                            u0 = L.wt[h][0]; u1 = L.wt[h][1]; u2 = L.wt[h][2]; u3 = L.wt[h][3];
                        } else {
                            // raggedbayesrkernel.cpp:10-30 per unit dBeta
                            const double mean = L.mstat[h][0], sd = L.mstat[h][1];
                            u0 = (2.0 - mean) / sd; u1 = 0.0; u2 = (1.0 - mean) / sd; u3 = -mean / sd;
                        }
#pragma unroll
                        for (int r = 0; r < R; ++r) {
                            const uint32_t bits = L.codes[h][r][lane >> 1] >> (16 * (lane & 1));
#pragma unroll
                            for (int e = 0; e < 8; ++e) {
                                const uint32_t idx2 = lo + (uint32_t)r * kRun + 8u * (uint32_t)lane + (uint32_t)e;
                                uh[r][e] = idx2 < hi ? wt_t((bits >> (2 * e)) & 3u, (E)u0, (E)u1, (E)u2, (E)u3) : (E)0;
                            }
                        }
                    }
                    // the column values come from the pair table, two per LDS load; slots
                    // go in groups of four whose loads are issued together (a group past
                    // nb is skipped, a slot past nb within a group reads slot nb - 1 and
                    // contributes 0). Per slot the products are summed in individual
                    // order.
                    E part[16], gpart[kG ? 16 : 1];
#pragma unroll
                    for (int g4 = 0; g4 < 4; ++g4) {
                        if (pos + wave + kWaves * 4 * g4 < nb) {
                            int kc[4];
                            uint32_t bits[4][R];
#pragma unroll
                            for (int j = 0; j < 4; ++j) {
                                const int k = pos + wave + kWaves * (4 * g4 + j);
                                kc[j] = k < nb ? k : nb - 1;
#pragma unroll
                                for (int r = 0; r < R; ++r)
                                    bits[j][r] = L.codes[kc[j]][r][lane >> 1] >> (16 * (lane & 1));
                            }
#pragma unroll
                            for (int j = 0; j < 4; ++j) {
                                const int i = 4 * g4 + j;
                                E s = 0, gs = 0;
#pragma unroll
                                for (int r = 0; r < R; ++r) {
                                    decltype(pair_at<R>(L, 0, 0u, E{})) x[4];
#pragma unroll
                                    for (int q = 0; q < 4; ++q) x[q] = pair_at<R>(L, kc[j], (bits[j][r] >> (4 * q)) & 15u, E{});
#pragma unroll
                                    for (int q = 0; q < 4; ++q) {
                                        s = fma_t(x[q].x, eps[r][2 * q], s);
                                        s = fma_t(x[q].y, eps[r][2 * q + 1], s);
                                        if constexpr (kG) {
                                            gs = fma_t(x[q].x, uh[r][2 * q], gs);
                                            gs = fma_t(x[q].y, uh[r][2 * q + 1], gs);
                                        }
                                    }
                                }
                                const bool live = pos + wave + kWaves * i < nb;
                                part[i] = live ? s : (E)0;
                                if constexpr (kG) gpart[i] = live ? gs : (E)0;
                            }
                        } else {
#pragma unroll
                            for (int i = 4 * g4; i < 4 * g4 + 4; ++i) {
                                part[i] = 0;
                                if constexpr (kG) gpart[i] = 0;
                            }
                        }
                    }
                    const int i = (lane >> 2) & 15;
                    const int k = pos + wave + kWaves * i;
                    auto reduce16 = [&](E (&v)[16]) -> double {
                        if constexpr (kF32) return transpose_reduce16_c(v, lane);
                        else return transpose_reduce16(v, lane);
                    };
                    const double tot = reduce16(part);
                    if constexpr (kG) {
                        const double gt = reduce16(gpart);
                        if ((lane & 3) == 0 && k >= gpos && k < nb) publish_pair(p.xchg1, slot1(p, rnd, kMaxB + k, w), rnd, gt);
                    }
                    if ((lane & 3) == 0 && k < nb) publish_pair(p.xchg1, slot1(p, rnd, k, w), rnd, tot);
                };
                if (gpos < nb) dots(BoolC<true>{});
                else dots(BoolC<false>{});
                BRR_TICK(1);
                if (!exchange<R>(p, L, rnd, pos, nb, w, wave, lane, gpos)) goto done;
                BRR_TICK(2);
                if (first_round && has2 && tid < nb2) L.mk_next2[tid] = jnext2;

                // ---- decisions and commit (wave 0, lane = slot), BayesRBase.cpp:300-378.
                // Pass 0 decides slots [pos, nb) on the exchanged dots; when its first
                // change is slot h (sequential sweep, Gram entries present), pass 1
                // decides slots (h, nb) again on the dots corrected for that change.
                if (wave == 0) {
                    const int k = pos + lane;
                    int start = pos;
                    double eS = L.epsSum;
                    double dcorr = 0.0;
                    int c0 = nb, cfin = nb, c1 = nb;
                    for (int pass = 0; pass < (kAsync ? 1 : 2); ++pass) {
                        const bool act = k >= start && k < nb;
                        bool changed = false;
                        double bnew = 0.0, acum0 = 0.0, ed = 0.0, numk = 0.0;
                        int sel = -1, g = 0;
                        // the decision, with its per-component loops unrolled to 4 components
                        // when K <= 4 (the usual mixture) and to kMaxK otherwise
                        auto decide = [&](auto kc) {
                            constexpr int kK = decltype(kc)::value;
                            g = L.gr[k];
                            const double* gc = L.gc + g * K;
                            const double* gden = L.gden + g * K;
                            const double* gbsd = L.gbsd + g * K;
                            const double beta_old = L.beta_old[k];
                            const double dotk = pass ? L.dot[k] + dcorr * L.gram[k] : L.dot[k];
                            double num;
                            if (dense) {
                                num = dotk + beta_old * NM1;
                            } else {
                                const double mean = L.mstat[k][0], sd = L.mstat[k][1];
                                num = beta_old * (dN - 1.0) - mean * eS / sd + dotk / sd;
                            }
                            numk = num;
                            // d_c = logL_c - logL_0 = gdg_c + num^2 ghk_c. While every |d_c| <= 350
                            // no 700 guard can fire and no exponential overflows, and the
                            // probability terms 1 / sum_i exp(logL_i - logL_c) equal
                            // exp(d_c) / sum_i exp(d_i): one set of K - 1 exponentials.
                            // Otherwise the reference's own formulation, guards included.
                            const double n2 = num * num;
                            double dd[kK];
                            dd[0] = 0.0;
                            bool far = false;
#pragma unroll
                            for (int c = 1; c < kK; ++c) {
                                dd[c] = 0.0;
                                if (c < K) {
                                    dd[c] = L.gdg[g * K + c] + n2 * L.ghk[g * K + c];
                                    far = far || fabs(dd[c]) > 350.0;
                                }
                            }
                            const double uu = L.uu[k];
                            if (!far) {
                                double ee[kK];
                                double esum = 1.0;
                                ee[0] = 1.0;
#pragma unroll
                                for (int c = 1; c < kK; ++c) {
                                    ee[c] = 0.0;
                                    if (c < K) {
                                        ee[c] = exp(dd[c]);
                                        esum += ee[c];
                                    }
                                }
                                const double inv = 1.0 / esum;
                                double acum = inv;
                                acum0 = acum;
#pragma unroll
                                for (int c = 0; c < kK; ++c) {
                                    if (c < K && sel < 0) {
                                        if (uu <= acum) sel = c;
                                        else if (c + 1 < K) acum += ee[c + 1 < kK ? c + 1 : 0] * inv;
                                    }
                                }
                            } else {
                                double logL[kK], muk[kK];
                                logL[0] = gc[0];
                                muk[0] = 0.0;
#pragma unroll
                                for (int c = 1; c < kK; ++c) {
                                    muk[c] = 0.0;
                                    logL[c] = 0.0;
                                    if (c < K) {
                                        muk[c] = num / gden[c];
                                        logL[c] = gc[c] + 0.5 * (muk[c] * num) / L.sigmaE;
                                    }
                                }
                                double acum = guard700(logL, K, logL[0]) ? 0.0 : inv_sum_exp(logL, K, logL[0]);
                                acum0 = acum;
#pragma unroll
                                for (int c = 0; c < kK; ++c) {
                                    if (c < K && sel < 0) {
                                        if (uu <= acum) {
                                            sel = c;
                                        } else if (c + 1 < K) {
                                            const double ref = logL[c + 1 < kK ? c + 1 : 0];
                                            if (!guard700(logL, K, ref)) acum += inv_sum_exp(logL, K, ref);
                                        }
                                    }
                                }
                            }
                            // the beta draw of the chosen component (BayesRBase.cpp:343-352)
                            bnew = beta_old;
                            if (sel == 0) bnew = 0.0;
                            else if (sel > 0) {
                                bnew = bits_as_double(L.zw[k]) * gbsd[sel] + num / gden[sel];
                            }
                            changed = (bnew != beta_old);
                        };
                        if (act) {
                            if (K <= 4) decide(IntC<4>{});
                            else decide(IntC<kMaxK>{});
                        }
                        const unsigned long long cm = __ballot(changed);
                        // sequential sweep: the first change ends the pass; asynchronous
                        // window (pos = 0): every slot decides on the window-start residual
                        const int c = (cm && !kAsync) ? pos + (int)__ffsll((long long)cm) - 1 : nb;
                        const int last = (c < nb) ? c : nb - 1;
                        const bool accepted = act && k <= last;
                        // v counts (integer-valued doubles: exact in any order), one add per
                        // distinct (group, component) of the accepted slots
                        {
                            const int key = (accepted && sel >= 0) ? g * K + sel : -1;
                            unsigned long long todo = __ballot(key >= 0);
                            while (todo) {
                                const int kb = lane_read(key, (int)__ffsll((long long)todo) - 1);
                                const unsigned long long same = __ballot(key == kb);
                                if (lane == 0) L.v[kb] += (double)__popcll(same);
                                todo &= ~same;
                            }
                        }
                        // betasqnG += beta^2 in marker order (BayesRBase.cpp:353)
                        unsigned long long sq = __ballot(accepted && sel > 0);
                        while (sq) {
                            const int b = (int)__ffsll((long long)sq) - 1;
                            const double bb = lane_read(bnew, b);
                            const int gb = lane_read(g, b);
                            if (lane == 0) L.betasqn[gb] += bb * bb;
                            sq &= sq - 1;
                        }
                        // slot k's outputs are written by workgroup k % P (its reducer), so no
                        // workgroup holds a round's worth of stores before its next publish
                        const bool writer = (uint32_t)k < p.P ? (uint32_t)k == w : (uint32_t)k % p.P == w;
                        if (writer && accepted) {
                            const uint32_t j = L.mk[k];
                            st_agent((uint64_t*)(p.beta + j), double_as_bits(bnew));
                            // agent-scope stores: a marker's writer changes between iterations
                            // and the XCD L2s are not coherent with each other
                            if (sel >= 0) __hip_atomic_store(p.comp + j, sel, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            st_agent((uint64_t*)(p.acum + j), double_as_bits(acum0));
                            if (p.num_out) st_agent((uint64_t*)(p.num_out + j), double_as_bits(numk));
                        }
                        // residual change per 2-bit code of each slot applied below: the
                        // pass's first change (sequential) or every change of the window
                        // (asynchronous: BayesRBase::processColumnAsync, BayesRBase.cpp:394-491,
                        // with the window's updateGlobal calls, :503-510, in slot order)
                        const bool apply = act && (kAsync ? changed : k == c);
                        const double dBeta = apply ? L.beta_old[k] - bnew : 0.0;
                        if (apply) {
                            if (dense) {
#pragma unroll
                                for (int q = 0; q < 4; ++q) L.dl[k][q] = dBeta * L.wt[k][q];
                            } else {
                                // raggedbayesrkernel.cpp:10-30 and sparsebayesrkernel.cpp:29-34
                                const double mean = L.mstat[k][0], sd = L.mstat[k][1], Zsum = L.mstat[k][2];
                                const double meanAdjustment = dBeta * mean / sd;
                                const double oneAdjustment = dBeta / sd;
                                L.dl[k][3] = -meanAdjustment;
                                L.dl[k][2] = -meanAdjustment + oneAdjustment;
                                L.dl[k][0] = -meanAdjustment + 2 * oneAdjustment;
                                L.dl[k][1] = -meanAdjustment + meanAdjustment;
                                ed = dBeta * Zsum / sd - dBeta * mean * dN / sd;
                            }
                        }
                        if constexpr (kAsync) {
                            if (!dense) {
                                // the window's epsilonSum updates in slot order (SparseBayesRRG.cpp:94-103)
                                unsigned long long mset = cm;
                                while (mset) {
                                    const int b = (int)__ffsll((long long)mset) - 1;
                                    eS = eS + lane_read(ed, b);
                                    mset &= mset - 1;
                                }
                            }
                            if (lane == 0) L.cmask = cm << pos;
                        } else if (!dense && c < nb) {
                            eS = eS + lane_read(ed, c - pos);
                        }
                        cfin = c;
                        if (pass == 0) {
                            c0 = c;
                            if (c != h || gpos >= nb) break;
                            dcorr = lane_read(dBeta, c - pos);
                            start = c + 1;
                        } else {
                            c1 = c;
                        }
                    }
                    if (lane == 0) {
                        L.epsSum = eS;
                        L.change = c0;
                        L.change2 = c1;
                        L.next = cfin < nb ? cfin + 1 : nb;
                    }
                } else if (wave == 1 && first_round && has1 && p.mode == kModeChain) {
                    // the next block's u and z, while wave 0 decides (stored at its commit)
                    if (lane < nb1) {
                        double u;
                        uint64_t zw;
                        chain_draws(p, L.mk_next[lane], iteration, u, zw);
                        L.uu_next[lane] = u;
                        L.zw_next[lane] = zw;
                    }
                }
                __syncthreads();
                BRR_TICK(3);
                if (first_round) {
                    first_round = false;
                    // ragged stage 1 of block blk+2: segment bounds to registers,
                    // stored once block blk+1's segments have been consumed
                    if (!kCodes && has2 && tid < 3 * nb2) seg_issue(p, L.mk_next2[tid / 3], tid % 3, w, st2, ln2);
                }
                // the applied slots: every change of an asynchronous window, else the
                // changes of the round's passes in slot order
                auto apply_slot = [&](int k) {
                    ++n_changes;
                    const E d0 = (E)L.dl[k][0], d1 = (E)L.dl[k][1], d2 = (E)L.dl[k][2], d3 = (E)L.dl[k][3];
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const uint32_t bits = L.codes[k][r][lane >> 1] >> (16 * (lane & 1));
#pragma unroll
                        for (int e = 0; e < 8; ++e) {
                            const uint32_t idx2 = lo + (uint32_t)r * kRun + 8u * (uint32_t)lane + (uint32_t)e;
                            if (idx2 < hi) eps[r][e] = eps[r][e] + wt_t((bits >> (2 * e)) & 3u, d0, d1, d2, d3);
                        }
                    }
                    if constexpr (kF32) {
                        // the FP64 replay log (a block never holds more than refresh_blocks * B
                        // <= kMaxLog changes between two refreshes: one per marker at most)
                        if (tid == 0) {
                            const int n = RL.n;
                            if (n < kMaxLog) {
                                RL.j[n] = L.mk[k];
#pragma unroll
                                for (int q = 0; q < 4; ++q) RL.dl[n][q] = L.dl[k][q];
                            }
                            RL.n = n + 1;
                        }
                    }
                };
                if constexpr (kAsync) {
                    unsigned long long mset = L.cmask;
                    while (mset) {
                        const int k = (int)__ffsll((long long)mset) - 1;
                        mset &= mset - 1;
                        apply_slot(k);
                    }
                } else {
                    const int c0 = L.change, c1 = L.change2;
                    if (c0 < nb) apply_slot(c0);
                    if (c1 < nb) apply_slot(c1);
                }
                pos = L.next;
                BRR_TICK(4);
            }
            __syncthreads();

            // ---- commit the prefetch: block blk+1 becomes current
            if (has1) {
                if constexpr (kCodes) dense_store<R>(p, L, nb1, lo, tid, words);
                else ragged_store<R>(p, L, nb1, lo, tid, idx, itm);
                if (tid < nb1) meta_store<R, kLay>(p, L, tid, mnext, false);
                __syncthreads();
                pair_table_store<kF32, R>(L, nb1, tid);
                if (has2) {
                    if (tid < nb2) L.mk_next[tid] = jnext2;
                    if constexpr (!kCodes) seg_store_prefix<R>(L, nb2, tid, wave, lane, st2, ln2);
                    else __syncthreads();
                }
                if (!kCodes || !has2) __syncthreads();
            }
            BRR_TICK(0);
            if constexpr (kF32) {
                // FP64 refresh of the residual every refresh_blocks blocks and after the
                // last block (in a chain the end-of-iteration FP64 region does the last)
                const bool last = blk + 1 == nblocks;
                const uint32_t rb = p.refresh_blocks ? p.refresh_blocks : 1u;
                if ((last && p.mode != kModeChain) || (!last && (blk + 1) % rb == 0)) {
                    if (!fp64_region([](double (&)[R][8]) { return true; })) goto done;
                    BRR_TICK(5);
                }
            }
        }
        // other workgroups read this workgroup's beta stores only in the block
        // loads of the next iteration, which follow this workgroup's next publish
        __builtin_amdgcn_s_waitcnt(0);

        if (p.mode == kModeChain) {
            double e2 = 0.0;
            const bool ok = fp64_region([&](double (&ed)[R][8]) {
                // ---- fixed effects (BayesRBase.cpp:212-235)
                if (p.F && !fixed_effects<R>(p, L, rnd, n_rounds, ed, iteration, lo, hi, w, wave, lane))
                    return false;
                // ---- |eps|^2 of the sigmaE draw
                double part = 0.0;
#pragma unroll
                for (int r = 0; r < R; ++r)
#pragma unroll
                    for (int e = 0; e < 8; ++e) part += ed[r][e] * ed[r][e];
                if (!allreduce_scalar<R>(p, L, rnd, part, w, wave, lane, &e2)) return false;
                ++n_rounds;
                return true;
            });
            if (!ok) goto done;
            // ---- sigmaE, sigmaG, pi draws (BayesRBase.cpp:237-247)
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
            BRR_TICK(5);
        }
    }

done:
    __syncthreads();
    if (L.abort == 0) {
        // (FP32 mode: the last FP64 region has stored the residual)
        if (wave == 0 && !kF32) {
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const uint32_t idx = lo + (uint32_t)r * kRun + 8u * (uint32_t)lane + (uint32_t)e;
                    if (idx < hi) p.eps[idx] = (double)eps[r][e];
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
                for (int q = 0; q < kPhases; ++q) p.counters[2 + q] = tk[q];
            }
            for (uint32_t t = tid; t < p.G; t += kThreads) {
                p.scal[SL.sigmaG() + t] = L.sigmaG[t];
                p.scal[SL.betasqn() + t] = L.betasqn[t];
            }
            for (uint32_t t = tid; t < p.G * p.K; t += kThreads) {
                p.scal[SL.pi() + t] = L.pi[t];
                p.scal[SL.v() + t] = L.v[t];
            }
            if (p.F && tid == 0) p.fixed[0] = L.sigmaF;
            for (uint32_t t = tid; t < p.F; t += kThreads) {
                p.fixed[1 + t] = L.gamma[t];
                p.forder[t] = L.forder[t];
            }
        }
    }
}

template <int R, int kLay, bool kCodes, bool kAsync, bool kF32>
__global__ void __launch_bounds__(kThreads)
bayesr_sweep_kernel(SweepParams p)
{
    __shared__ Shared<R> L;
    __shared__ RefreshLog<kF32, R> RL;
    sweep_body<R, kLay, kCodes, kAsync, kF32>(p, L, RL, blockIdx.x);
}

// ------------------------------------------------------- cohorts in one launch --

// Independent single-chain contexts (cohorts of uneven N and M) in one
// persistent cooperative launch: cohort c owns workgroups [first[c], first[c+1])
// and runs its own context's sweep or chain on them -- its own slices,
// exchange words, state and counters -- so small cohorts, whose rounds cost as
// much as a large one's, run side by side instead of one after another. Every
// cohort of a launch has the same lane runs R and residual precision; the
// layout (dense / ragged / EigenSparse formulas over 2-bit codes) is per cohort.
constexpr int kMaxCohorts = 64;

struct CohortTable {
    const SweepParams* params;          // device array, one per cohort
    uint32_t n;
    uint32_t first[kMaxCohorts + 1];    // first workgroup of each cohort; first[n] = grid size
};

template <int R, bool kF32>
__global__ void __launch_bounds__(kThreads)
bayesr_cohort_kernel(CohortTable t)
{
    __shared__ Shared<R> L;
    __shared__ RefreshLog<kF32, R> RL;
    uint32_t c = 0;
    while (c + 1 < t.n && blockIdx.x >= t.first[c + 1]) ++c;
    const SweepParams p = t.params[c];
    const uint32_t w = blockIdx.x - t.first[c];
    if (p.layout == kLayoutDense) sweep_body<R, kLayoutDense, true, false, kF32>(p, L, RL, w);
    else if (p.layout == kLayoutEigen) sweep_body<R, kLayoutEigen, true, false, kF32>(p, L, RL, w);
    else sweep_body<R, kLayoutRagged, true, false, kF32>(p, L, RL, w);
}

} // namespace brr
