// bayesw_kernel.hpp -- persistent gfx950 kernel for the BayesW (Weibull
// survival) marker sweep on ragged markers held as 2-bit codes.
//
// What it computes: BayesWBase::processColumn (bayeswbase.cpp:604-704) with the
// RaggedBayesWKernel sums (raggedbayeswkernel.cpp:10-77) over a marker order,
// optionally wrapped in whole runGibbs iterations (bayeswbase.cpp:777-809:
// sampleMu, the sweep, sampleAlpha, the sigma_b and pi_L draws).
//
// How: one cooperative launch of P workgroups; a workgroup owns a slice of
// individuals and keeps eps and vi = exp(alpha eps - EulerGamma) of the slice in
// registers (each wave a full copy). Markers go in blocks of 32. A round
// publishes the slice's partial sum of vi and, for every undecided slot, its
// partial sums of vi over the genotype-1 and genotype-2 individuals; after one
// deterministic reduce-scatter / all-gather every workgroup evaluates the same
// adaptive Gauss-Hermite class likelihoods (one slot per lane) and accepts
// slots up to the first that changes eps: a slot drawn into a nonzero class (its
// beta from ARMS, after which eps and vi are updated and the rest re-summed), or
// a marker already in the model, whose step starts from eps + x beta_old: it
// gets a round of its own that applies that shift, re-forms vi and sums it.
// ARMS (BayesW_arms.cpp) runs identically in every wave of every workgroup on a
// per-wave LDS envelope; the alpha density needs one grid-wide sum per
// evaluation, the beta and mu densities are closed forms of exchanged sums.
#pragma once

#include "sweep_kernel.hpp"

namespace brr {

constexpr int kBwB = 32;                       // slots per block
constexpr int kBwV = 3 * kBwB + 8;             // values per round: vi_sum, the per-code sums of each slot
                                               // (2 ragged, 3 dense) and the shifted sums of a speculated slot
constexpr int kArmsN = 100;                    // ARMS envelope points (npoint of bayeswbase.cpp)
constexpr uint32_t kStreamArms = 4;
constexpr uint32_t kSiteMu = 0xFFFFFFF0u;
constexpr uint32_t kSiteAlpha = 0xFFFFFFF1u;
constexpr uint32_t kSiteTheta0 = 0xFFFFFF00u;   // + covariate index f < 128
constexpr uint32_t kBwDrawSigmaB = 0;
constexpr uint32_t kBwDrawPi0 = 1;
constexpr double kEuler = 0.577215664901532;   // EuMasc (bayeswbase.cpp:26)
constexpr double kSqrtPi = 1.77245385090552;   // sqrtPI (bayeswbase.cpp:25)

struct BwParams {
    uint32_t N, M, K, P, S, nq;
    int mode;
    uint32_t iterations, iter_base, chain;
    uint64_t seed;
    const uint32_t* codes;
    uint64_t col_words, col_bytes;
    const double* stats;          // ragged: M x 4 {mean, sd, sqrdZ, Zsum}
    const double* lut;            // dense: M x 4, the standardised value of dosage 0, 1, 2, missing
    const double* sumf;           // M: calculateSumFailure
    const double* fail;           // N
    const double* cls;            // K - 1 mixture variances C_k
    const uint32_t* order;
    const double* u;              // sweep mode: marker uniforms by id
    double* eps;
    double* beta;
    int32_t* comp;
    double* scal;                 // {mu, alpha, sigma_b, d, betasq, pi[K], v[K]}
    double* hist;                 // iterations x 5
    uint64_t* xchg1;              // [2][kBwV][P][2]
    uint64_t* xchg2;              // [2][kBwV][2]
    uint32_t* ctl;
    unsigned long long* counters;
    // fixed effects (bayeswbase.cpp:572-601): F covariates, column-major N x F
    // floats (Data::X), theta and sum_k X(k,f) failure_k per covariate
    const float* fixedX;
    double* theta;
    const double* sumf_fix;
    uint32_t F, pad_f;
};

enum : int { kBwMu = 0, kBwAlpha = 1, kBwSigmaB = 2, kBwD = 3, kBwBetaSq = 4, kBwPi = 5 };

// One sampler run, one per wave in LDS. The envelope is kept sorted by x:
// position 0 is the left bound, odd positions are evaluated points, even ones
// intersections of chords, position n - 1 the right bound (the index-linked
// list of BayesW_arms.cpp in its traversal order). Lane l works on positions l
// and l + 64.
constexpr int kArmsCap = 128;
struct ArmsEnv {
    double x[kArmsCap], y[kArmsCap], ey[kArmsCap], cum[kArmsCap], area[kArmsCap];
    double xinit[4];
    double ymax, yu, ynew;
    double wx, wy, wey;           // the working point
    double xout;                  // point to evaluate next, or the sample
    uint64_t seed;
    uint32_t site, iteration, chain, k;
    int n, state, kinit, q, wq, guard;
};

template <int R>
struct BwShared {
    double theta[kMaxF];          // fixed effects
    uint32_t codes[kBwB][R][32];
    uint32_t mk[kBwB];
    uint32_t ord[2][kBwB];        // marker ids of the next two blocks, by block parity
    double bold[kBwB], mean[kBwB], sd[kBwB], sumf[kBwB], pu[kBwB];
    double tab[kBwB][4];          // dense: column value by 2-bit code
    // per-code indicators of two adjacent individuals whose codes form the 4-bit
    // field n: ind[n][q] = ([code_lo == cq], [code_hi == cq]) for cq = 2, 0, 1
    // (genotype 1, genotype 2, missing), so a slot's per-code sums of vi are FMAs
    // with 0 / 1 (the same additions as a select, in the same order)
    alignas(16) double ind[16][3][2];
    double xv[2][kBwV];           // round totals, by round parity
    double cls[kMaxK], pi[kMaxK], v[kMaxK];
    double gv[kBwB][kMaxK - 1][14];   // weighted integrand terms per slot, slab class and node
    double ml[kBwB][kMaxK];            // class likelihoods per slot
    double ghx[14], ghw[14], ghwc;     // the rule's nodes in summation order
    double mu, alpha, sigma_b, d, visum, betasq;
    int stop;                     // first slot that changes eps this round
    int sel;                      // its class
    ArmsEnv env[kWaves];
    int abort;
    int arms_err;
};

// ------------------------------------------------------------------ ARMS --
// The sampler of BayesW_arms.cpp restated as in the CPU oracle: index-linked
// envelope points, nsamp = 1, no Metropolis step. Every lane of a wave runs it
// on the wave's envelope with identical data, so the lockstep lanes store
// identical values.

__device__ __forceinline__ double arms_u(ArmsEnv& e)
{
    const uint32_t k = e.k++;
    const U4 x = draw_block(e.seed, e.site, e.iteration, kStreamArms | ((k >> 1) << 8), e.chain);
    const uint32_t hi = (k & 1u) ? x.x[2] : x.x[0], lo = (k & 1u) ? x.x[3] : x.x[1];
    const uint64_t w = ((uint64_t)hi << 32) | lo;
    return ((double)(w >> 11) + 0.5) * 1.1102230246251565404e-16;
}

__device__ __forceinline__ double arms_expshift(double y, double y0)
{
    return (y - y0 > -100.0) ? exp(y - y0 + 50.0) : 0.0;
}

__device__ __forceinline__ double arms_logshift(double y, double y0) { return log(y) + y0 - 50.0; }

// meet (BayesW_arms.cpp:707-808) of the intersection at even position q: its
// x and y from the chords through the evaluated points q-3, q-1 and q+1, q+3.
// Nonzero on a convexity violation or a misplaced intersection. Reads only
// evaluated points, so every intersection may be recomputed at once: one whose
// neighbours did not change gets the same value again.
__device__ __forceinline__ int arms_meet_at(ArmsEnv& e, int q)
{
    const int n = e.n;
    double gl = 0, gr = 0, grl = 0, dl = 0, dr = 0;
    const bool il = q - 3 >= 0, ir = q + 3 <= n - 1, irl = q - 1 >= 0 && q + 1 <= n - 1;
    if (il) gl = (e.y[q - 1] - e.y[q - 3]) / (e.x[q - 1] - e.x[q - 3]);
    if (ir) gr = (e.y[q + 1] - e.y[q + 3]) / (e.x[q + 1] - e.x[q + 3]);
    if (irl) grl = (e.y[q + 1] - e.y[q - 1]) / (e.x[q + 1] - e.x[q - 1]);
    if (irl && il && gl < grl) return 1;
    if (irl && ir && gr > grl) return 1;
    if (il && irl) {
        dr = (gl - grl) * (e.x[q + 1] - e.x[q - 1]);
        if (dr < 0.1) dr = 0.1;
    }
    if (ir && irl) {
        dl = (grl - gr) * (e.x[q + 1] - e.x[q - 1]);
        if (dl < 0.1) dl = 0.1;
    }
    double x = e.x[q], y;
    if (il && ir && irl) {
        x = (dl * e.x[q + 1] + dr * e.x[q - 1]) / (dl + dr);
        y = (dl * e.y[q + 1] + dr * e.y[q - 1] + dl * dr) / (dl + dr);
    } else if (il && irl) {
        x = e.x[q + 1];
        y = e.y[q + 1] + dr;
    } else if (ir && irl) {
        x = e.x[q - 1];
        y = e.y[q - 1] + dl;
    } else if (il) {
        y = e.y[q - 1] + gl * (x - e.x[q - 1]);
    } else if (ir) {
        y = e.y[q + 1] - gr * (e.x[q + 1] - x);
    } else {
        return 4;
    }
    e.x[q] = x;
    e.y[q] = y;
    if ((q - 1 >= 0 && x < e.x[q - 1]) || (q + 1 <= n - 1 && x > e.x[q + 1])) return 5;
    return 0;
}

// every intersection of the wave's envelope; nonzero if any fails
__device__ __forceinline__ int arms_meet_all(ArmsEnv& e, int lane)
{
    bool bad = false;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int q = lane + 64 * h;
        if (q < e.n && (q & 1) == 0) bad |= arms_meet_at(e, q) != 0;
    }
    return __any(bad) ? 1 : 0;
}

// exponentiate and integrate the envelope (BayesW_arms.cpp:672-703, 812-834):
// ey and the areas in parallel, the running sum in position order as the
// reference accumulates it
__device__ __forceinline__ void arms_cumulate(ArmsEnv& e, int lane)
{
    const int n = e.n;
    double m = -__builtin_inf();
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int q = lane + 64 * h;
        if (q < n) m = fmax(m, e.y[q]);
    }
    // (the whole wave runs the sampler: lane_xor's moves need every lane active)
    m = fmax(m, lane_xor<32>(m, lane));
    m = fmax(m, lane_xor<16>(m, lane));
    m = fmax(m, lane_xor<8>(m, lane));
    m = fmax(m, lane_xor<4>(m, lane));
    m = fmax(m, lane_xor<2>(m, lane));
    m = fmax(m, lane_xor<1>(m, lane));
    const double ymax = m;
    double eyq[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int q = lane + 64 * h;
        eyq[h] = 0.0;
        if (q < n) {
            eyq[h] = arms_expshift(e.y[q], ymax);
            e.ey[q] = eyq[h];
        }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int q = lane + 64 * h;
        if (q >= 1 && q < n) {
            const double eyl = arms_expshift(e.y[q - 1], ymax);   // the value its left neighbour stored
            double a;
            if (e.x[q - 1] == e.x[q]) a = 0.0;
            else if (fabs(e.y[q] - e.y[q - 1]) < 0.1) a = 0.5 * (eyq[h] + eyl) * (e.x[q] - e.x[q - 1]);
            else a = ((eyq[h] - eyl) / (e.y[q] - e.y[q - 1])) * (e.x[q] - e.x[q - 1]);
            e.area[q] = a;
        }
    }
    double c = 0.0;
    e.cum[0] = 0.0;
    for (int q = 1; q < n; ++q) {
        c = c + e.area[q];
        e.cum[q] = c;
    }
    e.ymax = ymax;
}

// the x whose envelope cumulative probability is prob, into the working point
// (BayesW_arms.cpp:364-425); it falls between positions q - 1 and q, q = e.wq
__device__ __forceinline__ int arms_invert(ArmsEnv& e, double prob, int lane)
{
    const int n = e.n;
    const double u = prob * e.cum[n - 1];
    // q - 1 = the last position below n - 1 whose running sum is <= u
    // (the reference walks left from the end while the left neighbour's sum is > u)
    const unsigned long long b0 = __ballot(lane <= n - 2 && !(e.cum[lane] > u));
    const unsigned long long b1 = __ballot(lane + 64 <= n - 2 && !(e.cum[lane + 64] > u));
    if (!b0 && !b1) return 6;   // (cum[0] = 0 <= u always holds)
    const int last = b1 ? 64 + 63 - __clzll((long long)b1) : 63 - __clzll((long long)b0);
    const int q = last + 1, l = last;
    e.wq = q;
    const double prop = (u - e.cum[l]) / (e.cum[q] - e.cum[l]);
    double xl = e.x[q], xr = e.x[q], wx, wy, wey;
    if (e.x[l] == e.x[q]) {
        wx = e.x[q];
        wy = e.y[q];
        wey = e.ey[q];
    } else {
        xl = e.x[l];
        xr = e.x[q];
        const double yl = e.y[l], yr = e.y[q], eyl = e.ey[l], eyr = e.ey[q];
        if (fabs(yr - yl) < 0.1) {
            if (fabs(eyr - eyl) > 0.001 * fabs(eyr + eyl))
                wx = xl + ((xr - xl) / (eyr - eyl)) * (-eyl + sqrt((1. - prop) * eyl * eyl + prop * eyr * eyr));
            else
                wx = xl + (xr - xl) * prop;
            wey = ((wx - xl) / (xr - xl)) * (eyr - eyl) + eyl;
            wy = arms_logshift(wey, e.ymax);
        } else {
            wx = xl + ((xr - xl) / (yr - yl)) * (-yl + arms_logshift(((1. - prop) * eyl + prop * eyr), e.ymax));
            wy = ((wx - xl) / (xr - xl)) * (yr - yl) + yl;
            wey = arms_expshift(wy, e.ymax);
        }
    }
    e.wx = wx;
    e.wy = wy;
    e.wey = wey;
    if (wx < xl || wx > xr) return 6;
    return 0;
}

enum : int { kArmsEval = 0, kArmsDone = 1 };
enum : int { kAsInit = 0, kAsSample = 1, kAsEvalW = 2, kAsEvalQ = 3 };

// Sets up one draw (arms with ninit = 4, npoint = 100, nsamp = 1 and no
// Metropolis step, BayesW_arms.cpp:133-340); 0 or the reference's error code.
__device__ __forceinline__ int arms_begin(ArmsEnv& e, const double (&xinit)[4], double xl, double xr, uint64_t seed,
                                          uint32_t site, uint32_t iteration, uint32_t chain, int lane)
{
    constexpr int ninit = 4, mpoint = 2 * ninit + 1;
    if (xinit[0] <= xl || xinit[ninit - 1] >= xr) return 1003;
    for (int i = 1; i < ninit; ++i)
        if (xinit[i] <= xinit[i - 1]) return 1004;
    if (lane < mpoint) {
        e.x[lane] = (lane & 1) ? xinit[lane >> 1] : (lane == 0 ? xl : (lane == mpoint - 1 ? xr : 0.0));
        e.y[lane] = 0.0;
    }
    e.n = mpoint;
    e.seed = seed;
    e.site = site;
    e.iteration = iteration;
    e.chain = chain;
    e.k = 0;
    e.state = kAsInit;
    e.kinit = 0;
    e.guard = 0;
    return 0;
}

// Advances the sampler. `y` is the log density at the point the previous call
// asked for. Returns kArmsEval with e.xout the next point to evaluate, kArmsDone
// with e.xout the sample, or an error code (>= 1000). Every lane of a wave calls
// it with the same arguments on the wave's own envelope.
__device__ __forceinline__ int arms_step(ArmsEnv& e, double y, int lane)
{
    constexpr int mpoint = 9;
    if (e.state == kAsInit) {
        if (e.kinit > 0) e.y[2 * e.kinit - 1] = y;
        if (e.kinit < 4) {
            e.xout = e.x[2 * e.kinit + 1];
            e.kinit++;
            return kArmsEval;
        }
        if (arms_meet_all(e, lane)) return 2000;
        arms_cumulate(e, lane);
        e.state = kAsSample;
    } else {
        if (e.state == kAsEvalW) {
            // the working point was evaluated (BayesW_arms.cpp:505-536)
            e.ynew = y;
            e.wy = y;
            e.wey = arms_expshift(y, e.ymax);
            if (e.n <= kArmsN - 2) {
                // update: insert it with a new intersection (BayesW_arms.cpp:540-668).
                // It lies between positions wq - 1 and wq; the evaluated one of the two
                // keeps an intersection between itself and the new point.
                const int n = e.n, wq = e.wq;
                double sx[2], sy[2];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int q = lane + 64 * h;
                    sx[h] = (q < n) ? e.x[q] : 0.0;
                    sy[h] = (q < n) ? e.y[q] : 0.0;
                }
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int q = lane + 64 * h;
                    if (q >= wq && q < n) {
                        e.x[q + 2] = sx[h];
                        e.y[q + 2] = sy[h];
                    }
                }
                // positions wq - 1 (wl) and wq + 2 (wr, shifted): wl evaluated <=> wq - 1 odd
                const int qn = ((wq - 1) & 1) ? wq + 1 : wq;    // the new evaluated point
                e.x[qn] = e.wx;
                e.y[qn] = e.wy;
                e.n = n + 2;
                e.q = qn;
                const int ql = (qn - 2 >= 0) ? qn - 2 : qn - 1;
                const int qr = (qn + 2 <= e.n - 1) ? qn + 2 : qn + 1;
                double xa = 0.0;
                bool move = false;
                if (e.x[qn] < (1. - 0.00001) * e.x[ql] + 0.00001 * e.x[qr]) {
                    xa = (1. - 0.00001) * e.x[ql] + 0.00001 * e.x[qr];
                    move = true;
                } else if (e.x[qn] > 0.00001 * e.x[ql] + (1. - 0.00001) * e.x[qr]) {
                    xa = 0.00001 * e.x[ql] + (1. - 0.00001) * e.x[qr];
                    move = true;
                }
                if (move) {
                    e.x[qn] = xa;
                    e.xout = xa;
                    e.state = kAsEvalQ;
                    return kArmsEval;
                }
                e.state = kAsEvalQ;
                y = e.y[qn];
            }
        }
        if (e.state == kAsEvalQ) {
            e.y[e.q] = y;
            if (arms_meet_all(e, lane)) return 2000;
            arms_cumulate(e, lane);
        }
        // rejection test (BayesW_arms.cpp:529-535)
        if (!(e.yu >= e.ynew)) {
            e.xout = e.wx;
            return kArmsDone;
        }
        e.state = kAsSample;
    }
    // sample and squeeze (BayesW_arms.cpp:344-360, 429-503)
    if (++e.guard > 100000) return 2001;
    const int err = arms_invert(e, arms_u(e), lane);
    if (err) return err;
    e.yu = arms_logshift(arms_u(e) * e.wey, e.ymax);
    const int wq = e.wq, n = e.n;
    if (wq - 2 >= 0 && wq + 1 <= n - 1) {
        const int ql = ((wq - 1) & 1) ? wq - 1 : wq - 2;
        const int qr = (wq & 1) ? wq : wq + 1;
        const double ysq = (e.y[qr] * (e.wx - e.x[ql]) + e.y[ql] * (e.x[qr] - e.wx)) / (e.x[qr] - e.x[ql]);
        if (e.yu <= ysq) {
            e.xout = e.wx;
            return kArmsDone;
        }
    }
    e.xout = e.wx;
    e.state = kAsEvalW;
    return kArmsEval;
}

// The same step behind a call, for the draws of the per-iteration scalars (one
// copy of the sampler's code for them; the per-marker draws inline it).
__device__ __noinline__ int arms_step_call(ArmsEnv& e, double y, int lane) { return arms_step(e, y, lane); }

// One draw from log density f (inlined at the call site, where it may run
// grid-wide sums: every thread of the grid takes the same path). 0 or an error.
template <bool kInline, class F>
__device__ __forceinline__ int arms_draw(ArmsEnv& e, const double (&xinit)[4], double xl, double xr, uint64_t seed,
                                         uint32_t site, uint32_t iteration, uint32_t chain, F&& f, double* out)
{
    const int lane = threadIdx.x & 63;
    int st = arms_begin(e, xinit, xl, xr, seed, site, iteration, chain, lane);
    if (st) return st;
    double y = 0.0;
    while ((st = kInline ? arms_step(e, y, lane) : arms_step_call(e, y, lane)) == kArmsEval) y = f(e.xout);
    if (st != kArmsDone) return st;
    *out = e.xout;
    return 0;
}

// ------------------------------------------------- Gauss-Hermite rules ---
// The tabulated adaptive rules of gauss_hermite_adaptive_integral
// (bayeswbase.cpp:126-457): positive nodes, exp(x^2)-adjusted weights and the
// centre weight of n = 3, 5, ..., 15. For n >= 9 the reference evaluates its
// sixth node at -x3; bw_gh_node returns the node it evaluates.
__constant__ const double kGhX[7][7] = {
    {1.2247448713916},
    {2.0201828704561, 0.95857246461382},
    {2.6519613568352, 1.6735516287675, 0.81628788285897},
    {3.1909932017815, 2.2665805845318, 1.4685532892167, 0.72355101875284},
    {3.6684708465596, 2.7832900997817, 2.0259480158258, 1.3265570844949, 0.6568095668821},
    {4.1013375961786, 3.2466089783724, 2.5197356856782, 1.8531076516015, 1.2200550365908, 0.60576387917106},
    {4.4999907073094, 3.6699503734045, 2.9671669279056, 2.3257324861739, 1.7199925751865, 1.1361155852109,
     0.5650695832556}};
__constant__ const double kGhW[7][7] = {
    {1.3239311752136},
    {1.181488625536, 0.98658099675143},
    {1.1013307296103, 0.8971846002252, 0.8286873032836},
    {1.0470035809767, 0.84175270147867, 0.7646081250946, 0.73030245274509},
    {1.0065267861724, 0.802516868851, 0.721953624728, 0.6812118810667, 0.66096041944096},
    {0.97458039564, 0.7725808233517, 0.6906180348378, 0.6467594633158, 0.6217160552868, 0.60852958370332},
    {0.94836897082761, 0.7486073660169, 0.666166005109, 0.620662603527, 0.5930274497642, 0.5761933502835,
     0.5670211534466}};
__constant__ const double kGhWc[7] = {1.1816359006037, 0.94530872048294, 0.81026461755681, 0.72023521560605,
                                      0.65475928691459, 0.60439318792116, 0.56410030872642};

struct BwSums {
    double vi_sum, vi_0, vi_1, vi_2;
};

// integrand_adaptive (raggedbayeswkernel.cpp:70-77)
__device__ __forceinline__ double bw_integrand(double s, double alpha, double sq, double sumf, double mean_sd, double sd,
                                               const BwSums& v)
{
    const double temp = -alpha * s * sumf * sq + v.vi_sum -
                        exp(alpha * mean_sd * s * sq) *
                            (v.vi_0 + v.vi_1 * exp(-alpha * s * sq / sd) + v.vi_2 * exp(-2 * alpha * s * sq / sd)) -
                        s * s;
    return exp(temp);
}

// Dense layout (densebayeswkernel.cpp:30-42, densebayesw.cpp:25-33). A column
// holds at most four values c_d (dosage 0, 1, 2, missing), so every sum over
// individuals regroups into sums over those groups: V_d = sum of vi over group d.
struct BwGroups {
    double V[4], c[4];            // by dosage 0, 1, 2, missing
};

// exponent_sum: sum_i vi_i Cx_i Cx_i
__device__ __forceinline__ double bw_dense_exp_sum(const BwGroups& g)
{
    double e = 0.0;
#pragma unroll
    for (int d = 0; d < 4; ++d) e += g.V[d] * g.c[d] * g.c[d];
    return e;
}

// integrand_adaptive: exp(-alpha s sf sq + sum_i vi_i (1 - exp(-Cx_i s sq alpha)) - s^2)
__device__ __forceinline__ double bw_dense_integrand(double s, double alpha, double sq, double sumf,
                                                     const BwGroups& g)
{
    double t = 0.0;
#pragma unroll
    for (int d = 0; d < 4; ++d) t += g.V[d] * (1 - exp(-g.c[d] * s * sq * alpha));
    const double temp = -alpha * s * sumf * sq + t - s * s;
    return exp(temp);
}

// Nodes and weights of the n-point adaptive rule in the order the reference
// sums them (+x1, -x1, +x3, -x3, ...; the sixth node is -x3 for n >= 9), the
// centre weight last.
__device__ __forceinline__ void bw_gh_nodes(int nq, int n, double* x, double* w)
{
    const int r = (nq - 3) >> 1, i = n >> 1;
    const double xp = kGhX[r][i];
    *x = (n & 1) ? ((nq >= 9 && n == 5) ? -kGhX[r][1] : -xp) : xp;
    *w = kGhW[r][i];
}

// exp_sum and the slab scale sigma of class i (bayeswbase.cpp:624-640)
__device__ __forceinline__ double bw_sigma(double alpha, double sigma_b, double Ck, double mean, double sd,
                                           const BwSums& v)
{
    const double exp_sum = (v.vi_1 * (1 - 2 * mean) + 4 * (1 - mean) * v.vi_2 + v.vi_sum * mean * mean) / (sd * sd);
    return 1.0 / sqrt(1 + alpha * alpha * sigma_b * Ck * exp_sum);
}

// class picked by u from the class likelihoods ml[0..K) (bayeswbase.cpp:641-688;
// the last class gets acum = 1)
__device__ __forceinline__ int bw_pick(int K, const double* ml, double u)
{
    double ml_sum = 0.0;
    for (int i = 0; i < K; ++i) ml_sum += ml[i];
    double acum = ml[0] / ml_sum;
    for (int c = 0; c < K; ++c) {
        if (u <= acum) return c;
        if ((c + 1) == (K - 1)) acum = 1;
        else acum += ml[c + 1] / ml_sum;
    }
    return K - 1;
}

// ------------------------------------------------------------- exchange ---

__device__ __forceinline__ uint64_t* bw_slot1(const BwParams& p, uint32_t rnd, int v, uint32_t w)
{
    return p.xchg1 + ((((size_t)(rnd & 1u) * kBwV + (size_t)v) * p.P + w) * 2u);
}

// totals in kCopies copies, as the BayesR sweep's (copy w % kCopies is read by workgroup w)
constexpr size_t kBwCopyWords = 2 * kBwV * 2 + 32;

__device__ __forceinline__ uint64_t* bw_slot2(const BwParams& p, uint32_t rnd, int v, uint32_t copy)
{
    return p.xchg2 + copy * kBwCopyWords + (((size_t)(rnd & 1u) * kBwV + (size_t)v) * 2u);
}

// Values [0, V) published at round rnd: value v is reduced by workgroup v % P
// (lanes strided over the P partials, then a butterfly), totals land in
// L.xv[rnd & 1]; a thread reads them before it enters round rnd + 1, so round
// rnd + 2 may overwrite the buffer without another barrier. False on abort.
template <int R>
__device__ __forceinline__ bool bw_exchange(const BwParams& p, BwShared<R>& L, uint32_t rnd, int V, uint32_t w,
                                            int wave, int lane)
{
    bool ok = true;
    // values w, w + P, ...: the c-th reduced by wave c % kWaves
    int c = 0;
    for (int v = (int)w; v < V; v += (int)p.P, ++c) {
        if ((c % kWaves) != wave) continue;
        double acc = 0.0;
        ok = wait_sum_lanes(p.xchg1, [&](uint32_t w2) { return bw_slot1(p, rnd, v, w2); }, p.P, rnd, p.ctl, lane, &acc);
        if (__any(!ok)) { ok = false; break; }
        acc = wave_sum_full(acc, lane);
        if (lane < kCopies) publish_pair(p.xchg2, bw_slot2(p, rnd, v, (uint32_t)lane), rnd, acc);
    }
    if (!ok && lane == 0) { L.abort = 1; raise_abort(p, 21); }
    const int tid = wave * 64 + lane;
    if (tid < V) {
        double x;
        if (!wait_pair(p.xchg2, bw_slot2(p, rnd, tid, w % kCopies), rnd, p.ctl, &x)) { L.abort = 1; raise_abort(p, 22); }
        else L.xv[rnd & 1u][tid] = x;
    }
    __syncthreads();
    return L.abort == 0;
}

// ------------------------------------------------------------------ kernel --

// residual change of marker slot k per 2-bit code (calculateResidualUpdate,
// raggedbayeswkernel.cpp:27-43): -mean/sd beta everywhere, + beta/sd for
// genotype 1, + 2 beta/sd for genotype 2 (missing calls shift like genotype 0)
__device__ __forceinline__ void bw_delta(double mean, double sd, double beta, double (&dv)[4])
{
    const double mean_sd_ratio = mean / sd;
    const double meanAdjustment = mean_sd_ratio * beta;
    const double oneAdjustment = beta / sd;
    const double twoAdjustment = 2 * oneAdjustment;
    dv[0] = -meanAdjustment + twoAdjustment;   // code 0: genotype 2
    dv[1] = -meanAdjustment;                   // code 1: missing
    dv[2] = -meanAdjustment + oneAdjustment;   // code 2: genotype 1
    dv[3] = -meanAdjustment;                   // code 3: genotype 0
}

constexpr int kBwPhases = 8;   // block load, round sums, exchange, decisions, beta draws,
                               // residual update, iteration scalars, in-model shifts

// Phase timer of thread 0 (s_memrealtime: 100 MHz constant clock).
#define BW_TICK(slot)                                                            \
    do {                                                                         \
        if (tid == 0) {                                                          \
            const uint64_t t_now_ = __builtin_amdgcn_s_memrealtime();           \
            tk[(slot)] += t_now_ - t_last;                                       \
            t_last = t_now_;                                                     \
        }                                                                        \
    } while (0)

template <int R, bool kDense>
__global__ void __launch_bounds__(kThreads)
bayesw_sweep_kernel(BwParams p)
{
    constexpr int kPer = kDense ? 3 : 2;   // per-code sums published per slot (codes 2, 0 [, 1])
    __shared__ BwShared<R> L;
    const uint32_t w = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const uint32_t lo = w * p.S;
    const uint32_t hi = min(lo + p.S, p.N);
    const int K = (int)p.K;

    double eps[R][8], vi[R][8];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const uint32_t idx = lo + (uint32_t)r * kRun + 8u * (uint32_t)lane + (uint32_t)e;
            eps[r][e] = (idx < hi) ? p.eps[idx] : 0.0;
            vi[r][e] = 0.0;
        }
    if (tid < 16 * 3 * 2) {
        const int n = tid / 6, q = (tid / 2) % 3, half = tid & 1;
        const uint32_t code = half ? ((uint32_t)n >> 2) : ((uint32_t)n & 3u);
        const uint32_t want = q == 0 ? 2u : (q == 1 ? 0u : 1u);
        L.ind[n][q][half] = code == want ? 1.0 : 0.0;
    }
    if (tid == 0) {
        L.abort = 0;
        L.arms_err = 0;
        L.mu = p.scal[kBwMu];
        L.alpha = p.scal[kBwAlpha];
        L.sigma_b = p.scal[kBwSigmaB];
        L.d = p.scal[kBwD];
        L.betasq = 0.0;
    }
    if (tid < (int)p.nq - 1) bw_gh_nodes((int)p.nq, tid, &L.ghx[tid], &L.ghw[tid]);
    if (tid == 0) L.ghwc = kGhWc[(p.nq - 3) >> 1];
    for (uint32_t t = tid; t < p.F; t += kThreads) L.theta[t] = p.theta[t];
    if (tid < kMaxK) {
        L.cls[tid] = (tid < K - 1) ? p.cls[tid] : 0.0;
        L.pi[tid] = (tid < K) ? p.scal[kBwPi + tid] : 0.0;
        L.v[tid] = (tid < K) ? p.scal[kBwPi + K + tid] : 0.0;   // chain mode resets it per iteration
    }
    uint32_t rnd = __hip_atomic_load(p.ctl + kCtlRound, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned long long n_rounds = 0, n_changes = 0;
    uint64_t tk[kBwPhases] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint64_t t_last = __builtin_amdgcn_s_memrealtime();
    const uint64_t clk0 = __builtin_amdgcn_s_memtime(), rt0 = t_last;
    __syncthreads();

    // one value per workgroup: every wave holds the whole slice, so wave 0's sum
    auto allreduce = [&](double part, double* total) -> bool {
        if (L.abort) return false;
        part = wave_sum(part);
        ++rnd;
        ++n_rounds;
        if (tid == 0) publish_pair(p.xchg1, bw_slot1(p, rnd, 0, w), rnd, part);
        if (!bw_exchange<R>(p, L, rnd, 1, w, wave, lane)) return false;
        *total = L.xv[rnd & 1u][0];
        return true;
    };
    auto refresh_vi = [&]() {
        const double a = L.alpha;
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const uint32_t idx = lo + (uint32_t)r * kRun + 8u * (uint32_t)lane + (uint32_t)e;
                vi[r][e] = (idx < hi) ? exp(a * eps[r][e] - kEuler) : 0.0;
            }
    };
    // eps += sign * delta of slot k's column, then vi
    auto apply = [&](int k, double beta, double sign) {
        double dv[4];
        if constexpr (kDense) {
            // calculateResidualUpdate: Cx * beta (densebayeswkernel.cpp:20-23)
#pragma unroll
            for (int c = 0; c < 4; ++c) dv[c] = L.tab[k][c] * beta;
        } else {
            bw_delta(L.mean[k], L.sd[k], beta, dv);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const uint32_t bits = L.codes[k][r][lane >> 1] >> (16 * (lane & 1));
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const uint32_t idx = lo + (uint32_t)r * kRun + 8u * (uint32_t)lane + (uint32_t)e;
                const double d = wt_of((bits >> (2 * e)) & 3u, dv[0], dv[1], dv[2], dv[3]);
                if (idx < hi) eps[r][e] = (sign > 0.0) ? eps[r][e] + d : eps[r][e] - d;
            }
        }
        refresh_vi();
    };

    const uint32_t iters = (p.mode == kModeChain) ? p.iterations : 1u;
    const uint32_t nblocks = (p.M + kBwB - 1) / kBwB;
    for (uint32_t it = 0; it < iters; ++it) {
        const uint32_t iteration = p.iter_base + it;
        const uint32_t* order = p.order + (size_t)((p.mode == kModeChain) ? it : 0u) * p.M;

        if (p.mode == kModeChain || p.mode == kModeFixed) {
            // ---- sampleMu (bayeswbase.cpp:542-569; chain mode only), then sampleTheta
            // (bayeswbase.cpp:572-601, 783-788) of every covariate in turn, through one
            // sampler call site: task -1 is mu, task f the fixed effect f.
            // mu_dens (bayeswbase.cpp:66-77, sigma_mu = 100): the sum over individuals
            // factors as exp(-alpha x) sum_i exp(alpha (eps_i + mu) - EulerGamma).
            // theta_dens (bayeswbase.cpp:87-97): one grid sum per density value of the
            // residual shifted by X_f theta_f.
            double smu = 0.0;
            if (p.mode == kModeChain) {
                double part = 0.0;
#pragma unroll
                for (int r = 0; r < R; ++r)
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        const uint32_t idx = lo + (uint32_t)r * kRun + 8u * (uint32_t)lane + (uint32_t)e;
                        if (idx < hi) part += exp(L.alpha * (eps[r][e] + L.mu) - kEuler);
                    }
                if (!allreduce(part, &smu)) goto done;
            }
            for (int task = (p.mode == kModeChain) ? -1 : 0; task < (int)p.F; ++task) {
                const bool is_mu = task < 0;
                const float* xf = p.fixedX + (size_t)(is_mu ? 0 : task) * p.N;
                const double a = L.alpha, d = L.d;
                const double x0 = is_mu ? L.mu : L.theta[task];
                const double sf = is_mu ? 0.0 : p.sumf_fix[task];
                if (!is_mu) {
                    // params.epsilon = eps + X_f theta_f
#pragma unroll
                    for (int r = 0; r < R; ++r)
#pragma unroll
                        for (int e = 0; e < 8; ++e) {
                            const uint32_t idx = lo + (uint32_t)r * kRun + 8u * (uint32_t)lane + (uint32_t)e;
                            if (idx < hi) eps[r][e] = eps[r][e] + (double)xf[idx] * x0;
                        }
                }
                double xinit[4];
                if (is_mu) {
                    xinit[0] = 0.995 * x0; xinit[1] = x0; xinit[2] = 1.005 * x0; xinit[3] = 1.01 * x0;
                } else {
                    xinit[0] = x0 - 0.01; xinit[1] = x0; xinit[2] = x0 + 0.005; xinit[3] = x0 + 0.01;
                }
                double xs = x0;
                bool ok = true;
                const int err = arms_draw<false>(
                    L.env[wave], xinit, is_mu ? 2.0 : -2.0, is_mu ? 5.0 : 2.0, p.seed,
                    is_mu ? kSiteMu : kSiteTheta0 + (uint32_t)task, iteration, p.chain,
                    [&](double x) {
                        if (is_mu) return -a * x * d - exp(-a * x) * smu - x * x / (2 * 100.0);
                        double sum = 0.0;
#pragma unroll
                        for (int r = 0; r < R; ++r)
#pragma unroll
                            for (int e = 0; e < 8; ++e) {
                                const uint32_t idx = lo + (uint32_t)r * kRun + 8u * (uint32_t)lane + (uint32_t)e;
                                if (idx < hi) sum += exp((eps[r][e] - (double)xf[idx] * x) * a - kEuler);
                            }
                        double tot = 0.0;
                        if (!allreduce(sum, &tot)) ok = false;
                        return -a * x * sf - tot - x * x / (2 * 100.0);
                    },
                    &xs);
                if (!ok) goto done;
                if (err) {
                    if (tid == 0) { L.arms_err = err; raise_abort(p, is_mu ? 30 : 33); }
                    __syncthreads();
                    goto done;
                }
#pragma unroll
                for (int r = 0; r < R; ++r)
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        const uint32_t idx = lo + (uint32_t)r * kRun + 8u * (uint32_t)lane + (uint32_t)e;
                        if (idx < hi)
                            eps[r][e] = is_mu ? (eps[r][e] + x0) - xs : eps[r][e] - (double)xf[idx] * xs;
                    }
                __syncthreads();
                if (tid == 0) {
                    if (is_mu) L.mu = xs;
                    else L.theta[task] = xs;
                }
                if (is_mu) {
                    if (tid < kMaxK) L.v[tid] = (tid < K) ? 1.0 : 0.0;
                    if (tid == 0) L.betasq = 0.0;
                }
                __syncthreads();
            }
            if (p.mode == kModeFixed) break;
        }
        refresh_vi();
        __syncthreads();
        BW_TICK(6);

        // ---- the sweep (AnalysisGraph::exec over the order). Block b + 1's code
        // words and per-marker constants are loaded into registers while block b
        // runs and written to LDS when it ends; its marker ids come from LDS,
        // loaded one block earlier still.
        constexpr int kWords = kBwB * R * 32 / kThreads;
        uint32_t nw[kWords];
        uint32_t nmk = 0, noj = 0;
        double nbold = 0.0, nmean = 0.0, nsd = 0.0, nsumf = 0.0, npu = 0.0;
        double nt[4] = {0.0, 0.0, 0.0, 0.0};
        auto issue = [&](uint32_t nblk) {
            const uint32_t c0 = nblk * kBwB;
            const int cnb = (int)min((uint32_t)kBwB, p.M - c0);
            const uint32_t* ids = L.ord[nblk & 1u];
            if (tid < cnb) {
                const uint32_t j = ids[tid];
                nmk = j;
                nbold = bits_as_double(ld_agent((const uint64_t*)(p.beta + j)));
                if constexpr (kDense) {
                    // the column value of each 2-bit code (00 -> dosage 2, 01 missing, 10 -> 1, 11 -> 0)
                    const double* lt = p.lut + 4 * (size_t)j;
                    nt[0] = lt[2];
                    nt[1] = lt[3];
                    nt[2] = lt[1];
                    nt[3] = lt[0];
                } else {
                    nmean = p.stats[4 * (size_t)j];
                    nsd = p.stats[4 * (size_t)j + 1];
                }
                nsumf = p.sumf[j];
                if (p.mode == kModeChain) {
                    const U4 x = draw_block(p.seed, j, iteration, kStreamMarker, p.chain);
                    npu = unit_u53(x.x[0], x.x[1]);
                } else {
                    npu = p.u[j];
                }
            }
#pragma unroll
            for (int q = 0; q < kWords; ++q) {
                const uint32_t t = (uint32_t)tid + (uint32_t)q * kThreads;
                const uint32_t k = t / (R * 32u), rem = t % (R * 32u);
                const uint64_t byte_off = (uint64_t)(lo >> 2) + 4ull * rem;
                uint32_t word = 0xFFFFFFFFu;
                if ((int)k < cnb && byte_off < p.col_bytes)
                    word = p.codes[(size_t)ids[k] * p.col_words + (lo >> 4) + rem];
                nw[q] = word;
            }
        };
        auto commit = [&]() {
            if (tid < kBwB) {
                L.mk[tid] = nmk;
                L.bold[tid] = nbold;
                L.mean[tid] = nmean;
                L.sd[tid] = nsd;
#pragma unroll
                for (int c = 0; c < 4; ++c) L.tab[tid][c] = nt[c];
                L.sumf[tid] = nsumf;
                L.pu[tid] = npu;
            }
#pragma unroll
            for (int q = 0; q < kWords; ++q) (&L.codes[0][0][0])[tid + q * kThreads] = nw[q];
        };
        if (tid < 2 * kBwB) {
            const uint32_t idx = (uint32_t)tid;
            L.ord[idx / kBwB][idx % kBwB] = idx < p.M ? order[idx] : 0u;
        }
        __syncthreads();
        issue(0);
        commit();
        __syncthreads();
        BW_TICK(0);
        for (uint32_t blk = 0; blk < nblocks; ++blk) {
            const uint32_t b0 = blk * kBwB;
            const int nb = (int)min((uint32_t)kBwB, p.M - b0);
            if (blk + 1 < nblocks) issue(blk + 1);
            if (blk + 2 < nblocks && tid < kBwB) {
                const uint32_t idx = (blk + 2) * kBwB + (uint32_t)tid;
                noj = idx < p.M ? order[idx] : 0u;
            }

            int pos = 0;
            while (pos < nb) {
                // h: the first slot from pos whose marker is in the model (beta_old != 0).
                // Its step starts from eps + x beta_old. When h == pos the round is
                // that marker's own; otherwise the round publishes the sums of slots
                // [pos, h) and, speculatively, h's sums after the shift, used when no
                // slot before h changes eps.
                const unsigned long long hm = __ballot(pos + lane < nb && L.bold[pos + lane] != 0.0);
                const int h = __builtin_amdgcn_readfirstlane(hm ? pos + (int)__ffsll((long long)hm) - 1 : nb);
                const bool breaker = h == pos;
                const bool spec = !breaker && h < nb;
                const int V0 = breaker ? 0 : 1 + kPer * (h - pos);   // first value of the shifted sums
                ++rnd;
                ++n_rounds;
                int V;
                if (!breaker) {
                    // vi_sum of the slice and (vi_1, vi_2) of slots pos..h-1: wave q
                    // takes slots pos + q + 4 i, values 2 i (genotype 1) and 2 i + 1
                    // slots in groups of four whose code words are read together (a group
                    // at or past h is skipped; a slot at or past h within a group reads
                    // slot pos and contributes 0); the indicator pairs come from L.ind
                    double part[16], part8[8];
#pragma unroll
                    for (int g4 = 0; g4 < 2; ++g4) {
                        if (pos + wave + kWaves * 4 * g4 < h) {
                            int kc[4];
                            uint32_t bits[4][R];
#pragma unroll
                            for (int j = 0; j < 4; ++j) {
                                const int k = pos + wave + kWaves * (4 * g4 + j);
                                kc[j] = k < h ? k : pos;
#pragma unroll
                                for (int r = 0; r < R; ++r)
                                    bits[j][r] = L.codes[kc[j]][r][lane >> 1] >> (16 * (lane & 1));
                            }
#pragma unroll
                            for (int j = 0; j < 4; ++j) {
                                const int i = 4 * g4 + j;
                                double s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
                                for (int r = 0; r < R; ++r) {
#pragma unroll
                                    for (int q = 0; q < 4; ++q) {
                                        const uint32_t n = (bits[j][r] >> (4 * q)) & 15u;
                                        const dbl2 a = *reinterpret_cast<const dbl2*>(&L.ind[n][0][0]);
                                        const dbl2 b = *reinterpret_cast<const dbl2*>(&L.ind[n][1][0]);
                                        s1 = __builtin_fma(a.x, vi[r][2 * q], s1);
                                        s1 = __builtin_fma(a.y, vi[r][2 * q + 1], s1);
                                        s2 = __builtin_fma(b.x, vi[r][2 * q], s2);
                                        s2 = __builtin_fma(b.y, vi[r][2 * q + 1], s2);
                                        if constexpr (kDense) {
                                            const dbl2 m = *reinterpret_cast<const dbl2*>(&L.ind[n][2][0]);
                                            s3 = __builtin_fma(m.x, vi[r][2 * q], s3);
                                            s3 = __builtin_fma(m.y, vi[r][2 * q + 1], s3);
                                        }
                                    }
                                }
                                const bool live = pos + wave + kWaves * i < h;
                                part[2 * i] = live ? s1 : 0.0;
                                part[2 * i + 1] = live ? s2 : 0.0;
                                part8[i] = live ? s3 : 0.0;
                            }
                        } else {
#pragma unroll
                            for (int i = 4 * g4; i < 4 * g4 + 4; ++i) {
                                part[2 * i] = 0.0;
                                part[2 * i + 1] = 0.0;
                                part8[i] = 0.0;
                            }
                        }
                    }
                    const double tot = transpose_reduce16(part, lane);
                    const int vv = (lane >> 2) & 15;
                    const int k = pos + wave + kWaves * (vv >> 1);
                    if ((lane & 3) == 0 && k < h)
                        publish_pair(p.xchg1, bw_slot1(p, rnd, 1 + kPer * (k - pos) + (vv & 1), w), rnd, tot);
                    if constexpr (kDense) {
                        const double t8 = transpose_reduce8(part8, lane);
                        const int k8 = pos + wave + kWaves * ((lane >> 3) & 7);
                        if ((lane & 7) == 0 && k8 < h)
                            publish_pair(p.xchg1, bw_slot1(p, rnd, 1 + kPer * (k8 - pos) + 2, w), rnd, t8);
                    }
                    double sv = 0.0;
#pragma unroll
                    for (int r = 0; r < R; ++r)
#pragma unroll
                        for (int e = 0; e < 8; ++e) sv += vi[r][e];
                    sv = wave_sum_full(sv, lane);
                    if (tid == 0) publish_pair(p.xchg1, bw_slot1(p, rnd, 0, w), rnd, sv);
                    V = V0;
                    if (spec) {
                        // h's sums of vi' = exp(alpha (eps + delta_h) - EulerGamma) (wave 3)
                        if (wave == kWaves - 1) {
                            double dv[4];
                            if constexpr (kDense) {
#pragma unroll
                                for (int c = 0; c < 4; ++c) dv[c] = L.tab[h][c] * L.bold[h];
                            } else {
                                bw_delta(L.mean[h], L.sd[h], L.bold[h], dv);
                            }
                            const double a = L.alpha;
                            double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
                            for (int r = 0; r < R; ++r) {
                                const uint32_t bits = L.codes[h][r][lane >> 1] >> (16 * (lane & 1));
#pragma unroll
                                for (int e = 0; e < 8; ++e) {
                                    const uint32_t idx = lo + (uint32_t)r * kRun + 8u * (uint32_t)lane + (uint32_t)e;
                                    const uint32_t c = (bits >> (2 * e)) & 3u;
                                    const double es = eps[r][e] + wt_of(c, dv[0], dv[1], dv[2], dv[3]);
                                    const double x = (idx < hi) ? exp(a * es - kEuler) : 0.0;
                                    s0 += x;
                                    s1 += (c == 2u) ? x : 0.0;
                                    s2 += (c == 0u) ? x : 0.0;
                                    if constexpr (kDense) s3 += (c == 1u) ? x : 0.0;
                                }
                            }
                            s0 = wave_sum_full(s0, lane);
                            s1 = wave_sum_full(s1, lane);
                            s2 = wave_sum_full(s2, lane);
                            if constexpr (kDense) s3 = wave_sum_full(s3, lane);
                            if (lane == 0) {
                                publish_pair(p.xchg1, bw_slot1(p, rnd, V0, w), rnd, s0);
                                publish_pair(p.xchg1, bw_slot1(p, rnd, V0 + 1, w), rnd, s1);
                                publish_pair(p.xchg1, bw_slot1(p, rnd, V0 + 2, w), rnd, s2);
                                if constexpr (kDense) publish_pair(p.xchg1, bw_slot1(p, rnd, V0 + 3, w), rnd, s3);
                            }
                        }
                        V = V0 + 1 + kPer;
                    }
                } else {
                    // a marker in the model: eps += x beta_old, vi re-formed, then its sums
                    apply(pos, L.bold[pos], 1.0);
                    BW_TICK(7);
                    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const uint32_t bits = L.codes[pos][r][lane >> 1] >> (16 * (lane & 1));
#pragma unroll
                        for (int e = 0; e < 8; ++e) {
                            const uint32_t c = (bits >> (2 * e)) & 3u;
                            s0 += vi[r][e];
                            s1 += (c == 2u) ? vi[r][e] : 0.0;
                            s2 += (c == 0u) ? vi[r][e] : 0.0;
                            if constexpr (kDense) s3 += (c == 1u) ? vi[r][e] : 0.0;
                        }
                    }
                    s0 = wave_sum_full(s0, lane);
                    s1 = wave_sum_full(s1, lane);
                    s2 = wave_sum_full(s2, lane);
                    if constexpr (kDense) s3 = wave_sum_full(s3, lane);
                    if (tid == 0) {
                        publish_pair(p.xchg1, bw_slot1(p, rnd, 0, w), rnd, s0);
                        publish_pair(p.xchg1, bw_slot1(p, rnd, 1, w), rnd, s1);
                        publish_pair(p.xchg1, bw_slot1(p, rnd, 2, w), rnd, s2);
                        if constexpr (kDense) publish_pair(p.xchg1, bw_slot1(p, rnd, 3, w), rnd, s3);
                    }
                    V = 1 + kPer;
                }
                BW_TICK(1);
                if (!bw_exchange<R>(p, L, rnd, V, w, wave, lane)) goto done;
                const double* xv = L.xv[rnd & 1u];
                BW_TICK(2);

                // ---- decisions. Class likelihoods of the slots [pos, h] (or h alone in
                // its own round): first the weighted integrand terms (slot, slab class,
                // node) over all threads, then each (slot, class) sum in the reference's
                // order, then the pick (wave 0, lane = slot). Slot h always stops the
                // round: its step changes eps whatever its class.
                // Sums of slot sl: (vi_sum, vi_1, vi_2) at xv[0], xv[1 + 2 sl], xv[2 + 2 sl],
                // or the shifted sums at xv[V0 ..] for the speculative slot h.
                {
                    const int ncls = K - 1, nn = (int)p.nq - 1, per = ncls * nn;
                    const int ns = breaker ? 1 : (spec ? h + 1 - pos : nb - pos);
                    for (int t = tid; t < ns * per; t += kThreads) {
                        const int sl = t / per, rem = t - sl * per, i = rem / nn, n = rem - i * nn;
                        const int k = pos + sl;
                        const bool sh = spec && k == h;
                        const int b = sh ? V0 + 1 : 1 + kPer * sl;
                        const double vs = sh ? xv[V0] : xv[0];
                        const double s1 = xv[b], s2 = xv[b + 1];
                        const double Ck = L.cls[i];
                        const double sq = sqrt(2 * Ck * L.sigma_b);
                        double f;
                        if constexpr (kDense) {
                            const double s3 = xv[b + 2];
                            const BwGroups g{{vs - s1 - s2 - s3, s1, s2, s3},
                                             {L.tab[k][3], L.tab[k][2], L.tab[k][0], L.tab[k][1]}};
                            const double sigma =
                                1.0 / sqrt(1 + L.alpha * L.alpha * L.sigma_b * Ck * bw_dense_exp_sum(g));
                            f = bw_dense_integrand(sigma * L.ghx[n], L.alpha, sq, L.sumf[k], g);
                        } else {
                            const BwSums v{vs, vs - s1 - s2, s1, s2};
                            const double mean = L.mean[k], sd = L.sd[k];
                            const double sigma = bw_sigma(L.alpha, L.sigma_b, Ck, mean, sd, v);
                            f = bw_integrand(sigma * L.ghx[n], L.alpha, sq, L.sumf[k], mean / sd, sd, v);
                        }
                        L.gv[sl][i][n] = L.ghw[n] * f;
                    }
                    __syncthreads();
                    for (int t = tid; t < ns * ncls; t += kThreads) {
                        const int sl = t / ncls, i = t - sl * ncls;
                        const int k = pos + sl;
                        const bool sh = spec && k == h;
                        const int b = sh ? V0 + 1 : 1 + kPer * sl;
                        const double vs = sh ? xv[V0] : xv[0];
                        const double s1 = xv[b], s2 = xv[b + 1];
                        double sigma;
                        if constexpr (kDense) {
                            const double s3 = xv[b + 2];
                            const BwGroups g{{vs - s1 - s2 - s3, s1, s2, s3},
                                             {L.tab[k][3], L.tab[k][2], L.tab[k][0], L.tab[k][1]}};
                            sigma = 1.0 / sqrt(1 + L.alpha * L.alpha * L.sigma_b * L.cls[i] * bw_dense_exp_sum(g));
                        } else {
                            const BwSums v{vs, vs - s1 - s2, s1, s2};
                            sigma = bw_sigma(L.alpha, L.sigma_b, L.cls[i], L.mean[k], L.sd[k], v);
                        }
                        double temp = L.gv[sl][i][0];
                        for (int n = 1; n < nn; ++n) temp = temp + L.gv[sl][i][n];
                        temp = temp + L.ghwc;
                        L.ml[sl][i + 1] = L.pi[i + 1] * (sigma * temp);
                    }
                    __syncthreads();
                    if (wave == 0) {
                        const int k = pos + lane;
                        const bool active = lane < ns;
                        int sel = -1;
                        if (active) {
                            L.ml[lane][0] = L.pi[0] * kSqrtPi;
                            sel = bw_pick(K, L.ml[lane], L.pu[k]);
                        }
                        const bool stop = active && (breaker || k == h || sel > 0);
                        const unsigned long long sm = __ballot(stop);
                        const int c = sm ? pos + (int)__ffsll((long long)sm) - 1 : nb;
                        const int cs = lane_read(sel, (c - pos) & 63);
                        // slots before c keep beta = 0 in class 0; slot k's outputs are
                        // written by workgroup k % P
                        const bool keep0 = k < c && active;
                        const unsigned long long nk = __ballot(keep0);
                        if (lane == 0 && nk) L.v[0] += (double)__popcll(nk);
                        if (keep0 && ((uint32_t)k < p.P ? (uint32_t)k == w : (uint32_t)k % p.P == w)) {
                            st_agent((uint64_t*)(p.beta + L.mk[k]), double_as_bits(0.0));
                            __hip_atomic_store(p.comp + L.mk[k], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        }
                        if (lane == 0) {
                            L.stop = c;
                            L.sel = (c < nb) ? cs : -1;
                        }
                    }
                }
                __syncthreads();
                BW_TICK(3);
                {
                    const int c = __builtin_amdgcn_readfirstlane(L.stop);
                    const int sel = __builtin_amdgcn_readfirstlane(L.sel);
                    if (c >= nb) {
                        pos = nb;
                    } else {
                        // slot c's sums; a speculated in-model marker first takes its shift
                        const bool sh = spec && c == h;
                        if (sh) {
                            apply(c, L.bold[c], 1.0);
                            BW_TICK(7);
                        }
                        const int i1 = sh ? V0 + 1 : 1 + kPer * (c - pos);
                        // draw the slot's beta: 0 for class 0, else ARMS on beta_dens
                        double bnew = 0.0;
                        if (sel > 0) {
                            const double s1 = xv[i1];
                            const double s2 = xv[i1 + 1];
                            const double vs = xv[sh ? V0 : 0];
                            const double a = L.alpha, sb = L.sigma_b, Ck = L.cls[sel - 1];
                            const double sf = L.sumf[c];
                            const double bo = L.bold[c];
                            const double safe = 2 * sqrt(sb * Ck);
                            const double xinit[4] = {bo - safe / 10, bo, bo + safe / 20, bo + safe / 10};
                            int err;
                            if constexpr (kDense) {
                                const double s3 = xv[i1 + 2];
                                const BwGroups g{{vs - s1 - s2 - s3, s1, s2, s3},
                                                 {L.tab[c][3], L.tab[c][2], L.tab[c][0], L.tab[c][1]}};
                                err = arms_draw<true>(
                                    L.env[wave], xinit, bo - safe, bo + safe, p.seed, L.mk[c], iteration, p.chain,
                                    [&](double x) {
                                        // DenseBayesW beta_dens (densebayesw.cpp:25-33): sum_i exp((eps_i - Cx_i x)
                                        // alpha - EulerGamma) = sum over groups V_d exp(-Cx_d x alpha)
                                        double t = 0.0;
#pragma unroll
                                        for (int d = 0; d < 4; ++d) t += g.V[d] * exp(-(g.c[d] * x) * a);
                                        return -a * x * sf - t - x * x / (2 * Ck * sb);
                                    },
                                    &bnew);
                            } else {
                                const double v0 = vs - s1 - s2;
                                const double msr = L.mean[c] / L.sd[c], sd = L.sd[c];
                                err = arms_draw<true>(
                                    L.env[wave], xinit, bo - safe, bo + safe, p.seed, L.mk[c], iteration, p.chain,
                                    [&](double x) {
                                        // beta_dens (sparsebayesw.cpp:26-37)
                                        return -a * x * sf -
                                               exp(a * x * msr) *
                                                   (v0 + s1 * exp(-a * x / sd) + s2 * exp(-2 * a * x / sd)) -
                                               x * x / (2 * Ck * sb);
                                    },
                                    &bnew);
                            }
                            if (err) {
                                if (tid == 0) { L.arms_err = err; raise_abort(p, 31); }
                                __syncthreads();
                                goto done;
                            }
                            BW_TICK(4);
                        }
                        // skip iff beta_old == 0 && beta_new == 0 (bayeswbase.cpp:691-696)
                        if (!(L.bold[c] == 0.0 && bnew == 0.0)) {
                            apply(c, bnew, -1.0);
                            ++n_changes;
                        }
                        __syncthreads();
                        BW_TICK(5);
                        if (tid == 0) {
                            L.v[sel] += 1.0;
                            L.betasq += bnew * bnew;
                            if ((uint32_t)c < p.P ? (uint32_t)c == w : (uint32_t)c % p.P == w) {
                                st_agent((uint64_t*)(p.beta + L.mk[c]), double_as_bits(bnew));
                                __hip_atomic_store(p.comp + L.mk[c], sel, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            }
                        }
                        pos = c + 1;
                    }
                }
                __syncthreads();
            }
            __syncthreads();
            if (blk + 1 < nblocks) commit();
            if (blk + 2 < nblocks && tid < kBwB) L.ord[blk & 1u][tid] = noj;   // block blk + 2 has blk's parity
            __syncthreads();
            BW_TICK(0);
        }
        // other workgroups read this workgroup's beta stores only in the block
        // loads of the next iteration, which follow this workgroup's next publish
        __builtin_amdgcn_s_waitcnt(0);

        if (p.mode == kModeChain) {
            // ---- sampleAlpha (bayeswbase.cpp:707-734): one grid sum per density value
            double part = 0.0;
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const uint32_t idx = lo + (uint32_t)r * kRun + 8u * (uint32_t)lane + (uint32_t)e;
                    if (idx < hi) part += eps[r][e] * p.fail[idx];
                }
            double ef;
            if (!allreduce(part, &ef)) goto done;
            const double a0 = L.alpha, d = L.d;
            bool ok = true;
            const double xinit[4] = {a0 * 0.5, a0, a0 * 1.15, a0 * 1.5};
            double xs = a0;
            // alpha_dens (bayeswbase.cpp:109-119), alpha_0 = kappa_0 = 0.01
            const int err = arms_draw<false>(
                L.env[wave], xinit, 0.0, 20.0, p.seed, kSiteAlpha, iteration, p.chain,
                [&](double x) {
                    double s = 0.0;
#pragma unroll
                    for (int r = 0; r < R; ++r)
#pragma unroll
                        for (int e = 0; e < 8; ++e) {
                            const uint32_t idx = lo + (uint32_t)r * kRun + 8u * (uint32_t)lane + (uint32_t)e;
                            if (idx < hi) s += exp(eps[r][e] * x - kEuler);
                        }
                    double tot = 0.0;
                    if (!allreduce(s, &tot)) ok = false;
                    return (0.01 + d - 1) * log(x) + x * (ef - 0.01) - tot;
                },
                &xs);
            if (!ok) goto done;
            if (err) {
                if (tid == 0) { L.arms_err = err; raise_abort(p, 32); }
                __syncthreads();
                goto done;
            }
            __syncthreads();
            if (tid == 0) {
                L.alpha = xs;
                // sigma_b ~ InvGamma(1 + (M - v0 + 1)/2, 1e-4 + (M - v0 + 1)/2 |beta|^2), pi_L ~ Dir(v)
                const double nz = (double)p.M - L.v[0] + 1;
                const double shape = 1.0 + 0.5 * nz;
                const double scale = 0.0001 + 0.5 * nz * L.betasq;
                L.sigma_b = 1.0 / rgamma(p.seed, iteration, kBwDrawSigmaB, p.chain, shape, 1.0 / scale);
                double sum = 0.0;
                double gs[kMaxK];
                for (int k = 0; k < K; ++k) {
                    gs[k] = rgamma(p.seed, iteration, kBwDrawPi0 + (uint32_t)k, p.chain, L.v[k], 1.0);
                    sum += gs[k];
                }
                for (int k = 0; k < K; ++k) L.pi[k] = gs[k] / sum;
                if (w == 0 && p.hist) {
                    double* h = p.hist + (size_t)it * 5u;
                    h[0] = L.mu; h[1] = L.alpha; h[2] = L.sigma_b; h[3] = nz; h[4] = L.betasq;
                }
            }
            __syncthreads();
            BW_TICK(6);
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
        if (w == 0)
            for (uint32_t t = tid; t < p.F; t += kThreads) p.theta[t] = L.theta[t];
        if (w == 0 && tid == 0) {
            p.scal[kBwMu] = L.mu;
            p.scal[kBwAlpha] = L.alpha;
            p.scal[kBwSigmaB] = L.sigma_b;
            p.scal[kBwBetaSq] = L.betasq;
            for (int k = 0; k < K; ++k) {
                p.scal[kBwPi + k] = L.pi[k];
                p.scal[kBwPi + K + k] = L.v[k];
            }
            __hip_atomic_store(p.ctl + kCtlRound, rnd, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            p.counters[0] = n_rounds;
            p.counters[1] = n_changes;
            for (int q = 0; q < 6; ++q) p.counters[2 + q] = tk[q];
            for (int q = 6; q < kBwPhases; ++q) p.counters[4 + q] = tk[q];
            // shader clock and 100 MHz reference clock over the launch: the core clock
            p.counters[12] = __builtin_amdgcn_s_memtime() - clk0;
            p.counters[13] = __builtin_amdgcn_s_memrealtime() - rt0;
        }
    } else if (w == 0 && tid == 0) {
        p.counters[9] = (unsigned long long)L.arms_err;
    }
}

// calculateSumFailure (raggedbayeswkernel.cpp:18-25) from the 2-bit codes: one
// workgroup per marker; list sums pass through an int as in the reference.
static __global__ void bw_sum_failure_kernel(const uint32_t* codes, uint64_t col_words, uint32_t N, const double* fail,
                                      const double* stats, double* sumf)
{
    __shared__ double r1[kWaves], r2[kWaves], rf[kWaves];
    const uint32_t j = blockIdx.x;
    const int tid = threadIdx.x;
    const uint8_t* col = (const uint8_t*)(codes + (size_t)j * col_words);
    double s1 = 0.0, s2 = 0.0, sf = 0.0;
    for (uint32_t i = tid; i < N; i += kThreads) {
        const uint32_t c = (col[i >> 2] >> ((i & 3u) * 2u)) & 3u;
        const double f = fail[i];
        s1 += (c == 2u) ? f : 0.0;
        s2 += (c == 0u) ? f : 0.0;
        sf += f;
    }
    s1 = wave_sum(s1);
    s2 = wave_sum(s2);
    sf = wave_sum(sf);
    if ((tid & 63) == 0) { r1[tid >> 6] = s1; r2[tid >> 6] = s2; rf[tid >> 6] = sf; }
    __syncthreads();
    if (tid == 0) {
        const double a1 = ((r1[0] + r1[1]) + r1[2]) + r1[3];
        const double a2 = ((r2[0] + r2[1]) + r2[2]) + r2[3];
        const double af = ((rf[0] + rf[1]) + rf[2]) + rf[3];
        int temp_sum = 0;
        temp_sum += (int)a1;
        temp_sum += (int)(2 * a2);
        sumf[j] = (temp_sum - stats[4 * (size_t)j] * af) / stats[4 * (size_t)j + 1];
    }
}

// DenseBayesWKernel::calculateSumFailure (densebayeswkernel.cpp:15-18): sum_i Cx_i f_i.
static __global__ void bw_dense_sum_failure_kernel(const uint32_t* codes, uint64_t col_words, uint32_t N, const double* fail,
                                            const double* lut, double* sumf)
{
    __shared__ double red[kWaves];
    const uint32_t j = blockIdx.x;
    const int tid = threadIdx.x;
    const uint8_t* col = (const uint8_t*)(codes + (size_t)j * col_words);
    const double* lt = lut + 4 * (size_t)j;
    const double w[4] = {lt[2], lt[3], lt[1], lt[0]};   // by 2-bit code
    double s = 0.0;
    for (uint32_t i = tid; i < N; i += kThreads) {
        const uint32_t c = (col[i >> 2] >> ((i & 3u) * 2u)) & 3u;
        s += w[c] * fail[i];
    }
    s = wave_sum(s);
    if ((tid & 63) == 0) red[tid >> 6] = s;
    __syncthreads();
    if (tid == 0) sumf[j] = ((red[0] + red[1]) + red[2]) + red[3];
}

// RaggedBayesWKernel::setVi (raggedbayeswkernel.cpp:10-16) of marker j for
// sum_k X(k,f) failure_k of each covariate (BayesWBase::init, bayeswbase.cpp:532-537):
// one workgroup per covariate.
static __global__ void bw_fixed_sumf_kernel(const float* X, uint32_t N, const double* fail, double* out)
{
    __shared__ double red[kWaves];
    const uint32_t f = blockIdx.x;
    const int tid = threadIdx.x;
    const float* x = X + (size_t)f * N;
    double s = 0.0;
    for (uint32_t i = tid; i < N; i += kThreads) s += (double)x[i] * fail[i];
    s = wave_sum(s);
    if ((tid & 63) == 0) red[tid >> 6] = s;
    __syncthreads();
    if (tid == 0) out[f] = ((red[0] + red[1]) + red[2]) + red[3];
}

// vi = exp(alpha eps - EulerGamma): out = {vi_sum, vi_0, vi_1, vi_2}. One workgroup.
static __global__ void bw_set_vi_kernel(const uint32_t* codes, uint64_t col_words, uint32_t N, const double* eps,
                                 double alpha, uint32_t j, double* out)
{
    __shared__ double r0[kWaves], r1[kWaves], r2[kWaves];
    const int tid = threadIdx.x;
    const uint8_t* col = (const uint8_t*)(codes + (size_t)j * col_words);
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (uint32_t i = tid; i < N; i += kThreads) {
        const uint32_t c = (col[i >> 2] >> ((i & 3u) * 2u)) & 3u;
        const double v = exp(alpha * eps[i] - kEuler);
        s0 += v;
        s1 += (c == 2u) ? v : 0.0;
        s2 += (c == 0u) ? v : 0.0;
    }
    s0 = wave_sum(s0);
    s1 = wave_sum(s1);
    s2 = wave_sum(s2);
    if ((tid & 63) == 0) { r0[tid >> 6] = s0; r1[tid >> 6] = s1; r2[tid >> 6] = s2; }
    __syncthreads();
    if (tid == 0) {
        const double vs = ((r0[0] + r0[1]) + r0[2]) + r0[3];
        const double v1 = ((r1[0] + r1[1]) + r1[2]) + r1[3];
        const double v2 = ((r2[0] + r2[1]) + r2[2]) + r2[3];
        out[0] = vs;
        out[1] = vs - v1 - v2;
        out[2] = v1;
        out[3] = v2;
    }
}

// eps += sign * calculateResidualUpdate(beta) of marker j (raggedbayeswkernel.cpp:27-43).
// Dense layout (lut != null): Cx * beta (densebayeswkernel.cpp:20-23).
static __global__ void bw_residual_kernel(const uint32_t* codes, uint64_t col_words, uint32_t N, const double* stats,
                                   const double* lut, uint32_t j, double beta, int sign, double* eps)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const uint8_t* col = (const uint8_t*)(codes + (size_t)j * col_words);
    const uint32_t c = (col[i >> 2] >> ((i & 3u) * 2u)) & 3u;
    double dv[4];
    if (lut) {
        const double* lt = lut + 4 * (size_t)j;
        dv[0] = lt[2] * beta;
        dv[1] = lt[3] * beta;
        dv[2] = lt[1] * beta;
        dv[3] = lt[0] * beta;
    } else {
        bw_delta(stats[4 * (size_t)j], stats[4 * (size_t)j + 1], beta, dv);
    }
    const double d = dv[c];
    eps[i] = (sign > 0) ? eps[i] + d : eps[i] - d;
}

} // namespace brr
