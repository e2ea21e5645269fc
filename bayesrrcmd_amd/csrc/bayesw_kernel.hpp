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
constexpr int kBwV = 1 + 2 * kBwB;             // values per round: vi_sum, (vi_1, vi_2) per slot
constexpr int kArmsN = 100;                    // ARMS envelope points (npoint of bayeswbase.cpp)
constexpr uint32_t kStreamArms = 4;
constexpr uint32_t kSiteMu = 0xFFFFFFF0u;
constexpr uint32_t kSiteAlpha = 0xFFFFFFF1u;
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
    const double* stats;          // M x 4 {mean, sd, sqrdZ, Zsum}
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
};

enum : int { kBwMu = 0, kBwAlpha = 1, kBwSigmaB = 2, kBwD = 3, kBwBetaSq = 4, kBwPi = 5 };

struct ArmsPoint {
    double x, y, ey, cum;
    int f, pl, pr, pad;
};

// One sampler run: the envelope (kArmsN points + the working point), the
// resumption state and the random stream position. Lives in LDS, one per wave.
struct ArmsEnv {
    ArmsPoint p[kArmsN + 1];
    double xinit[4];
    double ymax, yu, ynew;
    double xout;                  // point to evaluate next, or the sample
    uint64_t seed;
    uint32_t site, iteration, chain, k;
    int cpoint, state, kinit, q, guard, pad;
};

template <int R>
struct BwShared {
    uint32_t codes[kBwB][R][32];
    uint32_t mk[kBwB];
    double bold[kBwB], mean[kBwB], sd[kBwB], sumf[kBwB], pu[kBwB];
    double xv[2][kBwV];           // round totals, by round parity
    double cls[kMaxK], pi[kMaxK], v[kMaxK];
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

#define AP(i) (e.p[(i)])

// intersection of the chords either side of intersection point q
// (BayesW_arms.cpp:707-808); nonzero on a convexity violation
__device__ __forceinline__ int arms_meet(ArmsEnv& e, int q)
{
    double gl = 0, gr = 0, grl = 0, dl = 0, dr = 0;
    int il = 0, ir = 0, irl = 0;
    if (AP(q).f) return 3;
    const int ql = AP(q).pl, qr = AP(q).pr;
    if (ql >= 0 && AP(ql).pl >= 0 && AP(AP(ql).pl).pl >= 0) {
        const int q3 = AP(AP(ql).pl).pl;
        gl = (AP(ql).y - AP(q3).y) / (AP(ql).x - AP(q3).x);
        il = 1;
    }
    if (qr >= 0 && AP(qr).pr >= 0 && AP(AP(qr).pr).pr >= 0) {
        const int q3 = AP(AP(qr).pr).pr;
        gr = (AP(qr).y - AP(q3).y) / (AP(qr).x - AP(q3).x);
        ir = 1;
    }
    if (ql >= 0 && qr >= 0) {
        grl = (AP(qr).y - AP(ql).y) / (AP(qr).x - AP(ql).x);
        irl = 1;
    }
    if (irl && il && gl < grl) return 1;
    if (irl && ir && gr > grl) return 1;
    if (il && irl) {
        dr = (gl - grl) * (AP(qr).x - AP(ql).x);
        if (dr < 0.1) dr = 0.1;
    }
    if (ir && irl) {
        dl = (grl - gr) * (AP(qr).x - AP(ql).x);
        if (dl < 0.1) dl = 0.1;
    }
    if (il && ir && irl) {
        AP(q).x = (dl * AP(qr).x + dr * AP(ql).x) / (dl + dr);
        AP(q).y = (dl * AP(qr).y + dr * AP(ql).y + dl * dr) / (dl + dr);
    } else if (il && irl) {
        AP(q).x = AP(qr).x;
        AP(q).y = AP(qr).y + dr;
    } else if (ir && irl) {
        AP(q).x = AP(ql).x;
        AP(q).y = AP(ql).y + dl;
    } else if (il) {
        AP(q).y = AP(ql).y + gl * (AP(q).x - AP(ql).x);
    } else if (ir) {
        AP(q).y = AP(qr).y - gr * (AP(qr).x - AP(q).x);
    } else {
        return 4;
    }
    if ((ql >= 0 && AP(q).x < AP(ql).x) || (qr >= 0 && AP(q).x > AP(qr).x)) return 5;
    return 0;
}

// exponentiate and integrate the envelope (BayesW_arms.cpp:672-703, 812-834)
__device__ __forceinline__ void arms_cumulate(ArmsEnv& e)
{
    int lm = 0;
    while (AP(lm).pl >= 0) lm = AP(lm).pl;
    double ymax = AP(lm).y;
    for (int q = AP(lm).pr; q >= 0; q = AP(q).pr)
        if (AP(q).y > ymax) ymax = AP(q).y;
    e.ymax = ymax;
    for (int q = lm; q >= 0; q = AP(q).pr) AP(q).ey = arms_expshift(AP(q).y, ymax);
    AP(lm).cum = 0.0;
    for (int q = AP(lm).pr; q >= 0; q = AP(q).pr) {
        const int l = AP(q).pl;
        double a;
        if (AP(l).x == AP(q).x) a = 0.0;
        else if (fabs(AP(q).y - AP(l).y) < 0.1) a = 0.5 * (AP(q).ey + AP(l).ey) * (AP(q).x - AP(l).x);
        else a = ((AP(q).ey - AP(l).ey) / (AP(q).y - AP(l).y)) * (AP(q).x - AP(l).x);
        AP(q).cum = AP(l).cum + a;
    }
}

// the x whose envelope cumulative probability is prob, into the working point
// w (BayesW_arms.cpp:364-425)
__device__ __forceinline__ int arms_invert(ArmsEnv& e, double prob, int w)
{
    int q = 0;
    while (AP(q).pr >= 0) q = AP(q).pr;
    const double u = prob * AP(q).cum;
    while (AP(AP(q).pl).cum > u) q = AP(q).pl;
    const int l = AP(q).pl;
    AP(w).pl = l;
    AP(w).pr = q;
    AP(w).f = 0;
    AP(w).cum = u;
    const double prop = (u - AP(l).cum) / (AP(q).cum - AP(l).cum);
    double xl = AP(q).x, xr = AP(q).x;
    if (AP(l).x == AP(q).x) {
        AP(w).x = AP(q).x;
        AP(w).y = AP(q).y;
        AP(w).ey = AP(q).ey;
    } else {
        xl = AP(l).x;
        xr = AP(q).x;
        const double yl = AP(l).y, yr = AP(q).y, eyl = AP(l).ey, eyr = AP(q).ey;
        if (fabs(yr - yl) < 0.1) {
            if (fabs(eyr - eyl) > 0.001 * fabs(eyr + eyl))
                AP(w).x = xl + ((xr - xl) / (eyr - eyl)) * (-eyl + sqrt((1. - prop) * eyl * eyl + prop * eyr * eyr));
            else
                AP(w).x = xl + (xr - xl) * prop;
            AP(w).ey = ((AP(w).x - xl) / (xr - xl)) * (eyr - eyl) + eyl;
            AP(w).y = arms_logshift(AP(w).ey, e.ymax);
        } else {
            AP(w).x = xl + ((xr - xl) / (yr - yl)) * (-yl + arms_logshift(((1. - prop) * eyl + prop * eyr), e.ymax));
            AP(w).y = ((AP(w).x - xl) / (xr - xl)) * (yr - yl) + yl;
            AP(w).ey = arms_expshift(AP(w).y, e.ymax);
        }
    }
    if (AP(w).x < xl || AP(w).x > xr) return 6;
    return 0;
}

enum : int { kArmsEval = 0, kArmsDone = 1 };
enum : int { kAsInit = 0, kAsSample = 1, kAsEvalW = 2, kAsEvalQ = 3 };

// Sets up one draw (arms with ninit = 4, npoint = 100, nsamp = 1 and no
// Metropolis step, BayesW_arms.cpp:133-340); 0 or the reference's error code.
__device__ __forceinline__ int arms_begin(ArmsEnv& e, const double (&xinit)[4], double xl, double xr, uint64_t seed,
                                          uint32_t site, uint32_t iteration, uint32_t chain)
{
    constexpr int ninit = 4, mpoint = 2 * ninit + 1;
    if (xinit[0] <= xl || xinit[ninit - 1] >= xr) return 1003;
    for (int i = 1; i < ninit; ++i)
        if (xinit[i] <= xinit[i - 1]) return 1004;
    for (int j = 0; j < mpoint; ++j) {
        AP(j).pl = j - 1;
        AP(j).pr = (j + 1 < mpoint) ? j + 1 : -1;
        AP(j).f = (j & 1);
        AP(j).y = 0.0;
        AP(j).x = (j & 1) ? xinit[j >> 1] : 0.0;
    }
    AP(0).x = xl;
    AP(mpoint - 1).x = xr;
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
__device__ __noinline__ int arms_step(ArmsEnv& e, double y)
{
    constexpr int mpoint = 9, w = kArmsN;
    if (e.state == kAsInit) {
        if (e.kinit > 0) AP(2 * e.kinit - 1).y = y;
        if (e.kinit < 4) {
            e.xout = AP(2 * e.kinit + 1).x;
            e.kinit++;
            return kArmsEval;
        }
        for (int j = 0; j < mpoint; j += 2)
            if (arms_meet(e, j)) return 2000;
        arms_cumulate(e);
        e.cpoint = mpoint;
        e.state = kAsSample;
    } else {
        if (e.state == kAsEvalW) {
            // the working point was evaluated (BayesW_arms.cpp:505-536)
            e.ynew = y;
            AP(w).y = y;
            AP(w).ey = arms_expshift(y, e.ymax);
            AP(w).f = 1;
            if (e.cpoint <= kArmsN - 2) {
                // update: insert it with a new intersection (BayesW_arms.cpp:540-668)
                const int q = e.cpoint++;
                AP(q).x = AP(w).x;
                AP(q).y = AP(w).y;
                AP(q).f = 1;
                const int m = e.cpoint++;
                AP(m).f = 0;
                const int wl = AP(w).pl, wr = AP(w).pr;
                if (AP(wl).f && !AP(wr).f) {
                    AP(m).pl = wl; AP(m).pr = q; AP(q).pl = m; AP(q).pr = wr; AP(wl).pr = m; AP(wr).pl = q;
                } else if (!AP(wl).f && AP(wr).f) {
                    AP(m).pr = wr; AP(m).pl = q; AP(q).pr = m; AP(q).pl = wl; AP(wr).pl = m; AP(wl).pr = q;
                } else {
                    return 2000;
                }
                e.q = q;
                const int ql = (AP(AP(q).pl).pl >= 0) ? AP(AP(q).pl).pl : AP(q).pl;
                const int qr = (AP(AP(q).pr).pr >= 0) ? AP(AP(q).pr).pr : AP(q).pr;
                double xa = 0.0;
                bool move = false;
                if (AP(q).x < (1. - 0.00001) * AP(ql).x + 0.00001 * AP(qr).x) {
                    xa = (1. - 0.00001) * AP(ql).x + 0.00001 * AP(qr).x;
                    move = true;
                } else if (AP(q).x > 0.00001 * AP(ql).x + (1. - 0.00001) * AP(qr).x) {
                    xa = 0.00001 * AP(ql).x + (1. - 0.00001) * AP(qr).x;
                    move = true;
                }
                if (move) {
                    AP(q).x = xa;
                    e.xout = xa;
                    e.state = kAsEvalQ;
                    return kArmsEval;
                }
                e.state = kAsEvalQ;
                y = AP(q).y;
            }
        }
        if (e.state == kAsEvalQ) {
            const int q = e.q;
            AP(q).y = y;
            if (arms_meet(e, AP(q).pl)) return 2000;
            if (arms_meet(e, AP(q).pr)) return 2000;
            if (AP(AP(q).pl).pl >= 0 && arms_meet(e, AP(AP(AP(q).pl).pl).pl)) return 2000;
            if (AP(AP(q).pr).pr >= 0 && arms_meet(e, AP(AP(AP(q).pr).pr).pr)) return 2000;
            arms_cumulate(e);
        }
        // rejection test (BayesW_arms.cpp:529-535)
        if (!(e.yu >= e.ynew)) {
            e.xout = AP(w).x;
            return kArmsDone;
        }
        e.state = kAsSample;
    }
    // sample and squeeze (BayesW_arms.cpp:344-360, 429-503)
    if (++e.guard > 100000) return 2001;
    const int err = arms_invert(e, arms_u(e), w);
    if (err) return err;
    e.yu = arms_logshift(arms_u(e) * AP(w).ey, e.ymax);
    const int wl = AP(w).pl, wr = AP(w).pr;
    if (AP(wl).pl >= 0 && AP(wr).pr >= 0) {
        const int ql = AP(wl).f ? wl : AP(wl).pl;
        const int qr = AP(wr).f ? wr : AP(wr).pr;
        const double ysq = (AP(qr).y * (AP(w).x - AP(ql).x) + AP(ql).y * (AP(qr).x - AP(w).x)) / (AP(qr).x - AP(ql).x);
        if (e.yu <= ysq) {
            e.xout = AP(w).x;
            return kArmsDone;
        }
    }
    e.xout = AP(w).x;
    e.state = kAsEvalW;
    return kArmsEval;
}

#undef AP

// One draw from log density f (inlined at the call site, where it may run
// grid-wide sums: every thread of the grid takes the same path). 0 or an error.
template <class F>
__device__ __forceinline__ int arms_draw(ArmsEnv& e, const double (&xinit)[4], double xl, double xr, uint64_t seed,
                                         uint32_t site, uint32_t iteration, uint32_t chain, F&& f, double* out)
{
    int st = arms_begin(e, xinit, xl, xr, seed, site, iteration, chain);
    if (st) return st;
    double y = 0.0;
    while ((st = arms_step(e, y)) == kArmsEval) y = f(e.xout);
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

__device__ __forceinline__ double bw_gh(int nq, double sigma, double Ck, double sigma_b, double alpha, double sumf,
                                        double mean, double sd, const BwSums& v)
{
    const int r = (nq - 3) >> 1, h = (nq - 1) >> 1;
    const double sq = sqrt(2 * Ck * sigma_b);
    const double msr = mean / sd;
    double temp = 0.0;
    for (int i = 0; i < h; ++i) {
        const double xp = kGhX[r][i];
        const double xm = (nq >= 9 && i == 2) ? -kGhX[r][1] : -xp;   // x6 = -x3 for n >= 9
        const double a = kGhW[r][i] * bw_integrand(sigma * xp, alpha, sq, sumf, msr, sd, v);
        temp = (i == 0) ? a : temp + a;
        temp = temp + kGhW[r][i] * bw_integrand(sigma * xm, alpha, sq, sumf, msr, sd, v);
    }
    temp = temp + kGhWc[r];
    return sigma * temp;
}

// class of processColumn (bayeswbase.cpp:624-688): -1 when K < 2 (never)
template <int R>
__device__ __forceinline__ int bw_select(const BwParams& p, const BwShared<R>& L, int k, const BwSums& v)
{
    const int K = (int)p.K;
    const double mean = L.mean[k], sd = L.sd[k];
    double ml[kMaxK];
    ml[0] = L.pi[0] * kSqrtPi;
    const double exp_sum = (v.vi_1 * (1 - 2 * mean) + 4 * (1 - mean) * v.vi_2 + v.vi_sum * mean * mean) / (sd * sd);
    for (int i = 0; i < K - 1; ++i) {
        const double sigma = 1.0 / sqrt(1 + L.alpha * L.alpha * L.sigma_b * L.cls[i] * exp_sum);
        ml[i + 1] = L.pi[i + 1] * bw_gh((int)p.nq, sigma, L.cls[i], L.sigma_b, L.alpha, L.sumf[k], mean, sd, v);
    }
    double ml_sum = 0.0;
    for (int i = 0; i < K; ++i) ml_sum += ml[i];
    double acum = ml[0] / ml_sum;
    const double u = L.pu[k];
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

__device__ __forceinline__ uint64_t* bw_slot2(const BwParams& p, uint32_t rnd, int v)
{
    return p.xchg2 + (((size_t)(rnd & 1u) * kBwV + (size_t)v) * 2u);
}

// Values [0, V) published at round rnd: value v is reduced by workgroup v % P
// (lanes strided over the P partials, then a butterfly), totals land in
// L.xv[rnd & 1]; a thread reads them before it enters round rnd + 1, so round
// rnd + 2 may overwrite the buffer without another barrier. False on abort.
template <int R>
__device__ __forceinline__ bool bw_exchange(const BwParams& p, BwShared<R>& L, uint32_t rnd, int V, uint32_t w,
                                            int wave, int lane)
{
    int cnt = 0;
    bool ok = true;
    for (int v = 0; v < V; ++v) {
        if ((uint32_t)v % p.P != w) continue;
        if ((cnt++ % kWaves) != wave) continue;
        double acc = 0.0;
        for (uint32_t w2 = (uint32_t)lane; w2 < p.P; w2 += 64u) {
            double x;
            if (!wait_pair(bw_slot1(p, rnd, v, w2), rnd, p.ctl, &x)) { ok = false; break; }
            acc += x;
        }
        if (__any(!ok)) { ok = false; break; }
        acc = wave_sum(acc);
        if (lane == 0) publish_pair(bw_slot2(p, rnd, v), rnd, acc);
    }
    if (!ok && lane == 0) { L.abort = 1; raise_abort(p, 21); }
    const int tid = wave * 64 + lane;
    if (tid < V) {
        double x;
        if (!wait_pair(bw_slot2(p, rnd, tid), rnd, p.ctl, &x)) { L.abort = 1; raise_abort(p, 22); }
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

template <int R>
__global__ void __launch_bounds__(kThreads)
bayesw_sweep_kernel(BwParams p)
{
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
    if (tid == 0) {
        L.abort = 0;
        L.arms_err = 0;
        L.mu = p.scal[kBwMu];
        L.alpha = p.scal[kBwAlpha];
        L.sigma_b = p.scal[kBwSigmaB];
        L.d = p.scal[kBwD];
        L.betasq = 0.0;
    }
    if (tid < kMaxK) {
        L.cls[tid] = (tid < K - 1) ? p.cls[tid] : 0.0;
        L.pi[tid] = (tid < K) ? p.scal[kBwPi + tid] : 0.0;
        L.v[tid] = (tid < K) ? p.scal[kBwPi + K + tid] : 0.0;   // chain mode resets it per iteration
    }
    uint32_t rnd = __hip_atomic_load(p.ctl + kCtlRound, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned long long n_rounds = 0, n_changes = 0;
    __syncthreads();

    // one value per workgroup: every wave holds the whole slice, so wave 0's sum
    auto allreduce = [&](double part, double* total) -> bool {
        if (L.abort) return false;
        part = wave_sum(part);
        ++rnd;
        ++n_rounds;
        if (tid == 0) publish_pair(bw_slot1(p, rnd, 0, w), rnd, part);
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
        bw_delta(L.mean[k], L.sd[k], beta, dv);
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

        if (p.mode == kModeChain) {
            // ---- sampleMu (bayeswbase.cpp:542-569): the density's sum over
            // individuals factors as exp(-alpha x) sum_i exp(alpha (eps_i + mu) - EulerGamma)
            double part = 0.0;
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const uint32_t idx = lo + (uint32_t)r * kRun + 8u * (uint32_t)lane + (uint32_t)e;
                    if (idx < hi) part += exp(L.alpha * (eps[r][e] + L.mu) - kEuler);
                }
            double smu;
            if (!allreduce(part, &smu)) goto done;
            {
                const double mu = L.mu, a = L.alpha, d = L.d;
                const double xinit[4] = {0.995 * mu, mu, 1.005 * mu, 1.01 * mu};
                double xs = mu;
                // mu_dens (bayeswbase.cpp:66-77), sigma_mu = 100
                const int err = arms_draw(L.env[wave], xinit, 2.0, 5.0, p.seed, kSiteMu, iteration, p.chain,
                                          [&](double x) { return -a * x * d - exp(-a * x) * smu - x * x / (2 * 100.0); },
                                          &xs);
                if (err) {
                    if (tid == 0) { L.arms_err = err; raise_abort(p, 30); }
                    __syncthreads();
                    goto done;
                }
#pragma unroll
                for (int r = 0; r < R; ++r)
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        const uint32_t idx = lo + (uint32_t)r * kRun + 8u * (uint32_t)lane + (uint32_t)e;
                        if (idx < hi) eps[r][e] = (eps[r][e] + mu) - xs;
                    }
                __syncthreads();
                if (tid == 0) L.mu = xs;
                if (tid < kMaxK) L.v[tid] = (tid < K) ? 1.0 : 0.0;
                if (tid == 0) L.betasq = 0.0;
            }
        }
        refresh_vi();
        __syncthreads();

        // ---- the sweep (AnalysisGraph::exec over the order)
        for (uint32_t blk = 0; blk < nblocks; ++blk) {
            const uint32_t b0 = blk * kBwB;
            const int nb = (int)min((uint32_t)kBwB, p.M - b0);
            // block load: codes of the slice and per-marker constants
            if (tid < nb) {
                const uint32_t j = order[b0 + tid];
                L.mk[tid] = j;
                L.bold[tid] = bits_as_double(ld_agent((const uint64_t*)(p.beta + j)));
                L.mean[tid] = p.stats[4 * (size_t)j];
                L.sd[tid] = p.stats[4 * (size_t)j + 1];
                L.sumf[tid] = p.sumf[j];
                if (p.mode == kModeChain) {
                    const U4 x = draw_block(p.seed, j, iteration, kStreamMarker, p.chain);
                    L.pu[tid] = unit_u53(x.x[0], x.x[1]);
                } else {
                    L.pu[tid] = p.u[j];
                }
            }
            __syncthreads();
            for (uint32_t t = tid; t < (uint32_t)nb * R * 32u; t += kThreads) {
                const uint32_t k = t / (R * 32u), rem = t % (R * 32u);
                const uint64_t byte_off = (uint64_t)(lo >> 2) + 4ull * rem;
                uint32_t word = 0xFFFFFFFFu;
                if (byte_off < p.col_bytes) word = p.codes[(size_t)L.mk[k] * p.col_words + (lo >> 4) + rem];
                (&L.codes[0][0][0])[t] = word;
            }
            __syncthreads();

            int pos = 0;
            while (pos < nb) {
                const bool breaker = L.bold[pos] != 0.0;
                ++rnd;
                ++n_rounds;
                int V;
                if (!breaker) {
                    // vi_sum of the slice and (vi_1, vi_2) of slots pos..nb-1: wave q
                    // takes slots pos + q + 4 i, values 2 i (genotype 1) and 2 i + 1
                    double part[16];
#pragma unroll
                    for (int i = 0; i < 8; ++i) {
                        const int k = pos + wave + kWaves * i;
                        double s1 = 0.0, s2 = 0.0;
                        if (k < nb) {
#pragma unroll
                            for (int r = 0; r < R; ++r) {
                                const uint32_t bits = L.codes[k][r][lane >> 1] >> (16 * (lane & 1));
#pragma unroll
                                for (int e = 0; e < 8; ++e) {
                                    const uint32_t c = (bits >> (2 * e)) & 3u;
                                    s1 += (c == 2u) ? vi[r][e] : 0.0;
                                    s2 += (c == 0u) ? vi[r][e] : 0.0;
                                }
                            }
                        }
                        part[2 * i] = s1;
                        part[2 * i + 1] = s2;
                    }
                    const double tot = transpose_reduce16(part, lane);
                    const int vv = (lane >> 2) & 15;
                    const int k = pos + wave + kWaves * (vv >> 1);
                    if ((lane & 3) == 0 && k < nb)
                        publish_pair(bw_slot1(p, rnd, 1 + 2 * (k - pos) + (vv & 1), w), rnd, tot);
                    double sv = 0.0;
#pragma unroll
                    for (int r = 0; r < R; ++r)
#pragma unroll
                        for (int e = 0; e < 8; ++e) sv += vi[r][e];
                    sv = wave_sum(sv);
                    if (tid == 0) publish_pair(bw_slot1(p, rnd, 0, w), rnd, sv);
                    V = 1 + 2 * (nb - pos);
                } else {
                    // a marker in the model: eps += x beta_old, vi re-formed, then its sums
                    apply(pos, L.bold[pos], 1.0);
                    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const uint32_t bits = L.codes[pos][r][lane >> 1] >> (16 * (lane & 1));
#pragma unroll
                        for (int e = 0; e < 8; ++e) {
                            const uint32_t c = (bits >> (2 * e)) & 3u;
                            s0 += vi[r][e];
                            s1 += (c == 2u) ? vi[r][e] : 0.0;
                            s2 += (c == 0u) ? vi[r][e] : 0.0;
                        }
                    }
                    s0 = wave_sum(s0);
                    s1 = wave_sum(s1);
                    s2 = wave_sum(s2);
                    if (tid == 0) {
                        publish_pair(bw_slot1(p, rnd, 0, w), rnd, s0);
                        publish_pair(bw_slot1(p, rnd, 1, w), rnd, s1);
                        publish_pair(bw_slot1(p, rnd, 2, w), rnd, s2);
                    }
                    V = 3;
                }
                if (!bw_exchange<R>(p, L, rnd, V, w, wave, lane)) goto done;
                const double* xv = L.xv[rnd & 1u];

                // ---- decisions (wave 0, lane = slot)
                if (wave == 0) {
                    const int k = pos + lane;
                    // a marker in the model stops a normal round: its step starts from eps + x beta_old
                    const bool held = !breaker && k < nb && L.bold[k] != 0.0;
                    const bool active = breaker ? lane == 0 : (k < nb && !held);
                    int sel = -1;
                    if (active) {
                        const int vi1 = breaker ? 1 : 1 + 2 * lane;
                        const double s1 = xv[vi1], s2 = xv[vi1 + 1];
                        const BwSums v{xv[0], xv[0] - s1 - s2, s1, s2};
                        sel = bw_select<R>(p, L, k, v);
                    }
                    const bool stop = held || (active && (breaker || sel > 0));
                    const unsigned long long sm = __ballot(stop);
                    const int c = sm ? pos + (int)__ffsll((long long)sm) - 1 : nb;
                    const int cs = __shfl(sel, (c - pos) & 63, 64);
                    // slots before c keep beta = 0 in class 0
                    if (k < c && k < nb) {
                        atomicAdd(&L.v[0], 1.0);
                        if (w == 0) {
                            st_agent((uint64_t*)(p.beta + L.mk[k]), double_as_bits(0.0));
                            p.comp[L.mk[k]] = 0;
                        }
                    }
                    if (lane == 0) {
                        L.stop = c;
                        L.sel = (c < nb) ? cs : -1;
                    }
                }
                __syncthreads();
                {
                    const int c = L.stop, sel = L.sel;
                    if (c >= nb) {
                        pos = nb;
                    } else if (!breaker && L.bold[c] != 0.0) {
                        pos = c;                       // next round: the marker's own round
                    } else {
                        // draw the slot's beta: 0 for class 0, else ARMS on beta_dens
                        double bnew = 0.0;
                        if (sel > 0) {
                            const double s1 = breaker ? xv[1] : xv[1 + 2 * (c - pos)];
                            const double s2 = breaker ? xv[2] : xv[2 + 2 * (c - pos)];
                            const double vs = xv[0];
                            const double v0 = vs - s1 - s2;
                            const double a = L.alpha, sb = L.sigma_b, Ck = L.cls[sel - 1];
                            const double sf = L.sumf[c], msr = L.mean[c] / L.sd[c], sd = L.sd[c];
                            const double bo = L.bold[c];
                            const double safe = 2 * sqrt(sb * Ck);
                            const double xinit[4] = {bo - safe / 10, bo, bo + safe / 20, bo + safe / 10};
                            const int err = arms_draw(
                                L.env[wave], xinit, bo - safe, bo + safe, p.seed, L.mk[c], iteration, p.chain,
                                [&](double x) {
                                    // beta_dens (sparsebayesw.cpp:26-37)
                                    return -a * x * sf -
                                           exp(a * x * msr) * (v0 + s1 * exp(-a * x / sd) + s2 * exp(-2 * a * x / sd)) -
                                           x * x / (2 * Ck * sb);
                                },
                                &bnew);
                            if (err) {
                                if (tid == 0) { L.arms_err = err; raise_abort(p, 31); }
                                __syncthreads();
                                goto done;
                            }
                        }
                        // skip iff beta_old == 0 && beta_new == 0 (bayeswbase.cpp:691-696)
                        if (!(L.bold[c] == 0.0 && bnew == 0.0)) {
                            apply(c, bnew, -1.0);
                            ++n_changes;
                        }
                        __syncthreads();
                        if (tid == 0) {
                            L.v[sel] += 1.0;
                            L.betasq += bnew * bnew;
                            if (w == 0) {
                                st_agent((uint64_t*)(p.beta + L.mk[c]), double_as_bits(bnew));
                                p.comp[L.mk[c]] = sel;
                            }
                        }
                        pos = c + 1;
                    }
                }
                __syncthreads();
            }
            if (w == 0) __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
        }

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
            const int err = arms_draw(
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
        }
    } else if (w == 0 && tid == 0) {
        p.counters[9] = (unsigned long long)L.arms_err;
    }
}

// calculateSumFailure (raggedbayeswkernel.cpp:18-25) from the 2-bit codes: one
// workgroup per marker; list sums pass through an int as in the reference.
__global__ void bw_sum_failure_kernel(const uint32_t* codes, uint64_t col_words, uint32_t N, const double* fail,
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

// RaggedBayesWKernel::setVi (raggedbayeswkernel.cpp:10-16) of marker j for
// vi = exp(alpha eps - EulerGamma): out = {vi_sum, vi_0, vi_1, vi_2}. One workgroup.
__global__ void bw_set_vi_kernel(const uint32_t* codes, uint64_t col_words, uint32_t N, const double* eps,
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
__global__ void bw_residual_kernel(const uint32_t* codes, uint64_t col_words, uint32_t N, const double* stats,
                                   uint32_t j, double beta, int sign, double* eps)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const uint8_t* col = (const uint8_t*)(codes + (size_t)j * col_words);
    const uint32_t c = (col[i >> 2] >> ((i & 3u) * 2u)) & 3u;
    double dv[4];
    bw_delta(stats[4 * (size_t)j], stats[4 * (size_t)j + 1], beta, dv);
    const double d = dv[c];
    eps[i] = (sign > 0) ? eps[i] + d : eps[i] - d;
}

} // namespace brr
