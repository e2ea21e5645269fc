// brr_rng.h -- counter-based random stream used by the MI355X Gibbs sweep.
//
// The reference draws its randomness from Boost mt19937 (distributions_boost.cpp)
// and glibc rand() (std::random_shuffle, BayesRBase.cpp:198). A sequential
// generator cannot be shared by thousands of device threads, so this build
// defines the stream by counter: every draw is Philox4x32-10 of a counter that
// names it. Host and device compile this same header, so a device chain and a
// host driver given the same seed consume identical numbers.
//
//   block(c0, c1, c2, c3)  = Philox4x32-10(counter, key = {seed lo32, seed hi32})
//   u53(a, b)              = ((a << 32 | b) >> 11) * 2^-53                 [0, 1)
//   normal(a, b)           = Phi^-1((((a << 32 | b) >> 11) + 0.5) * 2^-53)
//   per-marker pair        : x = block(marker, iteration, 1, chain);
//                            u = u53(x0, x1), z = normal(x2, x3)
//   shuffle step i         : x = block(i, iteration, 2, chain); r = x0 << 32 | x1;
//                            swap(order[i], order[r % (i + 1)])     (i = 1..M-1)
//   scalar draw d, try a   : x = block(d, iteration, 3 | a << 8, chain)
//   draws before the first iteration use iteration = 0xFFFFFFFF.
//
// Phi^-1 is Wichura's AS241 (PPND16). Gamma(shape) is Marsaglia & Tsang
// (ACM TOMS 26, 2000) without the squeeze; shape < 1 takes Gamma(shape + 1) *
// U^(1/shape). log and exp are evaluated from + - * / only (an atanh series
// and a Taylor series after exact range reduction), so results are identical
// on every IEEE-754 binary64 target when floating-point contraction is off.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define BRR_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#define BRR_HD inline
#endif

namespace brr {

// kStreamFixed (5): the covariate order of the fixed effects, shuffle rule as above.
enum : uint32_t { kStreamMarker = 1, kStreamShuffle = 2, kStreamScalar = 3, kStreamFixed = 5 };
enum : uint32_t { kDrawMu = 0, kDrawSigmaE = 1, kDrawSigmaG0 = 2 };
constexpr uint32_t kInitIteration = 0xFFFFFFFFu;
constexpr uint32_t kBoostAttempt = 0xFFFFFFu;

struct U4 { uint32_t x[4]; };

BRR_HD U4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                        uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    U4 o;
    o.x[0] = c0; o.x[1] = c1; o.x[2] = c2; o.x[3] = c3;
    return o;
}

BRR_HD U4 draw_block(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3)
{
    return philox4x32_10(c0, c1, c2, c3, (uint32_t)seed, (uint32_t)(seed >> 32));
}

BRR_HD double bits_as_double(uint64_t u)
{
    union { uint64_t u; double d; } v;
    v.u = u;
    return v.d;
}

BRR_HD uint64_t double_as_bits(double d)
{
    union { uint64_t u; double d; } v;
    v.d = d;
    return v.u;
}

BRR_HD double unit_u53(uint32_t a, uint32_t b)
{
    const uint64_t r = ((uint64_t)a << 32) | b;
    return (double)(r >> 11) * 1.1102230246251565404e-16; // 2^-53
}

BRR_HD double det_log(double x)
{
    if (x != x || x < 0.0) return bits_as_double(0x7FF8000000000000ull);
    if (x == 0.0) return -bits_as_double(0x7FF0000000000000ull);
    if (x == bits_as_double(0x7FF0000000000000ull)) return x;
    int shift = 0;
    uint64_t b = double_as_bits(x);
    if ((b >> 52) == 0) {                       // subnormal
        x *= 18014398509481984.0;               // 2^54
        shift = -54;
        b = double_as_bits(x);
    }
    int e = (int)((b >> 52) & 0x7FF) - 1023 + shift;
    double m = bits_as_double((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
    if (m > 1.4142135623730951) { m *= 0.5; e += 1; }
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    const double inv_odd[12] = {1.0 / 25.0, 1.0 / 23.0, 1.0 / 21.0, 1.0 / 19.0, 1.0 / 17.0, 1.0 / 15.0,
                                1.0 / 13.0, 1.0 / 11.0, 1.0 / 9.0,  1.0 / 7.0,  1.0 / 5.0,  1.0 / 3.0};
    double p = inv_odd[0];
    for (int i = 1; i < 12; ++i) p = p * s2 + inv_odd[i];
    const double logm = 2.0 * s + 2.0 * s * (s2 * p);
    const double de = (double)e;
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    return de * ln2_hi + (de * ln2_lo + logm);
}

BRR_HD double det_exp(double x)
{
    if (x != x) return x;
    if (x > 709.0) return bits_as_double(0x7FF0000000000000ull);
    if (x < -745.0) return 0.0;
    const double t = x * 1.44269504088896338700e+00;
    const double kd = (t >= 0.0) ? (double)(int64_t)(t + 0.5) : -(double)(int64_t)(-t + 0.5);
    const int k = (int)kd;
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    const double r = (x - kd * ln2_hi) - kd * ln2_lo;
    // 1/16! ... 1/2!, then 1, 1 (Horner, degree 16)
    const double c[15] = {1.0 / 20922789888000.0, 1.0 / 1307674368000.0, 1.0 / 87178291200.0,
                          1.0 / 6227020800.0,     1.0 / 479001600.0,     1.0 / 39916800.0,
                          1.0 / 3628800.0,        1.0 / 362880.0,        1.0 / 40320.0,
                          1.0 / 5040.0,           1.0 / 720.0,           1.0 / 120.0,
                          1.0 / 24.0,             1.0 / 6.0,             0.5};
    double p = c[0];
    for (int i = 1; i < 15; ++i) p = p * r + c[i];
    p = p * r + 1.0;
    p = p * r + 1.0;
    int k1 = k, k2 = 0;
    if (k1 < -1000) { k2 = k1 + 1000; k1 = -1000; }
    if (k1 > 1000) { k2 = k1 - 1000; k1 = 1000; }
    p *= bits_as_double((uint64_t)(k1 + 1023) << 52);
    if (k2 != 0) p *= bits_as_double((uint64_t)(k2 + 1023) << 52);
    return p;
}

BRR_HD double det_sqrt(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __dsqrt_rn(x);
#else
    return std::sqrt(x);
#endif
}

// Wichura, Algorithm AS241 PPND16 (Applied Statistics 37, 1988).
BRR_HD double inv_norm_cdf(double p)
{
    const double q = p - 0.5;
    const double aq = q < 0.0 ? -q : q;
    if (aq <= 0.425) {
        const double r = 0.180625 - q * q;
        const double num = (((((((2.5090809287301226727e+3 * r + 3.3430575583588128105e+4) * r
                    + 6.7265770927008700853e+4) * r + 4.5921953931549871457e+4) * r
                    + 1.3731693765509461125e+4) * r + 1.9715909503065514427e+3) * r
                    + 1.3314166789178437745e+2) * r + 3.3871328727963666080e+0);
        const double den = (((((((5.2264952788528545610e+3 * r + 2.8729085735721942674e+4) * r
                    + 3.9307895800092710610e+4) * r + 2.1213794301586595867e+4) * r
                    + 5.3941960214247511077e+3) * r + 6.8718700749205790830e+2) * r
                    + 4.2313330701600911252e+1) * r + 1.0);
        return q * num / den;
    }
    double r = (q < 0.0) ? p : 1.0 - p;
    r = det_sqrt(-det_log(r));
    double val;
    if (r <= 5.0) {
        r -= 1.6;
        const double num = (((((((7.74545014278341407640e-4 * r + 2.27238449892691845833e-2) * r
                    + 2.41780725177450611770e-1) * r + 1.27045825245236838258e+0) * r
                    + 3.64784832476320460504e+0) * r + 5.76949722146069140550e+0) * r
                    + 4.63033784615654529590e+0) * r + 1.42343711074968357734e+0);
        const double den = (((((((1.05075007164441684324e-9 * r + 5.47593808499534494600e-4) * r
                    + 1.51986665636164571966e-2) * r + 1.48103976427480074590e-1) * r
                    + 6.89767334985100004550e-1) * r + 1.67638483018380384940e+0) * r
                    + 2.05319162663775882187e+0) * r + 1.0);
        val = num / den;
    } else {
        r -= 5.0;
        const double num = (((((((2.01033439929228813265e-7 * r + 2.71155556874348757815e-5) * r
                    + 1.24266094738807843860e-3) * r + 2.65321895265761230930e-2) * r
                    + 2.96560571828504891230e-1) * r + 1.78482653991729133580e+0) * r
                    + 5.46378491116411436990e+0) * r + 6.65790464350110377720e+0);
        const double den = (((((((2.04426310338993978564e-15 * r + 1.42151175831644588870e-7) * r
                    + 1.84631831751005468180e-5) * r + 7.86869131145613259100e-4) * r
                    + 1.48753612908506148525e-2) * r + 1.36929880922735805310e-1) * r
                    + 5.99832206555887937690e-1) * r + 1.0);
        val = num / den;
    }
    return (q < 0.0) ? -val : val;
}

BRR_HD double unit_normal(uint32_t a, uint32_t b)
{
    const uint64_t r = ((uint64_t)a << 32) | b;
    const double p = ((double)(r >> 11) + 0.5) * 1.1102230246251565404e-16;
    return inv_norm_cdf(p);
}

// (u_j, z_j) of marker j at iteration t (BayesRBase.cpp:101-110 draws these
// per marker id before the sweep).
BRR_HD void marker_uz(uint64_t seed, uint32_t iteration, uint32_t marker, uint32_t chain,
                      double* u, double* z)
{
    const U4 x = draw_block(seed, marker, iteration, kStreamMarker, chain);
    *u = unit_u53(x.x[0], x.x[1]);
    *z = unit_normal(x.x[2], x.x[3]);
}

BRR_HD U4 scalar_block(uint64_t seed, uint32_t iteration, uint32_t draw, uint32_t chain,
                       uint32_t attempt)
{
    return draw_block(seed, draw, iteration, kStreamScalar | (attempt << 8), chain);
}

BRR_HD double uniform01(uint64_t seed, uint32_t iteration, uint32_t draw, uint32_t chain)
{
    const U4 x = scalar_block(seed, iteration, draw, chain, 0);
    return unit_u53(x.x[0], x.x[1]);
}

// norm_rng(mean, var): mean + sqrt(var) * N(0,1) (distributions_boost.cpp:61-65)
BRR_HD double norm_rng(uint64_t seed, uint32_t iteration, uint32_t draw, uint32_t chain,
                       double mean, double var)
{
    const U4 x = scalar_block(seed, iteration, draw, chain, 0);
    return unit_normal(x.x[0], x.x[1]) * det_sqrt(var) + mean;
}

BRR_HD double std_gamma(uint64_t seed, uint32_t iteration, uint32_t draw, uint32_t chain,
                        double shape)
{
    if (!(shape > 0.0)) return bits_as_double(0x7FF8000000000000ull);
    const double a = (shape < 1.0) ? shape + 1.0 : shape;
    const double d = a - 1.0 / 3.0;
    const double c = 1.0 / det_sqrt(9.0 * d);
    double g = 0.0;
    for (uint32_t attempt = 0; attempt < kBoostAttempt; ++attempt) {
        const U4 x = scalar_block(seed, iteration, draw, chain, attempt);
        const double zz = unit_normal(x.x[0], x.x[1]);
        const double t = 1.0 + c * zz;
        if (t <= 0.0) continue;
        const double v = t * t * t;
        const double uu = unit_u53(x.x[2], x.x[3]);
        if (det_log(uu) < 0.5 * zz * zz + d - d * v + d * det_log(v)) { g = d * v; break; }
    }
    if (shape < 1.0) {
        const U4 x = scalar_block(seed, iteration, draw, chain, kBoostAttempt);
        const double uu = unit_u53(x.x[0], x.x[1]);
        g *= det_exp(det_log(uu) / shape);
    }
    return g;
}

// rgamma(shape, scale) (distributions_boost.cpp:22-26)
BRR_HD double rgamma(uint64_t seed, uint32_t iteration, uint32_t draw, uint32_t chain,
                     double shape, double scale)
{
    return std_gamma(seed, iteration, draw, chain, shape) * scale;
}

// inv_scaled_chisq_rng(dof, s) = inv_gamma_rng(dof/2, dof*s/2) = 1 / rgamma(dof/2, 1/(dof*s/2))
// (distributions_boost.cpp:45-47, 58-60)
BRR_HD double inv_scaled_chisq_rng(uint64_t seed, uint32_t iteration, uint32_t draw,
                                   uint32_t chain, double dof, double scale)
{
    const double shape = 0.5 * dof;
    const double ig_scale = 0.5 * dof * scale;
    return 1.0 / rgamma(seed, iteration, draw, chain, shape, 1.0 / ig_scale);
}

} // namespace brr
