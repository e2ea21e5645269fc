/*
 * rng.c -- CPU ORACLE (test infrastructure only): the counter-based random
 * stream that replaces the reference's Boost mt19937 and glibc rand().
 *
 * Stream definition (the product implements the same definition on its own):
 *   block(ctr0, ctr1, ctr2, ctr3) = Philox4x32-10(ctr, key = {seed lo32, seed hi32})
 *   u53(a, b)    = ((a << 32 | b) >> 11) * 2^-53                 in [0, 1)
 *   normal(a, b) = AS241(((a << 32 | b) >> 11) + 0.5) * 2^-53)   (inverse CDF)
 *   marker (u_j, z_j), iteration t, chain c:
 *        x = block(j, t, STREAM_MARKER, c); u = u53(x0, x1); z = normal(x2, x3)
 *   shuffle step i (i = 1..M-1):
 *        x = block(i, t, STREAM_SHUFFLE, c); r = (x0 << 32 | x1); j = r % (i + 1)
 *   scalar draw d, attempt a:
 *        x = block(d, t, STREAM_SCALAR | a << 8, c)
 *   init draws use t = 0xFFFFFFFF.
 *
 * Reference parameterisations restated (distributions_boost.cpp):
 *   unif_rng            U[0,1)                                  (:28-33)
 *   norm_rng(mean, var) mean + sqrt(var) * N(0,1)               (:61-65)
 *   rgamma(shape,scale) Gamma(shape) * scale                    (:22-26)
 *   inv_gamma_rng(a,b)  1 / rgamma(a, 1/b)                      (:45-47)
 *   inv_scaled_chisq_rng(dof, s) = inv_gamma_rng(dof/2, dof*s/2) (:58-60)
 *   dirichlet_rng(alpha) = rgamma(alpha_k, 1) / sum              (:35-43)
 * Gamma(shape) for shape >= 1 is Marsaglia & Tsang (2000) without the squeeze;
 * shape < 1 uses Gamma(shape + 1) * U^(1/shape).
 */
#include "brr_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static void block(uint64_t seed, uint32_t a, uint32_t b, uint32_t c, uint32_t d, uint32_t out[4])
{
    const uint32_t ctr[4] = {a, b, c, d};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    orc_philox4x32_10(ctr, key, out);
}

static double u53(uint32_t a, uint32_t b)
{
    const uint64_t r = ((uint64_t)a << 32) | b;
    return (double)(r >> 11) * 0x1.0p-53;
}

/* ---- deterministic log / exp (only + - * / and exact bit operations) ---- */

static double bits_to_double(uint64_t u) { double d; memcpy(&d, &u, 8); return d; }
static uint64_t double_to_bits(double d) { uint64_t u; memcpy(&u, &d, 8); return u; }

static const double LN2_HI = 6.93147180369123816490e-01; /* 0x3FE62E42FEE00000 */
static const double LN2_LO = 1.90821492927058770002e-10; /* 0x3DEA39EF35793C76 */

double orc_det_log(double x)
{
    if (x != x || x < 0.0) return bits_to_double(0x7FF8000000000000ull);
    if (x == 0.0) return -bits_to_double(0x7FF0000000000000ull);
    if (x == bits_to_double(0x7FF0000000000000ull)) return x;
    int e_adj = 0;
    uint64_t b = double_to_bits(x);
    if ((b >> 52) == 0) {            /* subnormal: scale up by 2^54 */
        x *= 0x1.0p54;
        e_adj = -54;
        b = double_to_bits(x);
    }
    int e = (int)((b >> 52) & 0x7FF) - 1023 + e_adj;
    double m = bits_to_double((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull); /* [1,2) */
    if (m > 1.4142135623730951) { m *= 0.5; e += 1; }        /* [sqrt(2)/2, sqrt(2)] */
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    /* atanh series: log m = 2 s (1 + s2/3 + s2^2/5 + ... + s2^12/25) */
    double p = 1.0 / 25.0;
    p = p * s2 + 1.0 / 23.0;
    p = p * s2 + 1.0 / 21.0;
    p = p * s2 + 1.0 / 19.0;
    p = p * s2 + 1.0 / 17.0;
    p = p * s2 + 1.0 / 15.0;
    p = p * s2 + 1.0 / 13.0;
    p = p * s2 + 1.0 / 11.0;
    p = p * s2 + 1.0 / 9.0;
    p = p * s2 + 1.0 / 7.0;
    p = p * s2 + 1.0 / 5.0;
    p = p * s2 + 1.0 / 3.0;
    const double logm = 2.0 * s + 2.0 * s * (s2 * p);
    const double de = (double)e;
    return de * LN2_HI + (de * LN2_LO + logm);
}

double orc_det_exp(double x)
{
    if (x != x) return x;
    if (x > 709.0) return bits_to_double(0x7FF0000000000000ull);
    if (x < -745.0) return 0.0;
    /* k = nearest integer to x / ln 2 */
    const double t = x * 1.44269504088896338700e+00;
    const double kd = (t >= 0.0) ? (double)(int64_t)(t + 0.5) : -(double)(int64_t)(-t + 0.5);
    const int k = (int)kd;
    const double r = (x - kd * LN2_HI) - kd * LN2_LO;     /* |r| <= ~0.35 */
    /* Taylor to degree 16 */
    double p = 1.0 / 20922789888000.0;                    /* 1/16! */
    p = p * r + 1.0 / 1307674368000.0;                    /* 1/15! */
    p = p * r + 1.0 / 87178291200.0;
    p = p * r + 1.0 / 6227020800.0;
    p = p * r + 1.0 / 479001600.0;
    p = p * r + 1.0 / 39916800.0;
    p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0;
    p = p * r + 1.0 / 40320.0;
    p = p * r + 1.0 / 5040.0;
    p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    /* scale by 2^k in up to two exact steps */
    int k1 = k, k2 = 0;
    if (k1 < -1000) { k2 = k1 + 1000; k1 = -1000; }
    if (k1 > 1000) { k2 = k1 - 1000; k1 = 1000; }
    p *= bits_to_double((uint64_t)(k1 + 1023) << 52);
    if (k2 != 0) p *= bits_to_double((uint64_t)(k2 + 1023) << 52);
    return p;
}

/* Wichura (1988), Algorithm AS241 PPND16, Applied Statistics 37:477-484. */
double orc_inv_norm_cdf(double p)
{
    const double q = p - 0.5;
    if (fabs(q) <= 0.425) {
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
    r = sqrt(-orc_det_log(r));
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

static double normal_from(uint32_t a, uint32_t b)
{
    const uint64_t r = ((uint64_t)a << 32) | b;
    const double p = ((double)(r >> 11) + 0.5) * 0x1.0p-53;
    return orc_inv_norm_cdf(p);
}

void orc_marker_uz(uint64_t seed, uint32_t iteration, uint32_t marker, uint32_t chain,
                   double* u, double* z)
{
    uint32_t x[4];
    block(seed, marker, iteration, ORC_STREAM_MARKER, chain, x);
    *u = u53(x[0], x[1]);
    *z = normal_from(x[2], x[3]);
}

void orc_fill_marker_uz(uint64_t seed, uint32_t iteration, uint32_t chain, uint32_t M,
                        double* u, double* z)
{
    for (uint32_t j = 0; j < M; ++j) orc_marker_uz(seed, iteration, j, chain, &u[j], &z[j]);
}

void orc_shuffle_philox(uint64_t seed, uint32_t iteration, uint32_t chain, uint32_t M,
                        uint32_t* order)
{
    orc_shuffle_philox_stream(seed, iteration, chain, ORC_STREAM_SHUFFLE, M, order);
}

void orc_shuffle_philox_stream(uint64_t seed, uint32_t iteration, uint32_t chain, uint32_t stream,
                               uint32_t M, uint32_t* order)
{
    /* libstdc++ random_shuffle: for i in [1, M): swap(order[i], order[r % (i + 1)]) */
    for (uint32_t i = 1; i < M; ++i) {
        uint32_t x[4];
        block(seed, i, iteration, stream, chain, x);
        const uint64_t r = ((uint64_t)x[0] << 32) | x[1];
        const uint32_t j = (uint32_t)(r % ((uint64_t)i + 1));
        if (j != i) { const uint32_t t = order[i]; order[i] = order[j]; order[j] = t; }
    }
}

void orc_shuffle_libc(uint32_t M, uint32_t* order)
{
    for (uint32_t i = 1; i < M; ++i) {
        const uint32_t j = (uint32_t)(rand() % (int)(i + 1));
        if (j != i) { const uint32_t t = order[i]; order[i] = order[j]; order[j] = t; }
    }
}

static void scalar_block(uint64_t seed, uint32_t iteration, uint32_t draw_id, uint32_t chain,
                         uint32_t attempt, uint32_t out[4])
{
    block(seed, draw_id, iteration, ORC_STREAM_SCALAR | (attempt << 8), chain, out);
}

double orc_uniform01(uint64_t seed, uint32_t iteration, uint32_t draw_id, uint32_t chain)
{
    uint32_t x[4];
    scalar_block(seed, iteration, draw_id, chain, 0, x);
    return u53(x[0], x[1]);
}

double orc_norm_rng(uint64_t seed, uint32_t iteration, uint32_t draw_id, uint32_t chain,
                    double mean, double var)
{
    uint32_t x[4];
    scalar_block(seed, iteration, draw_id, chain, 0, x);
    return normal_from(x[0], x[1]) * sqrt(var) + mean;
}

#define ORC_BOOST_ATTEMPT 0xFFFFFFu

static double std_gamma(uint64_t seed, uint32_t iteration, uint32_t draw_id, uint32_t chain,
                        double shape)
{
    if (!(shape > 0.0)) return bits_to_double(0x7FF8000000000000ull);
    const double a = (shape < 1.0) ? shape + 1.0 : shape;
    const double d = a - 1.0 / 3.0;
    const double c = 1.0 / sqrt(9.0 * d);
    double g = 0.0;
    for (uint32_t attempt = 0; attempt < ORC_BOOST_ATTEMPT; ++attempt) {
        uint32_t x[4];
        scalar_block(seed, iteration, draw_id, chain, attempt, x);
        const double zz = normal_from(x[0], x[1]);
        const double t = 1.0 + c * zz;
        if (t <= 0.0) continue;
        const double v = t * t * t;
        const double uu = u53(x[2], x[3]);
        if (orc_det_log(uu) < 0.5 * zz * zz + d - d * v + d * orc_det_log(v)) { g = d * v; break; }
    }
    if (shape < 1.0) {
        uint32_t x[4];
        scalar_block(seed, iteration, draw_id, chain, ORC_BOOST_ATTEMPT, x);
        const double uu = u53(x[0], x[1]);
        g *= orc_det_exp(orc_det_log(uu) / shape);
    }
    return g;
}

double orc_rgamma(uint64_t seed, uint32_t iteration, uint32_t draw_id, uint32_t chain,
                  double shape, double scale)
{
    return std_gamma(seed, iteration, draw_id, chain, shape) * scale;
}

double orc_inv_scaled_chisq_rng(uint64_t seed, uint32_t iteration, uint32_t draw_id,
                                uint32_t chain, double dof, double scale)
{
    /* inv_gamma_rng(0.5*dof, 0.5*dof*scale) = 1 / rgamma(shape, 1/(0.5*dof*scale)) */
    const double shape = 0.5 * dof;
    const double ig_scale = 0.5 * dof * scale;
    return 1.0 / orc_rgamma(seed, iteration, draw_id, chain, shape, 1.0 / ig_scale);
}

void orc_dirichlet_rng(uint64_t seed, uint32_t iteration, uint32_t first_id, uint32_t chain,
                       int K, const double* alpha, double* out)
{
    double sum = 0.0;
    for (int k = 0; k < K; ++k) {
        out[k] = orc_rgamma(seed, iteration, first_id + (uint32_t)k, chain, alpha[k], 1.0);
    }
    for (int k = 0; k < K; ++k) sum += out[k];
    for (int k = 0; k < K; ++k) out[k] /= sum;
}
