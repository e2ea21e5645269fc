/*
 * bayesw.c -- CPU ORACLE (test infrastructure only): the BayesW Weibull
 * survival sweep of the reference on the ragged layout, FP64, sequential.
 *
 *   processColumn            bayeswbase.cpp:604-704 with RaggedBayesWKernel
 *                            (raggedbayeswkernel.cpp:10-77) and
 *                            SparseBayesW::estimateBeta / beta_dens
 *                            (sparsebayesw.cpp:26-37, 81-102)
 *   adaptive Gauss-Hermite   bayeswbase.cpp:126-457 (the tabulated rules,
 *                            including the x6 = -x3 entry of n = 9..15)
 *   init, mu, alpha, sigma_b, pi   bayeswbase.cpp:469-540, 542-569, 707-734,
 *                            777-809
 *   dense layout             DenseBayesWKernel (densebayeswkernel.cpp:10-42) and
 *                            DenseBayesW::beta_dens (densebayesw.cpp:25-33): the
 *                            standardised column Cx (the dense table of each
 *                            individual's call), every sum over individuals
 *                            evaluated element by element as the reference does
 *   ARMS                     BayesW_arms.cpp:133-906, the derivative-free
 *                            adaptive rejection sampler of Gilks, Best & Tan
 *                            (1995), restated here with index-linked envelope
 *                            points and without the Metropolis step (BayesW
 *                            always calls it with dometrop = 0).
 * Random numbers come from the counter-based stream of rng.c; the ARMS uniform
 * k of one call at (site, iteration, chain) is taken from
 *   block(site, iteration, STREAM_ARMS | (k >> 1) << 8, chain),
 * words (0,1) for even k and (2,3) for odd k, as ((a << 32 | b) >> 11 + 0.5) 2^-53,
 * which, like the reference's u_random (BayesW_arms.cpp:910-915), is never 0 or 1.
 */
#include "brr_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define BW_PI 3.14159
#define BW_PI2 6.283185
#define BW_SQRT_PI 1.77245385090552
#define BW_EULER 0.577215664901532

/* ---------------------------------------------------------------- ARMS ---- */

#define A_XEPS 0.00001
#define A_YEPS 0.1
#define A_EYEPS 0.001
#define A_YCEIL 50.0

typedef struct {
    double x, y, ey, cum;
    int f;          /* 1: evaluated point of the log density; 0: intersection / bound */
    int pl, pr;     /* neighbours, -1 at the ends */
} apoint;

typedef struct {
    apoint p[ORC_ARMS_NPOINT + 1];   /* + the working point */
    int cpoint, npoint;
    double ymax;
    int neval;
} aenv;

typedef struct {
    orc_logdens f;
    void* ctx;
    const orc_arms_rng* r;
    uint32_t k;
} afun;

static double arms_u(afun* a)
{
    const uint32_t k = a->k++;
    const uint32_t ctr[4] = {a->r->site, a->r->iteration, (uint32_t)ORC_STREAM_ARMS | ((k >> 1) << 8),
                             a->r->chain};
    const uint32_t key[2] = {(uint32_t)a->r->seed, (uint32_t)(a->r->seed >> 32)};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    const uint32_t hi = (k & 1u) ? o[2] : o[0], lo = (k & 1u) ? o[3] : o[1];
    const uint64_t w = ((uint64_t)hi << 32) | lo;
    return ((double)(w >> 11) + 0.5) * 1.1102230246251565404e-16;
}

static double eval(afun* a, aenv* e, double x)
{
    e->neval++;
    return a->f(x, a->ctx);
}

static double expshift(double y, double y0)
{
    return (y - y0 > -2.0 * A_YCEIL) ? exp(y - y0 + A_YCEIL) : 0.0;
}

static double logshift(double y, double y0) { return log(y) + y0 - A_YCEIL; }

#define P(i) (e->p[(i)])

/* intersection of the chords either side of intersection point q
 * (BayesW_arms.cpp:707-808); nonzero when the envelope would not bound the
 * density (a convexity violation, fatal without the Metropolis step) */
static int meet(aenv* e, int q)
{
    double gl = 0, gr = 0, grl = 0, dl = 0, dr = 0;
    int il = 0, ir = 0, irl = 0;
    if (P(q).f) return 3;
    const int ql = P(q).pl, qr = P(q).pr;
    if (ql >= 0 && P(ql).pl >= 0 && P(P(ql).pl).pl >= 0) {
        const int q3 = P(P(ql).pl).pl;
        gl = (P(ql).y - P(q3).y) / (P(ql).x - P(q3).x);
        il = 1;
    }
    if (qr >= 0 && P(qr).pr >= 0 && P(P(qr).pr).pr >= 0) {
        const int q3 = P(P(qr).pr).pr;
        gr = (P(qr).y - P(q3).y) / (P(qr).x - P(q3).x);
        ir = 1;
    }
    if (ql >= 0 && qr >= 0) {
        grl = (P(qr).y - P(ql).y) / (P(qr).x - P(ql).x);
        irl = 1;
    }
    if (irl && il && gl < grl) return 1;
    if (irl && ir && gr > grl) return 1;
    if (il && irl) {
        dr = (gl - grl) * (P(qr).x - P(ql).x);
        if (dr < A_YEPS) dr = A_YEPS;
    }
    if (ir && irl) {
        dl = (grl - gr) * (P(qr).x - P(ql).x);
        if (dl < A_YEPS) dl = A_YEPS;
    }
    if (il && ir && irl) {
        P(q).x = (dl * P(qr).x + dr * P(ql).x) / (dl + dr);
        P(q).y = (dl * P(qr).y + dr * P(ql).y + dl * dr) / (dl + dr);
    } else if (il && irl) {
        P(q).x = P(qr).x;
        P(q).y = P(qr).y + dr;
    } else if (ir && irl) {
        P(q).x = P(ql).x;
        P(q).y = P(ql).y + dl;
    } else if (il) {
        P(q).y = P(ql).y + gl * (P(q).x - P(ql).x);
    } else if (ir) {
        P(q).y = P(qr).y - gr * (P(qr).x - P(q).x);
    } else {
        return 4;
    }
    if ((ql >= 0 && P(q).x < P(ql).x) || (qr >= 0 && P(q).x > P(qr).x)) return 5;
    return 0;
}

/* area of the exponentiated envelope left of q (BayesW_arms.cpp:812-834) */
static double area(const aenv* e, int q)
{
    const int l = P(q).pl;
    if (P(l).x == P(q).x) return 0.0;
    if (fabs(P(q).y - P(l).y) < A_YEPS) return 0.5 * (P(q).ey + P(l).ey) * (P(q).x - P(l).x);
    return ((P(q).ey - P(l).ey) / (P(q).y - P(l).y)) * (P(q).x - P(l).x);
}

static int leftmost(const aenv* e)
{
    int q = 0;
    while (P(q).pl >= 0) q = P(q).pl;
    return q;
}

/* exponentiate and integrate the envelope (BayesW_arms.cpp:672-703) */
static void cumulate(aenv* e)
{
    const int lm = leftmost(e);
    e->ymax = P(lm).y;
    for (int q = P(lm).pr; q >= 0; q = P(q).pr)
        if (P(q).y > e->ymax) e->ymax = P(q).y;
    for (int q = lm; q >= 0; q = P(q).pr) P(q).ey = expshift(P(q).y, e->ymax);
    P(lm).cum = 0.0;
    for (int q = P(lm).pr; q >= 0; q = P(q).pr) P(q).cum = P(P(q).pl).cum + area(e, q);
}

/* initial envelope: bounds, the ninit evaluated points and the intersections
 * between them (BayesW_arms.cpp:244-340) */
static int initial(aenv* e, afun* a, const double* xinit, int ninit, double xl, double xr, int npoint)
{
    if (ninit < 3) return 1001;
    const int mpoint = 2 * ninit + 1;
    if (npoint < mpoint || npoint > ORC_ARMS_NPOINT) return 1002;
    if (xinit[0] <= xl || xinit[ninit - 1] >= xr) return 1003;
    for (int i = 1; i < ninit; ++i)
        if (xinit[i] <= xinit[i - 1]) return 1004;
    e->neval = 0;
    e->npoint = npoint;
    for (int j = 0; j < mpoint; ++j) {
        P(j).pl = j - 1;
        P(j).pr = (j + 1 < mpoint) ? j + 1 : -1;
        P(j).f = 0;
        P(j).y = 0.0;
    }
    P(0).x = xl;
    for (int j = 1, k = 0; j < mpoint - 1; ++j) {
        if (j % 2) {
            P(j).x = xinit[k++];
            P(j).y = eval(a, e, P(j).x);
            P(j).f = 1;
        }
    }
    P(mpoint - 1).x = xr;
    for (int j = 0; j < mpoint; j += 2)
        if (meet(e, j)) return 2000;
    cumulate(e);
    e->cpoint = mpoint;
    return 0;
}

/* the x whose envelope cumulative probability is prob, into the working point
 * w (BayesW_arms.cpp:364-425) */
static int invert(aenv* e, double prob, int w)
{
    int q = 0;
    while (P(q).pr >= 0) q = P(q).pr;
    const double u = prob * P(q).cum;
    while (P(P(q).pl).cum > u) q = P(q).pl;
    const int l = P(q).pl;
    P(w).pl = l;
    P(w).pr = q;
    P(w).f = 0;
    P(w).cum = u;
    const double prop = (u - P(l).cum) / (P(q).cum - P(l).cum);
    double xl = P(q).x, xr = P(q).x;
    if (P(l).x == P(q).x) {
        P(w).x = P(q).x;
        P(w).y = P(q).y;
        P(w).ey = P(q).ey;
    } else {
        xl = P(l).x;
        xr = P(q).x;
        const double yl = P(l).y, yr = P(q).y, eyl = P(l).ey, eyr = P(q).ey;
        if (fabs(yr - yl) < A_YEPS) {
            if (fabs(eyr - eyl) > A_EYEPS * fabs(eyr + eyl))
                P(w).x = xl + ((xr - xl) / (eyr - eyl)) * (-eyl + sqrt((1. - prop) * eyl * eyl + prop * eyr * eyr));
            else
                P(w).x = xl + (xr - xl) * prop;
            P(w).ey = ((P(w).x - xl) / (xr - xl)) * (eyr - eyl) + eyl;
            P(w).y = logshift(P(w).ey, e->ymax);
        } else {
            P(w).x = xl + ((xr - xl) / (yr - yl)) * (-yl + logshift(((1. - prop) * eyl + prop * eyr), e->ymax));
            P(w).y = ((P(w).x - xl) / (xr - xl)) * (yr - yl) + yl;
            P(w).ey = expshift(P(w).y, e->ymax);
        }
    }
    if (P(w).x < xl || P(w).x > xr) return 6;
    return 0;
}

/* add the evaluated working point to the envelope (BayesW_arms.cpp:540-668) */
static int update(aenv* e, afun* a, int w)
{
    if (!P(w).f || e->cpoint > e->npoint - 2) return 0;
    const int q = e->cpoint++;
    P(q).x = P(w).x;
    P(q).y = P(w).y;
    P(q).f = 1;
    const int m = e->cpoint++;
    P(m).f = 0;
    const int wl = P(w).pl, wr = P(w).pr;
    if (P(wl).f && !P(wr).f) {
        P(m).pl = wl;
        P(m).pr = q;
        P(q).pl = m;
        P(q).pr = wr;
        P(wl).pr = m;
        P(wr).pl = q;
    } else if (!P(wl).f && P(wr).f) {
        P(m).pr = wr;
        P(m).pl = q;
        P(q).pr = m;
        P(q).pl = wl;
        P(wr).pl = m;
        P(wl).pr = q;
    } else {
        return 7;
    }
    const int ql = (P(P(q).pl).pl >= 0) ? P(P(q).pl).pl : P(q).pl;
    const int qr = (P(P(q).pr).pr >= 0) ? P(P(q).pr).pr : P(q).pr;
    if (P(q).x < (1. - A_XEPS) * P(ql).x + A_XEPS * P(qr).x) {
        P(q).x = (1. - A_XEPS) * P(ql).x + A_XEPS * P(qr).x;
        P(q).y = eval(a, e, P(q).x);
    } else if (P(q).x > A_XEPS * P(ql).x + (1. - A_XEPS) * P(qr).x) {
        P(q).x = A_XEPS * P(ql).x + (1. - A_XEPS) * P(qr).x;
        P(q).y = eval(a, e, P(q).x);
    }
    if (meet(e, P(q).pl)) return 1;
    if (meet(e, P(q).pr)) return 1;
    if (P(P(q).pl).pl >= 0 && meet(e, P(P(P(q).pl).pl).pl)) return 1;
    if (P(P(q).pr).pr >= 0 && meet(e, P(P(P(q).pr).pr).pr)) return 1;
    cumulate(e);
    return 0;
}

int orc_arms(const double* xinit, int ninit, double xl, double xr, orc_logdens f, void* ctx, int npoint,
             const orc_arms_rng* rng, double* xsamp, int* neval)
{
    /* arms (BayesW_arms.cpp:133-240) with nsamp = 1, no Metropolis step; the
     * requested envelope centiles are not used by BayesW and are skipped */
    aenv* e = (aenv*)calloc(1, sizeof(aenv));
    if (!e) return 1006;
    afun a = {f, ctx, rng, 0};
    int err = initial(e, &a, xinit, ninit, xl, xr, npoint);
    const int w = ORC_ARMS_NPOINT;          /* the working point */
    while (!err) {
        /* sample (BayesW_arms.cpp:344-360) */
        if ((err = invert(e, arms_u(&a), w))) break;
        /* test (BayesW_arms.cpp:429-536), squeezing then rejection */
        const double u = arms_u(&a) * P(w).ey;
        const double y = logshift(u, e->ymax);
        const int wl = P(w).pl, wr = P(w).pr;
        if (P(wl).pl >= 0 && P(wr).pr >= 0) {
            const int ql = P(wl).f ? wl : P(wl).pl;
            const int qr = P(wr).f ? wr : P(wr).pr;
            const double ysqueez = (P(qr).y * (P(w).x - P(ql).x) + P(ql).y * (P(qr).x - P(w).x)) / (P(qr).x - P(ql).x);
            if (y <= ysqueez) { *xsamp = P(w).x; break; }
        }
        const double ynew = eval(&a, e, P(w).x);
        P(w).y = ynew;
        P(w).ey = expshift(P(w).y, e->ymax);
        P(w).f = 1;
        if (update(e, &a, w)) { err = 2000; break; }
        if (!(y >= ynew)) { *xsamp = P(w).x; break; }
    }
    if (neval) *neval = e->neval;
    free(e);
    return err;
}

#undef P

/* ------------------------------------------------- Gauss-Hermite rules ---- */

/* The adaptive rules of gauss_hermite_adaptive_integral (bayeswbase.cpp:126-457):
 * positive nodes x_1 > x_3 > ... (each used at +x and -x) with weights w that
 * already include exp(x^2), and the weight of the centre node. For n >= 9 the
 * reference evaluates the sixth node at -x3 instead of -x5; *_x6 holds the node
 * it uses. */
typedef struct {
    int n;
    double x[7], w[7], wc;
} gh_rule;

static const gh_rule GH_RULES[7] = {
    {3, {1.2247448713916}, {1.3239311752136}, 1.1816359006037},
    {5, {2.0201828704561, 0.95857246461382}, {1.181488625536, 0.98658099675143}, 0.94530872048294},
    {7, {2.6519613568352, 1.6735516287675, 0.81628788285897},
        {1.1013307296103, 0.8971846002252, 0.8286873032836}, 0.81026461755681},
    {9, {3.1909932017815, 2.2665805845318, 1.4685532892167, 0.72355101875284},
        {1.0470035809767, 0.84175270147867, 0.7646081250946, 0.73030245274509}, 0.72023521560605},
    {11, {3.6684708465596, 2.7832900997817, 2.0259480158258, 1.3265570844949, 0.6568095668821},
         {1.0065267861724, 0.802516868851, 0.721953624728, 0.6812118810667, 0.66096041944096}, 0.65475928691459},
    {13, {4.1013375961786, 3.2466089783724, 2.5197356856782, 1.8531076516015, 1.2200550365908, 0.60576387917106},
         {0.97458039564, 0.7725808233517, 0.6906180348378, 0.6467594633158, 0.6217160552868, 0.60852958370332},
         0.60439318792116},
    {15, {4.4999907073094, 3.6699503734045, 2.9671669279056, 2.3257324861739, 1.7199925751865, 1.1361155852109,
          0.5650695832556},
         {0.94836897082761, 0.7486073660169, 0.666166005109, 0.620662603527, 0.5930274497642, 0.5761933502835,
          0.5670211534466},
         0.56410030872642},
};

int orc_bw_gh_rule(int n, double* x, double* w, double* wc)
{
    if (n < 3 || n > 15 || (n % 2) == 0) return 0;
    const gh_rule* r = &GH_RULES[(n - 3) / 2];
    const int h = (n - 1) / 2;
    /* off-centre nodes in the reference's order: +x1, -x1, +x3, -x3, ... */
    for (int i = 0; i < h; ++i) {
        x[2 * i] = r->x[i];
        x[2 * i + 1] = -r->x[i];
        w[2 * i] = r->w[i];
        w[2 * i + 1] = r->w[i];
    }
    if (n >= 9) x[5] = -r->x[1];     /* x6 = -x3 (bayeswbase.cpp:230, 276, 331, 394) */
    *wc = r->wc;
    return n - 1;
}

/* ------------------------------------------------------ the marker step ---- */

typedef struct {
    double vi_sum, vi_0, vi_1, vi_2;
} vi_sums;

/* RaggedBayesWKernel::setVi (raggedbayeswkernel.cpp:10-16) */
static vi_sums set_vi(const orc_data* d, uint32_t j, const double* vi)
{
    vi_sums s;
    double a = 0.0;
    for (uint32_t i = 0; i < d->N; ++i) a += vi[i];
    s.vi_sum = a;
    double t2 = 0.0, t1 = 0.0;
    for (uint64_t q = d->twos_off[j]; q < d->twos_off[j + 1]; ++q) t2 += vi[d->twos[q]];
    for (uint64_t q = d->ones_off[j]; q < d->ones_off[j + 1]; ++q) t1 += vi[d->ones[q]];
    s.vi_2 = t2;
    s.vi_1 = t1;
    s.vi_0 = s.vi_sum - s.vi_1 - s.vi_2;
    return s;
}

/* The dense column Cx of marker j: the table entry of each individual's call. */
static void dense_cx(const orc_data* d, uint32_t j, double* cx)
{
    const uint8_t* col = d->bed + (size_t)j * ((d->N + 3u) >> 2);
    const double* lut = d->lut + 4 * (size_t)j;
    for (uint32_t i = 0; i < d->N; ++i) {
        const unsigned c = (col[i >> 2] >> ((i & 3u) * 2u)) & 3u;
        cx[i] = lut[(c == 0u) ? 2 : (c == 2u) ? 1 : (c == 3u) ? 0 : 3];
    }
}

double orc_bw_sum_failure(const orc_data* d, uint32_t j, const double* fail)
{
    if (d->layout == ORC_DENSE) {
        /* DenseBayesWKernel::calculateSumFailure (densebayeswkernel.cpp:15-18) */
        double* cx = (double*)malloc(sizeof(double) * d->N);
        dense_cx(d, j, cx);
        double s = 0.0;
        for (uint32_t i = 0; i < d->N; ++i) s += cx[i] * fail[i];
        free(cx);
        return s;
    }
    /* calculateSumFailure (raggedbayeswkernel.cpp:18-25): the list sums go
     * through an int accumulator */
    int temp_sum = 0;
    double s1 = 0.0, s2 = 0.0, sf = 0.0;
    for (uint64_t q = d->ones_off[j]; q < d->ones_off[j + 1]; ++q) s1 += fail[d->ones[q]];
    temp_sum += (int)s1;
    for (uint64_t q = d->twos_off[j]; q < d->twos_off[j + 1]; ++q) s2 += fail[d->twos[q]];
    temp_sum += (int)(2 * s2);
    for (uint32_t i = 0; i < d->N; ++i) sf += fail[i];
    const double mean = d->stats[4 * (size_t)j], sd = d->stats[4 * (size_t)j + 1];
    return (temp_sum - mean * sf) / sd;
}

/* RaggedBayesWKernel::integrand_adaptive (raggedbayeswkernel.cpp:70-77) */
static double integrand(double s, double alpha, double sq, double sum_failure, double mean_sd, double sd,
                        const vi_sums* v)
{
    const double temp = -alpha * s * sum_failure * sq + v->vi_sum -
                        exp(alpha * mean_sd * s * sq) *
                            (v->vi_0 + v->vi_1 * exp(-alpha * s * sq / sd) + v->vi_2 * exp(-2 * alpha * s * sq / sd)) -
                        s * s;
    return exp(temp);
}

double orc_bw_gh_integral(int n, double sigma, double C_k, double sigma_b, double alpha, double sum_failure,
                          double mean, double sd, const double vi[4])
{
    /* gauss_hermite_adaptive_integral (bayeswbase.cpp:126-457) */
    double x[14], w[14], wc;
    const int m = orc_bw_gh_rule(n, x, w, &wc);
    const double sq = sqrt(2 * C_k * sigma_b);
    const vi_sums v = {vi[0], vi[1], vi[2], vi[3]};
    double temp = 0.0;
    for (int i = 0; i < m; ++i) {
        const double t = w[i] * integrand(sigma * x[i], alpha, sq, sum_failure, mean / sd, sd, &v);
        temp = (i == 0) ? t : temp + t;
    }
    temp = temp + wc;
    return sigma * temp;
}

typedef struct {
    double alpha, sigma_b, sum_failure, used_mixture, mean_sd_ratio, sd, vi_0, vi_1, vi_2;
} beta_ctx;

/* beta_dens (sparsebayesw.cpp:26-37) */
static double beta_dens(double x, void* c)
{
    const beta_ctx* p = (const beta_ctx*)c;
    return -p->alpha * x * p->sum_failure -
           exp(p->alpha * x * p->mean_sd_ratio) *
               (p->vi_0 + p->vi_1 * exp(-p->alpha * x / p->sd) + p->vi_2 * exp(-2 * p->alpha * x / p->sd)) -
           x * x / (2 * p->used_mixture * p->sigma_b);
}

typedef struct {
    double alpha, sigma_b, sum_failure, used_mixture;
    const double *cx, *eps, *vi;
    uint32_t N;
} dense_ctx;

/* DenseBayesW beta_dens (densebayesw.cpp:25-33), every individual's term */
static double dense_beta_dens(double x, void* c)
{
    const dense_ctx* p = (const dense_ctx*)c;
    double s = 0.0;
    for (uint32_t i = 0; i < p->N; ++i) s += exp(((p->eps[i] - p->cx[i] * x) * p->alpha) - BW_EULER);
    return -p->alpha * x * p->sum_failure - s - x * x / (2 * p->used_mixture * p->sigma_b);
}

/* DenseBayesWKernel::integrand_adaptive (densebayeswkernel.cpp:36-42) */
static double dense_integrand(double s, double alpha, double sq, double sum_failure, const double* cx,
                              const double* vi, uint32_t N)
{
    double t = 0.0;
    for (uint32_t i = 0; i < N; ++i) t += vi[i] * (1 - exp(-cx[i] * s * sq * alpha));
    const double temp = -alpha * s * sum_failure * sq + t - s * s;
    return exp(temp);
}

static double dense_gh_integral(int n, double sigma, double C_k, double sigma_b, double alpha, double sum_failure,
                                const double* cx, const double* vi, uint32_t N)
{
    /* gauss_hermite_adaptive_integral (bayeswbase.cpp:126-457) with the dense integrand */
    double x[14], w[14], wc;
    const int m = orc_bw_gh_rule(n, x, w, &wc);
    const double sq = sqrt(2 * C_k * sigma_b);
    double temp = 0.0;
    for (int i = 0; i < m; ++i) {
        const double t = w[i] * dense_integrand(sigma * x[i], alpha, sq, sum_failure, cx, vi, N);
        temp = (i == 0) ? t : temp + t;
    }
    temp = temp + wc;
    return sigma * temp;
}

/* calculateResidualUpdate (raggedbayeswkernel.cpp:27-43; dense: Cx * beta,
 * densebayeswkernel.cpp:20-23), eps += sign * delta */
static void residual_update(const orc_data* d, uint32_t j, double* eps, double beta, int sign)
{
    if (d->layout == ORC_DENSE) {
        double* cx = (double*)malloc(sizeof(double) * d->N);
        dense_cx(d, j, cx);
        for (uint32_t i = 0; i < d->N; ++i) {
            const double delta = cx[i] * beta;
            eps[i] = (sign > 0) ? eps[i] + delta : eps[i] - delta;
        }
        free(cx);
        return;
    }
    const double mean = d->stats[4 * (size_t)j], sd = d->stats[4 * (size_t)j + 1];
    const double mean_sd_ratio = mean / sd;
    const double meanAdjustment = mean_sd_ratio * beta;
    const double oneAdjustment = beta / sd;
    const double twoAdjustment = 2 * oneAdjustment;
    double* delta = (double*)malloc(sizeof(double) * d->N);
    for (uint32_t i = 0; i < d->N; ++i) delta[i] = -meanAdjustment;
    for (uint64_t q = d->ones_off[j]; q < d->ones_off[j + 1]; ++q) delta[d->ones[q]] += oneAdjustment;
    for (uint64_t q = d->twos_off[j]; q < d->twos_off[j + 1]; ++q) delta[d->twos[q]] += twoAdjustment;
    if (sign > 0)
        for (uint32_t i = 0; i < d->N; ++i) eps[i] += delta[i];
    else
        for (uint32_t i = 0; i < d->N; ++i) eps[i] -= delta[i];
    free(delta);
}

static void refresh_vi(uint32_t N, const double* eps, double alpha, double* vi)
{
    for (uint32_t i = 0; i < N; ++i) vi[i] = exp(alpha * eps[i] - BW_EULER);
}

int orc_bw_process_column(const orc_data* d, const orc_bw_model* m, orc_bw_state* s, uint32_t j, double p,
                          uint64_t seed, uint32_t iteration, uint32_t chain, orc_bw_trace* tr)
{
    const double beta_old = s->beta[j];
    if (beta_old != 0.0) {
        residual_update(d, j, s->eps, beta_old, +1);
        refresh_vi(d->N, s->eps, s->alpha, s->vi);
    }
    const int dense = d->layout == ORC_DENSE;
    vi_sums v = {0, 0, 0, 0};
    double mean = 0.0, sd = 1.0;
    double* cx = NULL;
    const double sum_failure = m->sum_failure[j];
    const int K = m->K;
    double ml[ORC_BW_MAXK];
    ml[0] = s->pi[0] * BW_SQRT_PI;
    double exp_sum;
    if (dense) {
        /* DenseBayesWKernel::exponent_sum (densebayeswkernel.cpp:30-34) */
        cx = (double*)malloc(sizeof(double) * d->N);
        dense_cx(d, j, cx);
        exp_sum = 0.0;
        for (uint32_t i = 0; i < d->N; ++i) exp_sum += s->vi[i] * cx[i] * cx[i];
        for (uint32_t i = 0; i < d->N; ++i) v.vi_sum += s->vi[i];
    } else {
        v = set_vi(d, j, s->vi);
        mean = d->stats[4 * (size_t)j];
        sd = d->stats[4 * (size_t)j + 1];
        exp_sum = (v.vi_1 * (1 - 2 * mean) + 4 * (1 - mean) * v.vi_2 + v.vi_sum * mean * mean) / (sd * sd);
    }
    const double vv[4] = {v.vi_sum, v.vi_0, v.vi_1, v.vi_2};
    for (int i = 0; i < K - 1; ++i) {
        const double sigma = 1.0 / sqrt(1 + s->alpha * s->alpha * s->sigma_b * m->C[i] * exp_sum);
        ml[i + 1] = s->pi[i + 1] *
                    (dense ? dense_gh_integral(m->quad_points, sigma, m->C[i], s->sigma_b, s->alpha, sum_failure, cx,
                                               s->vi, d->N)
                           : orc_bw_gh_integral(m->quad_points, sigma, m->C[i], s->sigma_b, s->alpha, sum_failure,
                                                mean, sd, vv));
    }
    double ml_sum = 0.0;
    for (int i = 0; i < K; ++i) ml_sum += ml[i];
    double acum = ml[0] / ml_sum;
    const double acum0 = acum;
    int component = 0;
    double beta_new = beta_old;
    int err = 0;
    for (int k = 0; k < K; ++k) {
        if (p <= acum) {
            if (k == 0) {
                beta_new = 0;
            } else {
                beta_ctx c = {s->alpha, s->sigma_b, sum_failure, m->C[k - 1], mean / sd, sd, v.vi_0, v.vi_1, v.vi_2};
                dense_ctx dc = {s->alpha, s->sigma_b, sum_failure, m->C[k - 1], cx, s->eps, s->vi, d->N};
                const double safe_limit = 2 * sqrt(s->sigma_b * m->C[k - 1]);
                const double xinit[4] = {beta_old - safe_limit / 10, beta_old, beta_old + safe_limit / 20,
                                         beta_old + safe_limit / 10};
                const orc_arms_rng r = {seed, j, iteration, chain};
                double xs = beta_old;
                err = dense ? orc_arms(xinit, 4, beta_old - safe_limit, beta_old + safe_limit, dense_beta_dens, &dc,
                                       100, &r, &xs, NULL)
                            : orc_arms(xinit, 4, beta_old - safe_limit, beta_old + safe_limit, beta_dens, &c, 100,
                                       &r, &xs, NULL);
                beta_new = xs;
            }
            s->v[k] += 1.0;
            component = k;
            break;
        } else {
            /* the last class gets acum = 1 (bayeswbase.cpp:685-687) */
            if ((k + 1) == (K - 1)) acum = 1;
            else acum += ml[k + 1] / ml_sum;
        }
    }
    const int skip = beta_old == 0.0 && beta_new == 0.0;
    if (!skip) {
        residual_update(d, j, s->eps, beta_new, -1);
        refresh_vi(d->N, s->eps, s->alpha, s->vi);
    }
    s->comp[j] = component;
    s->beta[j] = beta_new;
    free(cx);
    if (tr) {
        tr->vi_sum = v.vi_sum; tr->vi_1 = v.vi_1; tr->vi_2 = v.vi_2; tr->acum = acum0;
        for (int i = 0; i < K; ++i) tr->ml[i] = ml[i];
    }
    return err;
}

int orc_bw_sweep(const orc_data* d, const orc_bw_model* m, orc_bw_state* s, const uint32_t* order,
                 const double* u, uint64_t seed, uint32_t iteration, uint32_t chain)
{
    /* m_v.setOnes() before exec (bayeswbase.cpp:789) is the caller's */
    for (uint32_t i = 0; i < d->M; ++i) {
        const uint32_t j = order[i];
        const int err = orc_bw_process_column(d, m, s, j, u[j], seed, iteration, chain, NULL);
        if (err) return err;
    }
    return 0;
}

/* ---------------------------------------------------- per iteration ---- */

void orc_bw_init(const orc_data* d, const orc_bw_model* m, orc_bw_state* s, const double* y,
                 const double* fail)
{
    /* BayesWBase::init (bayeswbase.cpp:469-540), no fixed effects */
    const uint32_t N = d->N, M = d->M;
    const int K = m->K, km1 = K - 1;
    s->pi[0] = 0.99;
    for (int k = 1; k < K; ++k) s->pi[k] = (1 - s->pi[0]) / km1;
    for (uint32_t j = 0; j < M; ++j) { s->beta[j] = 0.0; s->comp[j] = 0; }
    double sy = 0.0;
    for (uint32_t i = 0; i < N; ++i) sy += y[i];
    s->mu = sy / N;
    double dd = 0.0;
    for (uint32_t i = 0; i < N; ++i) dd += fail[i];
    s->d = dd;
    double ss = 0.0;
    for (uint32_t i = 0; i < N; ++i) ss += (y[i] - s->mu) * (y[i] - s->mu);
    const double denominator = (6 * ss / (N - 1));
    s->alpha = BW_PI / sqrt(denominator);
    for (uint32_t i = 0; i < N; ++i) s->eps[i] = y[i] - s->mu;
    s->sigma_b = BW_PI2 / (6 * (s->alpha * s->alpha) * M);
    for (int k = 0; k < K; ++k) s->v[k] = 1.0;
}

typedef struct {
    double alpha, d, mu;
    const double* eps;
    uint32_t N;
} mu_ctx;

/* mu_dens (bayeswbase.cpp:66-77) on epsilon + mu */
static double mu_dens(double x, void* c)
{
    const mu_ctx* p = (const mu_ctx*)c;
    double s = 0.0;
    for (uint32_t i = 0; i < p->N; ++i) s += exp(((p->eps[i] + p->mu) - x) * p->alpha - BW_EULER);
    return -p->alpha * x * p->d - s - x * x / (2 * 100.0);
}

int orc_bw_sample_mu(const orc_data* d, orc_bw_state* s, uint64_t seed, uint32_t iteration, uint32_t chain)
{
    /* sampleMu (bayeswbase.cpp:542-569) */
    mu_ctx c = {s->alpha, s->d, s->mu, s->eps, d->N};
    const double mu = s->mu;
    const double xinit[4] = {0.995 * mu, mu, 1.005 * mu, 1.01 * mu};
    const orc_arms_rng r = {seed, ORC_BW_SITE_MU, iteration, chain};
    double xs = mu;
    const int err = orc_arms(xinit, 4, 2.0, 5.0, mu_dens, &c, 100, &r, &xs, NULL);
    if (err) return err;
    for (uint32_t i = 0; i < d->N; ++i) s->eps[i] = (s->eps[i] + mu) - xs;
    s->mu = xs;
    return 0;
}

typedef struct {
    double d, ef;
    const double* eps;
    uint32_t N;
} alpha_ctx;

/* alpha_dens (bayeswbase.cpp:109-119) */
static double alpha_dens(double x, void* c)
{
    const alpha_ctx* p = (const alpha_ctx*)c;
    double s = 0.0;
    for (uint32_t i = 0; i < p->N; ++i) s += exp(p->eps[i] * x - BW_EULER);
    return (0.01 + p->d - 1) * log(x) + x * (p->ef - 0.01) - s;
}

int orc_bw_sample_alpha(const orc_data* d, orc_bw_state* s, const double* fail, uint64_t seed,
                        uint32_t iteration, uint32_t chain)
{
    /* sampleAlpha (bayeswbase.cpp:707-734) */
    double ef = 0.0;
    for (uint32_t i = 0; i < d->N; ++i) ef += s->eps[i] * fail[i];
    alpha_ctx c = {s->d, ef, s->eps, d->N};
    const double a = s->alpha;
    const double xinit[4] = {a * 0.5, a, a * 1.15, a * 1.5};
    const orc_arms_rng r = {seed, ORC_BW_SITE_ALPHA, iteration, chain};
    double xs = a;
    const int err = orc_arms(xinit, 4, 0.0, 20.0, alpha_dens, &c, 100, &r, &xs, NULL);
    if (err) return err;
    s->alpha = xs;
    return 0;
}

void orc_bw_post_sweep(const orc_data* d, const orc_bw_model* m, orc_bw_state* s, uint64_t seed,
                       uint32_t iteration, uint32_t chain)
{
    /* sigma_b and pi_L draws (bayeswbase.cpp:798-806) */
    const uint32_t M = d->M;
    double b2 = 0.0;
    for (uint32_t j = 0; j < M; ++j) b2 += s->beta[j] * s->beta[j];
    const double nz = (double)M - s->v[0] + 1;
    const double shape = 1.0 + 0.5 * nz;
    const double scale = 0.0001 + 0.5 * nz * b2;
    s->sigma_b = 1.0 / orc_rgamma(seed, iteration, ORC_BW_DRAW_SIGMAB, chain, shape, 1.0 / scale);
    orc_dirichlet_rng(seed, iteration, ORC_BW_DRAW_PI0, chain, m->K, s->v, s->pi);
}

typedef struct {
    double alpha, sumf;
    const double* eps;      /* eps + X_f theta_f */
    const float* x;
    uint32_t N;
} theta_ctx;

/* theta_dens (bayeswbase.cpp:87-97), sigma_mu = 100 */
static double theta_dens(double x, void* c)
{
    const theta_ctx* p = (const theta_ctx*)c;
    double s = 0.0;
    for (uint32_t i = 0; i < p->N; ++i) s += exp((p->eps[i] - (double)p->x[i] * x) * p->alpha - BW_EULER);
    return -p->alpha * x * p->sumf - s - x * x / (2 * 100.0);
}

void orc_bw_sum_failure_fixed(uint32_t N, const float* X, uint32_t F, const double* fail, double* sumf)
{
    /* BayesWBase::init (bayeswbase.cpp:532-537): sum_k X(k,f) failure_k */
    for (uint32_t f = 0; f < F; ++f) {
        double t = 0.0;
        for (uint32_t k = 0; k < N; ++k) t += (double)X[(size_t)f * N + k] * fail[k];
        sumf[f] = t;
    }
}

int orc_bw_sample_theta(const orc_data* d, orc_bw_state* s, const float* X, uint32_t f, double* theta,
                        const double* sumf, uint64_t seed, uint32_t iteration, uint32_t chain)
{
    /* sampleTheta (bayeswbase.cpp:572-601) */
    const uint32_t N = d->N;
    const float* x = X + (size_t)f * N;
    const double th = theta[f];
    double* eps_adj = (double*)malloc(sizeof(double) * N);
    for (uint32_t i = 0; i < N; ++i) eps_adj[i] = s->eps[i] + (double)x[i] * th;
    theta_ctx c = {s->alpha, sumf[f], eps_adj, x, N};
    const double xinit[4] = {th - 0.01, th, th + 0.005, th + 0.01};
    const orc_arms_rng r = {seed, ORC_BW_SITE_THETA0 + f, iteration, chain};
    double xs = th;
    const int err = orc_arms(xinit, 4, -2.0, 2.0, theta_dens, &c, 100, &r, &xs, NULL);
    if (!err) {
        theta[f] = xs;
        for (uint32_t i = 0; i < N; ++i) s->eps[i] = eps_adj[i] - (double)x[i] * xs;
    }
    free(eps_adj);
    return err;
}

int orc_bw_run_gibbs(const orc_data* d, const orc_bw_model* m, orc_bw_state* s, const double* y,
                     const double* fail, uint64_t seed, uint32_t chain, uint32_t iterations, uint32_t* order,
                     double* hist)
{
    return orc_bw_run_gibbs_fixed(d, m, s, y, fail, seed, chain, iterations, order, hist, NULL, 0, NULL);
}

int orc_bw_run_gibbs_fixed(const orc_data* d, const orc_bw_model* m, orc_bw_state* s, const double* y,
                           const double* fail, uint64_t seed, uint32_t chain, uint32_t iterations,
                           uint32_t* order, double* hist, const float* X, uint32_t F, double* theta)
{
    /* BayesWBase::runGibbs (bayeswbase.cpp:737-830), Philox streams; F > 0: the
     * fixed effects after mu (:783-788) */
    const uint32_t M = d->M;
    orc_bw_init(d, m, s, y, fail);
    double* u = (double*)malloc(sizeof(double) * M);
    double* z = (double*)malloc(sizeof(double) * M);
    double* sumf = (double*)malloc(sizeof(double) * (F ? F : 1));
    if (F) orc_bw_sum_failure_fixed(d->N, X, F, fail, sumf);
    int err = 0;
    for (uint32_t it = 0; it < iterations && !err; ++it) {
        if ((err = orc_bw_sample_mu(d, s, seed, it, chain))) break;
        for (uint32_t f = 0; f < F && !err; ++f) err = orc_bw_sample_theta(d, s, X, f, theta, sumf, seed, it, chain);
        if (err) break;
        refresh_vi(d->N, s->eps, s->alpha, s->vi);
        orc_shuffle_philox(seed, it, chain, M, order);
        for (int k = 0; k < m->K; ++k) s->v[k] = 1.0;
        orc_fill_marker_uz(seed, it, chain, M, u, z);
        if ((err = orc_bw_sweep(d, m, s, order, u, seed, it, chain))) break;
        if ((err = orc_bw_sample_alpha(d, s, fail, seed, it, chain))) break;
        orc_bw_post_sweep(d, m, s, seed, it, chain);
        if (hist) {
            double b2 = 0.0;
            for (uint32_t j = 0; j < M; ++j) b2 += s->beta[j] * s->beta[j];
            double* h = hist + (size_t)it * 5;
            h[0] = s->mu; h[1] = s->alpha; h[2] = s->sigma_b; h[3] = (double)M - s->v[0] + 1; h[4] = b2;
        }
    }
    free(u);
    free(z);
    free(sumf);
    return err;
}
