/*
 * bayesr.c -- CPU ORACLE (test infrastructure only): the BayesR ppbayes
 * sweep of the reference, FP64, sequential, in the reference's operation
 * order where the reference fixes one (Eigen reductions run in index order).
 */
#include "brr_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static inline unsigned code_at(const uint8_t* col, uint32_t i)
{
    return (col[i >> 2] >> ((i & 3u) * 2u)) & 3u;
}

/* PLINK 2-bit value -> dosage index of the lut (0,1,2; 3 = missing), the
 * decode rule of preprocessgraph.cpp:50-51. */
static inline int dosage_of_code(unsigned c)
{
    return (c == 0u) ? 2 : (c == 2u) ? 1 : (c == 3u) ? 0 : 3;
}

static inline const uint8_t* dense_col(const orc_data* d, uint32_t j)
{
    return d->bed + (size_t)j * ((d->N + 3u) >> 2);
}

double orc_compute_num(const orc_data* d, uint32_t j, const double* eps,
                       double beta_old, double epsSum)
{
    const double N = (double)d->N;
    if (d->layout == ORC_DENSE) {
        /* densebayesrkernel.cpp:10-15: Cx.dot(eps) + beta_old * (N - 1) */
        const uint8_t* col = dense_col(d, j);
        const double* lut = d->lut + 4 * (size_t)j;
        double dot = 0.0;
        for (uint32_t i = 0; i < d->N; ++i) dot += lut[dosage_of_code(code_at(col, i))] * eps[i];
        return dot + beta_old * (double)(d->N - 1u);
    }
    /* sparsebayesrkernel.cpp:24-27 with raggedbayesrkernel.cpp:32-35:
     * beta_old*(N-1) - mean*epsSum/sd + (sum_ones eps + 2*sum_twos eps)/sd */
    const double mean = d->stats[4 * (size_t)j + 0];
    const double sd = d->stats[4 * (size_t)j + 1];
    if (d->layout == ORC_EIGEN) {
        /* eigenbayesrkernel.cpp:17-20: Zg.dot(eps) / sd, Zg's nonzeros in index
         * order with the values 1, 2 and the mean for a missing call */
        uint64_t a = d->ones_off[j], b = d->twos_off[j], c = d->miss_off[j];
        const uint64_t ae = d->ones_off[j + 1], be = d->twos_off[j + 1], ce = d->miss_off[j + 1];
        double dot = 0.0;
        while (a < ae || b < be || c < ce) {
            const int64_t ia = a < ae ? d->ones[a] : INT64_MAX;
            const int64_t ib = b < be ? d->twos[b] : INT64_MAX;
            const int64_t ic = c < ce ? d->miss[c] : INT64_MAX;
            if (ia < ib && ia < ic) { dot += 1.0 * eps[ia]; ++a; }
            else if (ib < ic) { dot += 2.0 * eps[ib]; ++b; }
            else { dot += mean * eps[ic]; ++c; }
        }
        return beta_old * (N - 1.0) - mean * epsSum / sd + dot / sd;
    }
    double s1 = 0.0, s2 = 0.0;
    for (uint64_t q = d->ones_off[j]; q < d->ones_off[j + 1]; ++q) s1 += eps[d->ones[q]];
    for (uint64_t q = d->twos_off[j]; q < d->twos_off[j + 1]; ++q) s2 += eps[d->twos[q]];
    const double dot = (s1 + 2.0 * s2) / sd;
    return beta_old * (N - 1.0) - mean * epsSum / sd + dot;
}

void orc_apply_epsilon_change(const orc_data* d, uint32_t j, double* eps,
                              double* epsSum, double beta_old, double beta_new)
{
    if (d->layout == ORC_DENSE) {
        /* densebayesrkernel.cpp:17-20: delta = (beta_old - beta) * Cx; eps += delta */
        const uint8_t* col = dense_col(d, j);
        const double* lut = d->lut + 4 * (size_t)j;
        const double dBeta = beta_old - beta_new;
        for (uint32_t i = 0; i < d->N; ++i) {
            const double delta = dBeta * lut[dosage_of_code(code_at(col, i))];
            eps[i] = eps[i] + delta;
        }
        return;
    }
    /* raggedbayesrkernel.cpp:10-30. EigenBayesRKernel (eigenbayesrkernel.cpp:10-15)
     * gives the same element values: (dBeta * v) / sd - (dBeta * mean) / sd with
     * v = 1, 2 (2 dBeta / sd = 2 (dBeta / sd) exactly), 0 for a missing call
     * (v = mean: the two terms cancel exactly) and -(dBeta * mean) / sd elsewhere. */
    const double mean = d->stats[4 * (size_t)j + 0];
    const double sd = d->stats[4 * (size_t)j + 1];
    const double Zsum = d->stats[4 * (size_t)j + 3];
    const double dBeta = beta_old - beta_new;
    const double meanAdjustment = dBeta * mean / sd;
    const double oneAdjustment = dBeta / sd;
    double* delta = (double*)malloc(sizeof(double) * d->N);
    for (uint32_t i = 0; i < d->N; ++i) delta[i] = -meanAdjustment;
    for (uint64_t q = d->ones_off[j]; q < d->ones_off[j + 1]; ++q) delta[d->ones[q]] += oneAdjustment;
    for (uint64_t q = d->twos_off[j]; q < d->twos_off[j + 1]; ++q) delta[d->twos[q]] += 2 * oneAdjustment;
    for (uint64_t q = d->miss_off[j]; q < d->miss_off[j + 1]; ++q) delta[d->miss[q]] += meanAdjustment;
    for (uint32_t i = 0; i < d->N; ++i) eps[i] = eps[i] + delta[i];
    free(delta);
    /* sparsebayesrkernel.cpp:29-34, added to the global sum (SparseBayesRRG.cpp:91) */
    const double upd = dBeta * Zsum / sd - dBeta * mean * (double)d->N / sd;
    *epsSum += upd;
}

/* any_{i=1..K-1} |logL[i] - logL[r]| > 700 (BayesRBase.cpp:335,360) */
static int guard700(const double* logL, int K, int r)
{
    for (int i = 1; i < K; ++i) if (fabs(logL[i] - logL[r]) > 700.0) return 1;
    return 0;
}

/* 1 / sum_k exp(logL[k] - logL[r]) (BayesRBase.cpp:338,363) */
static double inv_sum_exp(const double* logL, int K, int r)
{
    double s = 0.0;
    for (int k = 0; k < K; ++k) s += exp(logL[k] - logL[r]);
    return 1.0 / s;
}

/* The component draw and beta of marker j given its num (BayesRBase.cpp:300-368;
 * processColumnAsync :420-483 is the same arithmetic). Writes acum, beta, comp,
 * v and betasqn; returns beta_old and beta_new. */
static void decide_column(const orc_data* d, const orc_model* m, orc_state* s, uint32_t j, double u, double z,
                          double num, double* acum_out, double* beta_old_out, double* beta_new_out)
{
    const int K = m->K;
    const int km1 = K - 1;
    const double NM1 = (double)(d->N - 1u);
    const int g = m->group ? m->group[j] : 0;
    const double* Sg = m->S + (size_t)g * km1;
    const double beta_old = s->beta[j];
    const double sigmaG = s->sigmaG[g];
    const double sigmaE = s->sigmaE;

    double denom[64], muk[64], logL[64];
    const double sigmaEOverSigmaG = sigmaE / sigmaG;
    for (int k = 0; k < km1; ++k) denom[k] = NM1 + sigmaEOverSigmaG * (1.0 / Sg[k]);

    muk[0] = 0.0;
    for (int k = 1; k < K; ++k) muk[k] = num / denom[k - 1];

    const double logLScale = sigmaG / sigmaE * NM1;
    for (int k = 0; k < K; ++k) logL[k] = log(s->pi[(size_t)g * K + k]);
    for (int k = 1; k < K; ++k) {
        logL[k] = logL[k] - 0.5 * log(logLScale * Sg[k - 1] + 1.0)
                + 0.5 * (muk[k] * num) / sigmaE;
    }

    double acum = guard700(logL, K, 0) ? 0.0 : inv_sum_exp(logL, K, 0);
    s->acum[j] = acum;
    if (acum_out) *acum_out = acum;

    double beta_new = beta_old;
    for (int k = 0; k < K; ++k) {
        if (u <= acum) {
            if (k == 0) {
                beta_new = 0.0;
            } else {
                beta_new = z * sqrt(sigmaE / denom[k - 1]) + muk[k];
                s->betasqn[g] += beta_new * beta_new;
            }
            s->beta[j] = beta_new;
            s->v[(size_t)g * K + k] += 1.0;
            s->comp[j] = k;
            break;
        } else if (k + 1 < K) {
            /* (k = K-1 reads logL[K] in the reference; the value is never
             * compared again, so the loop simply ends) */
            if (!guard700(logL, K, k + 1)) acum += inv_sum_exp(logL, K, k + 1);
        }
    }
    *beta_old_out = beta_old;
    *beta_new_out = beta_new;
}

void orc_process_column(const orc_data* d, const orc_model* m, orc_state* s,
                        uint32_t j, double u, double z,
                        double* num_out, double* acum_out)
{
    /* BayesRBase::processColumn, BayesRBase.cpp:287-392 */
    const double num = orc_compute_num(d, j, s->eps, s->beta[j], s->epsSum);
    if (num_out) *num_out = num;
    double beta_old, beta_new;
    decide_column(d, m, s, j, u, z, num, acum_out, &beta_old, &beta_new);
    /* BayesRBase.cpp:370-378 */
    const int skip = (beta_old == 0.0 && beta_new == 0.0);
    if (!skip) orc_apply_epsilon_change(d, j, s->eps, &s->epsSum, beta_old, beta_new);
}

static void clear_sweep_counts(const orc_model* m, orc_state* s)
{
    const int K = m->K;
    for (int g = 0; g < m->G; ++g) {
        s->betasqn[g] = 0.0;
        for (int k = 0; k < K; ++k) s->v[(size_t)g * K + k] = 0.0;
    }
}

void orc_sweep(const orc_data* d, const orc_model* m, orc_state* s,
               const uint32_t* order, const double* u, const double* z,
               orc_trace* trace)
{
    clear_sweep_counts(m, s);
    for (uint32_t p = 0; p < d->M; ++p) {
        const uint32_t j = order[p];
        double num = 0.0, acum = 0.0;
        orc_process_column(d, m, s, j, u[j], z[j], &num, &acum);
        if (trace) {
            if (trace->num) trace->num[p] = num;
            if (trace->acum) trace->acum[p] = acum;
            if (trace->comp) trace->comp[p] = s->comp[j];
            if (trace->beta) trace->beta[p] = s->beta[j];
        }
    }
}

void orc_sweep_async(const orc_data* d, const orc_model* m, orc_state* s,
                     const uint32_t* order, const double* u, const double* z, uint32_t window)
{
    /* ParallelGraph (parallelgraph.cpp) runs up to analysisTokens markers at once;
     * processColumnAsync (BayesRBase.cpp:394-491) reads epsilon and epsilonSum under
     * a shared lock, decides, and queues its change; updateGlobal (:503-510,
     * SparseBayesRRG.cpp:94-103) applies the changes one at a time. This is the
     * schedule in which every marker of a window reads the window-start state. */
    clear_sweep_counts(m, s);
    if (window == 0) window = 1;
    double* num = (double*)malloc(sizeof(double) * window);
    double* bold = (double*)malloc(sizeof(double) * window);
    double* bnew = (double*)malloc(sizeof(double) * window);
    for (uint32_t p0 = 0; p0 < d->M; p0 += window) {
        const uint32_t n = (d->M - p0 < window) ? d->M - p0 : window;
        for (uint32_t q = 0; q < n; ++q) {
            const uint32_t j = order[p0 + q];
            num[q] = orc_compute_num(d, j, s->eps, s->beta[j], s->epsSum);
        }
        for (uint32_t q = 0; q < n; ++q) {
            const uint32_t j = order[p0 + q];
            decide_column(d, m, s, j, u[j], z[j], num[q], NULL, &bold[q], &bnew[q]);
        }
        for (uint32_t q = 0; q < n; ++q) {
            if (bold[q] == 0.0 && bnew[q] == 0.0) continue;
            orc_apply_epsilon_change(d, order[p0 + q], s->eps, &s->epsSum, bold[q], bnew[q]);
        }
    }
    free(num); free(bold); free(bnew);
}

void orc_init(const orc_data* d, const orc_model* m, orc_state* s,
              const double* y, uint64_t seed, uint32_t chain)
{
    /* BayesRBase::init, BayesRBase.cpp:39-99 */
    const int K = m->K;
    const uint32_t N = d->N;
    const double priorPi = 0.5 / K;
    for (int g = 0; g < m->G; ++g) {
        for (int k = 0; k < K; ++k) s->pi[(size_t)g * K + k] = priorPi;
        s->pi[(size_t)g * K] = 0.5;
        s->sigmaG[g] = orc_uniform01(seed, ORC_INIT_ITERATION, (uint32_t)g, chain);
        s->v[(size_t)g * K] = 0.0;
        s->betasqn[g] = 0.0;
    }
    s->mu = 0.0;
    for (uint32_t j = 0; j < d->M; ++j) { s->beta[j] = 0.0; s->acum[j] = 0.0; s->comp[j] = 0; }
    double ysum = 0.0;
    for (uint32_t i = 0; i < N; ++i) ysum += y[i];
    const double ymean = ysum / (double)N;
    for (uint32_t i = 0; i < N; ++i) s->eps[i] = y[i] - ymean;
    double sq = 0.0;
    for (uint32_t i = 0; i < N; ++i) sq += s->eps[i] * s->eps[i];
    const double scale = sqrt(sq / ((double)(N - 1u)));
    for (uint32_t i = 0; i < N; ++i) s->eps[i] /= scale;          /* m_y */
    for (uint32_t i = 0; i < N; ++i) s->eps[i] = s->eps[i] - s->mu; /* eps = y - mu */
    double e2 = 0.0, es = 0.0;
    for (uint32_t i = 0; i < N; ++i) e2 += s->eps[i] * s->eps[i];
    for (uint32_t i = 0; i < N; ++i) es += s->eps[i];
    s->sigmaE = e2 / (double)N * 0.5;
    s->epsSum = es;
}

void orc_update_mu(const orc_data* d, orc_state* s, uint64_t seed, uint32_t iteration,
                   uint32_t chain)
{
    const uint32_t N = d->N;
    const double dN = (double)N;
    const double old_mu = s->mu;
    if (d->layout == ORC_DENSE) {
        /* DenseBayesRRmz.cpp:36-42 */
        for (uint32_t i = 0; i < N; ++i) s->eps[i] = s->eps[i] + s->mu;
        double es = 0.0;
        for (uint32_t i = 0; i < N; ++i) es += s->eps[i];
        s->mu = orc_norm_rng(seed, iteration, ORC_DRAW_MU, chain, es / dN, s->sigmaE / dN);
        for (uint32_t i = 0; i < N; ++i) s->eps[i] = s->eps[i] - s->mu;
    } else {
        /* SparseBayesRRG.cpp:106-113 */
        for (uint32_t i = 0; i < N; ++i) s->eps[i] = s->eps[i] + s->mu;
        s->epsSum += old_mu * dN;
        s->mu = orc_norm_rng(seed, iteration, ORC_DRAW_MU, chain, s->epsSum / dN, s->sigmaE / dN);
        for (uint32_t i = 0; i < N; ++i) s->eps[i] = s->eps[i] - s->mu;
        s->epsSum -= s->mu * dN;
    }
}

void orc_post_sweep(const orc_data* d, const orc_model* m, orc_state* s,
                    uint64_t seed, uint32_t iteration, uint32_t chain)
{
    /* BayesRBase.cpp:237-247 */
    const uint32_t N = d->N;
    const int K = m->K;
    double e2 = 0.0;
    for (uint32_t i = 0; i < N; ++i) e2 += s->eps[i] * s->eps[i];
    s->sigmaE = orc_inv_scaled_chisq_rng(seed, iteration, ORC_DRAW_SIGMAE, chain,
                                         m->v0E + N, (e2 + m->v0E * m->s02E) / (m->v0E + N));
    for (int g = 0; g < m->G; ++g) {
        double vsum = 0.0;
        for (int k = 0; k < K; ++k) vsum += s->v[(size_t)g * K + k];
        const double m0 = vsum - s->v[(size_t)g * K];
        s->sigmaG[g] = orc_inv_scaled_chisq_rng(seed, iteration, ORC_DRAW_SIGMAG0 + (uint32_t)g, chain,
                                                m->v0G + m0,
                                                (s->betasqn[g] * m0 + m->v0G * m->s02G) / (m->v0G + m0));
        double alpha[64];
        for (int k = 0; k < K; ++k) alpha[k] = s->v[(size_t)g * K + k] + 1.0;
        orc_dirichlet_rng(seed, iteration, ORC_DRAW_PI0(m->G) + (uint32_t)(g * K), chain, K, alpha,
                          s->pi + (size_t)g * K);
    }
}

void orc_run_gibbs(const orc_data* d, const orc_model* m, orc_state* s,
                   const double* y, uint64_t seed, uint32_t chain,
                   uint32_t iterations, uint32_t* order, int shuffle_mode,
                   double* hist)
{
    orc_run_gibbs_window(d, m, s, y, seed, chain, iterations, order, shuffle_mode, hist, 0);
}

void orc_fixed_effects(const orc_data* d, const orc_model* m, orc_state* s, const float* X, uint32_t F,
                       double* gamma, uint32_t* forder, double* sigmaF, uint64_t seed, uint32_t iteration,
                       uint32_t chain)
{
    /* BayesRBase.cpp:212-235 */
    const uint32_t N = d->N;
    const double dNm1 = (double)(N - 1u);
    if (F == 0) return;
    orc_shuffle_philox_stream(seed, iteration, chain, ORC_STREAM_FIXED, F, forder);
    const double sigE_sigF = s->sigmaE / *sigmaF;
    for (uint32_t i = 0; i < F; ++i) {
        const uint32_t f = forder[i];
        const float* x = X + (size_t)f * N;
        const double gamma_old = gamma[f];
        double num_f = 0.0;
        for (uint32_t k = 0; k < N; ++k) num_f += (double)x[k] * (s->eps[k] + gamma_old * (double)x[k]);
        const double denom_f = dNm1 + sigE_sigF;
        gamma[f] = orc_norm_rng(seed, iteration, ORC_DRAW_FIXED0(m->G, m->K) + f, chain, num_f / denom_f,
                                s->sigmaE / denom_f);
        for (uint32_t k = 0; k < N; ++k) s->eps[k] = s->eps[k] + (gamma_old - gamma[f]) * (double)x[k];
    }
    *sigmaF = 1.0;   /* s02F (BayesRBase.hpp:90) */
}

void orc_run_gibbs_window(const orc_data* d, const orc_model* m, orc_state* s,
                          const double* y, uint64_t seed, uint32_t chain,
                          uint32_t iterations, uint32_t* order, int shuffle_mode,
                          double* hist, uint32_t window)
{
    orc_run_gibbs_fixed(d, m, s, y, seed, chain, iterations, order, shuffle_mode, hist, window, NULL, 0, NULL,
                        NULL, NULL);
}

void orc_run_gibbs_fixed(const orc_data* d, const orc_model* m, orc_state* s,
                         const double* y, uint64_t seed, uint32_t chain,
                         uint32_t iterations, uint32_t* order, int shuffle_mode,
                         double* hist, uint32_t window, const float* X, uint32_t F,
                         double* gamma, uint32_t* forder, double* sigmaF)
{
    /* BayesRBase::runGibbs, BayesRBase.cpp:130-285; window > 0: the sweep in
     * asynchronous windows (orc_sweep_async); F > 0: the fixed effects after it */
    const uint32_t M = d->M;
    const int G = m->G, K = m->K;
    orc_init(d, m, s, y, seed, chain);
    double* u = (double*)malloc(sizeof(double) * M);
    double* z = (double*)malloc(sizeof(double) * M);
    for (uint32_t it = 0; it < iterations; ++it) {
        orc_update_mu(d, s, seed, it, chain);
        orc_fill_marker_uz(seed, it, chain, M, u, z);
        if (shuffle_mode == 1) orc_shuffle_libc(M, order);
        else orc_shuffle_philox(seed, it, chain, M, order);
        if (window) orc_sweep_async(d, m, s, order, u, z, window);
        else orc_sweep(d, m, s, order, u, z, NULL);
        if (F) orc_fixed_effects(d, m, s, X, F, gamma, forder, sigmaF, seed, it, chain);
        double e2 = 0.0;
        for (uint32_t i = 0; i < d->N; ++i) e2 += s->eps[i] * s->eps[i];
        double m0_total = 0.0;
        for (int g = 0; g < G; ++g) {
            double vsum = 0.0;
            for (int k = 0; k < K; ++k) vsum += s->v[(size_t)g * K + k];
            m0_total += vsum - s->v[(size_t)g * K];
        }
        const double mu_it = s->mu;
        orc_post_sweep(d, m, s, seed, it, chain);
        if (hist) {
            double* h = hist + (size_t)it * (4 + G);
            h[0] = mu_it; h[1] = s->sigmaE; h[2] = m0_total; h[3] = e2;
            for (int g = 0; g < G; ++g) h[4 + g] = s->sigmaG[g];
        }
    }
    free(u); free(z);
}
