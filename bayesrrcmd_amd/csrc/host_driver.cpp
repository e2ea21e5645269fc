// host_driver.cpp -- the host half of BayesRBase::runGibbs around the device sweep.
//
// The reference keeps these on the host thread that calls exec():
//   init                 BayesRBase.cpp:39-99
//   random numbers       prepareForAnylsis, BayesRBase.cpp:101-110 (per marker id)
//   marker order         std::random_shuffle, BayesRBase.cpp:198
//   BayesW init          BayesWBase::init, bayeswbase.cpp:469-540
//   sample output        SampleWriter / Writer (samplewriter.cpp, writer.cpp): the CSV
//                        header and one row per kept iteration, printed like an
//                        Eigen row vector with IOFormat(StreamPrecision, DontAlignCols,
//                        ", ", ", ") on a default std::ofstream (6 significant digits)
// They are restated here over the counter-based stream of brr_rng.h so a host
// driver and the device-resident chain consume the same numbers.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "brr_hip.h"
#include "brr_rng.h"

extern "C" {

int brr_bayesr_init(const double* y, uint32_t N, uint32_t G, uint32_t K, uint64_t seed, uint32_t chain,
                    double* eps, brr_hip_scalars* s)
{
    if (!y || !eps || !s || !s->sigmaG || !s->pi || N < 2 || G < 1 || K < 2) return BRR_HIP_EINVAL;
    // m_priorPi: 0.5 for the zero component, 0.5/K for the others
    const double priorPi = 0.5 / K;
    for (uint32_t g = 0; g < G; ++g) {
        for (uint32_t k = 0; k < K; ++k) s->pi[g * K + k] = priorPi;
        s->pi[g * K] = 0.5;
        // m_sigmaG ~ U(0,1), the first draws of the chain
        s->sigmaG[g] = brr::uniform01(seed, brr::kInitIteration, g, chain);
        if (s->v)
            for (uint32_t k = 0; k < K; ++k) s->v[g * K + k] = 0.0;
        if (s->betasqn) s->betasqn[g] = 0.0;
    }
    s->mu = 0.0;
    // m_y = (y - mean) / sqrt(|y - mean|^2 / (N - 1)); eps = m_y - mu
    double ysum = 0.0;
    for (uint32_t i = 0; i < N; ++i) ysum += y[i];
    const double ymean = ysum / (double)N;
    for (uint32_t i = 0; i < N; ++i) eps[i] = y[i] - ymean;
    double sq = 0.0;
    for (uint32_t i = 0; i < N; ++i) sq += eps[i] * eps[i];
    const double scale = std::sqrt(sq / ((double)(N - 1u)));
    for (uint32_t i = 0; i < N; ++i) eps[i] /= scale;
    for (uint32_t i = 0; i < N; ++i) eps[i] = eps[i] - s->mu;
    double e2 = 0.0, es = 0.0;
    for (uint32_t i = 0; i < N; ++i) e2 += eps[i] * eps[i];
    for (uint32_t i = 0; i < N; ++i) es += eps[i];
    s->sigmaE = e2 / (double)N * 0.5;
    s->epsilon_sum = es;
    return BRR_HIP_OK;
}

int brr_shuffle(uint64_t seed, uint32_t iteration, uint32_t chain, uint32_t M, uint32_t* order, int mode)
{
    if (!order) return BRR_HIP_EINVAL;
    // libstdc++ random_shuffle: for i in [1, M): swap(order[i], order[r % (i + 1)])
    for (uint32_t i = 1; i < M; ++i) {
        uint64_t r;
        if (mode == 1) {
            r = (uint64_t)std::rand();
        } else {
            const brr::U4 x = brr::draw_block(seed, i, iteration, brr::kStreamShuffle, chain);
            r = ((uint64_t)x.x[0] << 32) | x.x[1];
        }
        const uint32_t j = (uint32_t)(r % ((uint64_t)i + 1));
        if (j != i) {
            const uint32_t t = order[i];
            order[i] = order[j];
            order[j] = t;
        }
    }
    return BRR_HIP_OK;
}

int brr_marker_uz(uint64_t seed, uint32_t iteration, uint32_t chain, uint32_t M, double* u, double* z)
{
    if (!u || !z) return BRR_HIP_EINVAL;
    for (uint32_t j = 0; j < M; ++j) brr::marker_uz(seed, iteration, j, chain, &u[j], &z[j]);
    return BRR_HIP_OK;
}

int brr_bayesw_init(const double* y, const double* fail, uint32_t N, uint32_t M, uint32_t K, double* eps,
                    brr_hip_bw_scalars* s)
{
    // BayesWBase::init (bayeswbase.cpp:469-540) without fixed effects
    if (!y || !fail || !eps || !s || !s->pi || N < 2 || M < 1 || K < 2 || K > 8) return BRR_HIP_EINVAL;
    const uint32_t km1 = K - 1;
    s->pi[0] = 0.99;
    for (uint32_t k = 1; k < K; ++k) s->pi[k] = (1 - s->pi[0]) / km1;
    if (s->v)
        for (uint32_t k = 0; k < K; ++k) s->v[k] = 1.0;
    double sy = 0.0;
    for (uint32_t i = 0; i < N; ++i) sy += y[i];
    s->mu = sy / N;
    double dd = 0.0;
    for (uint32_t i = 0; i < N; ++i) dd += fail[i];
    s->d = dd;
    double ss = 0.0;
    for (uint32_t i = 0; i < N; ++i) ss += (y[i] - s->mu) * (y[i] - s->mu);
    const double denominator = (6 * ss / (N - 1));
    s->alpha = 3.14159 / std::sqrt(denominator);     // PI of bayeswbase.cpp
    for (uint32_t i = 0; i < N; ++i) eps[i] = y[i] - s->mu;
    s->sigma_b = 6.283185 / (6 * (s->alpha * s->alpha) * M);   // PI2
    s->betasq = 0.0;
    return BRR_HIP_OK;
}

} // extern "C"

struct brr_samples {
    std::ofstream out;
    int kind = 0;
    bool binary = false;
    uint32_t M = 0, N = 0, G = 0, F = 0;
    std::vector<double> row;
};

namespace {

// One sample row: the values of the reference's `sample` vector in its order,
// as a default std::ostream prints them after ", " separators (CSV) or as raw
// little-endian doubles (binary).
int emit_row(brr_samples* w)
{
    std::ofstream& o = w->out;
    if (w->binary) {
        o.write(reinterpret_cast<const char*>(w->row.data()), (std::streamsize)(sizeof(double) * w->row.size()));
    } else {
        for (size_t i = 0; i < w->row.size(); ++i) {
            if (i) o << ", ";
            o << w->row[i];
        }
        o << std::endl;
    }
    return o ? BRR_HIP_OK : BRR_HIP_EINVAL;
}

} // namespace

extern "C" {

int brr_samples_open(const char* path, int kind, uint32_t M, uint32_t N, uint32_t G, uint32_t F, brr_samples** out)
{
    const bool binary = kind == BRR_SAMPLES_BAYESR_BINARY || kind == BRR_SAMPLES_BAYESW_BINARY;
    const int base = binary ? kind - 2 : kind;
    if (!path || !out || M == 0 || (base != BRR_SAMPLES_BAYESR && base != BRR_SAMPLES_BAYESW)) return BRR_HIP_EINVAL;
    if (base == BRR_SAMPLES_BAYESR && (N == 0 || G == 0)) return BRR_HIP_EINVAL;
    *out = nullptr;
    brr_samples* w = new brr_samples();
    w->kind = base;
    w->binary = binary;
    w->M = M; w->N = N; w->G = G; w->F = F;
    w->out.open(path, binary ? std::ios::out | std::ios::binary : std::ios::out);
    if (!w->out) { delete w; return BRR_HIP_EINVAL; }
    std::ofstream& o = w->out;
    if (binary) {
        // "BRRSAMP1", then u32 {kind, M, N, G, F, doubles per row}
        const uint64_t row_len = base == BRR_SAMPLES_BAYESR ? 3ull + 3ull * M + G + F + N : 4ull + 2ull * M + F;
        const uint32_t head[6] = {(uint32_t)kind, M, N, G, F, (uint32_t)row_len};
        o.write("BRRSAMP1", 8);
        o.write(reinterpret_cast<const char*>(head), sizeof(head));
    } else if (base == BRR_SAMPLES_BAYESR) {
        // SampleWriter::openGroups (samplewriter.cpp:82-118)
        o << "iteration," << "mu,";
        for (uint32_t i = 0; i < M; ++i) o << "beta[" << (i + 1) << "],";
        o << "sigmaE,";
        for (uint32_t i = 0; i < G; ++i) o << "sigmaG[" << (i + 1) << "],";
        for (uint32_t i = 0; i < F; ++i) o << "gamma[" << (i + 1) << "],";
        for (uint32_t i = 0; i < M; ++i) o << "comp[" << (i + 1) << "],";
        for (uint32_t i = 0; i < M; ++i) o << "acum[" << (i + 1) << "],";
        uint32_t i;
        for (i = 0; i < N - 1; ++i) o << "epsilon[" << (i + 1) << "],";
        o << "epsilon[" << (i + 1) << "]";
        o << std::endl;
    } else {
        // SampleWriter::open_bayesW / open_bayesW_fixed (samplewriter.cpp:34-80)
        o << "iteration," << "alpha," << "mu,";
        for (uint32_t i = 0; i < F; ++i) o << "theta[" << (i + 1) << "],";
        for (uint32_t i = 0; i < M; ++i) o << "beta[" << (i + 1) << "],";
        for (uint32_t i = 0; i < M; ++i) o << "comp[" << (i + 1) << "],";
        o << "sigma_b";
        o << std::endl;
    }
    o.flush();
    *out = w;
    return o ? BRR_HIP_OK : BRR_HIP_EINVAL;
}

int brr_samples_write_bayesr(brr_samples* w, uint32_t iteration, double mu, const double* beta, double sigmaE,
                             const double* sigmaG, const double* gamma, const int32_t* comp, const double* acum,
                             const double* eps)
{
    // BayesRBase.cpp:251: iteration, mu, beta, sigmaE, sigmaG, gamma, components, acum, epsilon
    if (!w || w->kind != BRR_SAMPLES_BAYESR || !beta || !sigmaG || !comp || !acum || !eps || (w->F && !gamma))
        return BRR_HIP_EINVAL;
    std::vector<double>& r = w->row;
    r.clear();
    r.push_back((double)iteration);
    r.push_back(mu);
    r.insert(r.end(), beta, beta + w->M);
    r.push_back(sigmaE);
    r.insert(r.end(), sigmaG, sigmaG + w->G);
    if (w->F) r.insert(r.end(), gamma, gamma + w->F);
    for (uint32_t i = 0; i < w->M; ++i) r.push_back((double)comp[i]);
    r.insert(r.end(), acum, acum + w->M);
    r.insert(r.end(), eps, eps + w->N);
    return emit_row(w);
}

int brr_samples_write_bayesw(brr_samples* w, uint32_t iteration, double alpha, double mu, const double* theta,
                             const double* beta, const int32_t* comp, double sigma_b)
{
    // bayeswbase.cpp:810-815: iteration, alpha, mu, (theta), beta, components, sigma_b
    if (!w || w->kind != BRR_SAMPLES_BAYESW || !beta || !comp || (w->F && !theta)) return BRR_HIP_EINVAL;
    std::vector<double>& r = w->row;
    r.clear();
    r.push_back((double)iteration);
    r.push_back(alpha);
    r.push_back(mu);
    if (w->F) r.insert(r.end(), theta, theta + w->F);
    r.insert(r.end(), beta, beta + w->M);
    for (uint32_t i = 0; i < w->M; ++i) r.push_back((double)comp[i]);
    r.push_back(sigma_b);
    return emit_row(w);
}

int brr_samples_close(brr_samples* w)
{
    if (!w) return BRR_HIP_EINVAL;
    w->out.flush();
    const bool ok = (bool)w->out;
    w->out.close();
    delete w;
    return ok ? BRR_HIP_OK : BRR_HIP_EINVAL;
}

int brr_sample_due(uint32_t iteration, uint32_t burn_in, uint32_t thinning)
{
    // BayesRBase.cpp:250, bayeswbase.cpp:810
    return thinning != 0 && iteration >= burn_in && iteration % thinning == 0;
}

} // extern "C"
