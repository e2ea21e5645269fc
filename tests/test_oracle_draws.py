"""Conditional draws of the chain scalars in the oracle, against their
reference parameterisations restated independently here.

- BayesW mu (sampleMu, bayeswbase.cpp:542-569 with mu_dens :66-77): ARMS on
  log f(x) = -alpha x d - sum_i exp((eps_i + mu - x) alpha - EulerGamma) - x^2 / (2 100)
  over (2, 5). The density is log-concave, so the sampler (no Metropolis
  step) draws exactly from it.
- BayesW alpha (sampleAlpha, :707-734 with alpha_dens :109-119):
  log f(x) = (0.01 + d - 1) log x + x (sum_i eps_i fail_i - 0.01) - sum_i exp(eps_i x - EulerGamma)
  over (0, 20), also log-concave.
- BayesW sigma_b (:805-806): inv_gamma_rng(1 + (M - v_0 + 1) / 2,
  1e-4 + (M - v_0 + 1) / 2 |beta|^2), where Distributions_boost::inv_gamma_rng(shape,
  scale) = 1 / Gamma(shape, scale = 1 / scale) (distributions_boost.cpp:45-47), i.e.
  scipy's invgamma(a = shape, scale = scale).
- BayesW pi_L (:809) ~ Dirichlet(v) and BayesR pi (BayesRBase.cpp:243-246) ~
  Dirichlet(v + 1); Distributions_boost::dirichlet_rng normalises Gamma(alpha_k, 1)
  (distributions_boost.cpp:35-43), so each component is Beta(alpha_k, sum alpha - alpha_k).

Each draw is made many times from one fixed state (a fresh iteration number per
draw selects a fresh random stream) and the sample is KS-tested against the
distribution above: mu and alpha against their densities integrated on a grid,
sigma_b and pi against scipy. Every test also checks that a plausible mistake
(a dropped prior term, a scale / rate swap, Dirichlet(v) for Dirichlet(v + 1))
would have been rejected by the same sample.
"""
import numpy as np
import pytest
from scipy import stats

import _oracle as orc
from _data import make_bw_problem, make_problem

EULER = 0.577215664901532
DRAWS = 3000


def grid_cdf(logd, lo, hi, n=200001):
    """CDF of exp(logd) on [lo, hi], zoomed onto the region holding the mass."""
    g = np.linspace(lo, hi, 20001)
    ld = logd(g)
    keep = g[ld > ld.max() - 60.0]
    a = max(lo, keep[0] - (hi - lo) / 20000)
    b = min(hi, keep[-1] + (hi - lo) / 20000)
    g = np.linspace(a, b, n)
    ld = logd(g)
    dens = np.exp(ld - ld.max())
    cdf = np.concatenate([[0.0], np.cumsum(0.5 * (dens[1:] + dens[:-1]) * np.diff(g))])
    cdf /= cdf[-1]
    return lambda t: np.interp(t, g, cdf)


@pytest.fixture(scope="module")
def bw_state():
    pr = make_bw_problem(1500, 30, seed=91, causal_frac=0.1)
    w = orc.BayesWProblem(pr.n, pr.m, pr.ragged, pr.fail)
    w.init(pr.y)
    # a residual other than the init one, as after some sweeps
    rng = np.random.default_rng(91)
    eps = w.eps.copy() + rng.normal(0.0, 0.05, pr.n)
    return pr, w, eps


def test_bw_mu_draws_follow_mu_dens(bw_state):
    pr, w, eps0 = bw_state
    st = w.state
    mu0, alpha, d = st.mu, st.alpha, st.d
    assert 2.0 < mu0 < 5.0
    base = eps0 + mu0          # y - X beta: what mu_dens sees

    def logd(x, prior=True):
        x = np.atleast_1d(x)
        out = np.empty(len(x))
        for s in range(0, len(x), 2048):
            xx = x[s:s + 2048]
            out[s:s + 2048] = (-alpha * xx * d - np.exp((base[None, :] - xx[:, None]) * alpha - EULER).sum(axis=1)
                               - (xx * xx / 200.0 if prior else 0.0))
        return out

    draws = np.empty(DRAWS)
    for it in range(DRAWS):
        w.eps[:] = eps0
        st.mu = mu0
        assert w.sample_mu(seed=5, iteration=it) == 0
        draws[it] = st.mu
        if it < 50:
            # the residual: (eps + mu_old) - mu_new element by element (bayeswbase.cpp:552,568)
            assert np.array_equal(w.eps, (eps0 + mu0) - st.mu)
    cdf = grid_cdf(logd, 2.0, 5.0)
    assert stats.kstest(draws, cdf).pvalue > 1e-3
    # a dropped alpha in the first term would be rejected by the same sample
    cdf_bad = grid_cdf(lambda x: logd(x) + alpha * x * d - x * d, 2.0, 5.0)
    assert stats.kstest(draws, cdf_bad).pvalue < 1e-6


def test_bw_alpha_draws_follow_alpha_dens(bw_state):
    pr, w, eps0 = bw_state
    st = w.state
    a0, d = st.alpha, st.d
    ef = float(eps0 @ pr.fail)

    def logd(x, kappa=0.01, alpha0=0.01):
        x = np.atleast_1d(x)
        out = np.empty(len(x))
        for s in range(0, len(x), 2048):
            xx = x[s:s + 2048]
            out[s:s + 2048] = ((alpha0 + d - 1.0) * np.log(xx) + xx * (ef - kappa)
                               - np.exp(eps0[None, :] * xx[:, None] - EULER).sum(axis=1))
        return out

    draws = np.empty(DRAWS)
    for it in range(DRAWS):
        w.eps[:] = eps0
        st.alpha = a0
        assert w.sample_alpha(seed=6, iteration=it) == 0
        draws[it] = st.alpha
    assert np.array_equal(w.eps, eps0)         # alpha moves no residual
    cdf = grid_cdf(logd, 1e-6, 20.0)
    assert stats.kstest(draws, cdf).pvalue > 1e-3
    # the failure term with its sign flipped would be rejected by the same sample
    cdf_bad = grid_cdf(lambda x: logd(x) - 2.0 * x * ef, 1e-6, 20.0)
    assert stats.kstest(draws, cdf_bad).pvalue < 1e-6


def test_bw_sigma_b_and_pi_l_draws(bw_state):
    pr, w, _ = bw_state
    st = w.state
    rng = np.random.default_rng(7)
    beta = np.zeros(pr.m)
    nz = rng.choice(pr.m, 6, replace=False)
    beta[nz] = rng.normal(0.0, 0.3, 6)
    v = np.array([pr.m - 6 + 1.0, 4.0, 2.0, 3.0])     # class counts + 1 (the (1, ..., 1) prior)
    sig = np.empty(DRAWS)
    pis = np.empty((DRAWS, 4))
    for it in range(DRAWS):
        w.beta[:] = beta
        w.v[:] = v
        w.post_sweep(seed=8, iteration=it)
        sig[it] = st.sigma_b
        pis[it] = w.pi
    assert np.array_equal(w.v, v) and np.array_equal(w.beta, beta)
    half = 0.5 * (pr.m - v[0] + 1.0)
    shape, scale = 1.0 + half, 1e-4 + half * float(beta @ beta)
    assert stats.kstest(sig, stats.invgamma(a=shape, scale=scale).cdf).pvalue > 1e-3
    # scale read as a rate would be rejected
    assert stats.kstest(sig, stats.invgamma(a=shape, scale=1.0 / scale).cdf).pvalue < 1e-6
    assert np.allclose(pis.sum(axis=1), 1.0, atol=1e-12)
    tot = v.sum()
    for k in range(4):
        assert stats.kstest(pis[:, k], stats.beta(v[k], tot - v[k]).cdf).pvalue > 1e-3
    # Dirichlet(v + 1) (the BayesR prior) would be rejected for a small class
    assert stats.kstest(pis[:, 2], stats.beta(v[2] + 1, tot + 4 - v[2] - 1).cdf).pvalue < 1e-6


def test_bayesr_pi_draw_is_dirichlet_v_plus_one():
    pr = make_problem(200, 40, seed=93)
    v = np.array([30.0, 6.0, 3.0, 1.0])
    K = 4
    pis = np.empty((DRAWS, K))
    for it in range(DRAWS):
        o = orc.OracleProblem(pr.n, pr.m, orc.DENSE, K, 1, np.array([0.01, 0.001, 0.0001]), bed=pr.bed,
                              lut=pr.lut)
        o.eps[:] = pr.y
        o.set_scalars(0.0, 0.8, 0.0, [0.5], [0.5, 0.125, 0.125, 0.125])
        o.v[:] = v
        o.betasqn[0] = 0.05
        o.post_sweep(seed=14, iteration=it)
        pis[it] = o.pi
    a = v + 1.0
    for k in range(K):
        assert stats.kstest(pis[:, k], stats.beta(a[k], a.sum() - a[k]).cdf).pvalue > 1e-3
    # Dirichlet(v) would be rejected by the same sample
    assert stats.kstest(pis[:, 3], stats.beta(v[3], v.sum() - v[3]).cdf).pvalue < 1e-6
