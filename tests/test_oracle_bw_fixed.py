"""BayesW fixed effects (sampleTheta, bayeswbase.cpp:572-601) in the oracle.

Each iteration, after mu, every covariate f in turn: the residual is shifted
by X_f theta_f, theta_f is drawn by ARMS from theta_dens(x) = -alpha x
sum_k X(k,f) fail_k - sum_k exp((eps_k - X(k,f) x) alpha - EulerGamma) -
x^2 / (2 100) on (-2, 2), and the residual is shifted back by X_f x.
Independent checks: the draws from one state against that density restated
in numpy and integrated numerically (KS); the residual shift and sum_failure_fix
recomputed with numpy; a chain recovers a simulated covariate effect on the
log-time scale.
"""
import numpy as np
import pytest
from scipy import stats

import _oracle as orc
from _data import make_bw_problem

EULER = 0.577215664901532


def setup(F=2, n=2000, m=40, seed=81):
    pr = make_bw_problem(n, m, seed)
    w = orc.BayesWProblem(pr.n, pr.m, pr.ragged, pr.fail)
    w.init(pr.y)
    rng = np.random.default_rng(seed)
    X = rng.normal(0.0, 0.5, size=(n, F)).astype(np.float32)
    w.set_fixed(X)
    return pr, w, X.astype(np.float64)


def test_sum_failure_fixed():
    pr, w, X = setup(F=3)
    assert np.allclose(w.sumf_fix, X.T @ pr.fail, rtol=1e-12, atol=1e-12)


def test_theta_draws_follow_theta_dens():
    pr, w, X = setup(F=1)
    x = X[:, 0]
    eps0 = w.eps.copy()
    a = w.state.alpha
    sf = w.sumf_fix[0]
    th0 = 0.05
    eadj = eps0 + x * th0

    def logd(t):
        return -a * t * sf - np.sum(np.exp((eadj[None, :] - np.outer(t, x)) * a - EULER), axis=1) - t * t / 200.0

    grid = np.linspace(-2.0, 2.0, 40001)
    ld = logd(grid)
    dens = np.exp(ld - ld.max())
    cdf = np.cumsum(dens)
    cdf /= cdf[-1]
    draws = []
    for it in range(2000):
        w.eps[:] = eps0
        w.theta[:] = [th0]
        assert w.sample_theta(0, 123, it) == 0
        draws.append(w.theta[0])
        # the residual: eps0 + x theta_old - x theta_new, element by element
        want = (eps0 + x * th0) - x * w.theta[0]
        assert np.array_equal(w.eps, want)
    d = stats.kstest(draws, lambda t: np.interp(t, grid, cdf))
    assert d.pvalue > 1e-3, d


def test_chain_recovers_theta():
    n, m = 3000, 40
    pr = make_bw_problem(n, m, 82, causal_frac=0.05)
    rng = np.random.default_rng(82)
    X = rng.normal(0.0, 1.0, size=(n, 2)).astype(np.float32).astype(np.float64)
    theta_true = np.array([0.3, -0.2])
    y = (pr.y + X @ theta_true).astype(np.float32).astype(np.float64)
    w = orc.BayesWProblem(n, m, pr.ragged, pr.fail)
    w.set_fixed(X)
    err, hist, _ = w.run_gibbs_fixed(y, 17, 40)
    assert err == 0
    assert np.allclose(w.theta, theta_true, atol=0.05), (w.theta, theta_true)
    assert abs(hist[-1, 1] - 3.0) < 0.4       # the Weibull shape is still recovered
