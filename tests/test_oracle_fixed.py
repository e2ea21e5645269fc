"""Fixed effects of BayesRBase::runGibbs (BayesRBase.cpp:212-235) in the oracle.

Per iteration, after the sweep: the covariate order is shuffled, then for each
covariate f: num = sum_k X(k,f) (eps_k + gamma_old X(k,f)), denom = (N - 1) +
sigmaE / sigmaF, gamma_f ~ N(num / denom, sigmaE / denom), and eps moves by
(gamma_old - gamma_f) X(., f). Independent checks: the draw's distribution
against that normal with num recomputed from its definition (KS over many
draws from the same state); the residual change recomputed with numpy; the
order a permutation carried from call to call; a chain recovers simulated
covariate effects.
"""
import numpy as np
import pytest
from scipy import stats

import _oracle as orc
from _data import make_problem

S = np.array([0.01, 0.001, 0.0001])


def problem(n=1500, m=80, seed=71, layout=orc.DENSE, F=3):
    pr = make_problem(n, m, seed=seed, h2=0.4, causal_frac=0.05)
    if layout == orc.DENSE:
        o = orc.OracleProblem(n, m, orc.DENSE, 4, 1, S, bed=pr.bed, lut=pr.lut)
    else:
        o = orc.OracleProblem(n, m, orc.RAGGED, 4, 1, S, ragged=pr.ragged)
    o.init(pr.y, 5)
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, F)).astype(np.float32)
    X[:, 0] = (rng.random(n) < 0.5).astype(np.float32)          # a 0/1 covariate, as sex
    o.set_fixed(X)
    return pr, o, X.astype(np.float64)


def test_gamma_draw_distribution():
    """gamma_f | rest ~ N(num / denom, sigmaE / denom), num and denom from their definitions."""
    pr, o, X = problem(F=1)
    o.gamma[:] = [0.37]
    o.sigmaF[:] = [2.5]
    eps0 = o.eps.copy()
    sig_e = o.state.sigmaE
    x = X[:, 0]
    num = np.sum(x * (eps0 + 0.37 * x))
    denom = (pr.n - 1) + sig_e / 2.5
    draws = []
    for it in range(3000):
        o.eps[:] = eps0
        o.gamma[:] = [0.37]
        o.sigmaF[:] = [2.5]
        o.fixed_effects(99, it)
        draws.append(o.gamma[0])
        assert o.sigmaF[0] == 1.0                               # s02F afterwards
    d = stats.kstest(draws, "norm", args=(num / denom, np.sqrt(sig_e / denom)))
    assert d.pvalue > 1e-3, d


@pytest.mark.parametrize("layout", [orc.DENSE, orc.RAGGED], ids=["dense", "ragged"])
def test_residual_change_and_eps_sum(layout):
    pr, o, X = problem(layout=layout, F=4)
    o.gamma[:] = [0.1, -0.2, 0.05, 0.3]
    eps0 = o.eps.copy()
    es0 = o.state.epsSum
    g0 = o.gamma.copy()
    o.fixed_effects(5, 3)
    want = eps0 + X @ (g0 - o.gamma)
    assert np.max(np.abs(o.eps - want)) < 1e-12 * max(1.0, np.max(np.abs(want)))
    assert o.state.epsSum == es0                              # not updated here, as in the reference
    assert not np.array_equal(o.gamma, g0)


def test_covariate_order_is_carried():
    pr, o, _ = problem(F=6)
    orders = []
    for it in range(6):
        o.fixed_effects(7, it)
        orders.append(o.forder.copy())
        assert sorted(o.forder.tolist()) == list(range(6))
    # each iteration shuffles the previous order: replaying the shuffles from 0..F-1 gives the same
    replay = np.arange(6, dtype=np.uint32)
    for it in range(6):
        orc.lib().orc_shuffle_philox_stream(7, it, 0, orc.ORC_STREAM_FIXED,
                                            6, orc.p(replay, orc.C.c_uint32))
        assert np.array_equal(replay, orders[it])
    assert len({tuple(x) for x in orders}) > 1


def test_chain_recovers_fixed_effects():
    n, m, F = 3000, 60, 3
    pr = make_problem(n, m, seed=72, h2=0.3, causal_frac=0.05)
    rng = np.random.default_rng(72)
    X = rng.normal(size=(n, F)).astype(np.float32).astype(np.float64)
    gamma_true = np.array([0.6, -0.4, 0.25])
    y = pr.y + X @ gamma_true
    o = orc.OracleProblem(n, m, orc.DENSE, 4, 1, S, bed=pr.bed, lut=pr.lut)
    o.set_fixed(X)
    o.run_gibbs_fixed(y, 11, 30)
    # the chain works on y standardised (BayesRBase::init): gamma is on that scale
    scale = np.sqrt(((y - y.mean()) ** 2).sum() / (n - 1))
    assert np.allclose(o.gamma * scale, gamma_true, atol=0.08), (o.gamma * scale, gamma_true)
