"""Pins the oracle's BayesR sampler (oracle/bayesr.c).

- a6: the mixture posterior of one marker against the marginal likelihood
  integrated numerically over beta (scipy quad), the conditional posterior of
  beta against its quadrature mean / sd, the inverse-CDF selection on u, the
  700 guard and the fall-through of BayesRBase.cpp:335-366.
- a5/a7: the ragged (sparse) num and epsilon update equal the dense ones on
  the standardised column when nothing is missing (the invariant of the
  reference's eigenSparseTest.cpp:69-81), and eps stays y - mu - X beta over a
  whole sweep.
- a8/a9: updateMu and the sigmaE / sigmaG / pi draws by their distributions.
- whole chain: recovery of the simulated heritability.
"""
import ctypes as C

import numpy as np
import pytest
from scipy import integrate, stats

import _oracle as orc
from _data import make_problem, spec_dosage

S = np.array([0.01, 0.001, 0.0001])


def oracle_for(pr, layout, K=4, cva=S, group=None, G=1):
    if layout == orc.DENSE:
        return orc.OracleProblem(pr.n, pr.m, orc.DENSE, K, G, np.tile(cva, G), group=group, bed=pr.bed,
                                 lut=pr.lut)
    return orc.OracleProblem(pr.n, pr.m, orc.RAGGED, K, G, np.tile(cva, G), group=group, ragged=pr.ragged)


# ---------------------------------------------------------------- a6 ----

def log_marginal(x, r, sigmaE, var_b):
    """log of  int exp((2 b x.r - b^2 |x|^2) / (2 sigmaE)) N(b; 0, var_b) db  by quadrature."""
    xr, xx = float(x @ r), float(x @ x)
    f = lambda b: (2 * b * xr - b * b * xx) / (2 * sigmaE) - 0.5 * b * b / var_b
    prec = xx / sigmaE + 1 / var_b
    mode = (xr / sigmaE) / prec
    sd = 1 / np.sqrt(prec)
    fmax = f(mode)
    val, _ = integrate.quad(lambda b: np.exp(f(b) - fmax), mode - 40 * sd, mode + 40 * sd,
                            epsabs=0, epsrel=1e-13, limit=200)
    m1, _ = integrate.quad(lambda b: b * np.exp(f(b) - fmax), mode - 40 * sd, mode + 40 * sd,
                           epsabs=0, epsrel=1e-13, limit=200)
    m2, _ = integrate.quad(lambda b: b * b * np.exp(f(b) - fmax), mode - 40 * sd, mode + 40 * sd,
                           epsabs=0, epsrel=1e-13, limit=200)
    mean = m1 / val
    var = m2 / val - mean * mean
    return fmax + np.log(val) - 0.5 * np.log(2 * np.pi * var_b), mean, np.sqrt(var)


@pytest.fixture(scope="module")
def small():
    return make_problem(200, 30, seed=11, h2=0.6, causal_frac=0.1)


@pytest.mark.parametrize("layout", [orc.DENSE, orc.RAGGED])
@pytest.mark.parametrize("case", range(6))
def test_mixture_posterior_by_quadrature(small, layout, case):
    pr = small
    o = oracle_for(pr, layout)
    rng = np.random.default_rng(case)
    X = pr.x_dense()
    j = int(pr.causal[0]) if case % 2 == 0 else 1 + case
    o.eps[:] = pr.y - pr.y.mean() + rng.normal(0, 0.1, pr.n)
    beta_old = [0.0, 0.3, -0.2, 0.0, 0.05, 0.0][case]
    o.beta[j] = beta_old
    sigmaE, sigmaG = [(0.5, 0.5), (0.3, 0.9), (0.9, 0.05), (0.6, 0.4), (1.2, 2.0), (0.2, 0.01)][case]
    pi = np.array([[0.5, 0.125, 0.125, 0.125], [0.9, 0.05, 0.03, 0.02], [0.4, 0.3, 0.2, 0.1],
                   [0.97, 0.01, 0.01, 0.01], [0.25, 0.25, 0.25, 0.25], [0.5, 0.3, 0.1, 0.1]][case])
    o.set_scalars(0.0, sigmaE, float(o.eps.sum()), [sigmaG], pi)
    x = X[:, j]
    r = o.eps + x * beta_old                    # residual without marker j
    logw = [np.log(pi[0])]
    post = []
    for k in range(3):
        lm, mean, sd = log_marginal(x, r, sigmaE, sigmaG * S[k])
        logw.append(np.log(pi[k + 1]) + lm)
        post.append((mean, sd))
    logw = np.array(logw)
    P = np.exp(logw - logw.max())
    P /= P.sum()
    num = C.c_double()
    acum = C.c_double()
    eps0 = o.eps.copy()
    # u = 0.5 * P0: selects the zero component
    orc.lib().orc_process_column(C.byref(o.data), C.byref(o.model), C.byref(o.state), j, 0.5 * P[0], 0.0,
                                 C.byref(num), C.byref(acum))
    assert abs(num.value - x @ r) < 1e-9 * max(1.0, abs(x @ r))
    assert abs(acum.value - P[0]) < 1e-9
    assert o.comp[j] == 0 and o.beta[j] == 0.0
    # each slab: u in the middle of its interval; z = 0 gives the posterior mean, z = 1 mean + sd
    cum = np.cumsum(P)
    for k in range(1, 4):
        for zz in (0.0, 1.0):
            if P[k] < 1e-6:
                continue
            o.eps[:] = eps0
            o.beta[j] = beta_old
            u = cum[k - 1] + 0.5 * P[k]
            orc.lib().orc_process_column(C.byref(o.data), C.byref(o.model), C.byref(o.state), j, u, zz,
                                         None, None)
            assert o.comp[j] == k
            mean, sd = post[k - 1]
            assert abs(o.beta[j] - (mean + zz * sd)) < 1e-7 * max(sd, abs(mean))
            # eps carries the change: eps = r - x * beta_new
            assert np.allclose(o.eps, r - x * o.beta[j], atol=1e-12)


def test_u_equal_acum_selects_zero_component(small):
    pr = small
    o = oracle_for(pr, orc.DENSE)
    o.eps[:] = pr.y - pr.y.mean()
    o.set_scalars(0.0, 0.5, float(o.eps.sum()), [0.5], [0.5, 0.125, 0.125, 0.125])
    acum = C.c_double()
    j = int(pr.causal[0])
    eps0 = o.eps.copy()
    orc.lib().orc_process_column(C.byref(o.data), C.byref(o.model), C.byref(o.state), j, 0.999999, 0.3,
                                 None, C.byref(acum))
    a = acum.value
    o.eps[:] = eps0
    o.beta[j] = 0.0
    orc.lib().orc_process_column(C.byref(o.data), C.byref(o.model), C.byref(o.state), j, a, 0.3, None, None)
    assert o.comp[j] == 0 and o.beta[j] == 0.0       # u <= acum is inclusive
    o.eps[:] = eps0
    orc.lib().orc_process_column(C.byref(o.data), C.byref(o.model), C.byref(o.state), j,
                                 np.nextafter(a, 2.0), 0.3, None, None)
    assert o.comp[j] != 0


def test_guard_700_fall_through_keeps_old_beta(small):
    """|logL_i - logL_r| > 700 for every r: no component is selected and beta /
    comp keep their previous values (BayesRBase.cpp:335-366)."""
    pr = small
    cva = np.array([1.0, 1e-7, 1e-14])
    o = oracle_for(pr, orc.DENSE, cva=cva)
    j = int(pr.causal[0])
    X = pr.x_dense()
    x = X[:, j]
    o.eps[:] = 60.0 * x                       # num = 60 (N - 1): a huge association
    o.beta[j] = 0.123
    o.comp[j] = 2
    o.set_scalars(0.0, 0.01, float(o.eps.sum()), [1.0], [0.25, 0.25, 0.25, 0.25])
    acum = C.c_double()
    orc.lib().orc_process_column(C.byref(o.data), C.byref(o.model), C.byref(o.state), j, 0.3, 0.0, None,
                                 C.byref(acum))
    assert acum.value == 0.0
    assert o.beta[j] == 0.123 and o.comp[j] == 2 and o.v.sum() == 0
    # u = 0 is <= acum = 0: the zero component is taken
    orc.lib().orc_process_column(C.byref(o.data), C.byref(o.model), C.byref(o.state), j, 0.0, 0.0, None,
                                 None)
    assert o.beta[j] == 0.0 and o.comp[j] == 0


# ------------------------------------------------------------- a5 / a7 ----

def test_ragged_equals_dense_without_missing():
    pr = make_problem(600, 120, seed=21, causal_frac=0.05)
    od = oracle_for(pr, orc.DENSE)
    og = oracle_for(pr, orc.RAGGED)
    rng = np.random.default_rng(2)
    eps = pr.y - pr.y.mean()
    for o in (od, og):
        o.eps[:] = eps
        o.set_scalars(0.0, 0.5, float(eps.sum()), [0.5], [0.5, 0.125, 0.125, 0.125])
    X = pr.x_dense()
    for j in rng.choice(pr.m, 20, replace=False):
        j = int(j)
        b_old = rng.normal()
        nd = od.compute_num(j, b_old)
        ng = og.compute_num(j, b_old)
        assert abs(nd - ng) < 1e-9 * max(1, abs(nd))
        assert abs(nd - (X[:, j] @ eps + b_old * (pr.n - 1))) < 1e-9 * max(1, abs(nd))
        b_new = rng.normal()
        od.apply_epsilon_change(j, b_old, b_new)
        og.apply_epsilon_change(j, b_old, b_new)
        eps = eps + (b_old - b_new) * X[:, j]
        assert np.allclose(od.eps, eps, atol=1e-11)
        assert np.allclose(og.eps, eps, atol=1e-11)
        assert abs(og.epsSum - og.eps.sum()) < 1e-8


@pytest.mark.parametrize("layout", [orc.DENSE, orc.RAGGED])
def test_sweep_keeps_residual_identity(layout):
    pr = make_problem(800, 300, seed=31, causal_frac=0.03)
    o = oracle_for(pr, layout)
    o.init(pr.y, seed=5)
    ystd = o.eps.copy()
    order = np.random.default_rng(1).permutation(pr.m).astype(np.uint32)
    u, z = np.random.default_rng(2).random(pr.m), np.random.default_rng(3).normal(size=pr.m)
    tr = o.sweep(order, u, z)
    X = pr.x_dense()
    assert np.allclose(o.eps, ystd - X @ o.beta, atol=1e-10)
    if layout == orc.RAGGED:
        assert abs(o.epsSum - o.eps.sum()) < 1e-8
    # trace is in sweep order and consistent with the state
    assert np.array_equal(tr["comp"], o.comp[order])
    assert np.array_equal(tr["beta"], o.beta[order])
    assert o.v.sum() == pr.m  # every marker selected a component
    assert abs(o.betasqn[0] - np.sum(o.beta ** 2)) < 1e-12


def test_missing_data_update_leaves_missing_individuals():
    """Ragged: a missing call gets delta 0 (raggedbayesrkernel.cpp:26-27)."""
    pr = make_problem(400, 20, seed=41, miss=0.1)
    og = oracle_for(pr, orc.RAGGED)
    og.eps[:] = pr.y
    og.state.epsSum = float(og.eps.sum())
    j = 3
    miss = pr.ragged.miss[int(pr.ragged.miss_off[j]): int(pr.ragged.miss_off[j + 1])]
    assert miss.size > 0
    e0 = og.eps.copy()
    og.apply_epsilon_change(j, 0.0, 0.7)
    assert np.array_equal(og.eps[miss], e0[miss])


# --------------------------------------------------------------- a8 / a9 ----

def test_update_mu_distribution():
    pr = make_problem(300, 10, seed=51)
    mus = []
    for it in range(3000):
        o = oracle_for(pr, orc.DENSE)
        o.eps[:] = pr.y - 0.2
        o.set_scalars(0.1, 0.8, 0.0, [0.5], [0.5, 0.125, 0.125, 0.125])
        o.update_mu(seed=9, iteration=it)
        mus.append(o.mu)
        assert np.allclose(o.eps, (pr.y - 0.2 + 0.1) - o.mu, atol=1e-14)
    mean = (pr.y - 0.1).sum() / pr.n
    zs = (np.array(mus) - mean) / np.sqrt(0.8 / pr.n)
    assert stats.kstest(zs, "norm").pvalue > 1e-4


def test_post_sweep_draw_parameterisation():
    pr = make_problem(300, 10, seed=52)
    draws_e, draws_g = [], []
    for it in range(3000):
        o = oracle_for(pr, orc.DENSE)
        o.eps[:] = pr.y
        o.set_scalars(0.0, 0.8, 0.0, [0.5], [0.5, 0.125, 0.125, 0.125])
        o.v[:] = [970, 20, 6, 4]
        o.betasqn[0] = 0.05
        o.post_sweep(seed=13, iteration=it)
        draws_e.append(o.sigmaE)
        draws_g.append(o.sigmaG[0])
    e2 = pr.y @ pr.y
    v0 = 1e-4
    # sigmaE = inv_scaled_chisq(v0E + N, (|eps|^2 + v0E s02E) / (v0E + N))
    dof = v0 + pr.n
    s2 = (e2 + v0 * v0) / dof
    assert stats.kstest(1 / np.array(draws_e), "gamma", args=(dof / 2, 0, 2 / (dof * s2))).pvalue > 1e-4
    # sigmaG = inv_scaled_chisq(v0G + m0, (betasqn * m0 + v0G s02G) / (v0G + m0)), m0 = 30
    m0 = 30.0
    dof = v0 + m0
    s2 = (0.05 * m0 + v0 * v0) / dof
    assert stats.kstest(1 / np.array(draws_g), "gamma", args=(dof / 2, 0, 2 / (dof * s2))).pvalue > 1e-4


def test_chain_recovers_heritability():
    pr = make_problem(3000, 1000, seed=61, h2=0.5, causal_frac=0.01)
    o = oracle_for(pr, orc.DENSE)
    hist, _ = o.run_gibbs(pr.y, seed=20261202, iterations=60)
    sigmaE, sigmaG = hist[20:, 1], hist[20:, 4]
    h2 = np.mean(sigmaG / (sigmaG + sigmaE))
    assert 0.35 < h2 < 0.65, h2
    # causal markers are the ones picked most
    assert np.mean(o.comp[pr.causal] > 0) > 0.5


def test_chain_is_deterministic_and_shuffle_modes():
    pr = make_problem(500, 200, seed=71)
    h1, o1 = oracle_for(pr, orc.RAGGED).run_gibbs(pr.y, seed=3, iterations=3)
    h2_, o2 = oracle_for(pr, orc.RAGGED).run_gibbs(pr.y, seed=3, iterations=3)
    assert np.array_equal(h1, h2_) and np.array_equal(o1, o2)
    libc = C.CDLL("libc.so.6")
    libc.srand(1)
    h3, o3 = oracle_for(pr, orc.RAGGED).run_gibbs(pr.y, seed=3, iterations=3, shuffle_mode=1)
    assert sorted(o3.tolist()) == list(range(pr.m)) and not np.array_equal(o3, o1)
