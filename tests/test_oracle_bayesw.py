"""The BayesW part of the CPU oracle against what the mathematics fixes.

- Gauss-Hermite rules: the tabulated nodes and exp(x^2)-adjusted weights of
  bayeswbase.cpp:126-457 against numpy's Golub-Welsch rule, and the x6 = -x3
  entry the reference carries for n = 9..15.
- The adaptive integral of processColumn against a brute-force evaluation of
  the same integrand from the individual-level Weibull log-likelihood.
- ARMS (BayesW_arms.cpp) draws against known log-concave densities
  (Kolmogorov-Smirnov), and its error codes.
- A BayesW sweep: eps = y - mu - X beta, vi = exp(alpha eps - EulerGamma),
  counts, and the chain's recovery of the Weibull shape.
"""
import ctypes as C

import numpy as np
import pytest
from scipy import stats

import bayesrrcmd_amd as brr
import _oracle as orc
from _data import make_bw_problem, spec_dosage

EULER = 0.577215664901532


def rule(n):
    x = np.zeros(14)
    w = np.zeros(14)
    wc = C.c_double()
    k = orc.lib().orc_bw_gh_rule(n, orc.p(x, C.c_double), orc.p(w, C.c_double), C.byref(wc))
    return k, x[:k], w[:k], wc.value


@pytest.mark.parametrize("n", [3, 5, 7, 9, 11, 13, 15])
def test_gh_rules_match_hermite(n):
    k, x, w, wc = rule(n)
    assert k == n - 1
    nodes, weights = np.polynomial.hermite.hermgauss(n)
    pos = nodes[nodes > 1e-12][::-1]                   # x1 > x3 > ...
    wpos = (weights * np.exp(nodes ** 2))[nodes > 1e-12][::-1]
    h = (n - 1) // 2
    np.testing.assert_allclose(x[0::2][:h], pos, rtol=0, atol=2e-11)
    np.testing.assert_allclose(w[0::2][:h], wpos, rtol=2e-10)
    np.testing.assert_allclose(w[1::2][:h], wpos, rtol=2e-10)
    assert wc == pytest.approx(weights[n // 2], rel=2e-10)
    for i in range(h):
        if n >= 9 and i == 2:
            # the reference evaluates the sixth node at -x3 (bayeswbase.cpp:230, 276, 331, 394)
            assert x[5] == -x[2]
        else:
            assert x[2 * i + 1] == -x[2 * i]


def test_gh_unsupported_order():
    assert rule(4)[0] == 0 and rule(17)[0] == 0


def brute_integrand(s, alpha, sq, dose, mean, sd, vi, fail):
    """exp(ell(beta) - ell(0) ... ) of one marker from individuals: the marker's
    Weibull log-likelihood terms at beta = s sq, plus sum(vi) - s^2."""
    x = (dose - mean) / sd
    beta = s * sq
    return np.exp(-alpha * beta * np.sum(x * fail) + vi.sum() - np.sum(vi * np.exp(-alpha * beta * x)) - s * s)


@pytest.mark.parametrize("n", [3, 7, 15])
def test_gh_integral_matches_brute_force(n):
    rng = np.random.default_rng(1)
    N = 400
    dose = rng.choice([0, 1, 2], N, p=[0.5, 0.4, 0.1]).astype(float)
    fail = (rng.random(N) < 0.9).astype(float)
    mean = dose.sum() / N
    sd = np.sqrt(np.sum((dose - mean) ** 2) / (N - 1))
    alpha, sigma_b, Ck = 3.0, 2e-4, 0.01
    eps = rng.normal(0, 0.4, N)
    vi = np.exp(alpha * eps - EULER)
    sums = np.array([vi.sum(), vi[dose == 0].sum(), vi[dose == 1].sum(), vi[dose == 2].sum()])
    x = (dose - mean) / sd
    sum_failure = np.sum(x * fail)
    exp_sum = (sums[2] * (1 - 2 * mean) + 4 * (1 - mean) * sums[3] + sums[0] * mean * mean) / (sd * sd)
    assert exp_sum == pytest.approx(np.sum(vi * x * x), rel=1e-10)   # closed form of sum vi x^2
    sigma = 1.0 / np.sqrt(1 + alpha * alpha * sigma_b * Ck * exp_sum)
    got = orc.lib().orc_bw_gh_integral(n, sigma, Ck, sigma_b, alpha, sum_failure, mean, sd,
                                       orc.p(np.ascontiguousarray(sums), C.c_double))
    # brute force: the same adaptive rule with numpy's nodes (for n >= 9 with the
    # reference's sixth node), the integrand summed over individuals
    k, xs, ws, wc = rule(n)
    sq = np.sqrt(2 * Ck * sigma_b)
    want = sigma * (sum(wi * brute_integrand(sigma * xi, alpha, sq, dose, mean, sd, vi, fail)
                        for xi, wi in zip(xs, ws)) + wc)
    assert got == pytest.approx(want, rel=1e-9)
    # and the rule integrates: for n <= 7 against adaptive quadrature of the integrand
    if n == 7:
        from scipy.integrate import quad
        exact = quad(lambda s: brute_integrand(s, alpha, sq, dose, mean, sd, vi, fail), -8, 8, limit=200)[0]
        assert got == pytest.approx(exact, rel=2e-3)


def arms_draws(logdens, xinit, xl, xr, n, seed=9, iteration=0):
    cb = orc.LOGDENS(lambda x, _: logdens(x))
    xi = np.ascontiguousarray(xinit, dtype=np.float64)
    out = np.zeros(n)
    for i in range(n):
        r = orc.OrcArmsRng(seed, i, iteration, 0)
        xs = C.c_double()
        err = orc.lib().orc_arms(orc.p(xi, C.c_double), len(xinit), xl, xr, cb, None, 100, C.byref(r),
                                 C.byref(xs), None)
        assert err == 0
        out[i] = xs.value
    return out


def test_arms_samples_a_normal():
    d = arms_draws(lambda x: -(x - 0.3) ** 2 / (2 * 0.5), [-1.0, 0.0, 0.5, 1.0], -6.0, 6.0, 3000)
    assert stats.kstest(d, "norm", args=(0.3, np.sqrt(0.5))).pvalue > 1e-3
    assert abs(d.mean() - 0.3) < 0.05


def test_arms_samples_a_gamma():
    d = arms_draws(lambda x: 2.0 * np.log(x) - x, [0.5, 2.0, 3.0, 5.0], 0.0, 30.0, 3000, seed=11)
    assert stats.kstest(d, "gamma", args=(3.0,)).pvalue > 1e-3


def test_arms_is_deterministic_per_site():
    f = lambda x: -x * x / 2
    a = arms_draws(f, [-1, 0, 0.5, 1], -5, 5, 20, seed=3, iteration=7)
    b = arms_draws(f, [-1, 0, 0.5, 1], -5, 5, 20, seed=3, iteration=7)
    c = arms_draws(f, [-1, 0, 0.5, 1], -5, 5, 20, seed=3, iteration=8)
    assert np.array_equal(a, b) and not np.array_equal(a, c)


def arms_err(logdens, xinit, xl, xr, ninit=None):
    cb = orc.LOGDENS(lambda x, _: logdens(x))
    xi = np.ascontiguousarray(xinit, dtype=np.float64)
    r = orc.OrcArmsRng(1, 0, 0, 0)
    xs = C.c_double()
    return orc.lib().orc_arms(orc.p(xi, C.c_double), ninit or len(xinit), xl, xr, cb, None, 100, C.byref(r),
                              C.byref(xs), None)


def test_arms_error_codes():
    f = lambda x: -x * x
    assert arms_err(f, [-1, 0], -5, 5) == 1001                 # too few initial points
    assert arms_err(f, [-5, 0, 1, 2], -5, 5) == 1003           # initial point on the bound
    assert arms_err(f, [-1, 1, 0, 2], -5, 5) == 1004           # not ascending
    assert arms_err(lambda x: x * x, [-0.5, 0.0, 0.3, 0.6], -1, 1) == 2000   # convex: no envelope


def residual_column(r, j, n):
    mean, sd = r.stats[j, 0], r.stats[j, 1]
    col = np.full(n, -(mean / sd))
    col[r.ones[r.ones_off[j]: r.ones_off[j + 1]]] += 1 / sd
    col[r.twos[r.twos_off[j]: r.twos_off[j + 1]]] += 2 / sd
    return col


def philox_order_u(seed, it, m):
    order = np.arange(m, dtype=np.uint32)
    orc.lib().orc_shuffle_philox(seed, it, 0, m, orc.p(order, C.c_uint32))
    u = np.zeros(m)
    z = np.zeros(m)
    orc.lib().orc_fill_marker_uz(seed, it, 0, m, orc.p(u, C.c_double), orc.p(z, C.c_double))
    return order, u


def test_bw_sweep_identities():
    pr = make_bw_problem(2000, 60, seed=31, causal_frac=0.1)
    w = orc.BayesWProblem(pr.n, pr.m, pr.ragged, pr.fail)
    w.init(pr.y)
    st = w.state
    assert st.d == pr.fail.sum()
    assert st.mu == pytest.approx(pr.y.mean(), rel=1e-12)
    assert orc.lib().orc_bw_sample_mu(C.byref(w.data), C.byref(st), 5, 0, 0) == 0
    w.refresh_vi()
    order, u = philox_order_u(5, 0, pr.m)
    w.v[:] = 1.0
    assert w.sweep(order, u, 5, 0) == 0
    # eps = y - mu - X beta with the residual-update columns
    xb = sum(w.beta[j] * residual_column(pr.ragged, j, pr.n) for j in np.flatnonzero(w.beta))
    np.testing.assert_allclose(w.eps, pr.y - st.mu - xb, rtol=0, atol=1e-10)
    np.testing.assert_allclose(w.vi, np.exp(st.alpha * w.eps - EULER), rtol=1e-12)
    assert np.all((w.beta == 0) == (w.comp == 0))
    np.testing.assert_array_equal(w.v, np.bincount(w.comp, minlength=4) + 1)
    assert (w.comp > 0).sum() > 0


def test_bw_process_column_sums_and_likelihoods():
    pr = make_bw_problem(1500, 20, seed=33, causal_frac=0.2)
    w = orc.BayesWProblem(pr.n, pr.m, pr.ragged, pr.fail, quad_points=7)
    w.init(pr.y)
    w.refresh_vi()
    st = w.state
    r = pr.ragged
    for j in range(5):
        tr = orc.OrcBwTrace()
        vi = w.vi.copy()
        pi = w.pi.copy()
        assert orc.lib().orc_bw_process_column(C.byref(w.data), C.byref(w.model), C.byref(st), j, 0.5, 7, 0, 0,
                                               C.byref(tr)) == 0
        ones = r.ones[r.ones_off[j]: r.ones_off[j + 1]]
        twos = r.twos[r.twos_off[j]: r.twos_off[j + 1]]
        assert tr.vi_sum == pytest.approx(vi.sum(), rel=1e-12)
        assert tr.vi_1 == pytest.approx(vi[ones].sum(), rel=1e-12)
        assert tr.vi_2 == pytest.approx(vi[twos].sum(), rel=1e-12)
        ml = np.array(tr.ml[:4])
        assert ml[0] == pytest.approx(pi[0] * 1.77245385090552, rel=1e-15)
        assert tr.acum == pytest.approx(ml[0] / ml.sum(), rel=1e-12)
        # class likelihoods: pi_k times the adaptive rule on the brute-force integrand
        dose = np.zeros(pr.n)
        dose[ones] = 1
        dose[twos] = 2
        mean, sd = r.stats[j, 0], r.stats[j, 1]
        x = (dose - mean) / sd
        k_, xs, ws, wc = rule(7)
        for c, Ck in enumerate(w.C):
            sigma = 1.0 / np.sqrt(1 + st.alpha ** 2 * st.sigma_b * Ck * np.sum(vi * x * x))
            sq = np.sqrt(2 * Ck * st.sigma_b)
            gh = sigma * (sum(wi * brute_integrand(sigma * xi, st.alpha, sq, dose, mean, sd, vi, pr.fail)
                              for xi, wi in zip(xs, ws)) + wc)
            assert ml[c + 1] == pytest.approx(pi[c + 1] * gh, rel=1e-8)
        if w.beta[j] != 0.0:
            break


@pytest.mark.slow
def test_bw_chain_recovers_weibull_shape():
    pr = make_bw_problem(3000, 30, seed=35, causal_frac=0.1, censor=0.1, alpha=3.0, mu=4.0)
    w = orc.BayesWProblem(pr.n, pr.m, pr.ragged, pr.fail)
    err, hist, _ = w.run_gibbs(pr.y, 17, 60)
    assert err == 0
    alpha = hist[20:, 1].mean()
    assert abs(alpha - 3.0) < 0.3, alpha
    assert abs(hist[20:, 0].mean() - 4.0) < 0.2


# ------------------------------------------------------------- dense layout --

def dense_problem(pr, **kw):
    lut = np.array([orc.dense_lut(pr.bed[j], pr.n) for j in range(pr.bed.shape[0])])
    return orc.BayesWProblem(pr.n, pr.m, None, pr.fail, bed=pr.bed, lut=lut, **kw), lut


def test_bw_dense_equals_ragged_without_missing():
    """DenseBayesWKernel and RaggedBayesWKernel are the same model on the same genotypes
    when no call is missing (the dense standardisation equals the ragged (x - mean) / sd):
    class likelihoods, classes and draws agree to rounding."""
    pr = make_bw_problem(1200, 24, seed=41, causal_frac=0.2)
    wr = orc.BayesWProblem(pr.n, pr.m, pr.ragged, pr.fail)
    wd, _ = dense_problem(pr)
    np.testing.assert_allclose(wd.sum_failure, wr.sum_failure, rtol=1e-10, atol=1e-9)
    for w in (wr, wd):
        w.init(pr.y)
        w.refresh_vi()
    for j in range(pr.m):
        tr = [orc.OrcBwTrace(), orc.OrcBwTrace()]
        for w, t in zip((wr, wd), tr):
            assert orc.lib().orc_bw_process_column(C.byref(w.data), C.byref(w.model), C.byref(w.state), j, 0.7, 3,
                                                   0, 0, C.byref(t)) == 0
        np.testing.assert_allclose(np.array(tr[1].ml[:4]), np.array(tr[0].ml[:4]), rtol=1e-9)
        assert wd.comp[j] == wr.comp[j]
        assert wd.beta[j] == pytest.approx(wr.beta[j], rel=1e-8, abs=1e-12)
        np.testing.assert_allclose(wd.eps, wr.eps, rtol=0, atol=1e-10)
    assert np.count_nonzero(wd.comp) > 0


def test_bw_dense_likelihoods_match_brute_force_with_missing():
    """With missing calls (mean-imputed in the dense column) the dense class likelihoods
    against the individual-level integrand over the dense column."""
    from _data import spec_dosage
    pr = make_bw_problem(900, 8, seed=43, causal_frac=0.25)
    bed, _ = brr.synth_bed(44, pr.n, pr.m, 0.05, 0.5, 0.05)     # 5% missing calls
    pr.bed = bed
    wd, lut = dense_problem(pr)
    wd.init(pr.y)
    wd.refresh_vi()
    st = wd.state
    d = spec_dosage(bed, pr.n)
    for j in range(4):
        cx = lut[j][np.where(d[j] < 0, 3, d[j])]
        vi = wd.vi.copy()
        pi = wd.pi.copy()
        tr = orc.OrcBwTrace()
        assert orc.lib().orc_bw_process_column(C.byref(wd.data), C.byref(wd.model), C.byref(st), j, 0.5, 5, 0, 0,
                                               C.byref(tr)) == 0
        sf = float(np.sum(cx * pr.fail))
        assert wd.sum_failure[j] == pytest.approx(sf, rel=1e-12, abs=1e-9)
        _, xs, ws, wc = rule(7)
        for c, Ck in enumerate(wd.C):
            sigma = 1.0 / np.sqrt(1 + st.alpha ** 2 * st.sigma_b * Ck * np.sum(vi * cx * cx))
            sq = np.sqrt(2 * Ck * st.sigma_b)
            f = lambda s: np.exp(-st.alpha * s * sf * sq + np.sum(vi * (1 - np.exp(-cx * s * sq * st.alpha))) - s * s)
            gh = sigma * (sum(wi * f(sigma * xi) for xi, wi in zip(xs, ws)) + wc)
            assert tr.ml[c + 1] == pytest.approx(pi[c + 1] * gh, rel=1e-9)
        if wd.beta[j] != 0.0:
            break


def test_bw_dense_sweep_residual_identity():
    pr = make_bw_problem(1500, 40, seed=45, causal_frac=0.1)
    wd, lut = dense_problem(pr)
    wd.init(pr.y)
    assert orc.lib().orc_bw_sample_mu(C.byref(wd.data), C.byref(wd.state), 5, 0, 0) == 0
    wd.refresh_vi()
    order, u = philox_order_u(5, 0, pr.m)
    wd.v[:] = 1.0
    assert wd.sweep(order, u, 5, 0) == 0
    from _data import spec_dosage
    d = spec_dosage(pr.bed, pr.n)
    x = np.take_along_axis(lut, np.where(d < 0, 3, d), axis=1)
    np.testing.assert_allclose(wd.eps, pr.y - wd.state.mu - wd.beta @ x, rtol=0, atol=1e-10)
    np.testing.assert_allclose(wd.vi, np.exp(wd.state.alpha * wd.eps - EULER), rtol=1e-12)
    assert np.count_nonzero(wd.beta) > 0
