"""BayesW fixed effects on the device (brr_hip_bw_set_fixed, brr_hip_bw_fixed_effects) against the oracle.

The device draws theta_0 .. theta_F-1 after mu with the same sampler, starting
points and uniforms as the oracle; its N-sums run per slice and workgroup and
its exp is the device library's, so the bar is that of the other BayesW chain
tests: classes and class counts identical, every float within 1e-7 relative
over whole chains.
"""
import numpy as np
import pytest

import bayesrrcmd_amd as brr
import _oracle as orc
from _data import make_bw_problem

pytestmark = pytest.mark.gpu

C = np.array([0.001, 0.01, 0.1])


def close(a, b, rtol, what=""):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(1.0, float(np.max(np.abs(b))) if b.size else 1.0)
    err = float(np.max(np.abs(a - b))) if a.size else 0.0
    assert err <= rtol * scale, f"{what}: max abs err {err:.3e} (scale {scale:.3e})"


def covariates(n, F, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(0.0, 1.0, size=(n, F)).astype(np.float32)
    X[:, 0] = (rng.random(n) < 0.5).astype(np.float32)
    return X


@pytest.mark.parametrize("layout,n,m,F,wg", [("ragged", 2500, 100, 1, 0), ("ragged", 3000, 80, 3, 7),
                                           ("dense", 2500, 90, 2, 0)])
def test_bw_chain_with_fixed_effects(gpu_lib, layout, n, m, F, wg):
    seed, iters = 5150 + F, 3
    pr = make_bw_problem(n, m, seed=seed, causal_frac=0.1)
    X = covariates(n, F, seed)
    y = (pr.y + X.astype(np.float64) @ np.linspace(0.3, -0.2, F)).astype(np.float32).astype(np.float64)
    lut = brr.build_dense(pr.bed, n)
    if layout == "ragged":
        w = orc.BayesWProblem(n, m, pr.ragged, pr.fail)
    else:
        w = orc.BayesWProblem(n, m, None, pr.fail, bed=pr.bed, lut=lut)
    w.set_fixed(X)
    err, hist_o, _ = w.run_gibbs_fixed(y, seed, iters)
    assert err == 0
    eps0, sc = brr.bayesw_init(y, pr.fail, m, 4)
    orders = []
    order = np.arange(m, dtype=np.uint32)
    for it in range(iters):
        brr.shuffle(seed, it, 0, order)
        orders.append(order.copy())
    with brr.HipBayesW(n, m, 4, quad_points=7, workgroups=wg) as h:
        h.upload_ragged(pr.ragged) if layout == "ragged" else h.upload_dense(pr.bed, lut)
        h.set_fixed(X)                                  # before the model: sum_failure_fix follows it
        h.set_model(C, pr.fail)
        h.set_state(eps0, np.zeros(m), np.zeros(m, dtype=np.int32))
        h.set_scalars(sc["mu"], sc["alpha"], sc["sigma_b"], sc["d"], sc["pi"], sc["v"])
        hist = h.run_chain(seed, 0, 0, np.stack(orders))
        np.testing.assert_array_equal(hist[:, 3], hist_o[:, 3])
        for c, name in enumerate(("mu", "alpha", "sigma_b", "nonzero", "sum beta^2")):
            close(hist[:, c], hist_o[:, c], 1e-7, f"chain {name}")
        d = h.get_state()
        np.testing.assert_array_equal(d["comp"], w.comp)
        close(d["beta"], w.beta, 1e-7, "beta")
        close(d["eps"], w.eps, 1e-7, "eps")
        close(h.get_theta(), w.theta, 1e-7, "theta")
        assert np.count_nonzero(w.theta) == F


def test_bw_fixed_effects_alone(gpu_lib):
    n, m, F = 3000, 30, 4
    pr = make_bw_problem(n, m, seed=5160)
    X = covariates(n, F, 5161)
    w = orc.BayesWProblem(n, m, pr.ragged, pr.fail)
    w.init(pr.y)
    w.set_fixed(X)
    w.theta[:] = [0.1, -0.05, 0.02, 0.0]
    with brr.HipBayesW(n, m, 4) as h:
        h.upload_ragged(pr.ragged)
        h.set_model(C, pr.fail)
        h.set_fixed(X)                                  # after the model
        h.set_theta(w.theta)
        h.set_state(w.eps, w.beta, w.comp)
        st = w.state
        h.set_scalars(st.mu, st.alpha, st.sigma_b, st.d, w.pi, np.ones(4))
        for it in range(2):
            for f in range(F):
                assert w.sample_theta(f, 31, it) == 0
            h.fixed_effects(31, it)
            close(h.get_theta(), w.theta, 1e-9, f"iteration {it} theta")
            close(h.get_state()["eps"], w.eps, 1e-9, f"iteration {it} eps")
        assert h.stats()["rounds"] > F                  # one grid sum per density value


def test_bw_fixed_bad_calls(gpu_lib):
    pr = make_bw_problem(500, 10, seed=5170)
    with brr.HipBayesW(pr.n, pr.m, 4) as h:
        h.upload_ragged(pr.ragged)
        h.set_model(C, pr.fail)
        with pytest.raises(brr.BrrHipError):
            h.fixed_effects(1, 0)
        with pytest.raises(brr.BrrHipError):
            h.set_fixed(np.zeros((pr.n, 129), np.float32))
        h.F = 2
        with pytest.raises(brr.BrrHipError):
            h.get_theta()
