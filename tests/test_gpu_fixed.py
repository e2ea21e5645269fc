"""Fixed effects on the device (brr_hip_set_fixed, brr_hip_fixed_effects) against the oracle.

Bar as the other BayesR chain tests: components and m0 identical; mu, sigmaE,
|eps|^2, sigmaG, beta, epsilon and gamma within 1e-8 relative over whole
chains (device sums per slice and workgroup, the oracle in index order); the
covariate order identical.
"""
import numpy as np
import pytest

import bayesrrcmd_amd as brr
import _oracle as orc
from _data import make_problem

pytestmark = pytest.mark.gpu

S = np.array([0.01, 0.001, 0.0001])


def close(a, b, rtol, what=""):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(1.0, float(np.max(np.abs(b))) if b.size else 1.0)
    err = float(np.max(np.abs(a - b))) if a.size else 0.0
    assert err <= rtol * scale, f"{what}: max abs err {err:.3e} (scale {scale:.3e})"


def covariates(n, F, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, F)).astype(np.float32)
    X[:, 0] = (rng.random(n) < 0.5).astype(np.float32)
    return X


@pytest.mark.parametrize("layout,storage,F,window", [
    (orc.DENSE, brr.STORAGE_AUTO, 1, 0),
    (orc.DENSE, brr.STORAGE_AUTO, 5, 0),
    (orc.RAGGED, brr.STORAGE_CODES, 3, 0),
    (orc.RAGGED, brr.STORAGE_LISTS, 2, 20),
])
def test_chain_with_fixed_effects(gpu_lib, layout, storage, F, window):
    pr = make_problem(2500, 300, seed=601 + F, h2=0.5, causal_frac=0.03)
    X = covariates(pr.n, F, 602)
    y = pr.y + X.astype(np.float64) @ np.linspace(0.5, -0.3, F)
    seed, iters = 20261203, 5
    if layout == orc.DENSE:
        o = orc.OracleProblem(pr.n, pr.m, orc.DENSE, 4, 1, S, bed=pr.bed, lut=pr.lut)
    else:
        o = orc.OracleProblem(pr.n, pr.m, orc.RAGGED, 4, 1, S, ragged=pr.ragged)
    o.set_fixed(X)
    hist_o, order_o = o.run_gibbs_fixed(y, seed, iters, window=window)
    eps, sc = brr.bayesr_init(y, 1, 4, seed)
    orders = np.empty((iters, pr.m), dtype=np.uint32)
    cur = np.arange(pr.m, dtype=np.uint32)
    for it in range(iters):
        brr.shuffle(seed, it, 0, cur)
        orders[it] = cur
    code = brr.LAYOUT_DENSE_2BIT if layout == orc.DENSE else brr.LAYOUT_RAGGED
    h = brr.HipSweep(pr.n, pr.m, 4, 1, code, ragged_storage=storage)
    h.upload_dense(pr.bed, pr.lut) if layout == orc.DENSE else h.upload_ragged(pr.ragged)
    h.set_model(S)
    h.set_fixed(X)
    if window:
        h.set_async(window)
    h.set_state(eps, np.zeros(pr.m), np.zeros(pr.m, np.int32))
    h.set_scalars(sc["mu"], sc["sigmaE"], sc["epsilon_sum"], sc["sigmaG"], sc["pi"])
    hist_h = h.run_chain(seed, 0, 0, orders)
    for c, what in ((0, "mu"), (1, "sigmaE"), (3, "|eps|^2"), (4, "sigmaG")):
        close(hist_h[:, c], hist_o[:, c], 1e-8, what)
    assert np.array_equal(hist_h[:, 2], hist_o[:, 2]), "m0"
    st = h.get_state()
    assert np.array_equal(st["comp"], o.comp)
    close(st["beta"], o.beta, 1e-8, "beta")
    close(st["eps"], o.eps, 1e-8, "eps")
    fx = h.get_fixed_state()
    close(fx["gamma"], o.gamma, 1e-8, "gamma")
    assert np.array_equal(fx["order"], o.forder)
    assert fx["sigma_f"] == 1.0
    if layout == orc.RAGGED:
        close([h.get_scalars()["epsilon_sum"]], [o.state.epsSum], 1e-8, "epsSum")
    h.close()


def test_fixed_effects_block_alone(gpu_lib):
    """brr_hip_fixed_effects: one block on the current state, as the host's runGibbs would call it."""
    pr = make_problem(3001, 50, seed=610)
    X = covariates(pr.n, 7, 611)
    o = orc.OracleProblem(pr.n, pr.m, orc.RAGGED, 4, 1, S, ragged=pr.ragged)
    o.init(pr.y, 5)
    o.set_fixed(X)
    o.gamma[:] = np.linspace(-0.2, 0.2, 7)
    o.sigmaF[:] = [3.0]
    h = brr.HipSweep(pr.n, pr.m, 4, 1, brr.LAYOUT_RAGGED)
    h.upload_ragged(pr.ragged)
    h.set_model(S)
    h.set_state(o.eps, o.beta, o.comp)
    h.set_scalars(o.mu, o.sigmaE, o.epsSum, o.sigmaG, o.pi)
    h.set_fixed(X)
    h.set_fixed_state(o.gamma, o.forder, 3.0)
    for it in range(3):
        o.fixed_effects(41, it)
        h.fixed_effects(41, it)
        fx = h.get_fixed_state()
        assert np.array_equal(fx["order"], o.forder), f"iteration {it} order"
        close(fx["gamma"], o.gamma, 1e-10, f"iteration {it} gamma")
        close(h.get_state()["eps"], o.eps, 1e-10, f"iteration {it} eps")
        assert fx["sigma_f"] == o.sigmaF[0] == 1.0
    assert h.stats()["rounds"] == 7
    h.close()


def test_fixed_effects_bad_calls(gpu_lib):
    pr = make_problem(400, 20, seed=612)
    h = brr.HipSweep(pr.n, pr.m, 4, 1, brr.LAYOUT_DENSE_2BIT)
    h.upload_dense(pr.bed, pr.lut)
    h.set_model(S)
    with pytest.raises(brr.BrrHipError):
        h.fixed_effects(1, 0)                                   # no covariates yet
    with pytest.raises(brr.BrrHipError):
        h.set_fixed(np.zeros((pr.n, 129), np.float32))
    h.set_fixed(np.ones((pr.n, 2), np.float32))
    with pytest.raises(brr.BrrHipError):
        h.set_fixed_state(order=np.array([1, 1]))
    with pytest.raises(brr.BrrHipError):
        h.set_fixed_state(sigma_f=0.0)
    h.close()
    t = brr.HipSweep(pr.n, pr.m, 4, 1, brr.LAYOUT_DENSE_2BIT, traits=2)
    with pytest.raises(brr.BrrHipError):
        t.set_fixed(np.ones((pr.n, 2), np.float32))
    t.close()
