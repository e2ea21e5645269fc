"""The asynchronous sweep (brr_hip_set_async) against the oracle's orc_sweep_async.

ParallelGraph's asyncppbayes in the schedule where every marker of a window of
W markers reads the window-start residual; the device runs one exchange round
per window. Bar as the sequential parity tests: components and v counts
identical, beta / acum / epsilon / betasqn within 1e-9 relative per sweep,
1e-8 over whole chains; window 1 equals the sequential device sweep with
blocks of one marker bit for bit.
"""
import numpy as np
import pytest

import bayesrrcmd_amd as brr
import _oracle as orc
from _data import make_problem

pytestmark = pytest.mark.gpu

S = np.array([0.01, 0.001, 0.0001])
RTOL = 1e-9

# (oracle layout, device layout, ragged storage)
CASES = [(orc.DENSE, brr.LAYOUT_DENSE_2BIT, brr.STORAGE_AUTO),
         (orc.RAGGED, brr.LAYOUT_RAGGED, brr.STORAGE_LISTS),
         (orc.RAGGED, brr.LAYOUT_RAGGED, brr.STORAGE_CODES),
         (orc.EIGEN, brr.LAYOUT_EIGEN_SPARSE, brr.STORAGE_AUTO)]
IDS = ["dense", "ragged-lists", "ragged-codes", "eigen"]


def close(a, b, rtol=RTOL, what=""):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(1.0, float(np.max(np.abs(b))) if b.size else 1.0)
    err = float(np.max(np.abs(a - b))) if a.size else 0.0
    assert err <= rtol * scale, f"{what}: max abs err {err:.3e} (scale {scale:.3e})"


def make_pair(pr, olay, dlay, storage, window, workgroups=0, seed=5):
    if olay == orc.DENSE:
        o = orc.OracleProblem(pr.n, pr.m, orc.DENSE, 4, 1, S, bed=pr.bed, lut=pr.lut)
    else:
        arena = brr.build_eigen(pr.bed, pr.n) if olay == orc.EIGEN else pr.ragged
        o = orc.OracleProblem(pr.n, pr.m, olay, 4, 1, S, ragged=arena)
    o.init(pr.y, seed)
    h = brr.HipSweep(pr.n, pr.m, 4, 1, dlay, workgroups=workgroups, ragged_storage=storage)
    if dlay == brr.LAYOUT_DENSE_2BIT:
        h.upload_dense(pr.bed, pr.lut)
    else:
        h.upload_ragged(o.ragged)
    h.set_model(S)
    h.set_state(o.eps, o.beta, o.comp)
    h.set_scalars(o.mu, o.sigmaE, o.epsSum, o.sigmaG, o.pi)
    if window is not None:
        h.set_async(window)
    return o, h


def compare_state(o, h, olay, what):
    st = h.get_state()
    sc = h.get_scalars()
    bad = np.flatnonzero(st["comp"] != o.comp)
    assert bad.size == 0, f"{what}: components differ at {bad[:10]}"
    close(st["beta"], o.beta, what=f"{what} beta")
    close(st["acum"], o.acum, what=f"{what} acum")
    close(st["eps"], o.eps, what=f"{what} eps")
    assert np.array_equal(sc["v"].reshape(-1), o.v), f"{what}: v"
    close(sc["betasqn"], o.betasqn, what=f"{what} betasqn")
    if olay != orc.DENSE:
        close([sc["epsilon_sum"]], [o.epsSum], rtol=1e-8, what=f"{what} epsSum")


def inputs(m, it, seed=77):
    order = np.arange(m, dtype=np.uint32)
    for t in range(it + 1):
        brr.shuffle(seed, t, 0, order)
    u, z = brr.marker_uz(seed, it, 0, m)
    return order, u, z


@pytest.mark.parametrize("olay,dlay,storage", CASES, ids=IDS)
@pytest.mark.parametrize("window,workgroups", [(7, 0), (20, 13), (64, 0)])
def test_async_sweep_matches_oracle(gpu_lib, olay, dlay, storage, window, workgroups):
    pr = make_problem(3001, 500, seed=501, h2=0.6, causal_frac=0.05, miss=0.02)
    o, h = make_pair(pr, olay, dlay, storage, window, workgroups=workgroups)
    for it in range(3):
        order, u, z = inputs(pr.m, it)
        o.sweep_async(order, u, z, window)
        h.sweep(order, u, z)
        compare_state(o, h, olay, f"window {window} iteration {it}")
        s = h.stats()
        assert s["rounds"] == (pr.m + window - 1) // window      # one exchange round per window
        assert s["block_markers"] == window
    h.close()


@pytest.mark.parametrize("olay,dlay,storage", CASES, ids=IDS)
def test_window_one_equals_sequential_blocks_of_one(gpu_lib, olay, dlay, storage):
    pr = make_problem(2001, 200, seed=502, h2=0.6, causal_frac=0.05, miss=0.02)
    _, a = make_pair(pr, olay, dlay, storage, 1)
    o, b = make_pair(pr, olay, dlay, storage, None)
    b.close()
    b = brr.HipSweep(pr.n, pr.m, 4, 1, dlay, block_markers=1, ragged_storage=storage)
    if dlay == brr.LAYOUT_DENSE_2BIT:
        b.upload_dense(pr.bed, pr.lut)
    else:
        b.upload_ragged(o.ragged)
    b.set_model(S)
    b.set_state(o.eps, o.beta, o.comp)
    b.set_scalars(o.mu, o.sigmaE, o.epsSum, o.sigmaG, o.pi)
    for it in range(2):
        order, u, z = inputs(pr.m, it)
        a.sweep(order, u, z)
        b.sweep(order, u, z)
    sa, sb = a.get_state(), b.get_state()
    for k in ("eps", "beta", "acum"):
        assert np.array_equal(sa[k].view(np.uint64), sb[k].view(np.uint64)), k
    assert np.array_equal(sa["comp"], sb["comp"])
    a.close()
    b.close()


@pytest.mark.parametrize("olay,dlay,storage", [CASES[0], CASES[2]], ids=["dense", "ragged-codes"])
def test_async_chain_matches_oracle(gpu_lib, olay, dlay, storage):
    pr = make_problem(2500, 400, seed=503, h2=0.5, causal_frac=0.03)
    seed, iters, window = 20261202, 5, 20
    if olay == orc.DENSE:
        o = orc.OracleProblem(pr.n, pr.m, orc.DENSE, 4, 1, S, bed=pr.bed, lut=pr.lut)
    else:
        o = orc.OracleProblem(pr.n, pr.m, olay, 4, 1, S, ragged=pr.ragged)
    hist_o, order_o = o.run_gibbs(pr.y, seed, iters, window=window)
    eps, sc = brr.bayesr_init(pr.y, 1, 4, seed)
    orders = np.empty((iters, pr.m), dtype=np.uint32)
    cur = np.arange(pr.m, dtype=np.uint32)
    for it in range(iters):
        brr.shuffle(seed, it, 0, cur)
        orders[it] = cur
    assert np.array_equal(orders[-1], order_o)
    h = brr.HipSweep(pr.n, pr.m, 4, 1, dlay, ragged_storage=storage)
    h.upload_dense(pr.bed, pr.lut) if olay == orc.DENSE else h.upload_ragged(pr.ragged)
    h.set_model(S)
    h.set_state(eps, np.zeros(pr.m), np.zeros(pr.m, np.int32))
    h.set_scalars(sc["mu"], sc["sigmaE"], sc["epsilon_sum"], sc["sigmaG"], sc["pi"])
    h.set_async(window)
    hist_h = h.run_chain(seed, 0, 0, orders)
    for c, what in ((0, "mu"), (1, "sigmaE"), (3, "|eps|^2"), (4, "sigmaG")):
        close(hist_h[:, c], hist_o[:, c], rtol=1e-8, what=what)
    assert np.array_equal(hist_h[:, 2], hist_o[:, 2]), "m0"
    st = h.get_state()
    assert np.array_equal(st["comp"], o.comp)
    close(st["beta"], o.beta, rtol=1e-8, what="beta")
    close(st["eps"], o.eps, rtol=1e-8, what="eps")
    # back to the sequential sweep on the same context
    h.set_async(0)
    order, u, z = inputs(pr.m, 9)
    o.sweep(order, u, z, trace=False)
    h.sweep(order, u, z)
    compare_state(o, h, olay, "sequential after asynchronous")
    h.close()


def test_async_bad_calls(gpu_lib):
    pr = make_problem(500, 40, seed=504)
    h = brr.HipSweep(pr.n, pr.m, 4, 1, brr.LAYOUT_DENSE_2BIT)
    with pytest.raises(brr.BrrHipError):
        h.set_async(65)
    h.close()
    t = brr.HipSweep(pr.n, pr.m, 4, 1, brr.LAYOUT_DENSE_2BIT, traits=3)
    with pytest.raises(brr.BrrHipError):
        t.set_async(8)
    t.close()
    w = brr.HipBayesW(pr.n, pr.m, 4)
    assert brr.lib().brr_hip_set_async(w.ctx, 8) != 0          # BayesW has no asynchronous graph here
    w.close()
