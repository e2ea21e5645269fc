"""Uneven cohorts side by side (brr_hip_run_cohorts): several independent
single-chain contexts in persistent launches captured into one hipGraph.

Each cohort is the reference's own runBayesRAnalysis on its data set
(analysisrunner.cpp:123-152), here one context each. Bars: every cohort's chain
equals the oracle's runGibbs for that cohort within the FP64 chain bar (1e-8,
identical components and model sizes), and is bit for bit the chain its
context runs alone (brr_hip_run_chain: the same kernel body on the same
inputs); the packing into launches of at most one workgroup per compute unit is
reported; cohorts of different lane runs share a launch; FP32-residual cohorts
equal their own FP32 chains bit for bit.
"""
import numpy as np
import pytest

import bayesrrcmd_amd as brr
import _oracle as orc
from _data import make_problem

pytestmark = pytest.mark.gpu

S = np.array([0.01, 0.001, 0.0001])
SEED = 20261202

# (N, M, layout): uneven sizes, dense and ragged (held as 2-bit codes)
COHORTS = [(300, 50, orc.DENSE), (2_001, 400, orc.RAGGED), (5_003, 120, orc.DENSE), (12_000, 300, orc.RAGGED),
           (20_011, 250, orc.DENSE), (700, 333, orc.RAGGED)]


def build(pr, layout, q, precision=brr.PRECISION_FP64, workgroups=0):
    code = brr.LAYOUT_DENSE_2BIT if layout == orc.DENSE else brr.LAYOUT_RAGGED
    h = brr.HipSweep(pr.n, pr.m, 4, 1, code, workgroups=workgroups, ragged_storage=brr.STORAGE_CODES)
    h.upload_dense(pr.bed, pr.lut) if layout == orc.DENSE else h.upload_ragged(pr.ragged)
    h.set_model(S)
    h.set_precision(precision, 1024)
    eps, sc = brr.bayesr_init(pr.y, 1, 4, SEED, chain=q)
    h.set_state(eps, np.zeros(pr.m), np.zeros(pr.m, np.int32))
    h.set_scalars(sc["mu"], sc["sigmaE"], sc["epsilon_sum"], sc["sigmaG"], sc["pi"])
    return h


def orders_for(m, q, iters):
    out = np.empty((iters, m), dtype=np.uint32)
    cur = np.arange(m, dtype=np.uint32)
    for it in range(iters):
        brr.shuffle(SEED, it, q, cur)
        out[it] = cur
    return out


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b)) / np.maximum(1.0, np.abs(np.asarray(b)))))


def test_cohorts_match_the_oracle_and_their_own_chains(gpu_lib):
    iters = 4
    probs = [make_problem(n, m, seed=600 + q, causal_frac=0.03) for q, (n, m, _) in enumerate(COHORTS)]
    hs = [build(pr, lay, q) for q, (pr, (_, _, lay)) in enumerate(zip(probs, COHORTS))]
    orders = [orders_for(pr.m, q, iters) for q, pr in enumerate(probs)]
    hists, rep = brr.run_cohorts(hs, SEED, 0, orders, chains=list(range(len(hs))))
    assert rep["launches"] == 1 and rep["device_ms"] > 0
    for q, (pr, (_, _, lay), h) in enumerate(zip(probs, COHORTS, hs)):
        if lay == orc.DENSE:
            o = orc.OracleProblem(pr.n, pr.m, orc.DENSE, 4, 1, S, bed=pr.bed, lut=pr.lut)
        else:
            o = orc.OracleProblem(pr.n, pr.m, orc.RAGGED, 4, 1, S, ragged=pr.ragged)
        hist_o, order_o = o.run_gibbs(pr.y, SEED, iters, chain=q)
        assert np.array_equal(order_o, orders[q][-1])
        hist = hists[q]
        for col, what in ((0, "mu"), (1, "sigmaE"), (3, "|eps|^2"), (4, "sigmaG")):
            assert rel(hist[:, col], hist_o[:, col]) <= 1e-8, (q, what)
        assert np.array_equal(hist[:, 2], hist_o[:, 2]), (q, "m0")
        st = h.get_state()
        assert np.array_equal(st["comp"], o.comp), q
        assert rel(st["beta"], o.beta) <= 1e-8 and rel(st["eps"], o.eps) <= 1e-8, q
        s = h.stats()
        assert s["markers"] == iters * pr.m and s["rounds"] > 0
        # the same chain by the context alone
        alone = build(pr, lay, q)
        hist_a = alone.run_chain(SEED, q, 0, orders[q])
        sa = alone.get_state()
        assert np.array_equal(hist_a, hist), q
        for k in ("eps", "beta", "comp", "acum"):
            assert np.array_equal(sa[k], st[k]), (q, k)
        alone.close()
    for h in hs:
        h.close()


def test_cohorts_pack_into_several_launches(gpu_lib):
    """Three cohorts of 100 workgroups each cannot share one launch of <= 256
    workgroups: two launches in one graph, same results as alone."""
    iters = 2
    probs = [make_problem(40_000 + 17 * q, 150, seed=620 + q, causal_frac=0.03) for q in range(3)]
    hs = [build(pr, orc.DENSE, q, workgroups=100) for q, pr in enumerate(probs)]
    P = [h.stats()["workgroups"] for h in hs]
    assert all(90 <= x <= 100 for x in P) and sum(P) > 256, P
    orders = [orders_for(pr.m, q, iters) for q, pr in enumerate(probs)]
    hists, rep = brr.run_cohorts(hs, SEED, 0, orders, chains=[0, 1, 2])
    assert rep["launches"] == 2, rep
    for q, (pr, h) in enumerate(zip(probs, hs)):
        alone = build(pr, orc.DENSE, q, workgroups=100)
        assert np.array_equal(alone.run_chain(SEED, q, 0, orders[q]), hists[q]), q
        assert np.array_equal(alone.get_state()["eps"], h.get_state()["eps"]), q
        alone.close()
        h.close()


def test_fp32_cohorts_equal_their_own_fp32_chains(gpu_lib):
    iters = 3
    sizes = [(3_001, 200, orc.RAGGED), (9_000, 150, orc.DENSE), (1_200, 90, orc.RAGGED)]
    probs = [make_problem(n, m, seed=640 + q, causal_frac=0.03) for q, (n, m, _) in enumerate(sizes)]
    hs = [build(pr, lay, q, brr.PRECISION_FP32_REFRESH) for q, (pr, (_, _, lay)) in enumerate(zip(probs, sizes))]
    orders = [orders_for(pr.m, q, iters) for q, pr in enumerate(probs)]
    hists, rep = brr.run_cohorts(hs, SEED, 0, orders, chains=[0, 1, 2])
    for q, (pr, (_, _, lay), h) in enumerate(zip(probs, sizes, hs)):
        alone = build(pr, lay, q, brr.PRECISION_FP32_REFRESH)
        assert np.array_equal(alone.run_chain(SEED, q, 0, orders[q]), hists[q]), q
        assert np.array_equal(alone.get_state()["eps"], h.get_state()["eps"]), q
        alone.close()
        h.close()


def test_cohort_arguments(gpu_lib):
    pr = make_problem(3_000, 60, seed=660)
    a = build(pr, orc.DENSE, 0)
    b = build(pr, orc.DENSE, 1, precision=brr.PRECISION_FP32_REFRESH)
    orders = [orders_for(pr.m, 0, 1), orders_for(pr.m, 1, 1)]
    with pytest.raises(brr.BrrHipError, match="precision"):
        brr.run_cohorts([a, b], SEED, 0, orders)
    bad = orders_for(pr.m, 0, 1)
    bad[0, 1] = bad[0, 0]
    with pytest.raises(brr.BrrHipError, match="permutation"):
        brr.run_cohorts([a], SEED, 0, [bad])
    for h in (a, b):
        h.close()


def test_cohorts_with_different_lane_runs(gpu_lib):
    """A cohort of two lane runs per slice beside one-run cohorts: the launch runs
    the two-run kernel, the short slices masked at their ends; each chain equals
    its context's own. A context of four runs is refused."""
    iters = 2
    probs = [make_problem(6_000, 120, seed=670, causal_frac=0.03), make_problem(900, 80, seed=671),
             make_problem(2_500, 100, seed=672)]
    hs = [build(probs[0], orc.DENSE, 0, workgroups=6), build(probs[1], orc.RAGGED, 1),
          build(probs[2], orc.DENSE, 2)]
    assert hs[0].stats()["slice"] == 1008 and hs[1].stats()["slice"] <= 512
    orders = [orders_for(pr.m, q, iters) for q, pr in enumerate(probs)]
    hists, rep = brr.run_cohorts(hs, SEED, 0, orders, chains=[0, 1, 2])
    assert rep["launches"] == 1
    for q, (pr, h) in enumerate(zip(probs, hs)):
        lay = orc.DENSE if q != 1 else orc.RAGGED
        alone = build(pr, lay, q, workgroups=6 if q == 0 else 0)
        assert np.array_equal(alone.run_chain(SEED, q, 0, orders[q]), hists[q]), q
        assert np.array_equal(alone.get_state()["eps"], h.get_state()["eps"]), q
        alone.close()
    four = build(probs[0], orc.DENSE, 0, workgroups=3)
    assert four.stats()["slice"] == 2000
    with pytest.raises(brr.BrrHipError, match="1024 individuals"):
        brr.run_cohorts([four, hs[1]], SEED, 0, [orders[0], orders[1]])
    four.close()
    for h in hs:
        h.close()


def test_cohorts_direct_launches_equal_the_graph(gpu_lib):
    """The same cohorts launched one after another without a graph: identical chains."""
    iters = 2
    probs = [make_problem(n, m, seed=680 + q, causal_frac=0.03) for q, (n, m) in enumerate([(4000, 90), (700, 60)])]
    runs = []
    for graph in (True, False):
        hs = [build(pr, orc.DENSE, q) for q, pr in enumerate(probs)]
        orders = [orders_for(pr.m, q, iters) for q, pr in enumerate(probs)]
        hists, _ = brr.run_cohorts(hs, SEED, 0, orders, chains=[0, 1], graph=graph)
        runs.append((hists, [h.get_state()["eps"] for h in hs]))
        for h in hs:
            h.close()
    for q in range(2):
        assert np.array_equal(runs[0][0][q], runs[1][0][q]) and np.array_equal(runs[0][1][q], runs[1][1][q]), q
