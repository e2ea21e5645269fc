"""The FP32-residual sweep (BRR_HIP_PRECISION_FP32_REFRESH) against the FP64 oracle.

The device holds each residual slice in FP32 and forms x_j . eps in FP32 with
compensated (TwoSum) wave reductions; the residual changes are replayed in FP64
on the context's FP64 residual every `refresh` markers and at every iteration
boundary, and the FP32 slices are re-rounded from it. The reference computes
everything in FP64 (BayesRBase.cpp:287-392), so the oracle is FP64 and the bars
are those of an FP32 residual, not of reduction order. Their size: replaying the
oracle's own first sweep at N = 3,001, M = 700 (232 changes, no refresh inside
the sweep) with the residual rounded to FP32 after every update and the same
decisions gives num errors up to 2.8e-4 against a largest |num| of 1,012
(2.8e-7 of it), from the drift of the FP32 residual alone. Hence:

- component choices and v counts identical (a flip needs u within ~1e-6 of an
  acum boundary);
- num (computeNum) within 1e-6 of the largest |num|, and above FP64 rounding
  (1e-12 of it), which shows the FP32 kernel ran; a refresh every 8 markers
  cuts the num error at least threefold against one per 1,024;
- beta within 2e-6, acum within 1e-5 (1e-4 at N = 100,000), the FP64 residual within 1e-5 of the
  largest |eps|; chain scalars within 1e-5 relative and h2 within 1e-5;
- the returned residual equals y_std - mu - X beta within 1e-9 of its scale
  after whole chains: the FP64 replay, not FP32 accumulation, carries it (an
  FP32-accumulated residual is off by ~1e-6).
"""
import numpy as np
import pytest

import bayesrrcmd_amd as brr
import _oracle as orc
from _data import make_problem

pytestmark = pytest.mark.gpu

S = np.array([0.01, 0.001, 0.0001])


def err_rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(1.0, float(np.max(np.abs(b))) if b.size else 1.0)
    return float(np.max(np.abs(a - b))) / scale if a.size else 0.0


def make_pair(pr, layout, workgroups=0, block=0, refresh=1024, seed=5):
    if layout == orc.DENSE:
        o = orc.OracleProblem(pr.n, pr.m, orc.DENSE, 4, 1, S, bed=pr.bed, lut=pr.lut)
        h = brr.HipSweep(pr.n, pr.m, 4, 1, brr.LAYOUT_DENSE_2BIT, workgroups=workgroups, block_markers=block)
        h.upload_dense(pr.bed, pr.lut)
    else:
        o = orc.OracleProblem(pr.n, pr.m, orc.RAGGED, 4, 1, S, ragged=pr.ragged)
        h = brr.HipSweep(pr.n, pr.m, 4, 1, brr.LAYOUT_RAGGED, workgroups=workgroups, block_markers=block,
                         ragged_storage=brr.STORAGE_CODES)
        h.upload_ragged(pr.ragged)
    o.init(pr.y, seed)
    h.set_model(S)
    h.set_precision(brr.PRECISION_FP32_REFRESH, refresh)
    h.set_state(o.eps, o.beta, o.comp)
    h.set_scalars(o.mu, o.sigmaE, o.epsSum, o.sigmaG, o.pi)
    h.record_num(True)
    return o, h


def sweep_inputs(m, it, seed=77):
    order = np.arange(m, dtype=np.uint32)
    for t in range(it + 1):
        brr.shuffle(seed, t, 0, order)
    u, z = brr.marker_uz(seed, it, 0, m)
    return order, u, z


def check_sweep(o, h, order, u, z, what, acum_bar=1e-5):
    tr = o.sweep(order, u, z, trace=True)
    h.sweep(order, u, z)
    st = h.get_state()
    sc = h.get_scalars()
    assert np.array_equal(st["comp"], o.comp), f"{what}: components differ at {np.flatnonzero(st['comp'] != o.comp)[:8]}"
    assert np.array_equal(sc["v"].reshape(-1), o.v), f"{what}: v"
    num = h.get_num()[order]
    e_num = err_rel(num, tr["num"])
    assert e_num <= 1e-6, f"{what}: num {e_num:.3e}"
    assert e_num > 1e-12, f"{what}: num error {e_num:.3e} is at FP64 level: the FP32 kernel did not run"
    assert err_rel(st["beta"], o.beta) <= 2e-6, f"{what}: beta"
    assert err_rel(st["acum"], o.acum) <= acum_bar, f"{what}: acum"
    assert err_rel(st["eps"], o.eps) <= 1e-5, f"{what}: eps"
    assert err_rel(sc["betasqn"], o.betasqn) <= 1e-5, f"{what}: betasqn"
    if o.layout == orc.RAGGED:
        assert err_rel([sc["epsilon_sum"]], [o.epsSum]) <= 1e-5, f"{what}: epsilon_sum"
    return e_num


@pytest.mark.parametrize("layout", [orc.DENSE, orc.RAGGED], ids=["dense", "ragged-codes"])
@pytest.mark.parametrize("workgroups,block,refresh", [(0, 0, 1024), (17, 7, 64), (0, 13, 1), (64, 64, 100),
                                                      (3, 1, 5)])
def test_fp32_sweep_matches_fp64_oracle(gpu_lib, layout, workgroups, block, refresh):
    pr = make_problem(3001, 700, seed=201, h2=0.5, causal_frac=0.02)
    o, h = make_pair(pr, layout, workgroups, block, refresh)
    for it in range(3):
        order, u, z = sweep_inputs(pr.m, it)
        check_sweep(o, h, order, u, z, f"iteration {it}")
    h.close()


def test_fp32_refresh_interval_bounds_the_drift(gpu_lib):
    """The same sweep with a refresh every 8 markers and every 1,024: the FP32
    residual drifts only between refreshes, so the short interval has a clearly
    smaller num error (the FP64 replay is what resets it)."""
    pr = make_problem(3001, 700, seed=201, h2=0.5, causal_frac=0.02)
    errs = {}
    for refresh in (8, 1024):
        o, h = make_pair(pr, orc.RAGGED, block=8, refresh=refresh)
        order, u, z = sweep_inputs(pr.m, 0)
        errs[refresh] = check_sweep(o, h, order, u, z, f"refresh {refresh}")
        h.close()
    assert errs[8] * 3 < errs[1024], errs


def test_fp32_sweep_with_missing_calls(gpu_lib):
    pr = make_problem(2503, 400, seed=203, miss=0.05, causal_frac=0.05)
    for layout in (orc.DENSE, orc.RAGGED):
        o, h = make_pair(pr, layout, block=32, refresh=256)
        for it in range(2):
            order, u, z = sweep_inputs(pr.m, it, seed=9)
            check_sweep(o, h, order, u, z, f"layout {layout} iteration {it}")
        h.close()


def test_fp64_num_record_matches_oracle(gpu_lib):
    """The per-marker num record of the FP64 sweep equals the oracle's computeNum
    trace to FP64 reduction order (1e-11 of the largest |num|)."""
    pr = make_problem(3001, 500, seed=205, causal_frac=0.03)
    for layout in (orc.DENSE, orc.RAGGED):
        o, h = make_pair(pr, layout)
        h.set_precision(brr.PRECISION_FP64, 0)
        for it in range(2):
            order, u, z = sweep_inputs(pr.m, it, seed=3)
            tr = o.sweep(order, u, z, trace=True)
            h.sweep(order, u, z)
            assert err_rel(h.get_num()[order], tr["num"]) <= 1e-11
            assert np.array_equal(h.get_state()["comp"], o.comp)
        h.close()


def xbeta_ragged(pr, beta):
    r = pr.ragged
    xb = np.zeros(pr.n)
    for j in np.flatnonzero(beta):
        mean, sd = r.stats[j, 0], r.stats[j, 1]
        col = np.full(pr.n, -mean / sd)
        col[r.ones[r.ones_off[j]: r.ones_off[j + 1]]] += 1 / sd
        col[r.twos[r.twos_off[j]: r.twos_off[j + 1]]] += 2 / sd
        col[r.miss[r.miss_off[j]: r.miss_off[j + 1]]] += mean / sd
        xb += beta[j] * col
    return xb


@pytest.mark.parametrize("layout", [orc.DENSE, orc.RAGGED], ids=["dense", "ragged-codes"])
def test_fp32_chain_matches_fp64_oracle_chain(gpu_lib, layout):
    pr = make_problem(2500, 500, seed=213, h2=0.5, causal_frac=0.02)
    seed, iters = 20261202, 6
    if layout == orc.DENSE:
        o = orc.OracleProblem(pr.n, pr.m, layout, 4, 1, S, bed=pr.bed, lut=pr.lut)
    else:
        o = orc.OracleProblem(pr.n, pr.m, layout, 4, 1, S, ragged=pr.ragged)
    hist_o, order_o = o.run_gibbs(pr.y, seed, iters)
    eps, sc = brr.bayesr_init(pr.y, 1, 4, seed)
    orders = np.empty((iters, pr.m), dtype=np.uint32)
    cur = np.arange(pr.m, dtype=np.uint32)
    for it in range(iters):
        brr.shuffle(seed, it, 0, cur)
        orders[it] = cur
    if layout == orc.DENSE:
        h = brr.HipSweep(pr.n, pr.m, 4, 1, brr.LAYOUT_DENSE_2BIT)
        h.upload_dense(pr.bed, pr.lut)
    else:
        h = brr.HipSweep(pr.n, pr.m, 4, 1, brr.LAYOUT_RAGGED, ragged_storage=brr.STORAGE_CODES)
        h.upload_ragged(pr.ragged)
    h.set_model(S)
    h.set_precision(brr.PRECISION_FP32_REFRESH, 1024)
    h.set_state(eps, np.zeros(pr.m), np.zeros(pr.m, np.int32))
    h.set_scalars(sc["mu"], sc["sigmaE"], sc["epsilon_sum"], sc["sigmaG"], sc["pi"])
    hist_h = h.run_chain(seed, 0, 0, orders)
    for col, what in ((0, "mu"), (1, "sigmaE"), (3, "|eps|^2"), (4, "sigmaG")):
        rel = np.max(np.abs(hist_h[:, col] - hist_o[:, col]) / np.abs(hist_o[:, col]))
        assert rel <= 1e-5, f"{what}: {rel:.3e}"
    assert np.array_equal(hist_h[:, 2], hist_o[:, 2]), "m0"
    h2_o = hist_o[:, 4] / (hist_o[:, 4] + hist_o[:, 1])
    h2_h = hist_h[:, 4] / (hist_h[:, 4] + hist_h[:, 1])
    assert np.max(np.abs(h2_h - h2_o)) <= 1e-5, "h2"
    st = h.get_state()
    assert np.array_equal(st["comp"], o.comp)
    assert err_rel(st["beta"], o.beta) <= 2e-6
    assert err_rel(st["eps"], o.eps) <= 1e-5
    h.close()


def test_fp32_c2_shape_against_oracle(gpu_lib):
    """N = 100,000 individuals on ragged markers held as 2-bit codes (the full-scale
    plain workload's slice shape: 196 workgroups of 512), M = 2,000.
    - One FP32 sweep against the oracle's sweep (the bars above; acum within 1e-4:
      at this N a marker's acum moves ~0.015 per unit of num).
    - A 4-iteration FP64 device chain against the oracle's chain within 1e-8 (the
      FP64 parity bar at this scale), identical components.
    - The same chain in FP32: identical to the FP64 chain in m0 and within 1e-5 in
      h2 up to the first iteration where the model sizes part (a component can
      flip when u lies within the FP32 error of an acum boundary), and its
      residual equal to y_std - mu - X beta within 1e-9."""
    pr = make_problem(100_000, 2000, seed=215, maf=(0.005, 0.5), h2=0.4, causal_frac=0.005)
    o, h = make_pair(pr, orc.RAGGED, seed=11)
    assert h.stats()["workgroups"] == 196 and h.stats()["slice"] == 512
    order, u, z = sweep_inputs(pr.m, 0, seed=11)
    check_sweep(o, h, order, u, z, "C2 sweep", acum_bar=1e-4)
    h.close()
    seed, iters = 11, 4
    o2 = orc.OracleProblem(pr.n, pr.m, orc.RAGGED, 4, 1, S, ragged=pr.ragged)
    hist_o, _ = o2.run_gibbs(pr.y, seed, iters)
    eps0, sc = brr.bayesr_init(pr.y, 1, 4, seed)
    orders = np.empty((iters, pr.m), dtype=np.uint32)
    cur = np.arange(pr.m, dtype=np.uint32)
    for it in range(iters):
        brr.shuffle(seed, it, 0, cur)
        orders[it] = cur
    h = brr.HipSweep(pr.n, pr.m, 4, 1, brr.LAYOUT_RAGGED, ragged_storage=brr.STORAGE_CODES)
    h.upload_ragged(pr.ragged)
    h.set_model(S)
    runs = {}
    for prec in (brr.PRECISION_FP64, brr.PRECISION_FP32_REFRESH):
        h.set_precision(prec, 1024)
        h.set_state(eps0, np.zeros(pr.m), np.zeros(pr.m, np.int32))
        h.set_scalars(sc["mu"], sc["sigmaE"], sc["epsilon_sum"], sc["sigmaG"], sc["pi"])
        hist = h.run_chain(seed, 0, 0, orders)
        runs[prec] = (hist, h.get_state(), h.get_scalars()["mu"])
    h64, st64, _ = runs[brr.PRECISION_FP64]
    for col, what in ((0, "mu"), (1, "sigmaE"), (3, "|eps|^2"), (4, "sigmaG")):
        rel = np.max(np.abs(h64[:, col] - hist_o[:, col]) / np.abs(hist_o[:, col]))
        assert rel <= 1e-8, f"FP64 chain {what}: {rel:.3e}"
    assert np.array_equal(h64[:, 2], hist_o[:, 2]) and np.array_equal(st64["comp"], o2.comp)
    assert err_rel(st64["beta"], o2.beta) <= 1e-8 and err_rel(st64["eps"], o2.eps) <= 1e-8
    h32, st32, mu32 = runs[brr.PRECISION_FP32_REFRESH]
    parted = np.flatnonzero(h32[:, 2] != h64[:, 2])
    upto = int(parted[0]) if parted.size else iters
    h2_32 = h32[:, 4] / (h32[:, 4] + h32[:, 1])
    h2_64 = h64[:, 4] / (h64[:, 4] + h64[:, 1])
    if upto:
        assert np.max(np.abs(h2_32[:upto] - h2_64[:upto])) <= 1e-5, "h2"
    resid = eps0 - mu32 - xbeta_ragged(pr, st32["beta"])
    assert err_rel(st32["eps"], resid) <= 1e-9, "residual identity"
    h.close()


def test_fp32_mode_arguments(gpu_lib):
    pr = make_problem(600, 20, seed=217)
    h = brr.HipSweep(pr.n, pr.m, 4, 1, brr.LAYOUT_RAGGED, ragged_storage=brr.STORAGE_LISTS)
    h.upload_ragged(pr.ragged)
    h.set_model(S)
    for bad in (0, 1025):
        with pytest.raises(brr.BrrHipError):
            h.set_precision(brr.PRECISION_FP32_REFRESH, bad)
    h.set_precision(brr.PRECISION_FP32_REFRESH, 64)
    order, u, z = sweep_inputs(pr.m, 0)
    with pytest.raises(brr.BrrHipError, match="2-bit codes"):
        h.sweep(order, u, z)            # index-list storage is refused
    with pytest.raises(brr.BrrHipError):
        h.set_async(8)                  # the asynchronous schedule stays FP64
    h.set_precision(brr.PRECISION_FP64, 0)
    h.sweep(order, u, z)
    h.close()
    hb = brr.HipSweep(pr.n, pr.m, 4, 1, brr.LAYOUT_DENSE_2BIT, traits=4)
    with pytest.raises(brr.BrrHipError):
        hb.set_precision(brr.PRECISION_FP32_REFRESH, 1024)
    hb.close()
