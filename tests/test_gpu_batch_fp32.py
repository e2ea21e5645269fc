"""Batched traits with FP32 residual tiles (BRR_HIP_PRECISION_FP32) against the
FP64 oracle, per trait.

The residual tiles are FP32 on the f32 matrix-core instruction; each
16-individual tile's FP32 dot and Gram products are summed over the tiles in
FP64; decisions and draws are FP64; the residual is not refreshed. The bars are
those of an FP32 residual (as for the single-chain FP32 mode): identical
components and v counts; beta within 2e-6, acum within 1e-5, the residual and
epsilonSum within 1e-5 of their scale; chain scalars within 1e-5 relative; and,
at N = 100,000 (slices of 400 individuals, 25 FP32 tiles: beyond the FP64
kernels' 256), the residual of every trait equal to y_std - mu - X beta within
1e-5 of its scale.
"""
import numpy as np
import pytest

import bayesrrcmd_amd as brr
import _oracle as orc
from _data import make_problem, spec_dosage

pytestmark = pytest.mark.gpu

S = np.array([0.01, 0.001, 0.0001])


def err_rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b))) / max(1.0, float(np.max(np.abs(b))))


def phenotypes(pr, T, seed):
    rng = np.random.default_rng(seed)
    d = spec_dosage(pr.bed, pr.n)
    x = np.take_along_axis(pr.lut, np.where(d < 0, 3, d), axis=1)
    ys = []
    for _ in range(T):
        h2 = rng.uniform(0.1, 0.8)
        mc = int(rng.integers(1, max(2, pr.m // 20)))
        causal = rng.choice(pr.m, mc, replace=False)
        b = rng.normal(0.0, np.sqrt(h2 / mc), mc)
        ys.append((b @ x[causal] + rng.normal(0.0, np.sqrt(1.0 - h2), pr.n)).astype(np.float32).astype(np.float64))
    return ys, x


def setup(pr, layout, T, seed, picked):
    code = brr.LAYOUT_DENSE_2BIT if layout == orc.DENSE else brr.LAYOUT_RAGGED
    ys, x = phenotypes(pr, T, seed + 1)
    h = brr.HipSweep(pr.n, pr.m, 4, 1, code, traits=T)
    h.upload_dense(pr.bed, pr.lut) if layout == orc.DENSE else h.upload_ragged(pr.ragged)
    h.set_model(S)
    h.set_precision(brr.PRECISION_FP32)
    eps0 = np.zeros((pr.n, T))
    os_ = {}
    for t, y in enumerate(ys):
        e, sc = brr.bayesr_init(y, 1, 4, seed, chain=t)
        eps0[:, t] = e
        h.set_scalars(sc["mu"], sc["sigmaE"], sc["epsilon_sum"], sc["sigmaG"], sc["pi"], trait=t)
        if t in picked:
            if layout == orc.DENSE:
                o = orc.OracleProblem(pr.n, pr.m, orc.DENSE, 4, 1, S, bed=pr.bed, lut=pr.lut)
            else:
                o = orc.OracleProblem(pr.n, pr.m, orc.RAGGED, 4, 1, S, ragged=pr.ragged)
            o.init(y, seed, chain=t)
            os_[t] = o
    h.set_state(eps0, np.zeros((pr.m, T)), np.zeros((pr.m, T), np.int32))
    return h, os_, eps0, x


def sweeps(h, os_, pr, T, seed, iters, layout):
    order = np.arange(pr.m, dtype=np.uint32)
    for it in range(iters):
        brr.shuffle(seed, it, 0, order)
        uz = [brr.marker_uz(seed, it, t, pr.m) for t in range(T)]
        u = np.stack([a for a, _ in uz], axis=1)
        z = np.stack([b for _, b in uz], axis=1)
        for t, o in os_.items():
            o.sweep(order, u[:, t], z[:, t], trace=False)
        h.sweep(order, u, z)
        st = h.get_state()
        for t, o in os_.items():
            w = f"iteration {it} trait {t}"
            assert np.array_equal(st["comp"][:, t], o.comp), w
            assert err_rel(st["beta"][:, t], o.beta) <= 2e-6, w
            assert err_rel(st["acum"][:, t], o.acum) <= 1e-5, w
            assert err_rel(st["eps"][:, t], o.eps) <= 1e-5, w
            sc = h.get_scalars(trait=t)
            assert np.array_equal(sc["v"].reshape(-1), o.v), w
            if layout == orc.RAGGED:
                assert err_rel([sc["epsilon_sum"]], [o.epsSum]) <= 1e-5, w
    return st


@pytest.mark.parametrize("layout", [orc.DENSE, orc.RAGGED], ids=["dense", "ragged"])
@pytest.mark.parametrize("T", [7, 64])
def test_fp32_batch_sweeps_match_oracle(gpu_lib, layout, T):
    pr = make_problem(2001, 333, seed=701, h2=0.5, causal_frac=0.03)
    picked = tuple(range(T)) if T <= 7 else (0, 17, 40, 63)
    h, os_, _, _ = setup(pr, layout, T, 31, picked)
    sweeps(h, os_, pr, T, 31, 3, layout)
    h.close()


def test_fp32_batch_chain_matches_oracle(gpu_lib):
    pr = make_problem(1800, 260, seed=703, h2=0.5, causal_frac=0.03)
    seed, T, iters = 20261202, 6, 4
    h, os_, _, _ = setup(pr, orc.DENSE, T, seed, tuple(range(T)))
    orders = np.empty((iters, pr.m), dtype=np.uint32)
    cur = np.arange(pr.m, dtype=np.uint32)
    for it in range(iters):
        brr.shuffle(seed, it, 0, cur)
        orders[it] = cur
    hist_h = h.run_chain(seed, 0, 0, orders)
    u = np.zeros(pr.m)
    z = np.zeros(pr.m)
    for t, o in os_.items():
        for it in range(iters):
            o.update_mu(seed, it, chain=t)
            orc.lib().orc_fill_marker_uz(seed, it, t, pr.m, orc.p(u, orc.C.c_double), orc.p(z, orc.C.c_double))
            o.sweep(orders[it], u, z, trace=False)
            e2 = float(np.sum(o.eps * o.eps))
            m0 = float(o.v.sum() - o.v[0])
            mu = o.mu
            o.post_sweep(seed, it, chain=t)
            want = [mu, o.sigmaE, m0, e2, o.sigmaG[0]]
            got = hist_h[it, t]
            assert got[2] == want[2], (t, it, "m0")
            for col in (0, 1, 3, 4):
                assert abs(got[col] - want[col]) <= 1e-5 * max(1.0, abs(want[col])), (t, it, col)
    h.close()


def test_fp32_batch_n100k_slices_of_400(gpu_lib):
    """N = 100,000 and T = 64: slices of 400 individuals (25 FP32 tiles) on 250
    workgroups, past the FP64 kernels' 256; traits 0 and 63 against the oracle,
    and every trait's residual equal to y_std - mu - X beta."""
    pr = make_problem(100_000, 160, seed=705, h2=0.5, causal_frac=0.05)
    T, seed = 64, 41
    h, os_, eps0, x = setup(pr, orc.DENSE, T, seed, (0, 63))
    s = h.stats()
    assert s["slice"] == 400 and s["workgroups"] == 250, s
    st = sweeps(h, os_, pr, T, seed, 2, orc.DENSE)
    for t in range(T):
        mu = h.get_scalars(trait=t)["mu"]
        assert err_rel(st["eps"][:, t], eps0[:, t] - mu - st["beta"][:, t] @ x) <= 1e-5, t
    h.close()


def test_fp64_batch_refuses_wide_slices(gpu_lib):
    pr = make_problem(100_000, 20, seed=707)
    h = brr.HipSweep(pr.n, pr.m, 4, 1, brr.LAYOUT_DENSE_2BIT, traits=4)
    h.upload_dense(pr.bed, pr.lut)
    h.set_model(S)
    assert h.stats()["slice"] == 400
    order = np.arange(pr.m, dtype=np.uint32)
    u = np.full((pr.m, 4), 0.5)
    with pytest.raises(brr.BrrHipError, match="FP32"):
        h.sweep(order, u, np.zeros((pr.m, 4)))
    with pytest.raises(brr.BrrHipError):
        h.set_precision(brr.PRECISION_FP32_REFRESH, 1024)      # the single-chain mode
    h.close()
