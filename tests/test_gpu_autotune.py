"""Launch shape on a live context (brr_hip_set_launch) and the autotuner.

Bars: after a change of shape the sweep still matches the oracle at the parity
bars of the sweep tests (identical components, 1e-9 relative); the autotuner
leaves the chain state bit for bit as it found it, reports one entry per
candidate shape, and keeps the fastest one.
"""
import numpy as np
import pytest

import bayesrrcmd_amd as brr
import _oracle as orc
from _data import make_problem

pytestmark = pytest.mark.gpu

S = np.array([0.01, 0.001, 0.0001])


def close(a, b, rtol=1e-9, what=""):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(1.0, float(np.max(np.abs(b))))
    err = float(np.max(np.abs(a - b)))
    assert err <= rtol * scale, f"{what}: {err:.3e}"


def pair(pr, layout, storage):
    if layout == orc.DENSE:
        o = orc.OracleProblem(pr.n, pr.m, orc.DENSE, 4, 1, S, bed=pr.bed, lut=pr.lut)
        h = brr.HipSweep(pr.n, pr.m, 4, 1, brr.LAYOUT_DENSE_2BIT)
        h.upload_dense(pr.bed, pr.lut)
    else:
        o = orc.OracleProblem(pr.n, pr.m, orc.RAGGED, 4, 1, S, ragged=pr.ragged)
        h = brr.HipSweep(pr.n, pr.m, 4, 1, brr.LAYOUT_RAGGED, ragged_storage=storage)
        h.upload_ragged(pr.ragged)
    o.init(pr.y, 3)
    h.set_model(S)
    h.set_state(o.eps, o.beta, o.comp)
    h.set_scalars(o.mu, o.sigmaE, o.epsSum, o.sigmaG, o.pi)
    return o, h


def step(o, h, it, what):
    order = np.arange(o.m, dtype=np.uint32)
    brr.shuffle(11, it, 0, order)
    u, z = brr.marker_uz(11, it, 0, o.m)
    o.sweep(order, u, z, trace=False)
    h.sweep(order, u, z)
    st = h.get_state()
    assert np.array_equal(st["comp"], o.comp), what
    close(st["beta"], o.beta, what=f"{what} beta")
    close(st["eps"], o.eps, what=f"{what} eps")


@pytest.mark.parametrize("layout,storage", [(orc.DENSE, 0), (orc.RAGGED, brr.STORAGE_LISTS),
                                            (orc.RAGGED, brr.STORAGE_CODES)],
                         ids=["dense", "ragged-lists", "ragged-codes"])
def test_set_launch_keeps_parity(gpu_lib, layout, storage):
    pr = make_problem(9001, 500, seed=501, causal_frac=0.02)
    o, h = pair(pr, layout, storage)
    shapes = [(0, 0), (40, 16), (5, 64), (18, 7), (0, 64)]
    for it, (wg, blk) in enumerate(shapes):
        h.set_launch(wg, blk)
        s = h.stats()
        assert s["workgroups"] * s["slice"] >= pr.n and (wg == 0 or s["workgroups"] <= wg)
        if blk:
            assert s["block_markers"] == blk
        step(o, h, it, f"shape {wg}x{blk}")
    h.close()


def test_autotune_plain(gpu_lib):
    pr = make_problem(30_001, 800, seed=503, causal_frac=0.01, maf=(0.005, 0.5))
    o, h = pair(pr, orc.RAGGED, brr.STORAGE_CODES)
    step(o, h, 0, "before")
    st0, sc0 = h.get_state(), h.get_scalars()
    tab = h.autotune(seed=7, iterations=2)
    assert len(tab) >= 6, tab
    ran = [e for e in tab if e["ms_per_iteration"] > 0]
    assert len(ran) == len(tab), tab
    best = min(ran, key=lambda e: e["ms_per_iteration"])
    s = h.stats()
    assert (s["workgroups"], s["block_markers"]) == (best["workgroups"], best["block_markers"]), (s, best)
    st1, sc1 = h.get_state(), h.get_scalars()
    for k in ("eps", "beta", "comp", "acum"):
        assert np.array_equal(st0[k], st1[k]), k
    for k in ("mu", "sigmaE", "epsilon_sum"):
        assert sc0[k] == sc1[k], k
    step(o, h, 1, "after autotune")
    h.close()


def test_autotune_batched(gpu_lib):
    pr = make_problem(20_001, 200, seed=505, causal_frac=0.03)
    T = 8
    rng = np.random.default_rng(506)
    ys = [pr.y + rng.normal(0.0, 0.1, pr.n) for _ in range(T)]
    os_ = []
    h = brr.HipSweep(pr.n, pr.m, 4, 1, brr.LAYOUT_DENSE_2BIT, traits=T)
    h.upload_dense(pr.bed, pr.lut)
    h.set_model(S)
    for t, y in enumerate(ys):
        o = orc.OracleProblem(pr.n, pr.m, orc.DENSE, 4, 1, S, bed=pr.bed, lut=pr.lut)
        o.init(y, 9, chain=t)
        os_.append(o)
        h.set_scalars(o.mu, o.sigmaE, o.epsSum, o.sigmaG, o.pi, trait=t)
    h.set_state(np.stack([o.eps for o in os_], axis=1), np.zeros((pr.m, T)), np.zeros((pr.m, T), np.int32))
    tab = h.autotune(seed=5, iterations=1)
    assert len(tab) >= 3 and all(e["ms_per_iteration"] > 0 for e in tab), tab
    for it in range(2):
        if it == 1:
            h.set_launch(101, 0)          # slices of 199 -> 208 individuals
        order = np.arange(pr.m, dtype=np.uint32)
        brr.shuffle(13, it, 0, order)
        uz = [brr.marker_uz(13, it, t, pr.m) for t in range(T)]
        u = np.stack([a for a, _ in uz], axis=1)
        z = np.stack([b for _, b in uz], axis=1)
        for t, o in enumerate(os_):
            o.sweep(order, u[:, t], z[:, t], trace=False)
        h.sweep(order, u, z)
        st = h.get_state()
        for t, o in enumerate(os_):
            assert np.array_equal(st["comp"][:, t], o.comp), (it, t)
            close(st["eps"][:, t], o.eps, what=f"iteration {it} trait {t} eps")
    h.close()


def test_set_launch_refuses_bad_shapes(gpu_lib):
    pr = make_problem(5000, 20, seed=507)
    h = brr.HipSweep(pr.n, pr.m, 4, 1, brr.LAYOUT_DENSE_2BIT, traits=4)
    with pytest.raises(brr.BrrHipError):
        h.set_launch(8, 0)                # 625 individuals per workgroup > 256
    assert h.stats()["slice"] <= 256      # the context keeps its shape
    h.close()
    h = brr.HipSweep(pr.n, pr.m, 4, 1, brr.LAYOUT_DENSE_2BIT)
    with pytest.raises(brr.BrrHipError):
        h.set_launch(0, 65)
    h.close()
