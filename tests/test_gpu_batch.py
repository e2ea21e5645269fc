"""Batched traits on the device (num_traits = T) against T oracle chains.

T phenotypes share the genotype columns and the marker order; trait t is an
independent BayesR chain. The device forms the dot of a later marker in a
block as C + sum(delta * Gram) rather than a fresh x . eps, so floating-point
state matches the FP64 oracle to reduction-order rounding (bar 1e-9
relative), and component choices and v counts must be identical.
"""
import numpy as np
import pytest

import bayesrrcmd_amd as brr
import _oracle as orc
from _data import make_problem, spec_dosage

pytestmark = pytest.mark.gpu

S = np.array([0.01, 0.001, 0.0001])
RTOL = 1e-9


def close(a, b, rtol=RTOL, what=""):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(1.0, float(np.max(np.abs(b))) if b.size else 1.0)
    err = float(np.max(np.abs(a - b))) if a.size else 0.0
    assert err <= rtol * scale, f"{what}: max abs err {err:.3e} (scale {scale:.3e})"


def phenotypes(pr, T, seed):
    """T phenotypes over the problem's genotypes, h2 and causal count per trait."""
    rng = np.random.default_rng(seed)
    d = spec_dosage(pr.bed, pr.n)
    d = np.where(d < 0, 3, d)
    x = np.take_along_axis(pr.lut, d, axis=1)          # (m, n) standardised
    ys = []
    for _ in range(T):
        h2 = rng.uniform(0.1, 0.8)
        mc = int(rng.integers(1, max(2, pr.m // 20)))
        causal = rng.choice(pr.m, mc, replace=False)
        b = rng.normal(0.0, np.sqrt(h2 / mc), mc)
        y = b @ x[causal] + rng.normal(0.0, np.sqrt(1.0 - h2), pr.n)
        ys.append(y.astype(np.float32).astype(np.float64))
    return ys


def oracle_traits(pr, layout, ys, seed, K=4):
    out = []
    for t, y in enumerate(ys):
        if layout == orc.DENSE:
            o = orc.OracleProblem(pr.n, pr.m, orc.DENSE, K, 1, S[: K - 1], bed=pr.bed, lut=pr.lut)
        else:
            o = orc.OracleProblem(pr.n, pr.m, orc.RAGGED, K, 1, S[: K - 1], ragged=pr.ragged)
        o.init(y, seed, chain=t)
        out.append(o)
    return out


def device_traits(pr, layout, os_, workgroups=0, K=4):
    T = len(os_)
    code = brr.LAYOUT_DENSE_2BIT if layout == orc.DENSE else brr.LAYOUT_RAGGED
    h = brr.HipSweep(pr.n, pr.m, K, 1, code, workgroups=workgroups, traits=T)
    h.upload_dense(pr.bed, pr.lut) if layout == orc.DENSE else h.upload_ragged(pr.ragged)
    h.set_model(S[: K - 1])
    h.set_state(np.stack([o.eps for o in os_], axis=1), np.zeros((pr.m, T)), np.zeros((pr.m, T), np.int32))
    for t, o in enumerate(os_):
        h.set_scalars(o.mu, o.sigmaE, o.epsSum, o.sigmaG, o.pi, trait=t)
    return h


def compare(os_, h, layout, what):
    st = h.get_state()
    for t, o in enumerate(os_):
        w = f"{what} trait {t}"
        assert np.array_equal(st["comp"][:, t], o.comp), \
            f"{w}: components differ at {np.flatnonzero(st['comp'][:, t] != o.comp)[:10]}"
        close(st["beta"][:, t], o.beta, what=f"{w} beta")
        close(st["acum"][:, t], o.acum, what=f"{w} acum")
        close(st["eps"][:, t], o.eps, what=f"{w} eps")
        sc = h.get_scalars(trait=t)
        assert np.array_equal(sc["v"].reshape(-1), o.v), f"{w}: v"
        close(sc["betasqn"], o.betasqn, what=f"{w} betasqn")
        if layout == orc.RAGGED:
            close([sc["epsilon_sum"]], [o.epsSum], rtol=1e-8, what=f"{w} epsSum")


@pytest.mark.parametrize("layout", [orc.DENSE, orc.RAGGED], ids=["dense", "ragged"])
@pytest.mark.parametrize("T,workgroups", [(1, 0), (7, 13), (64, 0), (64, 9), (20, 16)])
def test_batch_sweep_matches_oracle_per_trait(gpu_lib, layout, T, workgroups):
    pr = make_problem(2001, 333, seed=201, h2=0.5, causal_frac=0.03)
    seed = 31
    ys = phenotypes(pr, T, seed=202)
    os_ = oracle_traits(pr, layout, ys, seed)
    h = device_traits(pr, layout, os_, workgroups=workgroups)
    order = np.arange(pr.m, dtype=np.uint32)
    for it in range(3):
        brr.shuffle(seed, it, 0, order)
        uz = [brr.marker_uz(seed, it, t, pr.m) for t in range(T)]
        u = np.stack([a for a, _ in uz], axis=1)
        z = np.stack([b for _, b in uz], axis=1)
        for t, o in enumerate(os_):
            o.sweep(order, u[:, t], z[:, t], trace=False)
        h.sweep(order, u, z)
        compare(os_, h, layout, f"iteration {it}")
    s = h.stats()
    assert s["markers"] == pr.m * T and s["block_markers"] == 32
    assert s["rounds"] == (pr.m + 31) // 32
    assert s["changes"] > 0
    h.close()


@pytest.mark.parametrize("K", [2, 3, 8])
def test_batch_component_counts(gpu_lib, K):
    pr = make_problem(1500, 190, seed=203, h2=0.6, causal_frac=0.05)
    seed = 33
    cva = np.array([0.1, 0.03, 0.01, 0.003, 0.001, 0.0003, 0.0001])[: K - 1]
    ys = phenotypes(pr, 9, seed=204)
    os_ = []
    for t, y in enumerate(ys):
        o = orc.OracleProblem(pr.n, pr.m, orc.DENSE, K, 1, cva, bed=pr.bed, lut=pr.lut)
        o.init(y, seed, chain=t)
        os_.append(o)
    h = brr.HipSweep(pr.n, pr.m, K, 1, brr.LAYOUT_DENSE_2BIT, traits=9)
    h.upload_dense(pr.bed, pr.lut)
    h.set_model(cva)
    h.set_state(np.stack([o.eps for o in os_], axis=1), np.zeros((pr.m, 9)), np.zeros((pr.m, 9), np.int32))
    for t, o in enumerate(os_):
        h.set_scalars(o.mu, o.sigmaE, o.epsSum, o.sigmaG, o.pi, trait=t)
    order = np.arange(pr.m, dtype=np.uint32)
    for it in range(2):
        brr.shuffle(seed, it, 0, order)
        uz = [brr.marker_uz(seed, it, t, pr.m) for t in range(9)]
        u = np.stack([a for a, _ in uz], axis=1)
        z = np.stack([b for _, b in uz], axis=1)
        for t, o in enumerate(os_):
            o.sweep(order, u[:, t], z[:, t], trace=False)
        h.sweep(order, u, z)
        compare(os_, h, orc.DENSE, f"K={K} iteration {it}")
    h.close()


@pytest.mark.parametrize("layout", [orc.DENSE, orc.RAGGED], ids=["dense", "ragged"])
def test_batch_chain_matches_oracle_gibbs(gpu_lib, layout):
    """Device chains (updateMu, sweep, sigma / pi draws per trait) against the
    oracle's runGibbs steps per trait with the shared orders."""
    pr = make_problem(1800, 260, seed=205, h2=0.5, causal_frac=0.03)
    seed, T, iters = 20261202, 6, 5
    ys = phenotypes(pr, T, seed=206)
    os_ = oracle_traits(pr, layout, ys, seed)
    h = device_traits(pr, layout, os_)
    orders = np.empty((iters, pr.m), dtype=np.uint32)
    cur = np.arange(pr.m, dtype=np.uint32)
    for it in range(iters):
        brr.shuffle(seed, it, 0, cur)
        orders[it] = cur
    hist_h = h.run_chain(seed, 0, 0, orders)
    hist_o = np.zeros((iters, T, 5))
    u = np.zeros(pr.m)
    z = np.zeros(pr.m)
    for t, o in enumerate(os_):
        for it in range(iters):
            o.update_mu(seed, it, chain=t)
            orc.lib().orc_fill_marker_uz(seed, it, t, pr.m, orc.p(u, orc.C.c_double), orc.p(z, orc.C.c_double))
            o.sweep(orders[it], u, z, trace=False)
            e2 = float(np.sum(o.eps * o.eps))
            m0 = float(o.v.sum() - o.v[0])
            mu = o.mu
            o.post_sweep(seed, it, chain=t)
            hist_o[it, t] = [mu, o.sigmaE, m0, e2, o.sigmaG[0]]
    close(hist_h[..., 0], hist_o[..., 0], rtol=1e-8, what="mu")
    close(hist_h[..., 1], hist_o[..., 1], rtol=1e-8, what="sigmaE")
    assert np.array_equal(hist_h[..., 2], hist_o[..., 2]), "m0"
    close(hist_h[..., 3], hist_o[..., 3], rtol=1e-8, what="|eps|^2")
    close(hist_h[..., 4], hist_o[..., 4], rtol=1e-8, what="sigmaG")
    st = h.get_state()
    for t, o in enumerate(os_):
        assert np.array_equal(st["comp"][:, t], o.comp), f"trait {t} comp"
        close(st["beta"][:, t], o.beta, rtol=1e-8, what=f"trait {t} beta")
        close(st["eps"][:, t], o.eps, rtol=1e-8, what=f"trait {t} eps")
    s = h.stats()
    assert s["markers"] == iters * pr.m * T
    h.close()


def test_batch_large_slice_residual_identity(gpu_lib):
    """N = 50,000 on the 13-tile kernel: after a device chain, each trait's eps
    equals y_std - mu - X beta (size-independent identity)."""
    pr = make_problem(50_000, 96, seed=207, h2=0.5, causal_frac=0.1)
    seed, T = 5, 16
    ys = phenotypes(pr, T, seed=208)
    eps0 = np.zeros((pr.n, T))
    h = brr.HipSweep(pr.n, pr.m, 4, 1, brr.LAYOUT_DENSE_2BIT, traits=T)
    h.upload_dense(pr.bed, pr.lut)
    h.set_model(S)
    for t, y in enumerate(ys):
        e, sc = brr.bayesr_init(y, 1, 4, seed, chain=t)
        eps0[:, t] = e
        h.set_scalars(sc["mu"], sc["sigmaE"], sc["epsilon_sum"], sc["sigmaG"], sc["pi"], trait=t)
    h.set_state(eps0, np.zeros((pr.m, T)), np.zeros((pr.m, T), np.int32))
    orders = np.empty((3, pr.m), dtype=np.uint32)
    cur = np.arange(pr.m, dtype=np.uint32)
    for it in range(3):
        brr.shuffle(seed, it, 0, cur)
        orders[it] = cur
    h.run_chain(seed, 0, 0, orders)
    assert h.stats()["slice"] == 208
    st = h.get_state()
    d = spec_dosage(pr.bed, pr.n)
    x = np.take_along_axis(pr.lut, np.where(d < 0, 3, d), axis=1)    # (m, n)
    for t in range(T):
        mu = h.get_scalars(trait=t)["mu"]
        close(st["eps"][:, t], eps0[:, t] - mu - st["beta"][:, t] @ x, rtol=1e-9, what=f"trait {t} identity")
    h.close()


def test_batch_t64_n50k_against_oracle(gpu_lib):
    """The batched workload's shape: T = 64 traits, N = 50,000 (the 13-tile kernel,
    slices of 208 on 241 workgroups, four waves of 16 traits), 3 sweeps; traits 0,
    21, 42 and 63 (one per wave) against their oracle chains at the batched bars,
    and every trait's residual equal to y_std - mu - X beta."""
    pr = make_problem(50_000, 400, seed=211, h2=0.5, causal_frac=0.02)
    seed, T = 37, 64
    ys = phenotypes(pr, T, seed=212)
    eps0 = np.zeros((pr.n, T))
    h = brr.HipSweep(pr.n, pr.m, 4, 1, brr.LAYOUT_DENSE_2BIT, traits=T)
    h.upload_dense(pr.bed, pr.lut)
    h.set_model(S)
    picked = (0, 21, 42, 63)
    os_ = {}
    for t, y in enumerate(ys):
        e, sc = brr.bayesr_init(y, 1, 4, seed, chain=t)
        eps0[:, t] = e
        h.set_scalars(sc["mu"], sc["sigmaE"], sc["epsilon_sum"], sc["sigmaG"], sc["pi"], trait=t)
        if t in picked:
            o = orc.OracleProblem(pr.n, pr.m, orc.DENSE, 4, 1, S, bed=pr.bed, lut=pr.lut)
            o.init(y, seed, chain=t)
            assert np.array_equal(o.eps, e)
            os_[t] = o
    h.set_state(eps0, np.zeros((pr.m, T)), np.zeros((pr.m, T), np.int32))
    s = h.stats()
    assert s["slice"] == 208 and s["workgroups"] == 241
    order = np.arange(pr.m, dtype=np.uint32)
    for it in range(3):
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
            close(st["beta"][:, t], o.beta, what=f"{w} beta")
            close(st["acum"][:, t], o.acum, what=f"{w} acum")
            close(st["eps"][:, t], o.eps, what=f"{w} eps")
            sc = h.get_scalars(trait=t)
            assert np.array_equal(sc["v"].reshape(-1), o.v), f"{w}: v"
    d = spec_dosage(pr.bed, pr.n)
    x = np.take_along_axis(pr.lut, np.where(d < 0, 3, d), axis=1)    # (m, n)
    for t in range(T):
        mu = h.get_scalars(trait=t)["mu"]
        close(st["eps"][:, t], eps0[:, t] - mu - st["beta"][:, t] @ x, rtol=1e-9, what=f"trait {t} identity")
    h.close()


def test_batch_bad_configs(gpu_lib):
    pr = make_problem(4000, 20, seed=209)
    with pytest.raises(brr.BrrHipError):
        brr.HipSweep(pr.n, pr.m, 4, 1, brr.LAYOUT_DENSE_2BIT, traits=65)
    with pytest.raises(brr.BrrHipError):
        brr.HipSweep(pr.n, pr.m, 4, 2, brr.LAYOUT_DENSE_2BIT, traits=4)
    with pytest.raises(brr.BrrHipError):   # 4000 / 7 -> 576 individuals per workgroup > 512
        brr.HipSweep(pr.n, pr.m, 4, 1, brr.LAYOUT_DENSE_2BIT, workgroups=7, traits=4)
    h = brr.HipSweep(pr.n, pr.m, 4, 1, brr.LAYOUT_DENSE_2BIT, traits=4)
    h.upload_dense(pr.bed, pr.lut)
    h.set_model(S)
    with pytest.raises(brr.BrrHipError):
        h.compute_num(0, 0.0)
    with pytest.raises(ValueError):
        h.set_state(np.zeros(pr.n))
    h.close()
