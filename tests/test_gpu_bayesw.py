"""BayesW (Weibull survival) sweep on the device against the CPU oracle.

The device evaluates the same processColumn steps as the oracle (class
likelihoods by the tabulated Gauss-Hermite rules, beta by adaptive rejection
sampling on the same counter-based uniforms), but its N-sums run in a
different order and its exp / log are the device library's, so floating-point
state agrees to rounding (bar RTOL relative to the array's scale) while every
discrete outcome (class, class counts, which markers changed) must be equal.
"""
import numpy as np
import pytest

import bayesrrcmd_amd as brr
import _oracle as orc
from _data import make_bw_problem

pytestmark = pytest.mark.gpu

RTOL = 1e-8
EULER = 0.577215664901532


def close(a, b, rtol=RTOL, what=""):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(1.0, float(np.max(np.abs(b))) if b.size else 1.0)
    err = float(np.max(np.abs(a - b))) if a.size else 0.0
    assert err <= rtol * scale, f"{what}: max abs err {err:.3e} (scale {scale:.3e})"


def classes(K):
    return {2: (0.01,), 3: (0.001, 0.01), 4: (0.001, 0.01, 0.1),
            8: (0.0001, 0.0003, 0.001, 0.003, 0.01, 0.03, 0.1)}[K]


def setup(pr, K, nq, workgroups, seed):
    """Oracle after init + one mu draw, and a device context holding the same state."""
    w = orc.BayesWProblem(pr.n, pr.m, pr.ragged, pr.fail, K=K, C=classes(K), quad_points=nq)
    w.init(pr.y)
    assert orc.lib().orc_bw_sample_mu(orc.C.byref(w.data), orc.C.byref(w.state), seed, 0, 0) == 0
    w.refresh_vi()
    h = brr.HipBayesW(pr.n, pr.m, K, quad_points=nq, workgroups=workgroups)
    h.upload_ragged(pr.ragged)
    h.set_model(np.array(classes(K)), pr.fail)
    return w, h


def push(w, h):
    st = w.state
    h.set_state(w.eps, w.beta, w.comp)
    h.set_scalars(st.mu, st.alpha, st.sigma_b, st.d, w.pi, np.ones(w.K))


def order_u(seed, it, m):
    order = np.arange(m, dtype=np.uint32)
    brr.shuffle(seed, it, 0, order)
    u, _ = brr.marker_uz(seed, it, 0, m)
    return order, u


def compare(w, h, what):
    d = h.get_state()
    s = h.get_scalars()
    bad = np.flatnonzero(d["comp"] != w.comp)
    assert bad.size == 0, f"{what}: components differ at {bad[:10]} (device {d['comp'][bad[:5]]}, " \
                          f"oracle {w.comp[bad[:5]]})"
    np.testing.assert_array_equal(s["v"], w.v, err_msg=f"{what}: class counts")
    np.testing.assert_array_equal(d["beta"] == 0.0, w.beta == 0.0, err_msg=f"{what}: zero pattern of beta")
    close(d["beta"], w.beta, what=f"{what}: beta")
    close(d["eps"], w.eps, what=f"{what}: eps")
    close(s["betasq"], np.sum(w.beta ** 2), what=f"{what}: sum beta^2")
    return d, s


@pytest.mark.parametrize("n,m,K,nq,wg,frac", [
    (3000, 160, 4, 7, 0, 0.1),      # one run per lane (R = 1), 188 workgroups
    (3000, 160, 4, 9, 5, 0.1),      # R = 2, n = 9 rule (the x6 = -x3 entry)
    (6000, 96, 4, 15, 4, 0.15),     # R = 4
    (2500, 130, 2, 3, 13, 0.1),     # K = 2 (one slab: the last class gets acum = 1 at once)
    (2500, 100, 8, 11, 0, 0.2),     # K = 8
    (1999, 77, 3, 5, 7, 0.1),       # ragged N and M (partial last block and slice)
])
def test_bw_sweep_matches_oracle(n, m, K, nq, wg, frac):
    seed = 1000 + n + m
    pr = make_bw_problem(n, m, seed=seed, causal_frac=frac)
    w, h = setup(pr, K, nq, wg, seed)
    with h:
        np.testing.assert_array_equal(h.sum_failure(), w.sum_failure)
        for it in range(3):
            # sweep 0 starts from beta = 0; later sweeps carry markers in the model
            push(w, h)
            order, u = order_u(seed, it, m)
            w.v[:] = 1.0
            assert w.sweep(order, u, seed, it) == 0
            h.sweep(order, u, seed, it)
            compare(w, h, f"sweep {it}")
            st = h.stats()
            assert st["markers"] == m
        assert np.count_nonzero(w.beta) > 0


def test_bw_sweep_full_individuals():
    """N = 100,000 individuals (the survival workload's slice shape) against the oracle."""
    seed = 77
    pr = make_bw_problem(100_000, 300, seed=seed, causal_frac=0.05, maf=(0.005, 0.5))
    w, h = setup(pr, 4, 7, 0, seed)
    with h:
        for it in range(2):
            push(w, h)
            order, u = order_u(seed, it, pr.m)
            w.v[:] = 1.0
            assert w.sweep(order, u, seed, it) == 0
            h.sweep(order, u, seed, it)
            compare(w, h, f"N=100k sweep {it}")
        # eps = y - mu - X beta through the residual-update columns
        d = h.get_state()
        r = pr.ragged
        xb = np.zeros(pr.n)
        for j in np.flatnonzero(d["beta"]):
            mean, sd = r.stats[j, 0], r.stats[j, 1]
            col = np.full(pr.n, -(mean / sd))
            col[r.ones[r.ones_off[j]: r.ones_off[j + 1]]] += 1 / sd
            col[r.twos[r.twos_off[j]: r.twos_off[j + 1]]] += 2 / sd
            xb += d["beta"][j] * col
        close(d["eps"], pr.y - w.state.mu - xb, rtol=1e-9, what="eps identity")


@pytest.mark.parametrize("n,m,wg,iters", [(2500, 120, 0, 4), (3000, 90, 6, 3)])
def test_bw_chain_matches_oracle(n, m, wg, iters):
    """Whole runGibbs iterations (mu, sweep, alpha, sigma_b, pi) in one launch."""
    seed = 4242 + n
    pr = make_bw_problem(n, m, seed=seed, causal_frac=0.1)
    w = orc.BayesWProblem(pr.n, pr.m, pr.ragged, pr.fail)
    err, hist_o, _ = w.run_gibbs(pr.y, seed, iters)
    assert err == 0
    eps0, sc = brr.bayesw_init(pr.y, pr.fail, m, 4)
    orders = []
    order = np.arange(m, dtype=np.uint32)
    for it in range(iters):
        brr.shuffle(seed, it, 0, order)
        orders.append(order.copy())
    with brr.HipBayesW(n, m, 4, quad_points=7, workgroups=wg) as h:
        h.upload_ragged(pr.ragged)
        h.set_model(np.array(classes(4)), pr.fail)
        h.set_state(eps0, np.zeros(m), np.zeros(m, dtype=np.int32))
        h.set_scalars(sc["mu"], sc["alpha"], sc["sigma_b"], sc["d"], sc["pi"], sc["v"])
        hist = h.run_chain(seed, 0, 0, np.stack(orders))
        np.testing.assert_array_equal(hist[:, 3], hist_o[:, 3])
        for c, name in enumerate(("mu", "alpha", "sigma_b", "nonzero", "sum beta^2")):
            close(hist[:, c], hist_o[:, c], rtol=1e-7, what=f"chain {name}")
        d = h.get_state()
        s = h.get_scalars()
        np.testing.assert_array_equal(d["comp"], w.comp)
        close(d["beta"], w.beta, rtol=1e-7, what="chain beta")
        close(d["eps"], w.eps, rtol=1e-7, what="chain eps")
        close(s["pi"], w.pi, rtol=1e-7, what="chain pi")
        close([s["mu"], s["alpha"], s["sigma_b"]], [w.state.mu, w.state.alpha, w.state.sigma_b], rtol=1e-7,
              what="chain scalars")


def test_bw_per_marker_calls():
    """setVi, calculateResidualUpdate and calculateSumFailure, one marker at a time."""
    pr = make_bw_problem(4000, 12, seed=5, causal_frac=0.2)
    r = pr.ragged
    rng = np.random.default_rng(3)
    eps = rng.normal(0.0, 0.3, pr.n)
    alpha = 2.7
    w = orc.BayesWProblem(pr.n, pr.m, pr.ragged, pr.fail)
    with brr.HipBayesW(pr.n, pr.m, 4) as h:
        h.upload_ragged(pr.ragged)
        h.set_model(np.array(classes(4)), pr.fail)
        np.testing.assert_array_equal(h.sum_failure(), w.sum_failure)
        h.set_state(eps=eps)
        vi = np.exp(alpha * eps - EULER)
        for j in range(pr.m):
            ones = r.ones[r.ones_off[j]: r.ones_off[j + 1]]
            twos = r.twos[r.twos_off[j]: r.twos_off[j + 1]]
            out = h.set_vi(j, alpha)
            want = [vi.sum(), vi.sum() - vi[ones].sum() - vi[twos].sum(), vi[ones].sum(), vi[twos].sum()]
            close(out, want, rtol=1e-12, what=f"set_vi marker {j}")
        e = eps.copy()
        for j, b, sgn in [(0, 0.3, 1), (3, -0.7, -1), (7, 0.05, 1)]:
            mean, sd = r.stats[j, 0], r.stats[j, 1]
            col = np.full(pr.n, -(mean / sd) * b)
            col[r.ones[r.ones_off[j]: r.ones_off[j + 1]]] += b / sd
            col[r.twos[r.twos_off[j]: r.twos_off[j + 1]]] += 2 * (b / sd)
            e = e + sgn * col
            h.residual_update(j, b, sgn)
        close(h.get_state()["eps"], e, rtol=1e-14, what="residual updates")


def test_bw_bad_calls():
    pr = make_bw_problem(800, 10, seed=9)
    for nq in (2, 8, 17, 1):
        with pytest.raises(brr.BrrHipError):
            brr.HipBayesW(pr.n, pr.m, 4, quad_points=nq)
    with pytest.raises(brr.BrrHipError):
        brr.HipBayesW(pr.n, pr.m, 1)
    with brr.HipBayesW(pr.n, pr.m, 4) as h:
        order = np.arange(pr.m, dtype=np.uint32)
        with pytest.raises(brr.BrrHipError, match="markers not uploaded"):
            h.sweep(order, np.zeros(pr.m), 1, 0)
        with pytest.raises(brr.BrrHipError, match="markers not uploaded"):
            h.set_model(np.array(classes(4)), pr.fail)
        h.upload_ragged(pr.ragged)
        with pytest.raises(brr.BrrHipError, match="model not set"):
            h.sweep(order, np.zeros(pr.m), 1, 0)
        bad = pr.fail.copy()
        bad[0] = 0.5
        with pytest.raises(brr.BrrHipError, match="0 or 1"):
            h.set_model(np.array(classes(4)), bad)
        with pytest.raises(brr.BrrHipError, match="> 0"):
            h.set_model(np.array([0.001, 0.0, 0.1]), pr.fail)
        h.set_model(np.array(classes(4)), pr.fail)
        with pytest.raises(brr.BrrHipError, match="permutation"):
            h.sweep(np.zeros(pr.m, dtype=np.uint32), np.zeros(pr.m), 1, 0)
        # BayesR calls on a BayesW context are refused
        rc = brr.lib().brr_hip_compute_num(h.ctx, 0, 0.0, None)
        assert rc == brr._abi.ESTATE
        with pytest.raises(brr.BrrHipError, match="sign"):
            h.residual_update(0, 0.1, 0)


# ------------------------------------------------------------- dense layout --

def dense_setup(pr, K, nq, workgroups, seed):
    lut = brr.build_dense(pr.bed, pr.n)
    w = orc.BayesWProblem(pr.n, pr.m, None, pr.fail, K=K, C=classes(K), quad_points=nq, bed=pr.bed, lut=lut)
    w.init(pr.y)
    assert orc.lib().orc_bw_sample_mu(orc.C.byref(w.data), orc.C.byref(w.state), seed, 0, 0) == 0
    w.refresh_vi()
    h = brr.HipBayesW(pr.n, pr.m, K, quad_points=nq, workgroups=workgroups)
    h.upload_dense(pr.bed, lut)
    h.set_model(np.array(classes(K)), pr.fail)
    return w, h


@pytest.mark.parametrize("n,m,K,nq,wg,frac,miss", [
    (3000, 160, 4, 7, 0, 0.1, 0.0),
    (2600, 120, 4, 9, 5, 0.1, 0.04),     # missing calls (mean-imputed in the dense column), R = 2
    (6000, 64, 3, 15, 4, 0.15, 0.02),    # R = 4
    (1999, 77, 8, 5, 7, 0.2, 0.0),
])
def test_bw_dense_sweep_matches_oracle(n, m, K, nq, wg, frac, miss):
    """DenseBayesWKernel on the device: the O(N) sums of the reference regrouped by
    the column's four values, against the oracle's element-by-element sums."""
    seed = 2000 + n + m
    pr = make_bw_problem(n, m, seed=seed, causal_frac=frac, miss=miss)
    w, h = dense_setup(pr, K, nq, wg, seed)
    with h:
        close(h.sum_failure(), w.sum_failure, rtol=1e-12, what="sum_failure")
        for it in range(3):
            push(w, h)
            order, u = order_u(seed, it, m)
            w.v[:] = 1.0
            assert w.sweep(order, u, seed, it) == 0
            h.sweep(order, u, seed, it)
            compare(w, h, f"dense sweep {it}")
        assert np.count_nonzero(w.beta) > 0


def test_bw_dense_chain_matches_oracle():
    seed = 4343
    n, m, iters = 2500, 100, 3
    pr = make_bw_problem(n, m, seed=seed, causal_frac=0.1)
    lut = brr.build_dense(pr.bed, n)
    w = orc.BayesWProblem(n, m, None, pr.fail, bed=pr.bed, lut=lut)
    err, hist_o, _ = w.run_gibbs(pr.y, seed, iters)
    assert err == 0
    eps0, sc = brr.bayesw_init(pr.y, pr.fail, m, 4)
    orders = []
    order = np.arange(m, dtype=np.uint32)
    for it in range(iters):
        brr.shuffle(seed, it, 0, order)
        orders.append(order.copy())
    with brr.HipBayesW(n, m, 4, quad_points=7) as h:
        h.upload_dense(pr.bed, lut)
        h.set_model(np.array(classes(4)), pr.fail)
        h.set_state(eps0, np.zeros(m), np.zeros(m, dtype=np.int32))
        h.set_scalars(sc["mu"], sc["alpha"], sc["sigma_b"], sc["d"], sc["pi"], sc["v"])
        hist = h.run_chain(seed, 0, 0, np.stack(orders))
        np.testing.assert_array_equal(hist[:, 3], hist_o[:, 3])
        close(hist, hist_o, rtol=1e-7, what="dense chain history")
        d = h.get_state()
        np.testing.assert_array_equal(d["comp"], w.comp)
        close(d["beta"], w.beta, rtol=1e-7, what="dense chain beta")
        close(d["eps"], w.eps, rtol=1e-7, what="dense chain eps")


def test_bw_dense_per_marker_calls():
    pr = make_bw_problem(3000, 10, seed=7, causal_frac=0.2)
    lut = brr.build_dense(pr.bed, pr.n)
    from _data import spec_dosage
    d = spec_dosage(pr.bed, pr.n)
    x = np.take_along_axis(lut, np.where(d < 0, 3, d), axis=1)
    rng = np.random.default_rng(4)
    eps = rng.normal(0.0, 0.3, pr.n)
    with brr.HipBayesW(pr.n, pr.m, 4) as h:
        h.upload_dense(pr.bed, lut)
        h.set_model(np.array(classes(4)), pr.fail)
        close(h.sum_failure(), x @ pr.fail, rtol=1e-12, what="dense sum_failure")
        h.set_state(eps=eps)
        e = eps.copy()
        for j, b, sgn in [(1, 0.4, 1), (6, -0.2, -1)]:
            e = e + sgn * (x[j] * b)
            h.residual_update(j, b, sgn)
        close(h.get_state()["eps"], e, rtol=1e-15, what="dense residual updates")
