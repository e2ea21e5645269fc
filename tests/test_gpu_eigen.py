"""The EigenSparse layout on the device (BRR_HIP_LAYOUT_EIGEN_SPARSE) against the oracle.

EigenBayesRKernel (eigenbayesrkernel.cpp) is the ragged kernel with the missing
calls counted at the marker mean in the dot, on EigenSparse statistics. Bar as
the other BayesR parity tests: components and v counts identical; beta, acum,
epsilon, betasqn within 1e-9 relative over sweeps, 1e-8 over whole chains (the
device sums slices per workgroup, the oracle in index order).
"""
import numpy as np
import pytest

import bayesrrcmd_amd as brr
import _oracle as orc
from _data import make_problem

pytestmark = pytest.mark.gpu

S = np.array([0.01, 0.001, 0.0001])
RTOL = 1e-9


def close(a, b, rtol=RTOL, what=""):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(1.0, float(np.max(np.abs(b))) if b.size else 1.0)
    err = float(np.max(np.abs(a - b))) if a.size else 0.0
    assert err <= rtol * scale, f"{what}: max abs err {err:.3e} (scale {scale:.3e})"


def eigen_problem(n, m, seed, miss=0.04, **kw):
    pr = make_problem(n, m, seed=seed, miss=miss, **kw)
    pr.eigen = brr.build_eigen(pr.bed, n)
    return pr


def variances(K):
    return 10.0 ** -np.arange(2, K + 1)          # K - 1 slab variances 0.01, 0.001, ...


def make_pair(pr, K=4, G=1, group=None, workgroups=0, block=0, seed=5, decoded=False):
    cv = np.tile(variances(K), G)
    o = orc.OracleProblem(pr.n, pr.m, orc.EIGEN, K, G, cv, group=group, ragged=pr.eigen)
    o.init(pr.y, seed)
    h = brr.HipSweep(pr.n, pr.m, K, G, brr.LAYOUT_EIGEN_SPARSE, workgroups=workgroups, block_markers=block)
    if decoded:
        h.load_bed(brr.bed_file_image(pr.bed), brr.LAYOUT_EIGEN_SPARSE)
    else:
        h.upload_ragged(pr.eigen)
    if group is not None:
        h.set_groups(group)
    h.set_model(cv)
    h.set_state(o.eps, o.beta, o.comp)
    h.set_scalars(o.mu, o.sigmaE, o.epsSum, o.sigmaG, o.pi)
    return o, h


def compare_state(o, h, what):
    st = h.get_state()
    sc = h.get_scalars()
    bad = np.flatnonzero(st["comp"] != o.comp)
    assert bad.size == 0, f"{what}: components differ at {bad[:10]}"
    close(st["beta"], o.beta, what=f"{what} beta")
    close(st["acum"], o.acum, what=f"{what} acum")
    close(st["eps"], o.eps, what=f"{what} eps")
    assert np.array_equal(sc["v"].reshape(-1), o.v), f"{what}: v"
    close(sc["betasqn"], o.betasqn, what=f"{what} betasqn")
    close([sc["epsilon_sum"]], [o.epsSum], rtol=1e-8, what=f"{what} epsSum")


def sweep_inputs(m, it, seed=77):
    order = np.arange(m, dtype=np.uint32)
    for t in range(it + 1):
        brr.shuffle(seed, t, 0, order)
    u, z = brr.marker_uz(seed, it, 0, m)
    return order, u, z


@pytest.mark.parametrize("workgroups,block", [(0, 0), (3, 1), (17, 7), (0, 63)])
def test_eigen_sweep_parity(gpu_lib, workgroups, block):
    pr = eigen_problem(3001, 600, seed=401, h2=0.5, causal_frac=0.02)
    o, h = make_pair(pr, workgroups=workgroups, block=block)
    for it in range(3):
        order, u, z = sweep_inputs(pr.m, it)
        o.sweep(order, u, z, trace=False)
        h.sweep(order, u, z)
        compare_state(o, h, f"iteration {it}")
    assert h.stats()["storage"] == 2
    h.close()


def test_eigen_differs_from_ragged_with_missing(gpu_lib):
    """With missing calls the EigenSparse dot is not the ragged one: the layouts give different chains."""
    pr = eigen_problem(2001, 300, seed=402, miss=0.1, h2=0.5, causal_frac=0.05)
    o, h = make_pair(pr)
    r = brr.HipSweep(pr.n, pr.m, 4, 1, brr.LAYOUT_RAGGED, ragged_storage=brr.STORAGE_CODES)
    r.upload_ragged(pr.ragged)
    r.set_model(S)
    r.set_state(o.eps, o.beta, o.comp)
    r.set_scalars(o.mu, o.sigmaE, o.epsSum, o.sigmaG, o.pi)
    order, u, z = sweep_inputs(pr.m, 0)
    h.sweep(order, u, z)
    r.sweep(order, u, z)
    o.sweep(order, u, z, trace=False)
    compare_state(o, h, "eigen")
    assert not np.array_equal(h.get_state()["acum"], r.get_state()["acum"])
    h.close()
    r.close()


def test_eigen_groups_and_components(gpu_lib):
    pr = eigen_problem(2500, 400, seed=403, h2=0.5, causal_frac=0.03)
    group = (np.arange(pr.m) % 3).astype(np.int32)
    o, h = make_pair(pr, K=5, G=3, group=group)
    for it in range(2):
        order, u, z = sweep_inputs(pr.m, it)
        o.sweep(order, u, z, trace=False)
        h.sweep(order, u, z)
        compare_state(o, h, f"iteration {it}")
    h.close()


def test_eigen_chain_matches_oracle_gibbs(gpu_lib):
    pr = eigen_problem(2500, 500, seed=404, h2=0.5, causal_frac=0.02)
    seed, iters = 20261202, 6
    o = orc.OracleProblem(pr.n, pr.m, orc.EIGEN, 4, 1, S, ragged=pr.eigen)
    hist_o, order_o = o.run_gibbs(pr.y, seed, iters)
    eps, sc = brr.bayesr_init(pr.y, 1, 4, seed)
    orders = np.empty((iters, pr.m), dtype=np.uint32)
    cur = np.arange(pr.m, dtype=np.uint32)
    for it in range(iters):
        brr.shuffle(seed, it, 0, cur)
        orders[it] = cur
    assert np.array_equal(orders[-1], order_o)
    h = brr.HipSweep(pr.n, pr.m, 4, 1, brr.LAYOUT_EIGEN_SPARSE)
    h.upload_ragged(pr.eigen)
    h.set_model(S)
    h.set_state(eps, np.zeros(pr.m), np.zeros(pr.m, np.int32))
    h.set_scalars(sc["mu"], sc["sigmaE"], sc["epsilon_sum"], sc["sigmaG"], sc["pi"])
    hist_h = h.run_chain(seed, 0, 0, orders)
    for c, what in ((0, "mu"), (1, "sigmaE"), (3, "|eps|^2"), (4, "sigmaG")):
        close(hist_h[:, c], hist_o[:, c], rtol=1e-8, what=what)
    assert np.array_equal(hist_h[:, 2], hist_o[:, 2]), "m0"
    st = h.get_state()
    assert np.array_equal(st["comp"], o.comp)
    close(st["beta"], o.beta, rtol=1e-8, what="beta")
    close(st["eps"], o.eps, rtol=1e-8, what="eps")
    h.close()


def test_eigen_per_marker_kernels(gpu_lib):
    pr = eigen_problem(3000, 50, seed=405, miss=0.05)
    o, h = make_pair(pr)
    rng = np.random.default_rng(0)
    for j in rng.choice(pr.m, 12, replace=False):
        j = int(j)
        b_old, b_new = rng.normal(), rng.normal()
        close([h.compute_num(j, b_old)], [o.compute_num(j, b_old)], what="num")
        h.epsilon_change(j, b_old, b_new)
        o.apply_epsilon_change(j, b_old, b_new)
        assert np.array_equal(h.get_state()["eps"], o.eps)
        close([h.get_scalars()["epsilon_sum"]], [o.epsSum], rtol=1e-8, what="epsSum")
    h.close()


def test_eigen_batched_per_trait(gpu_lib):
    pr = eigen_problem(2001, 333, seed=406, h2=0.5, causal_frac=0.03)
    T, seed = 7, 31
    rng = np.random.default_rng(202)
    os_ = []
    for t in range(T):
        y = (pr.y * rng.uniform(0.5, 1.5) + rng.normal(0.0, 0.3, pr.n)).astype(np.float32).astype(np.float64)
        o = orc.OracleProblem(pr.n, pr.m, orc.EIGEN, 4, 1, S, ragged=pr.eigen)
        o.init(y, seed, chain=t)
        os_.append(o)
    h = brr.HipSweep(pr.n, pr.m, 4, 1, brr.LAYOUT_EIGEN_SPARSE, traits=T)
    h.upload_ragged(pr.eigen)
    h.set_model(S)
    h.set_state(np.stack([o.eps for o in os_], axis=1), np.zeros((pr.m, T)), np.zeros((pr.m, T), np.int32))
    for t, o in enumerate(os_):
        h.set_scalars(o.mu, o.sigmaE, o.epsSum, o.sigmaG, o.pi, trait=t)
    order = np.arange(pr.m, dtype=np.uint32)
    for it in range(2):
        brr.shuffle(seed, it, 0, order)
        uz = [brr.marker_uz(seed, it, t, pr.m) for t in range(T)]
        u = np.stack([a for a, _ in uz], axis=1)
        z = np.stack([b for _, b in uz], axis=1)
        for t, o in enumerate(os_):
            o.sweep(order, u[:, t], z[:, t], trace=False)
        h.sweep(order, u, z)
        st = h.get_state()
        for t, o in enumerate(os_):
            assert np.array_equal(st["comp"][:, t], o.comp), f"trait {t} components"
            close(st["beta"][:, t], o.beta, what=f"trait {t} beta")
            close(st["eps"][:, t], o.eps, what=f"trait {t} eps")
            sc = h.get_scalars(trait=t)
            close([sc["epsilon_sum"]], [o.epsSum], rtol=1e-8, what=f"trait {t} epsSum")
    h.close()


def test_eigen_decoded_file(gpu_lib):
    """brr_hip_load_bed with the EigenSparse layout: statistics bit for bit the host builder's, and the
    same sweep as from the host-built upload."""
    pr = eigen_problem(2003, 150, seed=407, miss=0.05, h2=0.5, causal_frac=0.05)
    o, hu = make_pair(pr)
    _, hd = make_pair(pr, decoded=True)
    assert np.array_equal(hd.marker_tables().view(np.uint64), pr.eigen.stats.view(np.uint64))
    for it in range(2):
        order, u, z = sweep_inputs(pr.m, it)
        hu.sweep(order, u, z)
        hd.sweep(order, u, z)
    a, b = hu.get_state(), hd.get_state()
    for k in ("eps", "beta", "acum"):
        assert np.array_equal(a[k].view(np.uint64), b[k].view(np.uint64)), k
    # a column with only missing calls is valid here (rejected by the ragged layout)
    bed = pr.bed.copy()
    bed[3, :] = 0x55
    hd.load_bed(brr.bed_file_image(bed), brr.LAYOUT_EIGEN_SPARSE)
    hom = pr.bed.copy()
    hom[6, :] = 0xFF
    with pytest.raises(brr.BrrHipError, match="marker 6 has no nonzero call"):
        hd.load_bed(brr.bed_file_image(hom), brr.LAYOUT_EIGEN_SPARSE)
    hu.close()
    hd.close()


def test_eigen_bad_calls(gpu_lib):
    pr = eigen_problem(500, 20, seed=408)
    h = brr.HipSweep(pr.n, pr.m, 4, 1, brr.LAYOUT_EIGEN_SPARSE)
    with pytest.raises(brr.BrrHipError):
        h.upload_dense(pr.bed, pr.lut)                 # not a dense context
    with pytest.raises(brr.BrrHipError):
        h.load_bed(brr.bed_file_image(pr.bed), brr.LAYOUT_RAGGED)
    h.close()
    w = brr.HipBayesW(pr.n, pr.m, 4)
    with pytest.raises(brr.BrrHipError):
        w.load_bed(brr.bed_file_image(pr.bed), brr.LAYOUT_EIGEN_SPARSE)   # BayesW has no EigenSparse kernel
    w.close()
