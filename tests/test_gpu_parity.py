"""Parity of the HIP sweep (through the C ABI) with the CPU oracle.

Bar: component choices and v counts identical; beta, acum, epsilon, betasqn,
epsilon_sum and the chain scalars within 1e-9 relative (both sides are FP64;
they differ only in reduction order: the device sums a slice per workgroup and
combines the partials in a fixed tree, the oracle sums in index order).
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


def layout_code(layout):
    return brr.LAYOUT_DENSE_2BIT if layout == orc.DENSE else brr.LAYOUT_RAGGED


# (oracle layout, device ragged storage): dense, ragged kept as index lists,
# ragged converted to 2-bit codes on the device
LAYOUTS = [(orc.DENSE, brr.STORAGE_AUTO), (orc.RAGGED, brr.STORAGE_LISTS), (orc.RAGGED, brr.STORAGE_CODES)]
LAYOUT_IDS = ["dense", "ragged-lists", "ragged-codes"]


def make_pair(pr, layout, K=4, G=1, group=None, cva=S, workgroups=0, block=0, seed=5, storage=0):
    cv = np.tile(cva, G)
    if layout == orc.DENSE:
        o = orc.OracleProblem(pr.n, pr.m, orc.DENSE, K, G, cv, group=group, bed=pr.bed, lut=pr.lut)
    else:
        o = orc.OracleProblem(pr.n, pr.m, orc.RAGGED, K, G, cv, group=group, ragged=pr.ragged)
    o.init(pr.y, seed)
    h = brr.HipSweep(pr.n, pr.m, K, G, layout_code(layout), workgroups=workgroups, block_markers=block,
                     ragged_storage=storage)
    if layout == orc.DENSE:
        h.upload_dense(pr.bed, pr.lut)
    else:
        h.upload_ragged(pr.ragged)
    if group is not None:
        h.set_groups(group)
    h.set_model(cv)
    h.set_state(o.eps, o.beta, o.comp)
    h.set_scalars(o.mu, o.sigmaE, o.epsSum, o.sigmaG, o.pi)
    return o, h


def compare_state(o, h, layout, what):
    st = h.get_state()
    sc = h.get_scalars()
    assert np.array_equal(st["comp"], o.comp), f"{what}: components differ at " \
        f"{np.flatnonzero(st['comp'] != o.comp)[:10]}"
    close(st["beta"], o.beta, what=f"{what} beta")
    close(st["acum"], o.acum, what=f"{what} acum")
    close(st["eps"], o.eps, what=f"{what} eps")
    assert np.array_equal(sc["v"].reshape(-1), o.v), f"{what}: v"
    close(sc["betasqn"], o.betasqn, what=f"{what} betasqn")
    if layout == orc.RAGGED:
        close([sc["epsilon_sum"]], [o.epsSum], rtol=1e-8, what=f"{what} epsSum")


def sweep_inputs(m, it, seed=77):
    order = np.arange(m, dtype=np.uint32)
    for t in range(it + 1):
        brr.shuffle(seed, t, 0, order)
    u, z = brr.marker_uz(seed, it, 0, m)
    return order, u, z


@pytest.mark.parametrize("layout,storage", LAYOUTS, ids=LAYOUT_IDS)
@pytest.mark.parametrize("workgroups,block", [(0, 0), (3, 1), (17, 7), (64, 64), (0, 13), (0, 63)])
def test_sweep_parity(gpu_lib, layout, storage, workgroups, block):
    pr = make_problem(3001, 700, seed=101, h2=0.5, causal_frac=0.02)
    o, h = make_pair(pr, layout, workgroups=workgroups, block=block, storage=storage)
    for it in range(3):
        order, u, z = sweep_inputs(pr.m, it)
        o.sweep(order, u, z, trace=False)
        h.sweep(order, u, z)
        compare_state(o, h, layout, f"iteration {it}")
    s = h.stats()
    assert s["markers"] == pr.m and s["rounds"] >= (pr.m + s["block_markers"] - 1) // s["block_markers"]
    assert s["storage"] == (2 if layout == orc.DENSE else storage)
    h.close()


@pytest.mark.parametrize("layout,storage", LAYOUTS, ids=LAYOUT_IDS)
def test_sweep_parity_missing_and_groups(gpu_lib, layout, storage):
    pr = make_problem(2503, 400, seed=103, miss=0.05, causal_frac=0.05)
    G = 3
    group = (np.arange(pr.m) % G).astype(np.int32)
    cva = np.array([0.05, 0.005, 0.0005])
    o, h = make_pair(pr, layout, G=G, group=group, cva=cva, storage=storage)
    for it in range(2):
        order, u, z = sweep_inputs(pr.m, it, seed=9)
        o.sweep(order, u, z, trace=False)
        h.sweep(order, u, z)
        compare_state(o, h, layout, f"iteration {it}")
    h.close()


@pytest.mark.parametrize("layout,storage", LAYOUTS, ids=LAYOUT_IDS)
@pytest.mark.parametrize("K", [2, 3, 5, 8])
def test_sweep_parity_component_counts(gpu_lib, layout, storage, K):
    pr = make_problem(1500, 300, seed=105, causal_frac=0.05)
    cva = np.array([0.02, 0.002, 0.0002, 0.00002, 2e-6, 2e-7, 2e-8])[: K - 1]
    o, h = make_pair(pr, layout, K=K, cva=cva, storage=storage)
    order, u, z = sweep_inputs(pr.m, 0)
    o.sweep(order, u, z, trace=False)
    h.sweep(order, u, z)
    compare_state(o, h, layout, f"K={K}")
    h.close()


@pytest.mark.parametrize("layout,storage", LAYOUTS, ids=LAYOUT_IDS)
def test_sweep_parity_four_lane_runs(gpu_lib, layout, storage):
    """Slices of 2,000 individuals (three workgroups for N = 6,000): every lane holds
    four runs of eight individuals, the widest single-chain shape."""
    pr = make_problem(6000, 300, seed=119, causal_frac=0.03)
    o, h = make_pair(pr, layout, workgroups=3, storage=storage)
    assert h.stats()["slice"] == 2000
    for it in range(2):
        order, u, z = sweep_inputs(pr.m, it)
        o.sweep(order, u, z, trace=False)
        h.sweep(order, u, z)
        compare_state(o, h, layout, f"R = 4 iteration {it}")
    h.close()


@pytest.mark.parametrize("n,m", [(37, 5), (513, 1), (1025, 64), (2048, 9)])
def test_sweep_parity_edge_sizes(gpu_lib, n, m):
    pr = make_problem(n, m, seed=107 + n, causal_frac=0.2)
    for layout, storage in LAYOUTS:
        o, h = make_pair(pr, layout, storage=storage)
        order, u, z = sweep_inputs(pr.m, 0)
        o.sweep(order, u, z, trace=False)
        h.sweep(order, u, z)
        compare_state(o, h, layout, f"N={n} M={m}")
        h.close()


@pytest.mark.parametrize("storage", [brr.STORAGE_LISTS, brr.STORAGE_CODES])
def test_sweep_parity_every_marker_changes(gpu_lib, storage):
    """pi_0 tiny: nearly every decision changes beta, one exchange round each."""
    pr = make_problem(2000, 300, seed=109)
    o, h = make_pair(pr, orc.RAGGED, block=32, storage=storage)
    pi = np.array([[1e-6, 0.4, 0.3, 0.3]])
    o.set_scalars(o.mu, o.sigmaE, o.epsSum, o.sigmaG, pi)
    h.set_scalars(o.mu, o.sigmaE, o.epsSum, o.sigmaG, pi)
    order, u, z = sweep_inputs(pr.m, 0)
    o.sweep(order, u, z, trace=False)
    h.sweep(order, u, z)
    compare_state(o, h, orc.RAGGED, "all change")
    s = h.stats()
    assert s["changes"] >= 0.9 * pr.m
    h.close()


@pytest.mark.parametrize("layout,storage", LAYOUTS, ids=LAYOUT_IDS)
def test_per_marker_kernels(gpu_lib, layout, storage):
    pr = make_problem(3000, 50, seed=111, miss=0.02)
    o, h = make_pair(pr, layout, storage=storage)
    rng = np.random.default_rng(0)
    for j in rng.choice(pr.m, 12, replace=False):
        j = int(j)
        b_old, b_new = rng.normal(), rng.normal()
        close([h.compute_num(j, b_old)], [o.compute_num(j, b_old)], what="num")
        h.epsilon_change(j, b_old, b_new)
        o.apply_epsilon_change(j, b_old, b_new)
        st = h.get_state()
        if layout == orc.DENSE:
            assert np.array_equal(st["eps"], o.eps)   # same per-element operations
        else:
            assert np.array_equal(st["eps"], o.eps)
            close([h.get_scalars()["epsilon_sum"]], [o.epsSum], rtol=1e-8, what="epsSum")
    h.close()


@pytest.mark.parametrize("layout,storage", LAYOUTS, ids=LAYOUT_IDS)
def test_device_chain_matches_oracle_gibbs(gpu_lib, layout, storage):
    pr = make_problem(2500, 500, seed=113, h2=0.5, causal_frac=0.02)
    seed, iters = 20261202, 6
    o = orc.OracleProblem(pr.n, pr.m, layout, 4, 1, S, bed=pr.bed, lut=pr.lut) if layout == orc.DENSE \
        else orc.OracleProblem(pr.n, pr.m, layout, 4, 1, S, ragged=pr.ragged)
    hist_o, order_o = o.run_gibbs(pr.y, seed, iters)
    # device: host init (product) + the same cumulative Philox shuffles
    eps, sc = brr.bayesr_init(pr.y, 1, 4, seed)
    orders = np.empty((iters, pr.m), dtype=np.uint32)
    cur = np.arange(pr.m, dtype=np.uint32)
    for it in range(iters):
        brr.shuffle(seed, it, 0, cur)
        orders[it] = cur
    assert np.array_equal(orders[-1], order_o)
    h = brr.HipSweep(pr.n, pr.m, 4, 1, layout_code(layout), ragged_storage=storage)
    h.upload_dense(pr.bed, pr.lut) if layout == orc.DENSE else h.upload_ragged(pr.ragged)
    h.set_model(S)
    h.set_state(eps, np.zeros(pr.m), np.zeros(pr.m, np.int32))
    h.set_scalars(sc["mu"], sc["sigmaE"], sc["epsilon_sum"], sc["sigmaG"], sc["pi"])
    hist_h = h.run_chain(seed, 0, 0, orders)
    close(hist_h[:, 0], hist_o[:, 0], rtol=1e-8, what="mu")
    close(hist_h[:, 1], hist_o[:, 1], rtol=1e-8, what="sigmaE")
    assert np.array_equal(hist_h[:, 2], hist_o[:, 2]), "m0"
    close(hist_h[:, 3], hist_o[:, 3], rtol=1e-8, what="|eps|^2")
    close(hist_h[:, 4], hist_o[:, 4], rtol=1e-8, what="sigmaG")
    st = h.get_state()
    assert np.array_equal(st["comp"], o.comp)
    close(st["beta"], o.beta, rtol=1e-8, what="beta")
    close(st["eps"], o.eps, rtol=1e-8, what="eps")
    h.close()


@pytest.mark.parametrize("storage", [brr.STORAGE_LISTS, brr.STORAGE_CODES])
def test_large_ragged_residual_identity(gpu_lib, storage):
    """At N = 100,000: after a device chain eps equals y_std - mu - X beta
    (size-independent identity), recomputed on the host from the ragged lists."""
    pr = make_problem(100_000, 400, seed=115, maf=(0.005, 0.5), h2=0.4, causal_frac=0.02)
    seed = 3
    eps0, sc = brr.bayesr_init(pr.y, 1, 4, seed)
    h = brr.HipSweep(pr.n, pr.m, 4, 1, brr.LAYOUT_RAGGED, ragged_storage=storage)
    h.upload_ragged(pr.ragged)
    h.set_model(S)
    h.set_state(eps0, np.zeros(pr.m), np.zeros(pr.m, np.int32))
    h.set_scalars(sc["mu"], sc["sigmaE"], sc["epsilon_sum"], sc["sigmaG"], sc["pi"])
    orders = np.empty((4, pr.m), dtype=np.uint32)
    cur = np.arange(pr.m, dtype=np.uint32)
    for it in range(4):
        brr.shuffle(seed, it, 0, cur)
        orders[it] = cur
    h.run_chain(seed, 0, 0, orders)
    st = h.get_state()
    mu = h.get_scalars()["mu"]
    r = pr.ragged
    xb = np.zeros(pr.n)
    for j in np.flatnonzero(st["beta"]):
        mean, sd = r.stats[j, 0], r.stats[j, 1]
        col = np.full(pr.n, -mean / sd)
        col[r.ones[r.ones_off[j]: r.ones_off[j + 1]]] += 1 / sd
        col[r.twos[r.twos_off[j]: r.twos_off[j + 1]]] += 2 / sd
        xb += st["beta"][j] * col
    close(st["eps"], eps0 - mu - xb, rtol=1e-9, what="eps identity")
    close([h.get_scalars()["epsilon_sum"]], [st["eps"].sum()], rtol=1e-7, what="epsSum identity")
    h.close()


def test_bad_inputs_raise(gpu_lib):
    pr = make_problem(600, 20, seed=117)
    o, h = make_pair(pr, orc.RAGGED)
    bad = np.arange(pr.m, dtype=np.uint32)
    bad[3] = bad[4]
    u, z = brr.marker_uz(1, 0, 0, pr.m)
    with pytest.raises(brr.BrrHipError):
        h.sweep(bad, u, z)
    with pytest.raises(brr.BrrHipError):
        h.set_groups(np.full(pr.m, 5, np.int32))
    h.close()


@pytest.mark.parametrize("n", [300, 3001, 100_000, 200_000])
def test_default_launch_shape(gpu_lib, n):
    """Single-chain contexts take the fewest lane runs R, then the fewest workgroups holding N at R runs
    (slices a multiple of 16, at most 512 R), reported from creation on; an explicit count is honoured."""
    for make in (lambda **kw: brr.HipSweep(n, 10, 4, 1, brr.LAYOUT_DENSE_2BIT, **kw),
                 lambda **kw: brr.HipBayesW(n, 10, 4, **kw)):
        h = make()
        s = h.stats()
        P, S = s["workgroups"], s["slice"]
        R = -(-S // 512)
        assert S % 16 == 0 and (P - 1) * S < n <= P * S, (P, S)
        assert P == -(-n // (512 * R)), (P, S)
        if n <= 100_000:
            assert R == 1, (P, S)
        h.close()
        req = max(7, -(-n // 2048))                  # a lane holds at most 4 runs: 2048 per workgroup
        h = make(workgroups=req)
        s = h.stats()
        assert s["workgroups"] <= req and s["workgroups"] * s["slice"] >= n
        h.close()
