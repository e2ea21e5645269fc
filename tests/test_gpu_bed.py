"""A whole .bed file decoded on the device (brr_hip_load_bed) against the host builders.

The device decode replaces the host preprocess pass (PreprocessGraph with the
dense / ragged marker builders, src/preprocessgraph.cpp:24-60). Bar:
- per-marker tables (dense) and statistics (ragged) equal to brr_dense_luts /
  brr_ragged_fill bit for bit (counts are exact integers; the dense column sums
  run in individual order as on the host);
- every sweep (plain, batched traits, BayesW) from a decoded context equal bit
  for bit to the same sweep from the host-built upload, and the plain sweep
  within the parity bar of the oracle;
- per-marker ragged kernels on codes alone (no index lists held) equal to the
  list path: epsilon update bit for bit, computeNum to 1e-12 relative
  (summation order differs).
"""
import json
from pathlib import Path

import numpy as np
import pytest

import bayesrrcmd_amd as brr
import _oracle as orc
from _data import make_problem, make_bw_problem

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
S = np.array([0.01, 0.001, 0.0001])


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def same_values(a, b, what):
    """Bitwise equality; NaNs only need to sit in the same places (their payload is the platform's)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    assert a.shape == b.shape, f"{what}: shapes {a.shape} vs {b.shape}"
    a, b = a.reshape(-1), b.reshape(-1)
    na, nb = np.isnan(a), np.isnan(b)
    assert np.array_equal(na, nb), f"{what}: NaN positions differ"
    diff = np.flatnonzero((bits(a) != bits(b)) & ~na)
    assert diff.size == 0, f"{what}: {diff.size} values differ, first at flat index {diff[:5]}: " \
                           f"{a[diff[:3]]} vs {b[diff[:3]]}"


def garble_pad(bed, n, seed):
    """Random bits in the pad positions (individuals >= N) of every column's last byte."""
    out = bed.copy()
    r = n % 4
    if r:
        rng = np.random.default_rng(seed)
        keep = (1 << (2 * r)) - 1
        noise = rng.integers(0, 256, out.shape[0], dtype=np.uint16).astype(np.uint8)
        out[:, -1] = (out[:, -1] & keep) | (noise & ~np.uint8(keep))
    return out


@pytest.mark.parametrize("n,m,miss", [(2, 7, 0.0), (3, 5, 0.2), (37, 40, 0.0), (1001, 97, 0.05),
                                      (4099, 64, 0.0), (20003, 33, 0.02), (100003, 24, 0.01)])
def test_tables_bit_exact(gpu_lib, n, m, miss):
    bed, _ = brr.synth_bed(300 + n, n, m, 0.05, 0.5, miss)
    bed = garble_pad(bed, n, n)
    f = brr.bed_file_image(bed)
    # dense: every marker, degenerate ones included (host and device both give NaN / inf there)
    with brr.HipSweep(n, m, 4, layout=brr.LAYOUT_DENSE_2BIT) as h:
        h.load_bed(f, brr.LAYOUT_DENSE_2BIT)
        same_values(h.marker_tables(), brr.build_dense(bed, n), f"dense tables N={n}")
    # ragged: the host rejects a marker without genotype 1 / 2 calls, and so must the device
    try:
        host = brr.build_ragged(bed, n)
    except brr.BrrHipError:
        host = None
    with brr.HipSweep(n, m, 4, layout=brr.LAYOUT_RAGGED) as h:
        if host is None:
            with pytest.raises(brr.BrrHipError, match="no genotype 1 or 2"):
                h.load_bed(f, brr.LAYOUT_RAGGED)
        else:
            h.load_bed(f, brr.LAYOUT_RAGGED)
            same_values(h.marker_tables(), host.stats, f"ragged statistics N={n}")


def oracle_tables(bed, n, layout):
    """The oracle's per-marker tables: orc_dense_lut (DenseMarkerBuilder) or the
    statistics of orc_ragged_build (RaggedSparseMarkerBuilder)."""
    if layout == brr.LAYOUT_DENSE_2BIT:
        return np.stack([orc.dense_lut(bed[j], n) for j in range(bed.shape[0])])
    return np.stack([orc.ragged_build(bed[j], n)[4] for j in range(bed.shape[0])])


@pytest.mark.parametrize("n,m,miss", [(37, 40, 0.0), (1001, 97, 0.05), (20003, 33, 0.02), (100003, 24, 0.01)])
def test_tables_equal_the_oracle_builders(gpu_lib, n, m, miss):
    """The device decode against the oracle's builders directly (not through the
    product's host builders): dense tables and ragged statistics bit for bit."""
    bed, _ = brr.synth_bed(700 + n, n, m, 0.05, 0.5, miss)
    bed = garble_pad(bed, n, n)
    f = brr.bed_file_image(bed)
    for layout in (brr.LAYOUT_DENSE_2BIT, brr.LAYOUT_RAGGED):
        with brr.HipSweep(n, m, 4, layout=layout) as h:
            h.load_bed(f, layout)
            same_values(h.marker_tables(), oracle_tables(bed, n, layout), f"layout {layout} N={n} vs oracle")


def test_degenerate_columns(gpu_lib):
    """All missing, all one dosage, a single non-missing call: tables as the host computes them."""
    n, m = 101, 6
    cb = (n + 3) // 4
    bed, _ = brr.synth_bed(5, n, m, 0.1, 0.5, 0.0)
    bed[0, :] = 0x55          # every call missing
    bed[1, :] = 0xFF          # every individual dosage 0
    bed[2, :] = 0x00          # every individual dosage 2
    bed[3, :] = 0x55
    bed[3, 0] = 0x56          # one heterozygote among missing calls
    assert bed.shape == (m, cb)
    with brr.HipSweep(n, m, 4, layout=brr.LAYOUT_DENSE_2BIT) as h:
        h.load_bed(brr.bed_file_image(bed), brr.LAYOUT_DENSE_2BIT)
        same_values(h.marker_tables(), brr.build_dense(bed, n), "degenerate dense tables")


def test_real_fixture(gpu_lib):
    """The reference's own uk10k genotypes, read from the file itself."""
    meta = json.loads((GOLDEN / "bed_fixtures.json").read_text())["uk10k_chr1_1mb"]
    path = GOLDEN / "uk10k_chr1_1mb.slice.bed"
    n, m = meta["num_inds"], meta["slice_cols"]
    raw = path.read_bytes()
    cols = np.frombuffer(raw[3:], dtype=np.uint8).reshape(m, (n + 3) // 4)
    with brr.HipSweep(n, m, 4, layout=brr.LAYOUT_DENSE_2BIT) as h:
        h.load_bed(str(path), brr.LAYOUT_DENSE_2BIT)
        same_values(h.marker_tables(), brr.build_dense(cols, n), "uk10k dense tables")
    with brr.HipSweep(n, m, 4, layout=brr.LAYOUT_RAGGED) as h:
        h.load_bed(str(path), brr.LAYOUT_RAGGED)
        same_values(h.marker_tables(), brr.build_ragged(cols, n).stats, "uk10k ragged statistics")
        same_values(h.marker_tables(), oracle_tables(cols, n, brr.LAYOUT_RAGGED), "uk10k ragged vs oracle")
    with brr.HipSweep(n, m, 4, layout=brr.LAYOUT_DENSE_2BIT) as h:
        h.load_bed(str(path), brr.LAYOUT_DENSE_2BIT)
        same_values(h.marker_tables(), oracle_tables(cols, n, brr.LAYOUT_DENSE_2BIT), "uk10k dense vs oracle")


def sweep_inputs(m, it, seed=77):
    order = np.arange(m, dtype=np.uint32)
    for t in range(it + 1):
        brr.shuffle(seed, t, 0, order)
    u, z = brr.marker_uz(seed, it, 0, m)
    return order, u, z


def plain_ctx(pr, layout, o, decoded, workgroups=0):
    storage = brr.STORAGE_CODES if layout == brr.LAYOUT_RAGGED else brr.STORAGE_AUTO
    h = brr.HipSweep(pr.n, pr.m, 4, 1, layout, workgroups=workgroups, ragged_storage=storage)
    if decoded:
        h.load_bed(brr.bed_file_image(pr.bed_file_cols), layout)
    elif layout == brr.LAYOUT_DENSE_2BIT:
        h.upload_dense(pr.bed, pr.lut)
    else:
        h.upload_ragged(pr.ragged)
    h.set_model(S)
    h.set_state(o.eps, o.beta, o.comp)
    h.set_scalars(o.mu, o.sigmaE, o.epsSum, o.sigmaG, o.pi)
    return h


def same_state(a, b, what):
    sa, sb = a.get_state(), b.get_state()
    for k in ("eps", "beta", "acum"):
        same_values(sa[k], sb[k], f"{what} {k}")
    assert np.array_equal(sa["comp"], sb["comp"]), f"{what} comp"


@pytest.mark.parametrize("layout", [brr.LAYOUT_DENSE_2BIT, brr.LAYOUT_RAGGED], ids=["dense", "ragged"])
@pytest.mark.parametrize("n,m,wg", [(2001, 150, 0), (1003, 77, 7)])
def test_plain_sweep_from_decoded_file(gpu_lib, layout, n, m, wg):
    pr = make_problem(n, m, seed=n + m, h2=0.5, causal_frac=0.05, miss=0.02)
    pr.bed_file_cols = garble_pad(pr.bed, n, 3)
    olay = orc.DENSE if layout == brr.LAYOUT_DENSE_2BIT else orc.RAGGED
    if olay == orc.DENSE:
        o = orc.OracleProblem(n, m, orc.DENSE, 4, 1, S, bed=pr.bed, lut=pr.lut)
    else:
        o = orc.OracleProblem(n, m, orc.RAGGED, 4, 1, S, ragged=pr.ragged)
    o.init(pr.y, 5)
    hu = plain_ctx(pr, layout, o, decoded=False, workgroups=wg)
    hd = plain_ctx(pr, layout, o, decoded=True, workgroups=wg)
    for it in range(2):
        order, u, z = sweep_inputs(m, it)
        hu.sweep(order, u, z)
        hd.sweep(order, u, z)
        o.sweep(order, u, z, trace=False)
    same_state(hd, hu, "decoded vs uploaded")
    st = hd.get_state()
    assert np.array_equal(st["comp"], o.comp)
    scale = max(1.0, float(np.max(np.abs(o.eps))))
    assert float(np.max(np.abs(st["eps"] - o.eps))) <= 1e-9 * scale
    hu.close()
    hd.close()


def test_per_marker_ragged_on_codes(gpu_lib):
    n, m = 3001, 40
    pr = make_problem(n, m, seed=17, miss=0.03)
    o = orc.OracleProblem(n, m, orc.RAGGED, 4, 1, S, ragged=pr.ragged)
    o.init(pr.y, 5)
    lists = brr.HipSweep(n, m, 4, 1, brr.LAYOUT_RAGGED, ragged_storage=brr.STORAGE_LISTS)
    lists.upload_ragged(pr.ragged)
    codes = brr.HipSweep(n, m, 4, 1, brr.LAYOUT_RAGGED)
    codes.load_bed(brr.bed_file_image(pr.bed), brr.LAYOUT_RAGGED)
    for h in (lists, codes):
        h.set_model(S)
        h.set_state(o.eps, o.beta, o.comp)
        h.set_scalars(o.mu, o.sigmaE, o.epsSum, o.sigmaG, o.pi)
    for j in (0, 7, 39):
        a, b = lists.compute_num(j, 0.013), codes.compute_num(j, 0.013)
        assert abs(a - b) <= 1e-12 * max(1.0, abs(a)), f"computeNum marker {j}: {a} vs {b}"
        lists.epsilon_change(j, 0.0, 0.021)
        codes.epsilon_change(j, 0.0, 0.021)
        same_values(codes.get_state()["eps"], lists.get_state()["eps"], f"epsilon after marker {j}")
        assert codes.get_scalars()["epsilon_sum"] == lists.get_scalars()["epsilon_sum"]
    lists.close()
    codes.close()


@pytest.mark.parametrize("layout", [brr.LAYOUT_DENSE_2BIT, brr.LAYOUT_RAGGED], ids=["dense", "ragged"])
def test_batched_sweep_from_decoded_file(gpu_lib, layout):
    n, m, T = 2003, 120, 5
    pr = make_problem(n, m, seed=41, miss=0.02)
    rng = np.random.default_rng(3)
    eps = rng.normal(0.0, 1.0, (n, T))
    hs = []
    for decoded in (False, True):
        h = brr.HipSweep(n, m, 4, 1, layout, traits=T)
        if decoded:
            h.load_bed(brr.bed_file_image(garble_pad(pr.bed, n, 9)), layout)
        elif layout == brr.LAYOUT_DENSE_2BIT:
            h.upload_dense(pr.bed, pr.lut)
        else:
            h.upload_ragged(pr.ragged)
        h.set_model(S)
        h.set_state(eps, np.zeros((m, T)), np.zeros((m, T), np.int32))
        for t in range(T):
            h.set_scalars(0.1 * t, 0.8, float(eps[:, t].sum()), np.array([0.2]), np.array([0.5, 0.3, 0.1, 0.1]),
                          trait=t)
        hs.append(h)
    for it in range(2):
        order, u, z = sweep_inputs(m, it)
        uu = np.repeat(u[:, None], T, axis=1)
        zz = np.repeat(z[:, None], T, axis=1)
        for h in hs:
            h.sweep(order, uu, zz)
    same_state(hs[1], hs[0], "batched decoded vs uploaded")
    for h in hs:
        h.close()


@pytest.mark.parametrize("layout", [brr.LAYOUT_DENSE_2BIT, brr.LAYOUT_RAGGED], ids=["dense", "ragged"])
def test_bayesw_sweep_from_decoded_file(gpu_lib, layout):
    n, m, K = 3001, 96, 4
    pr = make_bw_problem(n, m, seed=8, miss=0.02)
    C = np.array([0.001, 0.01, 0.1])
    lut = brr.build_dense(pr.bed, n)
    eps0, init = brr.bayesw_init(pr.y, pr.fail, m, K)
    hs = []
    for decoded in (False, True):
        h = brr.HipBayesW(n, m, K, quad_points=7)
        if decoded:
            h.load_bed(brr.bed_file_image(garble_pad(pr.bed, n, 4)), layout)
        elif layout == brr.LAYOUT_DENSE_2BIT:
            h.upload_dense(pr.bed, lut)
        else:
            h.upload_ragged(pr.ragged)
        h.set_model(C, pr.fail)
        h.set_state(eps0, np.zeros(m), np.zeros(m, np.int32))
        h.set_scalars(init["mu"], init["alpha"], init["sigma_b"], init["d"], init["pi"], np.ones(K))
        hs.append(h)
    for it in range(2):
        order = np.arange(m, dtype=np.uint32)
        brr.shuffle(11, it, 0, order)
        u, _ = brr.marker_uz(11, it, 0, m)
        for h in hs:
            h.sweep(order, u, 11, it)
    same_state_bw = [h.get_state() for h in hs]
    for k in ("eps", "beta"):
        same_values(same_state_bw[1][k], same_state_bw[0][k], f"BayesW decoded vs uploaded {k}")
    assert np.array_equal(same_state_bw[1]["comp"], same_state_bw[0]["comp"])
    assert np.count_nonzero(same_state_bw[0]["beta"]) > 0
    for h in hs:
        h.close()


def test_bad_files(gpu_lib):
    n, m = 203, 12
    bed, _ = brr.synth_bed(2, n, m, 0.1, 0.5, 0.0)
    f = brr.bed_file_image(bed)
    h = brr.HipSweep(n, m, 4, layout=brr.LAYOUT_RAGGED)
    with pytest.raises(brr.BrrHipError, match="magic"):
        h.load_bed(np.concatenate([[0x6C, 0x1B, 0x00], f[3:]]).astype(np.uint8), brr.LAYOUT_RAGGED)
    with pytest.raises(brr.BrrHipError, match="file size"):
        h.load_bed(f[:-1], brr.LAYOUT_RAGGED)
    with pytest.raises(brr.BrrHipError, match="layout"):
        h.load_bed(f, brr.LAYOUT_DENSE_2BIT)
    mono = bed.copy()
    mono[5, :] = 0xFF
    mono[9, :] = 0x55
    with pytest.raises(brr.BrrHipError, match="marker 5 has no genotype"):
        h.load_bed(brr.bed_file_image(mono), brr.LAYOUT_RAGGED)
    h.set_model(S)
    with pytest.raises(brr.BrrHipError, match="markers not uploaded"):
        order, u, z = sweep_inputs(m, 0)
        h.sweep(order, u, z)
    h.load_bed(f, brr.LAYOUT_RAGGED)                 # a good file afterwards is accepted
    same_values(h.marker_tables(), brr.build_ragged(bed, n).stats, "statistics after a rejected file")
    h.close()
