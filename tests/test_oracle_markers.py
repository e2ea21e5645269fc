"""Pins the oracle's genotype decode and marker builders (oracle/markers.c), and
checks the product's host builders against it bit for bit.

Independent checks: a textbook PLINK 1 .bed decoder written from the format
specification (SNP-major, 2 bits per call, first individual in the low bits;
00 = homozygous first allele, 01 = missing, 10 = heterozygous, 11 = homozygous
second allele), the real .bed fixtures of the reference's tests (sliced into
tests/golden by make_fixtures.py), and closed forms of the standardisation
(sum x = 0, |x|^2 = N - 1) and of the ragged statistics.
"""
import ctypes as C
import json
from pathlib import Path

import numpy as np
import pytest

import _oracle as orc
from _oracle import p

GOLDEN = Path(__file__).resolve().parent / "golden"


def spec_decode(col: np.ndarray, n: int):
    """PLINK spec decode -> dosage of the first (counted) allele, -1 = missing."""
    b = np.unpackbits(col[:, None], axis=1, bitorder="little").reshape(-1)  # bit k of byte
    lo = b[0::2][:n].astype(int)
    hi = b[1::2][:n].astype(int)
    code = lo | (hi << 1)          # value of the two bits, first bit least significant
    table = np.array([2, -1, 1, 0])  # 00 hom first allele, 01 missing, 10 het, 11 hom second
    return table[code]


def random_columns(rng, n, m, miss=0.0):
    stride = (n + 3) // 4
    dos = rng.integers(0, 3, size=(m, n))
    if miss:
        dos[rng.random((m, n)) < miss] = -1
    code = np.select([dos == 2, dos == -1, dos == 1, dos == 0], [0, 1, 2, 3])
    padded = np.zeros((m, stride * 4), dtype=np.uint8)
    padded[:, :n] = code
    # pad calls beyond N with random bits: the decoders must ignore them
    padded[:, n:] = rng.integers(0, 4, size=(m, stride * 4 - n))
    q = padded.reshape(m, stride, 4)
    bed = (q[..., 0] | (q[..., 1] << 2) | (q[..., 2] << 4) | (q[..., 3] << 6)).astype(np.uint8)
    return bed, dos


def load_fixture(name):
    meta = json.loads((GOLDEN / "bed_fixtures.json").read_text())[name]
    raw = (GOLDEN / f"{name}.slice.bed").read_bytes()
    n = meta["num_inds"]
    stride = (n + 3) // 4
    cols = np.frombuffer(raw[3:], dtype=np.uint8).reshape(meta["slice_cols"], stride)
    return meta, raw[:3], cols


# ------------------------------------------------------------------ a1 ----

@pytest.mark.parametrize("n", [1, 3, 4, 5, 17, 1000, 1003])
def test_decode_matches_spec(n):
    rng = np.random.default_rng(n)
    bed, dos = random_columns(rng, n, 20, miss=0.1)
    for j in range(20):
        a1, a2 = orc.decode(bed[j], n)
        d = np.where((a1 == 0) & (a2 == 1), -1, a1.astype(int) + a2)
        assert np.array_equal(d, dos[j])
        assert np.array_equal(spec_decode(bed[j], n), dos[j])


def test_decode_code_table():
    # one byte holding the four codes 00, 01, 10, 11 for individuals 0..3
    byte = np.array([0b11_10_01_00], dtype=np.uint8)
    a1, a2 = orc.decode(byte, 4)
    assert list(zip(a1.tolist(), a2.tolist())) == [(1, 1), (0, 1), (1, 0), (0, 0)]


@pytest.mark.parametrize("name", ["uk10k_chr1_1mb", "gauss_data"])
def test_real_bed_fixture(name):
    meta, magic, cols = load_fixture(name)
    assert orc.lib().orc_check_bed_magic(p(np.frombuffer(magic, np.uint8).copy(), C.c_uint8)) == 0
    n = meta["num_inds"]
    assert meta["bytes"] == 3 + meta["num_snps"] * ((n + 3) // 4)
    for j in range(cols.shape[0]):
        a1, a2 = orc.decode(cols[j], n)
        d = np.where((a1 == 0) & (a2 == 1), -1, a1.astype(int) + a2)
        assert np.array_equal(d, spec_decode(cols[j], n))


def test_bad_magic_rejected():
    assert orc.lib().orc_check_bed_magic(p(np.array([0x6c, 0x1b, 0x00], np.uint8), C.c_uint8)) != 0
    import bayesrrcmd_amd as brr
    assert brr.lib().brr_bed_check_magic(p(np.array([0x6c, 0x1b, 0x01], np.uint8), C.c_uint8)) == 0
    assert brr.lib().brr_bed_check_magic(p(np.array([0x6c, 0x1c, 0x01], np.uint8), C.c_uint8)) != 0


# ------------------------------------------------------------------ a3 ----

@pytest.mark.parametrize("miss", [0.0, 0.05])
def test_dense_standardisation(miss):
    rng = np.random.default_rng(3)
    n = 2001
    bed, dos = random_columns(rng, n, 10, miss=miss)
    for j in range(10):
        cx = orc.dense_build(bed[j], n)
        # closed form: impute missing with the non-missing mean, centre, scale to |x|^2 = N - 1
        g = dos[j].astype(float)
        obs = g >= 0
        g[~obs] = g[obs].mean()
        x = g - g.mean()
        x /= np.sqrt((x @ x) / (n - 1))
        assert np.allclose(cx, x, rtol=0, atol=1e-13)
        assert abs(cx.sum()) < 1e-9
        assert abs(cx @ cx - (n - 1)) < 1e-9
        # the 4-entry table reproduces every element exactly
        lut = orc.dense_lut(bed[j], n)
        d = np.where(dos[j] < 0, 3, dos[j])
        assert np.array_equal(lut[d], cx)


# ------------------------------------------------------------------ a2 ----

@pytest.mark.parametrize("miss", [0.0, 0.05])
def test_ragged_lists_and_stats(miss):
    rng = np.random.default_rng(4)
    n = 1501
    bed, dos = random_columns(rng, n, 10, miss=miss)
    for j in range(10):
        ok, ones, twos, ms, st = orc.ragged_build(bed[j], n)
        assert ok
        d = dos[j]
        assert np.array_equal(ones, np.flatnonzero(d == 1))
        assert np.array_equal(twos, np.flatnonzero(d == 2))
        assert np.array_equal(ms, np.flatnonzero(d == -1))
        zsum = d[d > 0].sum()
        sq = (d[d > 0] ** 2).sum()
        mean = zsum / n                                   # divides by N, missing included
        sd = np.sqrt((sq - 2 * mean * zsum + n * mean * mean) / (n - 1))
        assert st[2] == sq and st[3] == zsum              # integer sums are exact
        assert st[0] == mean
        assert abs(st[1] - sd) < 1e-15 * sd


def test_dense_and_ragged_mean_differ_with_missing():
    """Dense imputes the mean over non-missing calls (densemarkerbuilder.cpp:50);
    ragged divides by N (raggedsparsemarkerbuilder.cpp:49)."""
    n = 8
    dos = np.array([2, 1, -1, 0, 2, -1, 1, 1])
    code = np.select([dos == 2, dos == -1, dos == 1, dos == 0], [0, 1, 2, 3]).astype(np.uint8)
    col = np.array([code[0] | code[1] << 2 | code[2] << 4 | code[3] << 6,
                    code[4] | code[5] << 2 | code[6] << 4 | code[7] << 6], dtype=np.uint8)
    ok, ones, twos, ms, st = orc.ragged_build(col, n)
    assert st[0] == 7 / 8
    cx = orc.dense_build(col, n)
    g = np.array([2, 1, 7 / 6, 0, 2, 7 / 6, 1, 1])
    x = (g - g.mean())
    x /= np.sqrt(x @ x / (n - 1))
    assert np.allclose(cx, x, atol=1e-14)


def test_ragged_rejects_monomorphic():
    n = 9
    col = np.full((n + 3) // 4, 0xFF, dtype=np.uint8)   # all homozygous second allele (dosage 0)
    ok, *_ = orc.ragged_build(col, n)
    assert not ok


def test_ragged_serialised_layout():
    st = np.array([0.5, 1.25, 3.0, 2.0])
    ones = np.array([1, 4], np.int32)
    twos = np.array([7], np.int32)
    ms = np.array([], np.int32)
    size = orc.lib().orc_ragged_serialise(p(st, C.c_double), p(ones, C.c_int32), 2, p(twos, C.c_int32), 1,
                                          p(ms, C.c_int32), 0, None)
    assert size == 4 * 8 + 3 * 8 + 3 * 4
    buf = np.zeros(size, np.uint8)
    orc.lib().orc_ragged_serialise(p(st, C.c_double), p(ones, C.c_int32), 2, p(twos, C.c_int32), 1,
                                   p(ms, C.c_int32), 0, p(buf, C.c_uint8))
    b = buf.tobytes()
    assert np.frombuffer(b[:32], np.float64).tolist() == st.tolist()
    assert np.frombuffer(b[32:40], np.uint64)[0] == 2
    assert np.frombuffer(b[40:48], np.int32).tolist() == [1, 4]
    assert np.frombuffer(b[48:56], np.uint64)[0] == 1
    assert np.frombuffer(b[56:60], np.int32).tolist() == [7]
    assert np.frombuffer(b[60:68], np.uint64)[0] == 0


# ------------------------------------------ product host builders == oracle

@pytest.mark.parametrize("miss", [0.0, 0.03])
def test_product_builders_match_oracle(miss):
    import bayesrrcmd_amd as brr
    rng = np.random.default_rng(5)
    n, m = 1234, 40
    bed, dos = random_columns(rng, n, m, miss=miss)
    lut = brr.build_dense(bed, n, threads=4)
    arena = brr.build_ragged(bed, n, threads=3)
    for j in range(m):
        assert np.array_equal(lut[j], orc.dense_lut(bed[j], n))
        ok, ones, twos, ms, st = orc.ragged_build(bed[j], n)
        sl = slice(int(arena.ones_off[j]), int(arena.ones_off[j + 1]))
        assert np.array_equal(arena.ones[sl], ones)
        sl = slice(int(arena.twos_off[j]), int(arena.twos_off[j + 1]))
        assert np.array_equal(arena.twos[sl], twos)
        sl = slice(int(arena.miss_off[j]), int(arena.miss_off[j + 1]))
        assert np.array_equal(arena.miss[sl], ms)
        assert np.array_equal(arena.stats[j], st)


def test_product_builders_on_real_fixture():
    import bayesrrcmd_amd as brr
    meta, _, cols = load_fixture("uk10k_chr1_1mb")
    n = meta["num_inds"]
    lut = brr.build_dense(cols, n)
    for j in range(cols.shape[0]):
        assert np.array_equal(lut[j], orc.dense_lut(cols[j], n))


def test_synthetic_generator_properties():
    import bayesrrcmd_amd as brr
    n, m = 4000, 64
    bed, maf = brr.synth_bed(7, n, m, 0.05, 0.5, 0.0)
    bed2, maf2 = brr.synth_bed(7, n, m, 0.05, 0.5, 0.0, threads=1)
    assert np.array_equal(bed, bed2) and np.array_equal(maf, maf2)     # thread-count independent
    for j in range(m):
        d = spec_decode(bed[j], n)
        assert (d >= 0).all()
        f = d.mean() / 2
        assert abs(f - maf[j]) < 4 * np.sqrt(maf[j] * (1 - maf[j]) / (2 * n)) + 1e-3
        # Hardy-Weinberg: P(het) = 2p(1-p)
        het = (d == 1).mean()
        assert abs(het - 2 * maf[j] * (1 - maf[j])) < 0.04
