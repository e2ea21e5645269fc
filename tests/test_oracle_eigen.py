"""The EigenSparse layout (PreprocessDataType::SparseEigen): oracle pins and host checks.

EigenSparseMarkerBuilder (eigensparsemarkerbuilder.cpp:22-78) keeps the nonzero
calls of a column as a sparse vector Zg with the values 1 and 2, and the marker
mean (sum of the non-missing values / N) for a missing call; its statistics are
SparseMarker's with the missing calls added at that mean. EigenBayesRKernel
(eigenbayesrkernel.cpp) uses Zg in the dot and the ragged formulas otherwise.

Independent checks: the statistics, the dot and the residual update recomputed
from an explicit N-vector Zg built from the PLINK spec decode; without missing
calls every EigenSparse quantity equals the ragged one; the serialised layout
read back field by field.
"""
import ctypes as C
import zlib

import numpy as np
import pytest

import bayesrrcmd_amd as brr
import _oracle as orc
from _oracle import p
from test_oracle_markers import random_columns


def eigen_build(col, n):
    col = np.ascontiguousarray(col, dtype=np.uint8)
    o = np.zeros(n, dtype=np.int32); t = np.zeros(n, dtype=np.int32); m = np.zeros(n, dtype=np.int32)
    no, nt, nm = C.c_uint64(), C.c_uint64(), C.c_uint64()
    st = np.zeros(4, dtype=np.float64)
    ok = orc.lib().orc_eigen_build(p(col, C.c_uint8), n, p(o, C.c_int32), C.byref(no), p(t, C.c_int32),
                                   C.byref(nt), p(m, C.c_int32), C.byref(nm), p(st, C.c_double))
    return bool(ok), o[: no.value].copy(), t[: nt.value].copy(), m[: nm.value].copy(), st


def zg_of(dos):
    """Zg as a dense N-vector: 1 or 2 for those calls, the mean over N for a missing call, 0 otherwise."""
    n = dos.size
    mean = dos[dos > 0].sum() / n
    return np.where(dos == -1, mean, np.where(dos > 0, dos, 0)).astype(np.float64), mean


def oracle_arena(bed, n):
    """RaggedArena-like lists + EigenSparse statistics from the oracle, column by column."""
    cols = [eigen_build(bed[j], n) for j in range(bed.shape[0])]
    assert all(c[0] for c in cols)

    class A:
        pass
    a = A()
    for k, name in ((1, "ones"), (2, "twos"), (3, "miss")):
        lists = [c[k] for c in cols]
        setattr(a, name, np.ascontiguousarray(np.concatenate(lists), dtype=np.int32))
        setattr(a, name + "_off", np.concatenate([[0], np.cumsum([l.size for l in lists])]).astype(np.uint64))
    a.stats = np.ascontiguousarray(np.array([c[4] for c in cols]))
    return a


@pytest.mark.parametrize("miss", [0.0, 0.05, 0.3])
def test_eigen_statistics_against_definition(miss):
    rng = np.random.default_rng(11)
    n = 1203
    bed, dos = random_columns(rng, n, 12, miss=miss)
    for j in range(12):
        ok, ones, twos, ms, st = eigen_build(bed[j], n)
        assert ok
        d = dos[j]
        assert np.array_equal(ones, np.flatnonzero(d == 1))
        assert np.array_equal(twos, np.flatnonzero(d == 2))
        assert np.array_equal(ms, np.flatnonzero(d == -1))
        zg, mean = zg_of(d)
        assert st[0] == mean
        assert abs(st[2] - zg @ zg) <= 1e-12 * (zg @ zg)
        assert abs(st[3] - zg.sum()) <= 1e-12 * zg.sum()
        sd = np.sqrt(np.sum((zg - mean) ** 2) / (n - 1))      # the sd formula's closed form
        assert abs(st[1] - sd) <= 1e-12 * sd


def test_eigen_equals_ragged_without_missing():
    rng = np.random.default_rng(12)
    n = 777
    bed, _ = random_columns(rng, n, 8, miss=0.0)
    for j in range(8):
        e = eigen_build(bed[j], n)
        r = orc.ragged_build(bed[j], n)
        for a, b in zip(e[1:4], r[1:4]):
            assert np.array_equal(a, b)
        assert np.array_equal(e[4].view(np.uint64), r[4].view(np.uint64))


def test_eigen_validity():
    n = 9
    hom = np.full((n + 3) // 4, 0xFF, dtype=np.uint8)       # every call dosage 0: no nonzero
    assert not eigen_build(hom, n)[0]
    allmiss = np.full((n + 3) // 4, 0x55, dtype=np.uint8)    # every call missing: nonzeros at the mean 0
    ok, _, _, ms, st = eigen_build(allmiss, n)
    assert ok and ms.size == n and st[0] == 0.0
    assert not orc.ragged_build(allmiss, n)[0]               # the ragged builder rejects it


def test_eigen_dot_and_update_against_definition():
    rng = np.random.default_rng(13)
    n, m = 1501, 10
    bed, dos = random_columns(rng, n, m, miss=0.08)
    a = oracle_arena(bed, n)
    d = orc.OrcData()
    d.N, d.M, d.layout = n, m, orc.EIGEN
    d.ones, d.ones_off = p(a.ones, C.c_int32), p(a.ones_off, C.c_uint64)
    d.twos, d.twos_off = p(a.twos, C.c_int32), p(a.twos_off, C.c_uint64)
    d.miss, d.miss_off = p(a.miss, C.c_int32), p(a.miss_off, C.c_uint64)
    d.stats = p(a.stats, C.c_double)
    dr = orc.OrcData()
    C.memmove(C.byref(dr), C.byref(d), C.sizeof(d))
    dr.layout = orc.RAGGED
    L = orc.lib()
    eps = rng.normal(size=n)
    epsSum = float(eps.sum())
    for j in range(m):
        zg, mean = zg_of(dos[j])
        sd = a.stats[j, 1]
        got = L.orc_compute_num(C.byref(d), j, p(eps, C.c_double), 0.03, epsSum)
        want = 0.03 * (n - 1) - mean * epsSum / sd + (zg @ eps) / sd      # sparsebayesrkernel.cpp:24-27
        assert abs(got - want) <= 1e-12 * max(1.0, abs(want))
        # the missing calls are what separates it from the ragged dot
        rag = L.orc_compute_num(C.byref(dr), j, p(eps, C.c_double), 0.03, epsSum)
        miss_term = mean * eps[dos[j] == -1].sum() / sd
        assert abs((got - rag) - miss_term) <= 1e-10 * max(1.0, abs(miss_term))
        # eigenbayesrkernel.cpp:10-15, element by element in the reference's operation order
        e = eps.copy()
        es = np.array([epsSum])
        L.orc_apply_epsilon_change(C.byref(d), j, p(e, C.c_double), p(es, C.c_double), 0.02, -0.05)
        dBeta = 0.02 - (-0.05)
        want_e = eps + ((dBeta * zg) / sd - (dBeta * mean) / sd)
        assert np.array_equal(e.view(np.uint64), want_e.view(np.uint64))


@pytest.mark.parametrize("miss", [0.0, 0.04])
def test_host_builder_matches_oracle(miss):
    rng = np.random.default_rng(14)
    n, m = 2049, 30
    bed, _ = random_columns(rng, n, m, miss=miss)
    host = brr.build_eigen(bed, n, threads=3)
    ora = oracle_arena(bed, n)
    for f in ("ones", "ones_off", "twos", "twos_off", "miss", "miss_off"):
        assert np.array_equal(getattr(host, f), getattr(ora, f)), f
    assert np.array_equal(host.stats.view(np.uint64), ora.stats.view(np.uint64))
    mono = bed.copy()
    mono[4, :] = 0xFF
    with pytest.raises(brr.BrrHipError):
        brr.build_eigen(mono, n)


def eigen_file(arena, compress=False):
    L = orc.lib()
    parts, index, pos = [], [], 0
    for j in range(arena.stats.shape[0]):
        lists = []
        for lst, off in ((arena.ones, arena.ones_off), (arena.twos, arena.twos_off), (arena.miss, arena.miss_off)):
            lists.append(np.ascontiguousarray(lst[int(off[j]): int(off[j + 1])], dtype=np.int32))
        st = np.ascontiguousarray(arena.stats[j], dtype=np.float64)
        args = []
        for a in lists:
            args += [p(a, C.c_int32), a.size]
        size = L.orc_eigen_serialise(p(st, C.c_double), *args, None)
        buf = np.zeros(size, dtype=np.uint8)
        L.orc_eigen_serialise(p(st, C.c_double), *args, p(buf, C.c_uint8))
        raw = buf.tobytes()
        blob = zlib.compress(raw) if compress else raw
        parts.append(blob)
        index.append((pos, len(blob), len(raw)))
        pos += len(blob)
    return b"".join(parts), np.array(index, dtype=np.uint64)


def test_eigen_serialised_layout():
    st = np.array([0.75, 1.5, 3.0, 2.5])
    ones = np.array([1, 6], np.int32)
    twos = np.array([3], np.int32)
    miss = np.array([0, 4], np.int32)
    L = orc.lib()
    size = L.orc_eigen_serialise(p(st, C.c_double), p(ones, C.c_int32), 2, p(twos, C.c_int32), 1,
                                 p(miss, C.c_int32), 2, None)
    assert size == 32 + 8 + 5 * 12
    buf = np.zeros(size, np.uint8)
    L.orc_eigen_serialise(p(st, C.c_double), p(ones, C.c_int32), 2, p(twos, C.c_int32), 1,
                          p(miss, C.c_int32), 2, p(buf, C.c_uint8))
    raw = buf.tobytes()
    assert np.array_equal(np.frombuffer(raw[:32], np.float64), st)
    assert np.frombuffer(raw[32:40], np.int64)[0] == 5
    assert np.array_equal(np.frombuffer(raw[40:80], np.float64), [0.75, 1.0, 2.0, 0.75, 1.0])
    assert np.array_equal(np.frombuffer(raw[80:100], np.int32), [0, 1, 3, 4, 6])


@pytest.mark.parametrize("compress", [False, True])
def test_eigen_sparsebed_round_trip(compress):
    n, m = 1003, 41
    bed, _ = brr.synth_bed(31, n, m, 0.01, 0.5, 0.04)
    arena = brr.build_eigen(bed, n)
    data, index = eigen_file(arena, compress)
    got = brr.read_eigen_sparsebed(data, index, n, compressed=compress, threads=3)
    for f in ("ones", "ones_off", "twos", "twos_off", "miss", "miss_off"):
        assert np.array_equal(getattr(got, f), getattr(arena, f)), f
    assert np.array_equal(got.stats.view(np.uint64), arena.stats.view(np.uint64))
    with pytest.raises(brr.BrrHipError):
        brr.read_eigen_sparsebed(data[:-3], index, n, compressed=compress)
    bad = index.copy()
    bad[1, 1] += 8
    bad[1, 2] += 8
    with pytest.raises(brr.BrrHipError):
        brr.read_eigen_sparsebed(data, bad, n)
