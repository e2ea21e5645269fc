"""The reference's preprocessed marker files read straight into the device arenas.

A preprocessed file (`.ragged.sparsebed`, `.ppbed`) is the markers' serialised
bytes back to back, with an index of {pos, compressedSize, originalSize} per
marker (preprocessgraph.cpp:92-126, compression.cpp:137-154); with --compress
each marker is one zlib stream. The files here are written by the oracle's
restatement of the reference's writers (RaggedSparseMarker::write, the dense
builder's column) and, for compression, by zlib itself; the library's readers
must give back exactly the arena that the .bed builders produce.
"""
import ctypes as C
import json
import zlib
from pathlib import Path

import numpy as np
import pytest

import bayesrrcmd_amd as brr
import _oracle as orc

GOLDEN = Path(__file__).resolve().parent / "golden"


def ragged_file(arena, compress=False):
    """Serialise every marker of a RaggedArena the way the reference's preprocess writes it."""
    L = orc.lib()
    parts, index, pos = [], [], 0
    m = arena.stats.shape[0]
    for j in range(m):
        lists = []
        for lst, off in ((arena.ones, arena.ones_off), (arena.twos, arena.twos_off), (arena.miss, arena.miss_off)):
            lists.append(np.ascontiguousarray(lst[int(off[j]): int(off[j + 1])], dtype=np.int32))
        st = np.ascontiguousarray(arena.stats[j], dtype=np.float64)
        args = []
        for a in lists:
            args += [orc.p(a, C.c_int32), a.size]
        size = L.orc_ragged_serialise(orc.p(st, C.c_double), *args, None)
        buf = np.zeros(size, dtype=np.uint8)
        L.orc_ragged_serialise(orc.p(st, C.c_double), *args, orc.p(buf, C.c_uint8))
        raw = buf.tobytes()
        blob = zlib.compress(raw) if compress else raw
        parts.append(blob)
        index.append((pos, len(blob), len(raw)))
        pos += len(blob)
    return b"".join(parts), np.array(index, dtype=np.uint64)


def dense_file(bed, n, compress=False):
    """Every column standardised by the oracle's DenseMarkerBuilder, N doubles each."""
    parts, index, pos, cols = [], [], 0, []
    for j in range(bed.shape[0]):
        cx = orc.dense_build(bed[j], n)
        cols.append(cx)
        raw = cx.tobytes()
        blob = zlib.compress(raw) if compress else raw
        parts.append(blob)
        index.append((pos, len(blob), len(raw)))
        pos += len(blob)
    return b"".join(parts), np.array(index, dtype=np.uint64), np.array(cols)


def same_arena(a, b):
    for f in ("ones", "ones_off", "twos", "twos_off", "miss", "miss_off"):
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f), err_msg=f)
    np.testing.assert_array_equal(a.stats.view(np.uint64), b.stats.view(np.uint64), err_msg="stats bits")


def lut_of_codes(bed, lut, n):
    """Per individual, the table entry its 2-bit code selects (the device's weight rule)."""
    bits = np.unpackbits(bed[..., None], axis=-1, bitorder="little").reshape(bed.shape[0], -1)
    code = (bits[:, 0::2] | (bits[:, 1::2] << 1))[:, :n]
    slot = np.array([2, 3, 1, 0])[code]          # code 00 -> dosage 2, 01 -> missing, 10 -> 1, 11 -> 0
    return np.take_along_axis(lut, slot, axis=1)


@pytest.mark.parametrize("compress", [False, True])
def test_ragged_sparsebed_round_trip(compress):
    n, m = 1003, 57
    bed, _ = brr.synth_bed(21, n, m, 0.01, 0.5, 0.03)
    arena = brr.build_ragged(bed, n)
    data, index = ragged_file(arena, compress)
    got = brr.read_ragged_sparsebed(data, index, n, compressed=compress, threads=3)
    same_arena(got, arena)


def test_ragged_sparsebed_real_fixture():
    """The reference's own test genotypes (uk10k slice) through build -> write -> read."""
    meta = json.loads((GOLDEN / "bed_fixtures.json").read_text())["uk10k_chr1_1mb"]
    raw = (GOLDEN / "uk10k_chr1_1mb.slice.bed").read_bytes()
    n = meta["num_inds"]
    stride = (n + 3) // 4
    cols = np.frombuffer(raw[3:], dtype=np.uint8).reshape(meta["slice_cols"], stride)
    arena = brr.build_ragged(cols, n)
    for compress in (False, True):
        data, index = ragged_file(arena, compress)
        same_arena(brr.read_ragged_sparsebed(data, index, n, compressed=compress), arena)


@pytest.mark.parametrize("compress", [False, True])
def test_ppbed_tables_reproduce_columns(compress):
    n, m = 1001, 40
    bed, _ = brr.synth_bed(22, n, m, 0.01, 0.5, 0.05)
    data, index, cols = dense_file(bed, n, compress)
    codes, lut = brr.read_ppbed(data, index, n, compressed=compress, threads=2)
    assert codes.shape == (m, (n + 3) // 4)
    got = lut_of_codes(codes, lut, n)
    np.testing.assert_array_equal(got.view(np.uint64), cols.view(np.uint64))
    # and the same column values as the .bed path's tables
    np.testing.assert_array_equal(lut_of_codes(bed, brr.build_dense(bed, n), n).view(np.uint64),
                                  cols.view(np.uint64))


def test_malformed_files_are_rejected():
    n, m = 200, 6
    bed, _ = brr.synth_bed(23, n, m, 0.05, 0.5, 0.0)
    arena = brr.build_ragged(bed, n)
    data, index = ragged_file(arena)
    with pytest.raises(brr.BrrHipError):                       # truncated file
        brr.read_ragged_sparsebed(data[: len(data) - 5], index, n)
    bad = index.copy()
    bad[2, 1] += 4                                             # size disagrees with the layout
    bad[2, 2] += 4
    with pytest.raises(brr.BrrHipError):
        brr.read_ragged_sparsebed(data, bad, n)
    with pytest.raises(brr.BrrHipError):                       # index beyond N
        brr.read_ragged_sparsebed(data, index, 10)
    with pytest.raises(brr.BrrHipError):                       # not a zlib stream
        brr.read_ragged_sparsebed(data, index, n, compressed=True)
    ddata, dindex, _ = dense_file(bed, n)
    col = np.frombuffer(ddata, dtype=np.float64).copy()
    col[:5] = [1.0, 2.0, 3.0, 4.0, 5.0]                        # five distinct values in column 0
    with pytest.raises(brr.BrrHipError):
        brr.read_ppbed(col.tobytes(), dindex, n)
