"""MI355X-native BayesR marker sweep (BayesRRcmd `ppbayes`), Python face.

Everything here calls the C ABI of libbrr_hip.so (include/brr_hip.h); the
arrays are numpy. The device path has no CPU fallback: constructing a
HipSweep on a machine without a HIP device raises BrrHipError.

Reference correspondence (BayesRRcmd src/):
  HipSweep.sweep        AnalysisGraph::exec -> BayesRBase::processColumn
  HipSweep.run_chain    BayesRBase::runGibbs iterations, device-resident
  build_dense           DenseMarkerBuilder (as 2-bit codes + 4-entry table)
  build_ragged          RaggedSparseMarkerBuilder
  bayesr_init           BayesRBase::init
  shuffle               std::random_shuffle of the marker order
  marker_uz             BayesRBase::prepareForAnylsis
  HipBayesW.sweep       AnalysisGraph::exec -> BayesWBase::processColumn (ragged)
  HipBayesW.run_chain   BayesWBase::runGibbs iterations, device-resident
  bayesw_init           BayesWBase::init
"""
from __future__ import annotations

import ctypes as C
import os
import sys
from dataclasses import dataclass

import numpy as np

from . import _abi
from ._abi import (LAYOUT_DENSE_2BIT, LAYOUT_EIGEN_SPARSE, LAYOUT_RAGGED, PRECISION_FP32, PRECISION_FP32_REFRESH,
                   PRECISION_FP64,
                   STORAGE_AUTO, STORAGE_CODES, STORAGE_LISTS)

__all__ = [
    "BrrHipError", "HipSweep", "HipBayesW", "bayesw_init", "bed_file_image", "read_samples_binary", "RaggedArena", "read_ragged_sparsebed", "read_ppbed", "SampleWriter", "sample_due", "build_dense", "build_ragged", "synth_bed", "synth_xb",
    "bayesr_init", "shuffle", "marker_uz", "LAYOUT_DENSE_2BIT", "LAYOUT_RAGGED", "LAYOUT_EIGEN_SPARSE", "lib",
    "build_eigen", "read_eigen_sparsebed",
    "STORAGE_AUTO", "STORAGE_CODES", "STORAGE_LISTS", "PRECISION_FP64", "PRECISION_FP32_REFRESH", "PRECISION_FP32",
]


class BrrHipError(RuntimeError):
    pass


def lib():
    return _abi.load()


def _p(a: np.ndarray | None, ctype):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


def _threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


def _check(rc: int, what: str, ctx=None):
    if rc != 0:
        msg = ""
        if ctx is not None:
            m = lib().brr_hip_last_error(ctx)
            msg = m.decode() if m else ""
        raise BrrHipError(f"{what} failed with status {rc}: {msg}")


# --------------------------------------------------------------- builders --

def synth_bed(seed: int, n: int, m: int, maf_lo: float, maf_hi: float, miss_rate: float = 0.0,
              threads: int | None = None):
    """Synthetic SNP-major PLINK payload (m x ceil(n/4) bytes) and the MAF per SNP."""
    stride = (n + 3) // 4
    bed = np.empty((m, stride), dtype=np.uint8)
    maf = np.empty(m, dtype=np.float64)
    rc = lib().brr_synth_bed(seed, n, m, maf_lo, maf_hi, miss_rate, _p(bed, C.c_uint8), stride,
                             _p(maf, C.c_double), threads or _threads())
    if rc != 0:
        raise BrrHipError(f"brr_synth_bed failed with status {rc}")
    return bed, maf


def synth_xb(bed: np.ndarray, n: int, lut: np.ndarray, beta: np.ndarray, threads: int | None = None):
    """X beta over the standardised columns: (n,) for beta (m,), (n, T) for beta (m, T)."""
    bed = np.ascontiguousarray(bed, dtype=np.uint8)
    lut = np.ascontiguousarray(lut, dtype=np.float64)
    b = np.ascontiguousarray(beta, dtype=np.float64)
    T = 1 if b.ndim == 1 else b.shape[1]
    out = np.empty((n, T), dtype=np.float64)
    rc = lib().brr_synth_xb(_p(bed, C.c_uint8), bed.shape[1], n, bed.shape[0], _p(lut, C.c_double), T,
                            _p(b, C.c_double), _p(out, C.c_double), threads or _threads())
    if rc != 0:
        raise BrrHipError(f"brr_synth_xb failed with status {rc}")
    return out[:, 0] if b.ndim == 1 else out


def build_dense(bed: np.ndarray, n: int, threads: int | None = None) -> np.ndarray:
    """DenseMarkerBuilder tables: (m, 4) doubles for dosage 0, 1, 2, missing."""
    bed = np.ascontiguousarray(bed, dtype=np.uint8)
    m, stride = bed.shape
    lut = np.empty((m, 4), dtype=np.float64)
    _check(lib().brr_dense_luts(_p(bed, C.c_uint8), stride, n, m, _p(lut, C.c_double),
                                threads or _threads()), "brr_dense_luts")
    return lut


@dataclass
class RaggedArena:
    ones: np.ndarray
    ones_off: np.ndarray
    twos: np.ndarray
    twos_off: np.ndarray
    miss: np.ndarray
    miss_off: np.ndarray
    stats: np.ndarray      # (m, 4): mean, sd, sqrdZ, Zsum

    def nbytes(self) -> int:
        return int(4 * (self.ones.size + self.twos.size + self.miss.size))


def build_ragged(bed: np.ndarray, n: int, threads: int | None = None) -> RaggedArena:
    """RaggedSparseMarkerBuilder over every column of a .bed payload."""
    bed = np.ascontiguousarray(bed, dtype=np.uint8)
    m, stride = bed.shape
    t = threads or _threads()
    offs = [np.zeros(m + 1, dtype=np.uint64) for _ in range(3)]
    _check(lib().brr_ragged_count(_p(bed, C.c_uint8), stride, n, m, *(_p(o, C.c_uint64) for o in offs), t),
           "brr_ragged_count")
    lists = [np.empty(max(1, int(o[-1])), dtype=np.int32) for o in offs]
    stats = np.empty((m, 4), dtype=np.float64)
    _check(lib().brr_ragged_fill(_p(bed, C.c_uint8), stride, n, m, *(_p(o, C.c_uint64) for o in offs),
                                 *(_p(l, C.c_int32) for l in lists), _p(stats, C.c_double), t),
           "brr_ragged_fill (a column has no genotype 1 or 2)")
    lists = [l[: int(o[-1])] for l, o in zip(lists, offs)]
    return RaggedArena(lists[0], offs[0], lists[1], offs[1], lists[2], offs[2], stats)


def build_eigen(bed: np.ndarray, n: int, threads: int | None = None) -> RaggedArena:
    """EigenSparseMarkerBuilder over every column: the ragged lists with the EigenSparse statistics."""
    bed = np.ascontiguousarray(bed, dtype=np.uint8)
    m, stride = bed.shape
    t = threads or _threads()
    offs = [np.zeros(m + 1, dtype=np.uint64) for _ in range(3)]
    _check(lib().brr_ragged_count(_p(bed, C.c_uint8), stride, n, m, *(_p(o, C.c_uint64) for o in offs), t),
           "brr_ragged_count")
    lists = [np.empty(max(1, int(o[-1])), dtype=np.int32) for o in offs]
    stats = np.empty((m, 4), dtype=np.float64)
    _check(lib().brr_eigen_fill(_p(bed, C.c_uint8), stride, n, m, *(_p(o, C.c_uint64) for o in offs),
                                *(_p(l, C.c_int32) for l in lists), _p(stats, C.c_double), t),
           "brr_eigen_fill (a column has no nonzero call)")
    lists = [l[: int(o[-1])] for l, o in zip(lists, offs)]
    return RaggedArena(lists[0], offs[0], lists[1], offs[1], lists[2], offs[2], stats)


def _read_sparse(prefix: str, data, index, n: int, compressed: bool, threads) -> RaggedArena:
    buf = np.frombuffer(data, dtype=np.uint8) if isinstance(data, (bytes, bytearray)) else np.ascontiguousarray(data, np.uint8)
    idx = np.ascontiguousarray(index, dtype=np.uint64).reshape(-1, 3)
    m = idx.shape[0]
    t = threads or _threads()
    offs = [np.zeros(m + 1, dtype=np.uint64) for _ in range(3)]
    _check(getattr(lib(), f"{prefix}_read_counts")(_p(buf, C.c_uint8), buf.size, _p(idx, C.c_uint64), m,
                                                   int(compressed), *(_p(o, C.c_uint64) for o in offs), t),
           f"{prefix}_read_counts")
    lists = [np.empty(max(1, int(o[-1])), dtype=np.int32) for o in offs]
    stats = np.empty((m, 4), dtype=np.float64)
    _check(getattr(lib(), f"{prefix}_read_fill")(_p(buf, C.c_uint8), buf.size, _p(idx, C.c_uint64), n, m,
                                                 int(compressed), *(_p(o, C.c_uint64) for o in offs),
                                                 *(_p(l, C.c_int32) for l in lists), _p(stats, C.c_double), t),
           f"{prefix}_read_fill")
    lists = [l[: int(o[-1])] for l, o in zip(lists, offs)]
    return RaggedArena(lists[0], offs[0], lists[1], offs[1], lists[2], offs[2], stats)


def read_ragged_sparsebed(data: bytes | np.ndarray, index: np.ndarray, n: int, compressed: bool = False,
                          threads: int | None = None) -> RaggedArena:
    """A .ragged.sparsebed file's bytes and its index (M x 3 u64) -> RaggedArena."""
    return _read_sparse("brr_ragged", data, index, n, compressed, threads)


def read_eigen_sparsebed(data: bytes | np.ndarray, index: np.ndarray, n: int, compressed: bool = False,
                         threads: int | None = None) -> RaggedArena:
    """A .eigen.sparsebed file's bytes and its index (M x 3 u64) -> RaggedArena with EigenSparse statistics."""
    return _read_sparse("brr_eigen", data, index, n, compressed, threads)


def read_ppbed(data: bytes | np.ndarray, index: np.ndarray, n: int, compressed: bool = False,
               threads: int | None = None):
    """A .ppbed file's bytes and its index (M x 3 u64) -> (bed codes (M, ceil(n/4)), tables (M, 4))."""
    buf = np.frombuffer(data, dtype=np.uint8) if isinstance(data, (bytes, bytearray)) else np.ascontiguousarray(data, np.uint8)
    idx = np.ascontiguousarray(index, dtype=np.uint64).reshape(-1, 3)
    m = idx.shape[0]
    bed = np.empty((m, (n + 3) // 4), dtype=np.uint8)
    lut = np.empty((m, 4), dtype=np.float64)
    _check(lib().brr_dense_read_ppbed(_p(buf, C.c_uint8), buf.size, _p(idx, C.c_uint64), n, m, int(compressed),
                                      _p(bed, C.c_uint8), _p(lut, C.c_double), threads or _threads()),
           "brr_dense_read_ppbed")
    return bed, lut


class SampleWriter:
    """The reference's sample file (SampleWriter): kind "bayesr" or "bayesw"; CSV, or with binary=True the
    same rows as raw doubles (read back with read_samples_binary)."""

    def __init__(self, path: str, kind: str, m: int, n: int = 0, groups: int = 1, fixed: int = 0,
                 binary: bool = False):
        self.kind = {"bayesr": 0, "bayesw": 1}[kind] + (2 if binary else 0)
        self.m, self.n, self.g, self.f = m, n, groups, fixed
        h = C.c_void_p()
        _check(lib().brr_samples_open(str(path).encode(), self.kind, m, n, groups, fixed, C.byref(h)),
               "brr_samples_open")
        self.h = h

    def write_bayesr(self, iteration, mu, beta, sigmaE, sigmaG, comp, acum, eps, gamma=None):
        b = np.ascontiguousarray(beta, np.float64)
        sg = np.ascontiguousarray(sigmaG, np.float64).reshape(-1)
        gm = None if gamma is None else np.ascontiguousarray(gamma, np.float64)
        c = np.ascontiguousarray(comp, np.int32)
        a = np.ascontiguousarray(acum, np.float64)
        e = np.ascontiguousarray(eps, np.float64)
        _check(lib().brr_samples_write_bayesr(self.h, iteration, mu, _p(b, C.c_double), sigmaE, _p(sg, C.c_double),
                                              _p(gm, C.c_double), _p(c, C.c_int32), _p(a, C.c_double),
                                              _p(e, C.c_double)), "brr_samples_write_bayesr")

    def write_bayesw(self, iteration, alpha, mu, beta, comp, sigma_b, theta=None):
        b = np.ascontiguousarray(beta, np.float64)
        c = np.ascontiguousarray(comp, np.int32)
        th = None if theta is None else np.ascontiguousarray(theta, np.float64)
        _check(lib().brr_samples_write_bayesw(self.h, iteration, alpha, mu, _p(th, C.c_double), _p(b, C.c_double),
                                              _p(c, C.c_int32), sigma_b), "brr_samples_write_bayesw")

    def close(self):
        if getattr(self, "h", None):
            _check(lib().brr_samples_close(self.h), "brr_samples_close")
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def read_samples_binary(path: str):
    """A binary sample file: (header dict {kind, M, N, G, F}, rows as an (iterations, row length) array)."""
    raw = np.fromfile(path, dtype=np.uint8)
    if raw.size < 32 or raw[:8].tobytes() != b"BRRSAMP1":
        raise BrrHipError(f"{path}: not a binary sample file")
    kind, m, n, g, f, row_len = (int(v) for v in raw[8:32].view(np.uint32))
    body = raw[32:]
    if body.size % (8 * row_len):
        raise BrrHipError(f"{path}: truncated row")
    rows = body.view(np.float64).reshape(-1, row_len)
    return dict(kind=kind, M=m, N=n, G=g, F=f), rows


def sample_due(iteration: int, burn_in: int, thinning: int) -> bool:
    return bool(lib().brr_sample_due(iteration, burn_in, thinning))


# ---------------------------------------------------------- host driver --

def bayesr_init(y: np.ndarray, n_groups: int, n_components: int, seed: int, chain: int = 0):
    """BayesRBase::init: returns (eps, scalars dict)."""
    y = np.ascontiguousarray(y, dtype=np.float64)
    n = y.size
    eps = np.empty(n, dtype=np.float64)
    sigmaG = np.empty(n_groups, dtype=np.float64)
    pi = np.empty(n_groups * n_components, dtype=np.float64)
    s = _abi.Scalars(0.0, 0.0, 0.0, _p(sigmaG, C.c_double), _p(pi, C.c_double), None, None)
    _check(lib().brr_bayesr_init(_p(y, C.c_double), n, n_groups, n_components, seed, chain,
                                 _p(eps, C.c_double), C.byref(s)), "brr_bayesr_init")
    return eps, dict(mu=s.mu, sigmaE=s.sigmaE, epsilon_sum=s.epsilon_sum, sigmaG=sigmaG,
                     pi=pi.reshape(n_groups, n_components))


def bayesw_init(y: np.ndarray, fail: np.ndarray, m: int, n_components: int):
    """BayesWBase::init: returns (eps, scalars dict with mu, alpha, sigma_b, d, pi, v)."""
    y = np.ascontiguousarray(y, dtype=np.float64)
    f = np.ascontiguousarray(fail, dtype=np.float64)
    n = y.size
    eps = np.empty(n, dtype=np.float64)
    pi = np.empty(n_components, dtype=np.float64)
    v = np.empty(n_components, dtype=np.float64)
    s = _abi.BwScalars(0.0, 0.0, 0.0, 0.0, _p(pi, C.c_double), _p(v, C.c_double), 0.0)
    _check(lib().brr_bayesw_init(_p(y, C.c_double), _p(f, C.c_double), n, m, n_components, _p(eps, C.c_double),
                                 C.byref(s)), "brr_bayesw_init")
    return eps, dict(mu=s.mu, alpha=s.alpha, sigma_b=s.sigma_b, d=s.d, pi=pi, v=v)


def shuffle(seed: int, iteration: int, chain: int, order: np.ndarray, mode: int = 0) -> np.ndarray:
    """In-place std::random_shuffle of order (uint32) over the stream (mode 0) or rand() (mode 1)."""
    assert order.dtype == np.uint32 and order.flags.c_contiguous
    _check(lib().brr_shuffle(seed, iteration, chain, order.size, _p(order, C.c_uint32), mode), "brr_shuffle")
    return order


def marker_uz(seed: int, iteration: int, chain: int, m: int):
    u = np.empty(m, dtype=np.float64)
    z = np.empty(m, dtype=np.float64)
    _check(lib().brr_marker_uz(seed, iteration, chain, m, _p(u, C.c_double), _p(z, C.c_double)),
           "brr_marker_uz")
    return u, z


def run_cohorts(sweeps, seed: int, first_iteration: int, orders, chains=None, graph: bool = True):
    """Independent chains of several HipSweep contexts side by side
    (brr_hip_run_cohorts): orders[q] is an (iterations, M_q) array of marker
    orders. graph=False launches one after another instead of through a hipGraph.
    Returns the per-cohort histories (iterations x (4 + G)) and a report
    {launches, device_ms, launch_of}."""
    n = len(sweeps)
    ctxs = (C.c_void_p * n)(*[h.ctx for h in sweeps])
    ords = [np.ascontiguousarray(o, dtype=np.uint32) for o in orders]
    iters = ords[0].shape[0]
    if any(o.shape != (iters, h.m) for o, h in zip(ords, sweeps)):
        raise ValueError("orders[q] must have shape (iterations, M_q)")
    optrs = (_abi.P_u32 * n)(*[_p(o, C.c_uint32) for o in ords])
    hists = [np.zeros((iters, 4 + h.G)) for h in sweeps]
    hptrs = (_abi.P_f64 * n)(*[_p(x, C.c_double) for x in hists])
    ch = np.ascontiguousarray(chains if chains is not None else np.zeros(n), dtype=np.uint32)
    rep = _abi.CohortReport()
    rc = lib().brr_hip_run_cohorts(ctxs, n, int(seed), _p(ch, C.c_uint32), int(first_iteration), iters, optrs,
                                   hptrs, 0 if graph else 1, C.byref(rep))
    if rc != 0:
        raise BrrHipError(f"brr_hip_run_cohorts failed with status {rc}: "
                          f"{lib().brr_hip_last_error(sweeps[0].ctx).decode(errors='replace')}")
    return hists, dict(launches=rep.launches, device_ms=rep.device_ms, launch_of=list(rep.launch_of[: min(n, 64)]))


def calibration_copy(nbytes: int = 1 << 30, device: int = 0) -> float:
    """One device-to-device copy of nbytes by a kernel of this library (a reference
    for HBM counters); returns its device milliseconds."""
    ms = C.c_double()
    _check(lib().brr_hip_calibration_copy(int(device), int(nbytes), C.byref(ms)), "brr_hip_calibration_copy")
    return ms.value


def bed_file_image(bed: np.ndarray) -> np.ndarray:
    """A PLINK .bed file's bytes (magic 6c 1b 01, then the SNP-major columns) from M x ceil(N/4) column bytes."""
    bed = np.ascontiguousarray(bed, dtype=np.uint8)
    return np.concatenate([np.array([0x6C, 0x1B, 0x01], dtype=np.uint8), bed.reshape(-1)])


def _as_file_bytes(bed_file) -> np.ndarray:
    if isinstance(bed_file, (str, os.PathLike)):
        return np.fromfile(bed_file, dtype=np.uint8)
    if isinstance(bed_file, (bytes, bytearray, memoryview)):
        return np.frombuffer(bed_file, dtype=np.uint8)
    return np.ascontiguousarray(bed_file, dtype=np.uint8).reshape(-1)


class _DeviceMarkers:
    """Calls shared by the BayesR and BayesW contexts: the device-side .bed decode."""

    def load_bed(self, bed_file, layout: int):
        """Decode a whole .bed file (path, bytes or uint8 array, magic included) on the device into
        `layout` markers (PreprocessGraph + the marker builders, src/preprocessgraph.cpp:24-60).
        Returns the device time of the decode in ms: {"unpack": ..., "tables": ...}."""
        f = _as_file_bytes(bed_file)
        self._ck(lib().brr_hip_load_bed(self.ctx, _p(f, C.c_uint8), f.size, int(layout)), "brr_hip_load_bed")
        ph = np.zeros(12)
        lib().brr_hip_last_phase_ms(self.ctx, _p(ph, C.c_double), 12)
        return {"unpack": float(ph[0]), "tables": float(ph[1])}

    def marker_tables(self) -> np.ndarray:
        """M x 4: dense tables {dosage 0, 1, 2, missing} or ragged statistics {mean, sd, sqrdZ, Zsum}."""
        out = np.zeros((self.m, 4))
        self._ck(lib().brr_hip_get_marker_tables(self.ctx, _p(out, C.c_double)), "brr_hip_get_marker_tables")
        return out


# ---------------------------------------------------------- device sweep --

class HipSweep(_DeviceMarkers):
    """Device context: marker arena + chain state on one MI355X (HipSweepGraph)."""

    def __init__(self, n: int, m: int, n_components: int, n_groups: int = 1,
                 layout: int = LAYOUT_RAGGED, workgroups: int = 0, block_markers: int = 0,
                 device: int = 0, ragged_storage: int = 0, traits: int = 0):
        """traits = 0: one chain; traits = T >= 1: T chains over the same columns and
        order (state arrays gain a trailing trait axis)."""
        self.n, self.m, self.K, self.G = n, m, n_components, n_groups
        self.layout = layout
        self.traits = traits
        self.Tn = max(1, traits)
        cfg = _abi.Config(n, m, n_components, n_groups, layout, workgroups, block_markers, device,
                          ragged_storage, traits)
        ctx = _abi.CTX()
        rc = lib().brr_hip_create(C.byref(cfg), C.byref(ctx))
        if rc != 0:
            raise BrrHipError(f"brr_hip_create failed with status {rc} (N={n}, M={m}, K={n_components}, "
                              f"G={n_groups}, layout={layout}, workgroups={workgroups}, traits={traits})")
        self.ctx = ctx
        self._keep = []

    def close(self):
        if getattr(self, "ctx", None):
            lib().brr_hip_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        # no HIP calls while the interpreter shuts down: a profiler or the runtime
        # may already be torn down by then
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _ck(self, rc, what):
        _check(rc, what, self.ctx)

    # uploads
    def upload_dense(self, bed: np.ndarray, lut: np.ndarray):
        bed = np.ascontiguousarray(bed, dtype=np.uint8)
        lut = np.ascontiguousarray(lut, dtype=np.float64)
        self._ck(lib().brr_hip_upload_dense(self.ctx, _p(bed, C.c_uint8), bed.shape[1], _p(lut, C.c_double)),
                 "brr_hip_upload_dense")

    def upload_ragged(self, a: RaggedArena):
        self._ck(lib().brr_hip_upload_ragged(
            self.ctx, _p(a.ones, C.c_int32), _p(a.ones_off, C.c_uint64), _p(a.twos, C.c_int32),
            _p(a.twos_off, C.c_uint64), _p(a.miss, C.c_int32), _p(a.miss_off, C.c_uint64),
            _p(np.ascontiguousarray(a.stats), C.c_double)), "brr_hip_upload_ragged")

    def set_fixed(self, X: np.ndarray):
        """Fixed-effect covariates, (N, F): stored as the reference's float matrix (column-major)."""
        x = np.asarray(X, dtype=np.float32).reshape(self.n, -1)
        self.F = x.shape[1]
        self._fixedX = np.ascontiguousarray(x.T)          # column f contiguous
        self._ck(lib().brr_hip_set_fixed(self.ctx, _p(self._fixedX, C.c_float), self.F), "brr_hip_set_fixed")

    def set_fixed_state(self, gamma=None, order=None, sigma_f: float = 1.0):
        g = None if gamma is None else np.ascontiguousarray(gamma, dtype=np.float64).reshape(self.F)
        o = None if order is None else np.ascontiguousarray(order, dtype=np.uint32).reshape(self.F)
        self._ck(lib().brr_hip_set_fixed_state(self.ctx, _p(g, C.c_double), _p(o, C.c_uint32), sigma_f),
                 "brr_hip_set_fixed_state")

    def get_fixed_state(self):
        g = np.empty(self.F)
        o = np.empty(self.F, dtype=np.uint32)
        sf = C.c_double()
        self._ck(lib().brr_hip_get_fixed_state(self.ctx, _p(g, C.c_double), _p(o, C.c_uint32), C.byref(sf)),
                 "brr_hip_get_fixed_state")
        return dict(gamma=g, order=o, sigma_f=sf.value)

    def fixed_effects(self, seed: int, iteration: int, chain: int = 0):
        """The fixed-effects block alone on the device state (BayesRBase.cpp:212-235)."""
        self._ck(lib().brr_hip_fixed_effects(self.ctx, seed, iteration, chain), "brr_hip_fixed_effects")

    def set_async(self, window: int):
        """ParallelGraph's asynchronous sweep in windows of `window` markers (0: the exact sequential sweep)."""
        self._ck(lib().brr_hip_set_async(self.ctx, int(window)), "brr_hip_set_async")

    def set_precision(self, precision: int, refresh_markers: int = 1024):
        """PRECISION_FP64 (default) or PRECISION_FP32_REFRESH: FP32 residual slices and dots with
        compensated wave reductions, the FP64 residual brought up to date every refresh_markers."""
        self._ck(lib().brr_hip_set_precision(self.ctx, int(precision), int(refresh_markers)),
                 "brr_hip_set_precision")

    def kernel_compute_num(self, marker: int, epsilon: np.ndarray, epsilon_sum: float, beta_old: float) -> float:
        """BayesRKernel::computeNum(epsilon, beta_old) on a host residual (HipBayesRKernel)."""
        e = np.ascontiguousarray(epsilon, dtype=np.float64)
        if e.shape != (self.n,):
            raise ValueError(f"epsilon must have shape ({self.n},)")
        out = C.c_double()
        self._ck(lib().brr_hip_kernel_compute_num(self.ctx, int(marker), _p(e, C.c_double), float(epsilon_sum),
                                                  float(beta_old), C.byref(out)), "brr_hip_kernel_compute_num")
        return out.value

    def kernel_epsilon_change(self, marker: int, beta_old: float, beta: float):
        """BayesRKernel::calculateEpsilonChange(beta_old, beta): (residual change, epsilonSum change)."""
        d = np.zeros(self.n)
        es = C.c_double()
        self._ck(lib().brr_hip_kernel_epsilon_change(self.ctx, int(marker), float(beta_old), float(beta),
                                                     _p(d, C.c_double), C.byref(es)), "brr_hip_kernel_epsilon_change")
        return d, es.value

    def set_launch(self, workgroups: int = 0, block_markers: int = 0):
        """Launch shape: workgroups (0: default rule) and markers per block (0: keep)."""
        self._ck(lib().brr_hip_set_launch(self.ctx, int(workgroups), int(block_markers)), "brr_hip_set_launch")

    def autotune(self, seed: int, iterations: int = 2) -> list:
        """Times `iterations` chain iterations from the current state per candidate launch
        shape, keeps the fastest and the state; returns the table of trials."""
        tab = (_abi.TuneEntry * 32)()
        n = C.c_uint32()
        self._ck(lib().brr_hip_autotune(self.ctx, int(seed), int(iterations), tab, 32, C.byref(n)),
                 "brr_hip_autotune")
        return [dict(workgroups=e.workgroups, block_markers=e.block_markers, slice=e.slice,
                     ms_per_iteration=e.ms_per_iteration) for e in tab[: n.value]]

    def probe_exchange(self, rounds: int = 2000) -> float:
        """Device microseconds per exchange round of a full block with no other work."""
        out = C.c_double()
        self._ck(lib().brr_hip_probe_exchange(self.ctx, int(rounds), C.byref(out)), "brr_hip_probe_exchange")
        return out.value

    def record_num(self, enable: bool = True):
        """Record each marker's num (computeNum) in later sweeps; read with get_num()."""
        self._ck(lib().brr_hip_record_num(self.ctx, int(bool(enable))), "brr_hip_record_num")

    def get_num(self) -> np.ndarray:
        out = np.zeros(self.m)
        self._ck(lib().brr_hip_get_num(self.ctx, _p(out, C.c_double)), "brr_hip_get_num")
        return out

    def set_groups(self, group: np.ndarray):
        g = np.ascontiguousarray(group, dtype=np.int32)
        self._ck(lib().brr_hip_set_groups(self.ctx, _p(g, C.c_int32)), "brr_hip_set_groups")

    def set_model(self, cva: np.ndarray, v0E=1e-4, s02E=1e-4, v0G=1e-4, s02G=1e-4):
        cva = np.ascontiguousarray(cva, dtype=np.float64).reshape(self.G, self.K - 1)
        self._cva = cva
        mdl = _abi.Model(_p(cva, C.c_double), v0E, s02E, v0G, s02G)
        self._ck(lib().brr_hip_set_model(self.ctx, C.byref(mdl)), "brr_hip_set_model")

    # state
    def _shape(self, rows):
        return (rows, self.traits) if self.traits else (rows,)

    def _arr(self, a, rows, dtype, what):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=dtype)
        if a.shape != self._shape(rows):
            raise ValueError(f"{what}: shape {a.shape}, expected {self._shape(rows)}")
        return a

    def set_state(self, eps=None, beta=None, comp=None):
        """eps (N,), beta (M,), comp (M,); (N, T) / (M, T) for a batched context."""
        e = self._arr(eps, self.n, np.float64, "eps")
        b = self._arr(beta, self.m, np.float64, "beta")
        c = self._arr(comp, self.m, np.int32, "comp")
        self._ck(lib().brr_hip_set_state(self.ctx, _p(e, C.c_double), _p(b, C.c_double), _p(c, C.c_int32)),
                 "brr_hip_set_state")

    def set_scalars(self, mu, sigmaE, epsilon_sum, sigmaG, pi, trait: int = 0):
        sg = np.ascontiguousarray(sigmaG, dtype=np.float64).reshape(self.G)
        pp = np.ascontiguousarray(pi, dtype=np.float64).reshape(self.G * self.K)
        s = _abi.Scalars(mu, sigmaE, epsilon_sum, _p(sg, C.c_double), _p(pp, C.c_double), None, None)
        self._ck(lib().brr_hip_set_trait_scalars(self.ctx, trait, C.byref(s)), "brr_hip_set_trait_scalars")

    def get_state(self):
        eps = np.empty(self._shape(self.n), dtype=np.float64)
        beta = np.empty(self._shape(self.m), dtype=np.float64)
        comp = np.empty(self._shape(self.m), dtype=np.int32)
        acum = np.empty(self._shape(self.m), dtype=np.float64)
        self._ck(lib().brr_hip_get_state(self.ctx, _p(eps, C.c_double), _p(beta, C.c_double),
                                         _p(comp, C.c_int32), _p(acum, C.c_double)), "brr_hip_get_state")
        return dict(eps=eps, beta=beta, comp=comp, acum=acum)

    def get_scalars(self, trait: int = 0):
        sg = np.empty(self.G, dtype=np.float64)
        pi = np.empty(self.G * self.K, dtype=np.float64)
        v = np.empty(self.G * self.K, dtype=np.float64)
        bs = np.empty(self.G, dtype=np.float64)
        s = _abi.Scalars(0.0, 0.0, 0.0, _p(sg, C.c_double), _p(pi, C.c_double), _p(v, C.c_double),
                         _p(bs, C.c_double))
        self._ck(lib().brr_hip_get_trait_scalars(self.ctx, trait, C.byref(s)), "brr_hip_get_trait_scalars")
        return dict(mu=s.mu, sigmaE=s.sigmaE, epsilon_sum=s.epsilon_sum, sigmaG=sg,
                    pi=pi.reshape(self.G, self.K), v=v.reshape(self.G, self.K), betasqn=bs)

    # compute
    def sweep(self, order: np.ndarray, u: np.ndarray, z: np.ndarray):
        """u, z by marker id: (M,), or (M, T) for a batched context."""
        o = np.ascontiguousarray(order, dtype=np.uint32)
        uu = self._arr(u, self.m, np.float64, "u")
        zz = self._arr(z, self.m, np.float64, "z")
        self._ck(lib().brr_hip_sweep(self.ctx, _p(o, C.c_uint32), _p(uu, C.c_double), _p(zz, C.c_double)),
                 "brr_hip_sweep")

    def run_chain(self, seed: int, chain: int, first_iteration: int, orders: np.ndarray) -> np.ndarray:
        """hist: (iterations, 4 + G), or (iterations, T, 5) for a batched context."""
        o = np.ascontiguousarray(orders, dtype=np.uint32)
        iters = o.size // self.m
        assert iters * self.m == o.size
        shape = (iters, self.traits, 4 + self.G) if self.traits else (iters, 4 + self.G)
        hist = np.empty(shape, dtype=np.float64)
        self._ck(lib().brr_hip_run_chain(self.ctx, seed, chain, first_iteration, iters, _p(o, C.c_uint32),
                                         _p(hist, C.c_double)), "brr_hip_run_chain")
        return hist

    def compute_num(self, marker: int, beta_old: float) -> float:
        out = C.c_double()
        self._ck(lib().brr_hip_compute_num(self.ctx, marker, beta_old, C.byref(out)), "brr_hip_compute_num")
        return out.value

    def epsilon_change(self, marker: int, beta_old: float, beta: float):
        self._ck(lib().brr_hip_epsilon_change(self.ctx, marker, beta_old, beta), "brr_hip_epsilon_change")

    PHASES = ("block_load", "dots", "exchange", "decisions", "commit_update", "fp64_refresh_and_iteration_scalars")
    BATCH_PHASES = ("block_commit", "dot_tiles", "exchange_gather", "decision_draws", "residual_update",
                    "iteration_scalars", "gram_tiles", "uz_draws", "exchange_reduce", "accept_correct",
                    "block_end_barrier")

    def stats(self) -> dict:
        s = _abi.Stats()
        self._ck(lib().brr_hip_last_stats(self.ctx, C.byref(s)), "brr_hip_last_stats")
        out = {f: getattr(s, f) for f, _ in _abi.Stats._fields_}
        ph = np.zeros(12)
        self._ck(lib().brr_hip_last_phase_ms(self.ctx, _p(ph, C.c_double), 12), "brr_hip_last_phase_ms")
        names = self.BATCH_PHASES if self.traits else self.PHASES
        out["phase_ms"] = dict(zip(names, ph.tolist()))
        return out


# ------------------------------------------------------------ BayesW sweep --

class HipBayesW(_DeviceMarkers):
    """Device context of the BayesW (Weibull survival) sweep on ragged markers."""

    def __init__(self, n: int, m: int, n_components: int, quad_points: int = 7, workgroups: int = 0,
                 device: int = 0):
        self.n, self.m, self.K, self.nq = n, m, n_components, quad_points
        cfg = _abi.BwConfig(n, m, n_components, quad_points, workgroups, device)
        ctx = _abi.CTX()
        rc = lib().brr_hip_bw_create(C.byref(cfg), C.byref(ctx))
        if rc != 0:
            raise BrrHipError(f"brr_hip_bw_create failed with status {rc} (N={n}, M={m}, K={n_components}, "
                              f"quad_points={quad_points}, workgroups={workgroups})")
        self.ctx = ctx

    def close(self):
        if getattr(self, "ctx", None):
            lib().brr_hip_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        # no HIP calls while the interpreter shuts down: a profiler or the runtime
        # may already be torn down by then
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _ck(self, rc, what):
        _check(rc, what, self.ctx)

    def upload_ragged(self, a: RaggedArena):
        self._ck(lib().brr_hip_upload_ragged(
            self.ctx, _p(a.ones, C.c_int32), _p(a.ones_off, C.c_uint64), _p(a.twos, C.c_int32),
            _p(a.twos_off, C.c_uint64), _p(a.miss, C.c_int32), _p(a.miss_off, C.c_uint64),
            _p(np.ascontiguousarray(a.stats), C.c_double)), "brr_hip_upload_ragged")

    def upload_dense(self, bed: np.ndarray, lut: np.ndarray):
        """Dense markers (DenseBayesW): 2-bit columns and their standardised tables."""
        bed = np.ascontiguousarray(bed, dtype=np.uint8)
        lut = np.ascontiguousarray(lut, dtype=np.float64)
        self._ck(lib().brr_hip_upload_dense(self.ctx, _p(bed, C.c_uint8), bed.shape[1], _p(lut, C.c_double)),
                 "brr_hip_upload_dense")

    def set_fixed(self, X: np.ndarray):
        """Fixed-effect covariates, (N, F): stored as the reference's float matrix (column-major)."""
        x = np.asarray(X, dtype=np.float32).reshape(self.n, -1)
        self.F = x.shape[1]
        self._fixedX = np.ascontiguousarray(x.T)
        self._ck(lib().brr_hip_bw_set_fixed(self.ctx, _p(self._fixedX, C.c_float), self.F), "brr_hip_bw_set_fixed")

    def set_theta(self, theta):
        t = np.ascontiguousarray(theta, dtype=np.float64).reshape(self.F)
        self._ck(lib().brr_hip_bw_set_theta(self.ctx, _p(t, C.c_double)), "brr_hip_bw_set_theta")

    def get_theta(self) -> np.ndarray:
        t = np.empty(self.F)
        self._ck(lib().brr_hip_bw_get_theta(self.ctx, _p(t, C.c_double)), "brr_hip_bw_get_theta")
        return t

    def fixed_effects(self, seed: int, iteration: int, chain: int = 0):
        """Every sampleTheta of one iteration on the device state."""
        self._ck(lib().brr_hip_bw_fixed_effects(self.ctx, seed, iteration, chain), "brr_hip_bw_fixed_effects")

    def set_model(self, mixture_C: np.ndarray, fail: np.ndarray):
        cc = np.ascontiguousarray(mixture_C, dtype=np.float64).reshape(self.K - 1)
        f = np.ascontiguousarray(fail, dtype=np.float64).reshape(self.n)
        self._ck(lib().brr_hip_bw_set_model(self.ctx, _p(cc, C.c_double), _p(f, C.c_double)),
                 "brr_hip_bw_set_model")

    def set_state(self, eps=None, beta=None, comp=None):
        e = None if eps is None else np.ascontiguousarray(eps, dtype=np.float64).reshape(self.n)
        b = None if beta is None else np.ascontiguousarray(beta, dtype=np.float64).reshape(self.m)
        c = None if comp is None else np.ascontiguousarray(comp, dtype=np.int32).reshape(self.m)
        self._ck(lib().brr_hip_set_state(self.ctx, _p(e, C.c_double), _p(b, C.c_double), _p(c, C.c_int32)),
                 "brr_hip_set_state")

    def get_state(self):
        eps = np.empty(self.n, dtype=np.float64)
        beta = np.empty(self.m, dtype=np.float64)
        comp = np.empty(self.m, dtype=np.int32)
        self._ck(lib().brr_hip_get_state(self.ctx, _p(eps, C.c_double), _p(beta, C.c_double),
                                         _p(comp, C.c_int32), None), "brr_hip_get_state")
        return dict(eps=eps, beta=beta, comp=comp)

    def set_scalars(self, mu, alpha, sigma_b, d, pi, v=None):
        pp = np.ascontiguousarray(pi, dtype=np.float64).reshape(self.K)
        vv = None if v is None else np.ascontiguousarray(v, dtype=np.float64).reshape(self.K)
        s = _abi.BwScalars(mu, alpha, sigma_b, d, _p(pp, C.c_double), _p(vv, C.c_double), 0.0)
        self._ck(lib().brr_hip_bw_set_scalars(self.ctx, C.byref(s)), "brr_hip_bw_set_scalars")

    def get_scalars(self):
        pi = np.empty(self.K, dtype=np.float64)
        v = np.empty(self.K, dtype=np.float64)
        s = _abi.BwScalars(0.0, 0.0, 0.0, 0.0, _p(pi, C.c_double), _p(v, C.c_double), 0.0)
        self._ck(lib().brr_hip_bw_get_scalars(self.ctx, C.byref(s)), "brr_hip_bw_get_scalars")
        return dict(mu=s.mu, alpha=s.alpha, sigma_b=s.sigma_b, d=s.d, pi=pi, v=v, betasq=s.betasq)

    def sweep(self, order: np.ndarray, u: np.ndarray, seed: int, iteration: int, chain: int = 0):
        o = np.ascontiguousarray(order, dtype=np.uint32)
        uu = np.ascontiguousarray(u, dtype=np.float64).reshape(self.m)
        self._ck(lib().brr_hip_bw_sweep(self.ctx, _p(o, C.c_uint32), _p(uu, C.c_double), seed, iteration, chain),
                 "brr_hip_bw_sweep")

    def run_chain(self, seed: int, chain: int, first_iteration: int, orders: np.ndarray) -> np.ndarray:
        """hist: (iterations, 5) {mu, alpha, sigma_b, M - v0 + 1, sum beta^2}."""
        o = np.ascontiguousarray(orders, dtype=np.uint32)
        iters = o.size // self.m
        assert iters * self.m == o.size
        hist = np.empty((iters, 5), dtype=np.float64)
        self._ck(lib().brr_hip_bw_run_chain(self.ctx, seed, chain, first_iteration, iters, _p(o, C.c_uint32),
                                            _p(hist, C.c_double)), "brr_hip_bw_run_chain")
        return hist

    def set_vi(self, marker: int, alpha: float) -> np.ndarray:
        out = np.empty(4, dtype=np.float64)
        self._ck(lib().brr_hip_bw_set_vi(self.ctx, marker, alpha, _p(out, C.c_double)), "brr_hip_bw_set_vi")
        return out

    def residual_update(self, marker: int, beta: float, sign: int):
        self._ck(lib().brr_hip_bw_residual_update(self.ctx, marker, beta, sign), "brr_hip_bw_residual_update")

    def sum_failure(self) -> np.ndarray:
        out = np.empty(self.m, dtype=np.float64)
        self._ck(lib().brr_hip_bw_sum_failure(self.ctx, _p(out, C.c_double)), "brr_hip_bw_sum_failure")
        return out

    PHASES = ("block_load", "round_sums", "exchange", "decisions", "beta_draws", "residual_update",
              "iteration_scalars", "in_model_shift")

    def stats(self) -> dict:
        s = _abi.Stats()
        self._ck(lib().brr_hip_last_stats(self.ctx, C.byref(s)), "brr_hip_last_stats")
        out = {f: getattr(s, f) for f, _ in _abi.Stats._fields_}
        ph = np.zeros(12)
        self._ck(lib().brr_hip_last_phase_ms(self.ctx, _p(ph, C.c_double), 12), "brr_hip_last_phase_ms")
        out["phase_ms"] = dict(zip(self.PHASES, ph.tolist()))
        # phase slots 8 and 9 carry the shader-clock and reference-clock ticks of the launch
        out["core_clock_mhz"] = 100.0 * ph[8] / ph[9] if ph[9] > 0 else float("nan")
        return out
