"""ctypes face of the CPU oracle (oracle/liboracle.so), for the tests only.

The oracle is the checker: an FP64 C restatement of the reference's BayesR
sweep. Product code never imports this module.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE_DIR = ROOT / "oracle"

P_u8 = C.POINTER(C.c_uint8)
P_i32 = C.POINTER(C.c_int32)
P_u32 = C.POINTER(C.c_uint32)
P_u64 = C.POINTER(C.c_uint64)
P_f64 = C.POINTER(C.c_double)

DENSE, RAGGED, EIGEN = 1, 2, 3
ORC_STREAM_FIXED = 5


class OrcData(C.Structure):
    _fields_ = [("N", C.c_uint32), ("M", C.c_uint32), ("layout", C.c_int),
                ("bed", P_u8), ("lut", P_f64),
                ("ones", P_i32), ("ones_off", P_u64), ("twos", P_i32), ("twos_off", P_u64),
                ("miss", P_i32), ("miss_off", P_u64), ("stats", P_f64)]


class OrcModel(C.Structure):
    _fields_ = [("K", C.c_int), ("G", C.c_int), ("S", P_f64), ("group", P_i32),
                ("v0E", C.c_double), ("s02E", C.c_double), ("v0G", C.c_double), ("s02G", C.c_double)]


class OrcState(C.Structure):
    _fields_ = [("eps", P_f64), ("beta", P_f64), ("comp", P_i32), ("acum", P_f64),
                ("mu", C.c_double), ("sigmaE", C.c_double), ("epsSum", C.c_double),
                ("sigmaG", P_f64), ("pi", P_f64), ("v", P_f64), ("betasqn", P_f64)]


class OrcTrace(C.Structure):
    _fields_ = [("num", P_f64), ("acum", P_f64), ("comp", P_i32), ("beta", P_f64)]


LOGDENS = C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p)


class OrcArmsRng(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("site", C.c_uint32), ("iteration", C.c_uint32), ("chain", C.c_uint32)]


class OrcBwModel(C.Structure):
    _fields_ = [("K", C.c_int), ("C", P_f64), ("quad_points", C.c_int), ("sum_failure", P_f64)]


class OrcBwState(C.Structure):
    _fields_ = [("eps", P_f64), ("vi", P_f64), ("beta", P_f64), ("comp", P_i32), ("v", P_f64), ("pi", P_f64),
                ("mu", C.c_double), ("alpha", C.c_double), ("sigma_b", C.c_double), ("d", C.c_double)]


class OrcBwTrace(C.Structure):
    _fields_ = [("vi_sum", C.c_double), ("vi_1", C.c_double), ("vi_2", C.c_double), ("acum", C.c_double),
                ("ml", C.c_double * 8)]

SITE_MU, SITE_ALPHA = 0xFFFFFFF0, 0xFFFFFFF1


_LIB = None


def lib(fast: bool = False):
    global _LIB
    name = "liboracle_fast.so" if fast else "liboracle.so"
    path = ORACLE_DIR / name
    if not path.exists():
        subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)], check=True)
    if not fast and _LIB is not None:
        return _LIB
    L = C.CDLL(str(path))
    sig = {
        "orc_philox4x32_10": (None, [P_u32, P_u32, P_u32]),
        "orc_det_log": (C.c_double, [C.c_double]),
        "orc_det_exp": (C.c_double, [C.c_double]),
        "orc_inv_norm_cdf": (C.c_double, [C.c_double]),
        "orc_marker_uz": (None, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, P_f64, P_f64]),
        "orc_fill_marker_uz": (None, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, P_f64, P_f64]),
        "orc_shuffle_philox": (None, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, P_u32]),
        "orc_shuffle_libc": (None, [C.c_uint32, P_u32]),
        "orc_uniform01": (C.c_double, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32]),
        "orc_norm_rng": (C.c_double, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_double,
                                      C.c_double]),
        "orc_rgamma": (C.c_double, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_double,
                                    C.c_double]),
        "orc_inv_scaled_chisq_rng": (C.c_double, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32,
                                                  C.c_double, C.c_double]),
        "orc_dirichlet_rng": (None, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, P_f64,
                                     P_f64]),
        "orc_decode_bed_column": (None, [P_u8, C.c_uint32, P_u8, P_u8]),
        "orc_check_bed_magic": (C.c_int, [P_u8]),
        "orc_dense_build": (None, [P_u8, C.c_uint32, P_f64]),
        "orc_dense_lut": (None, [P_u8, C.c_uint32, P_f64]),
        "orc_ragged_build": (C.c_int, [P_u8, C.c_uint32, P_i32, P_u64, P_i32, P_u64, P_i32, P_u64, P_f64]),
        "orc_ragged_serialise": (C.c_size_t, [P_f64, P_i32, C.c_uint64, P_i32, C.c_uint64, P_i32,
                                              C.c_uint64, P_u8]),
        "orc_eigen_build": (C.c_int, [P_u8, C.c_uint32, P_i32, P_u64, P_i32, P_u64, P_i32, P_u64, P_f64]),
        "orc_eigen_serialise": (C.c_size_t, [P_f64, P_i32, C.c_uint64, P_i32, C.c_uint64, P_i32,
                                             C.c_uint64, P_u8]),
        "orc_compute_num": (C.c_double, [C.POINTER(OrcData), C.c_uint32, P_f64, C.c_double, C.c_double]),
        "orc_apply_epsilon_change": (None, [C.POINTER(OrcData), C.c_uint32, P_f64, P_f64, C.c_double,
                                            C.c_double]),
        "orc_process_column": (None, [C.POINTER(OrcData), C.POINTER(OrcModel), C.POINTER(OrcState),
                                      C.c_uint32, C.c_double, C.c_double, P_f64, P_f64]),
        "orc_sweep": (None, [C.POINTER(OrcData), C.POINTER(OrcModel), C.POINTER(OrcState), P_u32, P_f64,
                             P_f64, C.POINTER(OrcTrace)]),
        "orc_init": (None, [C.POINTER(OrcData), C.POINTER(OrcModel), C.POINTER(OrcState), P_f64,
                            C.c_uint64, C.c_uint32]),
        "orc_update_mu": (None, [C.POINTER(OrcData), C.POINTER(OrcState), C.c_uint64, C.c_uint32,
                                 C.c_uint32]),
        "orc_post_sweep": (None, [C.POINTER(OrcData), C.POINTER(OrcModel), C.POINTER(OrcState), C.c_uint64,
                                  C.c_uint32, C.c_uint32]),
        "orc_run_gibbs": (None, [C.POINTER(OrcData), C.POINTER(OrcModel), C.POINTER(OrcState), P_f64,
                                 C.c_uint64, C.c_uint32, C.c_uint32, P_u32, C.c_int, P_f64]),
        "orc_run_gibbs_window": (None, [C.POINTER(OrcData), C.POINTER(OrcModel), C.POINTER(OrcState), P_f64,
                                        C.c_uint64, C.c_uint32, C.c_uint32, P_u32, C.c_int, P_f64,
                                        C.c_uint32]),
        "orc_fixed_effects": (None, [C.POINTER(OrcData), C.POINTER(OrcModel), C.POINTER(OrcState),
                                     C.POINTER(C.c_float), C.c_uint32, P_f64, P_u32, P_f64, C.c_uint64,
                                     C.c_uint32, C.c_uint32]),
        "orc_run_gibbs_fixed": (None, [C.POINTER(OrcData), C.POINTER(OrcModel), C.POINTER(OrcState), P_f64,
                                       C.c_uint64, C.c_uint32, C.c_uint32, P_u32, C.c_int, P_f64,
                                       C.c_uint32, C.POINTER(C.c_float), C.c_uint32, P_f64, P_u32, P_f64]),
        "orc_sweep_async": (None, [C.POINTER(OrcData), C.POINTER(OrcModel), C.POINTER(OrcState), P_u32, P_f64,
                                   P_f64, C.c_uint32]),
        "orc_arms": (C.c_int, [P_f64, C.c_int, C.c_double, C.c_double, LOGDENS, C.c_void_p, C.c_int,
                               C.POINTER(OrcArmsRng), P_f64, C.POINTER(C.c_int)]),
        "orc_bw_gh_rule": (C.c_int, [C.c_int, P_f64, P_f64, P_f64]),
        "orc_bw_gh_integral": (C.c_double, [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                            C.c_double, C.c_double, P_f64]),
        "orc_bw_sum_failure": (C.c_double, [C.POINTER(OrcData), C.c_uint32, P_f64]),
        "orc_bw_init": (None, [C.POINTER(OrcData), C.POINTER(OrcBwModel), C.POINTER(OrcBwState), P_f64, P_f64]),
        "orc_bw_process_column": (C.c_int, [C.POINTER(OrcData), C.POINTER(OrcBwModel), C.POINTER(OrcBwState),
                                            C.c_uint32, C.c_double, C.c_uint64, C.c_uint32, C.c_uint32,
                                            C.POINTER(OrcBwTrace)]),
        "orc_bw_sweep": (C.c_int, [C.POINTER(OrcData), C.POINTER(OrcBwModel), C.POINTER(OrcBwState), P_u32, P_f64,
                                   C.c_uint64, C.c_uint32, C.c_uint32]),
        "orc_bw_sample_mu": (C.c_int, [C.POINTER(OrcData), C.POINTER(OrcBwState), C.c_uint64, C.c_uint32,
                                       C.c_uint32]),
        "orc_bw_sample_alpha": (C.c_int, [C.POINTER(OrcData), C.POINTER(OrcBwState), P_f64, C.c_uint64,
                                          C.c_uint32, C.c_uint32]),
        "orc_bw_post_sweep": (None, [C.POINTER(OrcData), C.POINTER(OrcBwModel), C.POINTER(OrcBwState), C.c_uint64,
                                     C.c_uint32, C.c_uint32]),
        "orc_bw_run_gibbs": (C.c_int, [C.POINTER(OrcData), C.POINTER(OrcBwModel), C.POINTER(OrcBwState), P_f64,
                                       P_f64, C.c_uint64, C.c_uint32, C.c_uint32, P_u32, P_f64]),
        "orc_bw_run_gibbs_fixed": (C.c_int, [C.POINTER(OrcData), C.POINTER(OrcBwModel), C.POINTER(OrcBwState),
                                             P_f64, P_f64, C.c_uint64, C.c_uint32, C.c_uint32, P_u32, P_f64,
                                             C.POINTER(C.c_float), C.c_uint32, P_f64]),
        "orc_bw_sample_theta": (C.c_int, [C.POINTER(OrcData), C.POINTER(OrcBwState), C.POINTER(C.c_float),
                                          C.c_uint32, P_f64, P_f64, C.c_uint64, C.c_uint32, C.c_uint32]),
        "orc_bw_sum_failure_fixed": (None, [C.c_uint32, C.POINTER(C.c_float), C.c_uint32, P_f64, P_f64]),
    }
    for n, (r, a) in sig.items():
        f = getattr(L, n)
        f.restype = r
        f.argtypes = a
    if not fast:
        _LIB = L
    return L


def p(a, t):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(p(c, C.c_uint32), p(k, C.c_uint32), p(out, C.c_uint32))
    return out


def decode(col: np.ndarray, n: int):
    col = np.ascontiguousarray(col, dtype=np.uint8)
    a1 = np.zeros(n, dtype=np.uint8)
    a2 = np.zeros(n, dtype=np.uint8)
    lib().orc_decode_bed_column(p(col, C.c_uint8), n, p(a1, C.c_uint8), p(a2, C.c_uint8))
    return a1, a2


def dense_build(col, n):
    col = np.ascontiguousarray(col, dtype=np.uint8)
    cx = np.zeros(n, dtype=np.float64)
    lib().orc_dense_build(p(col, C.c_uint8), n, p(cx, C.c_double))
    return cx


def dense_lut(col, n):
    col = np.ascontiguousarray(col, dtype=np.uint8)
    lut = np.zeros(4, dtype=np.float64)
    lib().orc_dense_lut(p(col, C.c_uint8), n, p(lut, C.c_double))
    return lut


def ragged_build(col, n):
    col = np.ascontiguousarray(col, dtype=np.uint8)
    o = np.zeros(n, dtype=np.int32); t = np.zeros(n, dtype=np.int32); m = np.zeros(n, dtype=np.int32)
    no, nt, nm = C.c_uint64(), C.c_uint64(), C.c_uint64()
    st = np.zeros(4, dtype=np.float64)
    ok = lib().orc_ragged_build(p(col, C.c_uint8), n, p(o, C.c_int32), C.byref(no), p(t, C.c_int32),
                                C.byref(nt), p(m, C.c_int32), C.byref(nm), p(st, C.c_double))
    return bool(ok), o[: no.value].copy(), t[: nt.value].copy(), m[: nm.value].copy(), st


@dataclass
class OracleProblem:
    """Marker data + model + chain state for the oracle, with owned numpy buffers."""
    n: int
    m: int
    layout: int
    K: int
    G: int
    cva: np.ndarray
    group: np.ndarray | None = None
    bed: np.ndarray | None = None
    lut: np.ndarray | None = None
    ragged: object = None          # bayesrrcmd_amd.RaggedArena-like (ones, ones_off, ...)
    priors: tuple = (1e-4, 1e-4, 1e-4, 1e-4)
    _keep: list = field(default_factory=list)

    def __post_init__(self):
        self.eps = np.zeros(self.n)
        self.beta = np.zeros(self.m)
        self.comp = np.zeros(self.m, dtype=np.int32)
        self.acum = np.zeros(self.m)
        self.sigmaG = np.zeros(self.G)
        self.pi = np.zeros(self.G * self.K)
        self.v = np.zeros(self.G * self.K)
        self.betasqn = np.zeros(self.G)
        self.cva = np.ascontiguousarray(self.cva, dtype=np.float64).reshape(-1)
        if self.group is not None:
            self.group = np.ascontiguousarray(self.group, dtype=np.int32)
        d = OrcData()
        d.N, d.M, d.layout = self.n, self.m, self.layout
        if self.layout == DENSE:
            self.bed = np.ascontiguousarray(self.bed, dtype=np.uint8)
            self.lut = np.ascontiguousarray(self.lut, dtype=np.float64)
            d.bed = p(self.bed, C.c_uint8)
            d.lut = p(self.lut, C.c_double)
        else:
            r = self.ragged
            self._keep += [r]
            d.ones, d.ones_off = p(r.ones, C.c_int32), p(r.ones_off, C.c_uint64)
            d.twos, d.twos_off = p(r.twos, C.c_int32), p(r.twos_off, C.c_uint64)
            d.miss, d.miss_off = p(r.miss, C.c_int32), p(r.miss_off, C.c_uint64)
            self._stats = np.ascontiguousarray(r.stats, dtype=np.float64)
            d.stats = p(self._stats, C.c_double)
        self.data = d
        mdl = OrcModel()
        mdl.K, mdl.G = self.K, self.G
        mdl.S = p(self.cva, C.c_double)
        mdl.group = p(self.group, C.c_int32)
        mdl.v0E, mdl.s02E, mdl.v0G, mdl.s02G = self.priors
        self.model = mdl
        st = OrcState()
        st.eps, st.beta, st.comp, st.acum = (p(self.eps, C.c_double), p(self.beta, C.c_double),
                                             p(self.comp, C.c_int32), p(self.acum, C.c_double))
        st.sigmaG, st.pi, st.v, st.betasqn = (p(self.sigmaG, C.c_double), p(self.pi, C.c_double),
                                              p(self.v, C.c_double), p(self.betasqn, C.c_double))
        self.state = st

    # scalar accessors
    @property
    def mu(self): return self.state.mu
    @property
    def sigmaE(self): return self.state.sigmaE
    @property
    def epsSum(self): return self.state.epsSum

    def set_scalars(self, mu, sigmaE, epsSum, sigmaG, pi):
        self.state.mu, self.state.sigmaE, self.state.epsSum = mu, sigmaE, epsSum
        self.sigmaG[:] = np.asarray(sigmaG).reshape(-1)
        self.pi[:] = np.asarray(pi).reshape(-1)

    def init(self, y, seed, chain=0):
        y = np.ascontiguousarray(y, dtype=np.float64)
        lib().orc_init(C.byref(self.data), C.byref(self.model), C.byref(self.state), p(y, C.c_double),
                       seed, chain)

    def sweep(self, order, u, z, trace=True):
        order = np.ascontiguousarray(order, dtype=np.uint32)
        u = np.ascontiguousarray(u, dtype=np.float64)
        z = np.ascontiguousarray(z, dtype=np.float64)
        tr = None
        out = None
        if trace:
            out = dict(num=np.zeros(self.m), acum=np.zeros(self.m), comp=np.zeros(self.m, dtype=np.int32),
                       beta=np.zeros(self.m))
            tr = OrcTrace(p(out["num"], C.c_double), p(out["acum"], C.c_double), p(out["comp"], C.c_int32),
                          p(out["beta"], C.c_double))
        lib().orc_sweep(C.byref(self.data), C.byref(self.model), C.byref(self.state), p(order, C.c_uint32),
                        p(u, C.c_double), p(z, C.c_double), C.byref(tr) if tr is not None else None)
        return out

    def sweep_async(self, order, u, z, window):
        order = np.ascontiguousarray(order, dtype=np.uint32)
        u = np.ascontiguousarray(u, dtype=np.float64)
        z = np.ascontiguousarray(z, dtype=np.float64)
        lib().orc_sweep_async(C.byref(self.data), C.byref(self.model), C.byref(self.state), p(order, C.c_uint32),
                              p(u, C.c_double), p(z, C.c_double), window)

    def compute_num(self, j, beta_old):
        return lib().orc_compute_num(C.byref(self.data), j, p(self.eps, C.c_double), beta_old,
                                     self.state.epsSum)

    def apply_epsilon_change(self, j, beta_old, beta_new):
        es = C.c_double(self.state.epsSum)
        lib().orc_apply_epsilon_change(C.byref(self.data), j, p(self.eps, C.c_double), C.byref(es),
                                       beta_old, beta_new)
        self.state.epsSum = es.value

    def update_mu(self, seed, iteration, chain=0):
        lib().orc_update_mu(C.byref(self.data), C.byref(self.state), seed, iteration, chain)

    def post_sweep(self, seed, iteration, chain=0):
        lib().orc_post_sweep(C.byref(self.data), C.byref(self.model), C.byref(self.state), seed, iteration,
                             chain)

    def run_gibbs(self, y, seed, iterations, order=None, shuffle_mode=0, chain=0, window=0):
        y = np.ascontiguousarray(y, dtype=np.float64)
        if order is None:
            order = np.arange(self.m, dtype=np.uint32)
        order = np.ascontiguousarray(order, dtype=np.uint32)
        hist = np.zeros((iterations, 4 + self.G))
        lib().orc_run_gibbs_window(C.byref(self.data), C.byref(self.model), C.byref(self.state), p(y, C.c_double),
                                   seed, chain, iterations, p(order, C.c_uint32), shuffle_mode,
                                   p(hist, C.c_double), window)
        return hist, order

    def set_fixed(self, X):
        """Covariates (N, F) as the reference's float matrix, column-major; gamma 0, order 0..F-1, sigmaF 1."""
        x = np.asarray(X, dtype=np.float32).reshape(self.n, -1)
        self.F = x.shape[1]
        self.fixedX = np.ascontiguousarray(x.T)
        self.gamma = np.zeros(self.F)
        self.forder = np.arange(self.F, dtype=np.uint32)
        self.sigmaF = np.array([1.0])

    def fixed_effects(self, seed, iteration, chain=0):
        lib().orc_fixed_effects(C.byref(self.data), C.byref(self.model), C.byref(self.state),
                                p(self.fixedX, C.c_float), self.F, p(self.gamma, C.c_double),
                                p(self.forder, C.c_uint32), p(self.sigmaF, C.c_double), seed, iteration, chain)

    def run_gibbs_fixed(self, y, seed, iterations, order=None, chain=0, window=0):
        y = np.ascontiguousarray(y, dtype=np.float64)
        if order is None:
            order = np.arange(self.m, dtype=np.uint32)
        order = np.ascontiguousarray(order, dtype=np.uint32)
        hist = np.zeros((iterations, 4 + self.G))
        lib().orc_run_gibbs_fixed(C.byref(self.data), C.byref(self.model), C.byref(self.state), p(y, C.c_double),
                                  seed, chain, iterations, p(order, C.c_uint32), 0, p(hist, C.c_double), window,
                                  p(self.fixedX, C.c_float), self.F, p(self.gamma, C.c_double),
                                  p(self.forder, C.c_uint32), p(self.sigmaF, C.c_double))
        return hist, order

    def snapshot(self):
        """Copies of the chain state arrays and scalars."""
        return dict(eps=self.eps.copy(), beta=self.beta.copy(), comp=self.comp.copy(), acum=self.acum.copy(),
                    v=self.v.copy(), betasqn=self.betasqn.copy(), epsSum=self.state.epsSum)

    def restore(self, snap):
        for k in ("eps", "beta", "comp", "acum", "v", "betasqn"):
            getattr(self, k)[:] = snap[k]
        self.state.epsSum = snap["epsSum"]


@dataclass
class BayesWProblem:
    """Ragged markers + failure indicators + BayesW state for the oracle."""
    n: int
    m: int
    ragged: object
    fail: np.ndarray
    K: int = 4
    C: tuple = (0.001, 0.01, 0.1)
    quad_points: int = 7

    bed: object = None      # dense layout: M x ceil(N/4) codes and M x 4 tables instead of ragged
    lut: object = None

    def __post_init__(self):
        d = OrcData()
        d.N, d.M = self.n, self.m
        if self.ragged is not None:
            r = self.ragged
            d.layout = RAGGED
            d.ones, d.ones_off = p(r.ones, C.c_int32), p(r.ones_off, C.c_uint64)
            d.twos, d.twos_off = p(r.twos, C.c_int32), p(r.twos_off, C.c_uint64)
            d.miss, d.miss_off = p(r.miss, C.c_int32), p(r.miss_off, C.c_uint64)
            self._stats = np.ascontiguousarray(r.stats, dtype=np.float64)
            d.stats = p(self._stats, C.c_double)
        else:
            d.layout = DENSE
            self._bed = np.ascontiguousarray(self.bed, dtype=np.uint8)
            self._lut = np.ascontiguousarray(self.lut, dtype=np.float64)
            d.bed, d.lut = p(self._bed, C.c_uint8), p(self._lut, C.c_double)
        self.data = d
        self.fail = np.ascontiguousarray(self.fail, dtype=np.float64)
        self.sum_failure = np.array([lib().orc_bw_sum_failure(C.byref(d), j, p(self.fail, C.c_double))
                                     for j in range(self.m)])
        self._C = np.ascontiguousarray(self.C, dtype=np.float64)
        self.model = OrcBwModel(self.K, p(self._C, C.c_double), self.quad_points, p(self.sum_failure, C.c_double))
        self.eps = np.zeros(self.n)
        self.vi = np.zeros(self.n)
        self.beta = np.zeros(self.m)
        self.comp = np.zeros(self.m, dtype=np.int32)
        self.v = np.ones(self.K)
        self.pi = np.zeros(self.K)
        st = OrcBwState()
        st.eps, st.vi, st.beta = p(self.eps, C.c_double), p(self.vi, C.c_double), p(self.beta, C.c_double)
        st.comp, st.v, st.pi = p(self.comp, C.c_int32), p(self.v, C.c_double), p(self.pi, C.c_double)
        self.state = st

    def init(self, y):
        y = np.ascontiguousarray(y, dtype=np.float64)
        lib().orc_bw_init(C.byref(self.data), C.byref(self.model), C.byref(self.state), p(y, C.c_double),
                          p(self.fail, C.c_double))

    def refresh_vi(self):
        self.vi[:] = np.exp(self.state.alpha * self.eps - 0.577215664901532)

    def sweep(self, order, u, seed, iteration, chain=0):
        order = np.ascontiguousarray(order, dtype=np.uint32)
        u = np.ascontiguousarray(u, dtype=np.float64)
        return lib().orc_bw_sweep(C.byref(self.data), C.byref(self.model), C.byref(self.state),
                                  p(order, C.c_uint32), p(u, C.c_double), seed, iteration, chain)

    def set_fixed(self, X):
        """Covariates (N, F) as the reference's float matrix (column-major); theta 0."""
        x = np.asarray(X, dtype=np.float32).reshape(self.n, -1)
        self.F = x.shape[1]
        self.fixedX = np.ascontiguousarray(x.T)
        self.theta = np.zeros(self.F)
        self.sumf_fix = np.zeros(self.F)
        lib().orc_bw_sum_failure_fixed(self.n, p(self.fixedX, C.c_float), self.F, p(self.fail, C.c_double),
                                       p(self.sumf_fix, C.c_double))

    def sample_mu(self, seed, iteration, chain=0):
        return lib().orc_bw_sample_mu(C.byref(self.data), C.byref(self.state), seed, iteration, chain)

    def sample_alpha(self, seed, iteration, chain=0):
        return lib().orc_bw_sample_alpha(C.byref(self.data), C.byref(self.state), p(self.fail, C.c_double),
                                         seed, iteration, chain)

    def post_sweep(self, seed, iteration, chain=0):
        lib().orc_bw_post_sweep(C.byref(self.data), C.byref(self.model), C.byref(self.state), seed, iteration,
                                chain)

    def sample_theta(self, f, seed, iteration, chain=0):
        return lib().orc_bw_sample_theta(C.byref(self.data), C.byref(self.state), p(self.fixedX, C.c_float), f,
                                         p(self.theta, C.c_double), p(self.sumf_fix, C.c_double), seed,
                                         iteration, chain)

    def run_gibbs_fixed(self, y, seed, iterations, chain=0):
        y = np.ascontiguousarray(y, dtype=np.float64)
        order = np.arange(self.m, dtype=np.uint32)
        hist = np.zeros((iterations, 5))
        err = lib().orc_bw_run_gibbs_fixed(C.byref(self.data), C.byref(self.model), C.byref(self.state),
                                           p(y, C.c_double), p(self.fail, C.c_double), seed, chain, iterations,
                                           p(order, C.c_uint32), p(hist, C.c_double), p(self.fixedX, C.c_float),
                                           self.F, p(self.theta, C.c_double))
        return err, hist, order

    def run_gibbs(self, y, seed, iterations, chain=0):
        y = np.ascontiguousarray(y, dtype=np.float64)
        order = np.arange(self.m, dtype=np.uint32)
        hist = np.zeros((iterations, 5))
        err = lib().orc_bw_run_gibbs(C.byref(self.data), C.byref(self.model), C.byref(self.state), p(y, C.c_double),
                                     p(self.fail, C.c_double), seed, chain, iterations, p(order, C.c_uint32),
                                     p(hist, C.c_double))
        return err, hist, order
