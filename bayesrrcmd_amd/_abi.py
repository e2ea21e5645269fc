"""ctypes declarations of include/brr_hip.h and include/brr_synth.h.

The shared library is built in-tree (bayesrrcmd_amd/libbrr_hip.so). Loading
fails loudly when it is missing: there is no CPU fallback for the sweep.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libbrr_hip.so"

OK, EINVAL, ENOMEM, EDEVICE, ESTATE, ETIMEOUT, ESAMPLER = 0, 1, 2, 3, 4, 5, 6
LAYOUT_DENSE_2BIT = 1
LAYOUT_RAGGED = 2
LAYOUT_EIGEN_SPARSE = 3
STORAGE_AUTO, STORAGE_LISTS, STORAGE_CODES = 0, 1, 2
PRECISION_FP64, PRECISION_FP32_REFRESH, PRECISION_FP32 = 0, 1, 2

P_u8 = C.POINTER(C.c_uint8)
P_i32 = C.POINTER(C.c_int32)
P_u32 = C.POINTER(C.c_uint32)
P_u64 = C.POINTER(C.c_uint64)
P_f64 = C.POINTER(C.c_double)


class Config(C.Structure):
    _fields_ = [
        ("num_inds", C.c_uint32), ("num_snps", C.c_uint32), ("num_components", C.c_uint32),
        ("num_groups", C.c_uint32), ("layout", C.c_uint32), ("workgroups", C.c_uint32),
        ("block_markers", C.c_uint32), ("device", C.c_uint32), ("ragged_storage", C.c_uint32),
        ("num_traits", C.c_uint32),
    ]


class Model(C.Structure):
    _fields_ = [("cva", P_f64), ("v0E", C.c_double), ("s02E", C.c_double),
                ("v0G", C.c_double), ("s02G", C.c_double)]


class Scalars(C.Structure):
    _fields_ = [("mu", C.c_double), ("sigmaE", C.c_double), ("epsilon_sum", C.c_double),
                ("sigmaG", P_f64), ("pi", P_f64), ("v", P_f64), ("betasqn", P_f64)]


class Stats(C.Structure):
    _fields_ = [("markers", C.c_uint64), ("rounds", C.c_uint64), ("changes", C.c_uint64),
                ("kernel_ms", C.c_double), ("workgroups", C.c_uint32), ("block_markers", C.c_uint32),
                ("slice", C.c_uint32), ("storage", C.c_uint32),
                ("decision_rounds", C.c_uint64)]


class TuneEntry(C.Structure):
    _fields_ = [("workgroups", C.c_uint32), ("block_markers", C.c_uint32), ("slice", C.c_uint32),
                ("ms_per_iteration", C.c_double)]


class CohortReport(C.Structure):
    _fields_ = [("launches", C.c_uint32), ("device_ms", C.c_double), ("launch_of", C.c_uint32 * 64)]


class BwConfig(C.Structure):
    _fields_ = [("num_inds", C.c_uint32), ("num_snps", C.c_uint32), ("num_components", C.c_uint32),
                ("quad_points", C.c_uint32), ("workgroups", C.c_uint32), ("device", C.c_uint32)]


class BwScalars(C.Structure):
    _fields_ = [("mu", C.c_double), ("alpha", C.c_double), ("sigma_b", C.c_double), ("d", C.c_double),
                ("pi", P_f64), ("v", P_f64), ("betasq", C.c_double)]


CTX = C.c_void_p

# name -> (restype, argtypes)
SIGNATURES = {
    "brr_hip_abi_version": (C.c_int, []),
    "brr_hip_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "brr_hip_create": (C.c_int, [C.POINTER(Config), C.POINTER(CTX)]),
    "brr_hip_destroy": (None, [CTX]),
    "brr_hip_last_error": (C.c_char_p, [CTX]),
    "brr_hip_upload_dense": (C.c_int, [CTX, P_u8, C.c_uint64, P_f64]),
    "brr_hip_upload_ragged": (C.c_int, [CTX, P_i32, P_u64, P_i32, P_u64, P_i32, P_u64, P_f64]),
    "brr_hip_load_bed": (C.c_int, [CTX, P_u8, C.c_uint64, C.c_int]),
    "brr_hip_get_marker_tables": (C.c_int, [CTX, P_f64]),
    "brr_hip_set_async": (C.c_int, [CTX, C.c_uint32]),
    "brr_hip_set_precision": (C.c_int, [CTX, C.c_int, C.c_uint32]),
    "brr_hip_record_num": (C.c_int, [CTX, C.c_int]),
    "brr_hip_probe_exchange": (C.c_int, [CTX, C.c_uint32, P_f64]),
    "brr_hip_calibration_copy": (C.c_int, [C.c_uint32, C.c_uint64, P_f64]),
    "brr_hip_set_launch": (C.c_int, [CTX, C.c_uint32, C.c_uint32]),
    "brr_hip_run_cohorts": (C.c_int, [C.POINTER(CTX), C.c_uint32, C.c_uint64, P_u32, C.c_uint32, C.c_uint32,
                                      C.POINTER(P_u32), C.POINTER(P_f64), C.c_uint32, C.POINTER(CohortReport)]),
    "brr_hip_autotune": (C.c_int, [CTX, C.c_uint64, C.c_uint32, C.POINTER(TuneEntry), C.c_uint32,
                                   C.POINTER(C.c_uint32)]),
    "brr_hip_kernel_compute_num": (C.c_int, [CTX, C.c_uint32, P_f64, C.c_double, C.c_double, P_f64]),
    "brr_hip_kernel_epsilon_change": (C.c_int, [CTX, C.c_uint32, C.c_double, C.c_double, P_f64, P_f64]),
    "brr_hip_sweep_noalloc": (C.c_int, [CTX, C.c_void_p, C.c_void_p, C.c_void_p]),
    "brr_hip_get_num": (C.c_int, [CTX, P_f64]),
    "brr_hip_set_fixed": (C.c_int, [CTX, C.POINTER(C.c_float), C.c_uint32]),
    "brr_hip_set_fixed_state": (C.c_int, [CTX, P_f64, P_u32, C.c_double]),
    "brr_hip_get_fixed_state": (C.c_int, [CTX, P_f64, P_u32, C.POINTER(C.c_double)]),
    "brr_hip_fixed_effects": (C.c_int, [CTX, C.c_uint64, C.c_uint32, C.c_uint32]),
    "brr_hip_set_groups": (C.c_int, [CTX, P_i32]),
    "brr_hip_set_model": (C.c_int, [CTX, C.POINTER(Model)]),
    "brr_hip_set_state": (C.c_int, [CTX, P_f64, P_f64, P_i32]),
    "brr_hip_set_scalars": (C.c_int, [CTX, C.POINTER(Scalars)]),
    "brr_hip_get_state": (C.c_int, [CTX, P_f64, P_f64, P_i32, P_f64]),
    "brr_hip_get_scalars": (C.c_int, [CTX, C.POINTER(Scalars)]),
    "brr_hip_set_trait_scalars": (C.c_int, [CTX, C.c_uint32, C.POINTER(Scalars)]),
    "brr_hip_get_trait_scalars": (C.c_int, [CTX, C.c_uint32, C.POINTER(Scalars)]),
    "brr_hip_sweep": (C.c_int, [CTX, P_u32, P_f64, P_f64]),
    "brr_hip_run_chain": (C.c_int, [CTX, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, P_u32, P_f64]),
    "brr_hip_compute_num": (C.c_int, [CTX, C.c_uint32, C.c_double, P_f64]),
    "brr_hip_epsilon_change": (C.c_int, [CTX, C.c_uint32, C.c_double, C.c_double]),
    "brr_hip_last_stats": (C.c_int, [CTX, C.POINTER(Stats)]),
    "brr_hip_last_phase_ms": (C.c_int, [CTX, P_f64, C.c_uint32]),
    "brr_hip_bw_create": (C.c_int, [C.POINTER(BwConfig), C.POINTER(CTX)]),
    "brr_hip_bw_set_model": (C.c_int, [CTX, P_f64, P_f64]),
    "brr_hip_bw_set_scalars": (C.c_int, [CTX, C.POINTER(BwScalars)]),
    "brr_hip_bw_get_scalars": (C.c_int, [CTX, C.POINTER(BwScalars)]),
    "brr_hip_bw_sweep": (C.c_int, [CTX, P_u32, P_f64, C.c_uint64, C.c_uint32, C.c_uint32]),
    "brr_hip_bw_run_chain": (C.c_int, [CTX, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, P_u32, P_f64]),
    "brr_hip_bw_set_vi": (C.c_int, [CTX, C.c_uint32, C.c_double, P_f64]),
    "brr_hip_bw_residual_update": (C.c_int, [CTX, C.c_uint32, C.c_double, C.c_int]),
    "brr_hip_bw_sum_failure": (C.c_int, [CTX, P_f64]),
    "brr_hip_bw_set_fixed": (C.c_int, [CTX, C.POINTER(C.c_float), C.c_uint32]),
    "brr_hip_bw_set_theta": (C.c_int, [CTX, P_f64]),
    "brr_hip_bw_get_theta": (C.c_int, [CTX, P_f64]),
    "brr_hip_bw_fixed_effects": (C.c_int, [CTX, C.c_uint64, C.c_uint32, C.c_uint32]),
    "brr_bayesw_init": (C.c_int, [P_f64, P_f64, C.c_uint32, C.c_uint32, C.c_uint32, P_f64,
                                  C.POINTER(BwScalars)]),
    "brr_eigen_fill": (C.c_int, [P_u8, C.c_uint64, C.c_uint32, C.c_uint32, P_u64, P_u64, P_u64,
                                 P_i32, P_i32, P_i32, P_f64, C.c_uint32]),
    "brr_eigen_read_counts": (C.c_int, [P_u8, C.c_uint64, P_u64, C.c_uint32, C.c_int, P_u64, P_u64, P_u64,
                                        C.c_uint32]),
    "brr_eigen_read_fill": (C.c_int, [P_u8, C.c_uint64, P_u64, C.c_uint32, C.c_uint32, C.c_int, P_u64, P_u64,
                                      P_u64, P_i32, P_i32, P_i32, P_f64, C.c_uint32]),
    "brr_ragged_read_counts": (C.c_int, [P_u8, C.c_uint64, P_u64, C.c_uint32, C.c_int, P_u64, P_u64, P_u64,
                                         C.c_uint32]),
    "brr_ragged_read_fill": (C.c_int, [P_u8, C.c_uint64, P_u64, C.c_uint32, C.c_uint32, C.c_int, P_u64, P_u64,
                                       P_u64, P_i32, P_i32, P_i32, P_f64, C.c_uint32]),
    "brr_dense_read_ppbed": (C.c_int, [P_u8, C.c_uint64, P_u64, C.c_uint32, C.c_uint32, C.c_int, P_u8, P_f64,
                                       C.c_uint32]),
    "brr_samples_open": (C.c_int, [C.c_char_p, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                   C.POINTER(C.c_void_p)]),
    "brr_samples_write_bayesr": (C.c_int, [C.c_void_p, C.c_uint32, C.c_double, P_f64, C.c_double, P_f64, P_f64,
                                           P_i32, P_f64, P_f64]),
    "brr_samples_write_bayesw": (C.c_int, [C.c_void_p, C.c_uint32, C.c_double, C.c_double, P_f64, P_f64, P_i32,
                                           C.c_double]),
    "brr_samples_close": (C.c_int, [C.c_void_p]),
    "brr_sample_due": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32]),
    "brr_bed_check_magic": (C.c_int, [P_u8]),
    "brr_dense_lut": (C.c_int, [P_u8, C.c_uint32, P_f64]),
    "brr_dense_luts": (C.c_int, [P_u8, C.c_uint64, C.c_uint32, C.c_uint32, P_f64, C.c_uint32]),
    "brr_ragged_count": (C.c_int, [P_u8, C.c_uint64, C.c_uint32, C.c_uint32, P_u64, P_u64, P_u64,
                                   C.c_uint32]),
    "brr_ragged_fill": (C.c_int, [P_u8, C.c_uint64, C.c_uint32, C.c_uint32, P_u64, P_u64, P_u64,
                                  P_i32, P_i32, P_i32, P_f64, C.c_uint32]),
    "brr_bayesr_init": (C.c_int, [P_f64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32,
                                  P_f64, C.POINTER(Scalars)]),
    "brr_shuffle": (C.c_int, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, P_u32, C.c_int]),
    "brr_marker_uz": (C.c_int, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, P_f64, P_f64]),
    "brr_synth_bed": (C.c_int, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double,
                                P_u8, C.c_uint64, P_f64, C.c_uint32]),
    "brr_synth_xb": (C.c_int, [P_u8, C.c_uint64, C.c_uint32, C.c_uint32, P_f64, C.c_uint32, P_f64, P_f64,
                               C.c_uint32]),
}

_LIB = None


def load() -> C.CDLL:
    """Load libbrr_hip.so and attach signatures; raise if it is not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is not built; run `python -m bayesrrcmd_amd.build` "
            "(hipcc, gfx950). The sweep has no CPU fallback.")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib
