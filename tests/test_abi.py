"""The C-ABI library loads and exports every symbol that include/*.h declares.

Runs without a GPU: no compute entry point is called here except the host-only
builders; brr_hip_create must fail cleanly (not crash) when no device exists.
"""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import bayesrrcmd_amd as brr
from bayesrrcmd_amd import _abi

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    names = set()
    for h in sorted((ROOT / "include").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[\w\s\*]+?\b(brr_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_every_declared_symbol_is_exported():
    lib = brr.lib()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    # and every declared function has a ctypes signature in the Python face
    assert names <= set(_abi.SIGNATURES), sorted(names - set(_abi.SIGNATURES))


def test_abi_version():
    assert brr.lib().brr_hip_abi_version() == 7


def test_create_rejects_bad_config():
    lib = brr.lib()
    ctx = _abi.CTX()
    bad = [
        _abi.Config(0, 10, 4, 1, 2, 0, 0, 0, 0),     # N < 2
        _abi.Config(10, 0, 4, 1, 2, 0, 0, 0, 0),     # M = 0
        _abi.Config(10, 10, 1, 1, 2, 0, 0, 0, 0),    # K < 2
        _abi.Config(10, 10, 9, 1, 2, 0, 0, 0, 0),    # K > 8
        _abi.Config(10, 10, 4, 0, 2, 0, 0, 0, 0),    # G = 0
        _abi.Config(10, 10, 4, 1, 7, 0, 0, 0, 0),    # unknown layout
        _abi.Config(10, 10, 4, 1, 2, 0, 65, 0, 0),   # block > 64
        _abi.Config(10, 10, 4, 1, 2, 0, 0, 0, 3),    # unknown ragged storage
        _abi.Config(10, 10, 4, 1, 2, 0, 0, 0, 0, 65),  # more than 64 traits
        _abi.Config(10, 10, 4, 2, 2, 0, 0, 0, 0, 4),   # traits need one group
    ]
    for cfg in bad:
        assert lib.brr_hip_create(C.byref(cfg), C.byref(ctx)) == _abi.EINVAL


def test_bw_create_rejects_bad_config():
    lib = brr.lib()
    ctx = _abi.CTX()
    bad = [
        _abi.BwConfig(1, 10, 4, 7, 0, 0),      # N < 2
        _abi.BwConfig(10, 0, 4, 7, 0, 0),      # M = 0
        _abi.BwConfig(10, 10, 1, 7, 0, 0),     # K < 2
        _abi.BwConfig(10, 10, 9, 7, 0, 0),     # K > 8
        _abi.BwConfig(10, 10, 4, 8, 0, 0),     # even number of quadrature points
        _abi.BwConfig(10, 10, 4, 17, 0, 0),    # more than 15 points
        _abi.BwConfig(10, 10, 4, 1, 0, 0),     # fewer than 3 points
    ]
    for cfg in bad:
        assert lib.brr_hip_bw_create(C.byref(cfg), C.byref(ctx)) == _abi.EINVAL
    assert lib.brr_hip_bw_sweep(None, None, None, 0, 0, 0) == _abi.EINVAL


def test_bayesw_init_matches_oracle():
    """BayesWBase::init of the host library against the oracle's restatement."""
    import sys
    sys.path.insert(0, str(ROOT / "tests"))
    import _oracle as orc
    from _data import make_bw_problem
    pr = make_bw_problem(1500, 40, seed=12)
    w = orc.BayesWProblem(pr.n, pr.m, pr.ragged, pr.fail)
    w.init(pr.y)
    eps, s = brr.bayesw_init(pr.y, pr.fail, pr.m, 4)
    st = w.state
    assert (s["mu"], s["alpha"], s["sigma_b"], s["d"]) == (st.mu, st.alpha, st.sigma_b, st.d)
    np.testing.assert_array_equal(eps, w.eps)
    np.testing.assert_array_equal(s["pi"], w.pi)
    np.testing.assert_array_equal(s["v"], np.ones(4))
    # the closed forms of bayeswbase.cpp:494-526
    assert s["alpha"] == pytest.approx(3.14159 / np.sqrt(6 * np.sum((pr.y - pr.y.mean()) ** 2) / (pr.n - 1)),
                                       rel=1e-13)
    assert s["sigma_b"] == pytest.approx(6.283185 / (6 * s["alpha"] ** 2 * pr.m), rel=1e-14)
    np.testing.assert_allclose(s["pi"], [0.99, 0.01 / 3, 0.01 / 3, 0.01 / 3], rtol=1e-15)


def test_create_without_device_fails_cleanly():
    lib = brr.lib()
    n = C.c_int(-1)
    assert lib.brr_hip_device_count(C.byref(n)) == 0
    if n.value == 0:
        ctx = _abi.CTX()
        cfg = _abi.Config(100, 10, 4, 1, 2, 0, 0, 0, 0)
        assert lib.brr_hip_create(C.byref(cfg), C.byref(ctx)) == _abi.EDEVICE
        assert not ctx.value


def test_null_context_calls_are_rejected():
    lib = brr.lib()
    assert lib.brr_hip_sweep(None, None, None, None) == _abi.EINVAL
    assert lib.brr_hip_get_state(None, None, None, None, None) == _abi.EINVAL
    assert lib.brr_hip_last_error(None) == b"null context"
    lib.brr_hip_destroy(None)


def test_host_init_matches_reference_formulas():
    rng = np.random.default_rng(0)
    y = rng.normal(3.0, 2.0, 1000).astype(np.float32).astype(np.float64)
    eps, s = brr.bayesr_init(y, 2, 4, seed=1)
    ys = (y - y.mean())
    ys /= np.sqrt(ys @ ys / (y.size - 1))
    assert np.allclose(eps, ys, atol=1e-14)
    assert abs(s["sigmaE"] - (eps @ eps) / y.size * 0.5) < 1e-14
    assert np.allclose(s["pi"], [[0.5, 0.125, 0.125, 0.125]] * 2)
    assert np.all((s["sigmaG"] >= 0) & (s["sigmaG"] < 1))


def test_synth_xb_matches_dense_product():
    """X beta of the phenotype generator against the explicit standardised matrix."""
    import sys
    sys.path.insert(0, str(ROOT / "tests"))
    from _data import spec_dosage
    n, m, T = 1234, 57, 5
    bed, _ = brr.synth_bed(3, n, m, 0.05, 0.5, 0.02)
    lut = brr.build_dense(bed, n)
    rng = np.random.default_rng(4)
    beta = rng.normal(size=(m, T)) * (rng.random((m, T)) < 0.3)
    d = spec_dosage(bed, n)
    x = np.take_along_axis(lut, np.where(d < 0, 3, d), axis=1)      # (m, n)
    got = brr.synth_xb(bed, n, lut, beta, threads=3)
    assert got.shape == (n, T)
    np.testing.assert_allclose(got, x.T @ beta, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(brr.synth_xb(bed, n, lut, beta[:, 2]), x.T @ beta[:, 2], rtol=1e-12, atol=1e-12)
