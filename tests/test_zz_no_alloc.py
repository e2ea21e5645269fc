"""brr_hip_sweep_noalloc: the plain sweep without device allocations.

It runs the kernels of brr_hip_sweep on the same context state, with the
exchange, control words, counters and the order check in the library's static
device scratch, and order / u / z passed as device pointers (here buffers the
test makes with the HIP runtime's own hipMalloc, through ctypes). Bar:
bit-identical state to brr_hip_sweep from the same inputs (the same kernel, the
same operations), on two calls in a row, the second on a smaller context; a bad
order is refused before the sweep kernel touches the state.
"""
import ctypes as C

import numpy as np
import pytest

import bayesrrcmd_amd as brr
from _data import make_problem

pytestmark = pytest.mark.gpu

S = np.array([0.01, 0.001, 0.0001])


def make(pr, storage, precision=brr.PRECISION_FP64):
    h = brr.HipSweep(pr.n, pr.m, 4, 1, brr.LAYOUT_RAGGED, ragged_storage=storage)
    h.upload_ragged(pr.ragged)
    h.set_model(S)
    h.set_precision(precision, 1024)
    eps, sc = brr.bayesr_init(pr.y, 1, 4, 7)
    h.set_state(eps, np.zeros(pr.m), np.zeros(pr.m, np.int32))
    h.set_scalars(sc["mu"], sc["sigmaE"], sc["epsilon_sum"], sc["sigmaG"], sc["pi"])
    return h


class DeviceBuffer:
    """A device copy of a host array, made with the HIP runtime (test plumbing)."""
    _hip = None

    def __init__(self, a):
        if DeviceBuffer._hip is None:
            DeviceBuffer._hip = C.CDLL("libamdhip64.so")
        hip = DeviceBuffer._hip
        a = np.ascontiguousarray(a)
        self.ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), C.c_size_t(a.nbytes)) == 0
        assert hip.hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0   # H2D
        assert hip.hipDeviceSynchronize() == 0

    def free(self):
        if self.ptr:
            DeviceBuffer._hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def noalloc(h, order, u, z):
    bufs = [DeviceBuffer(order), DeviceBuffer(u), DeviceBuffer(z)]
    try:
        return brr.lib().brr_hip_sweep_noalloc(h.ctx, *(b.ptr for b in bufs))
    finally:
        for b in bufs:
            b.free()


def inputs(m, it):
    order = np.arange(m, dtype=np.uint32)
    brr.shuffle(5, it, 0, order)
    u, z = brr.marker_uz(5, it, 0, m)
    return order, u, z


@pytest.mark.parametrize("storage,precision", [(brr.STORAGE_CODES, brr.PRECISION_FP64),
                                               (brr.STORAGE_LISTS, brr.PRECISION_FP64),
                                               (brr.STORAGE_CODES, brr.PRECISION_FP32_REFRESH)],
                         ids=["codes-fp64", "lists-fp64", "codes-fp32"])
def test_noalloc_sweep_equals_sweep(gpu_lib, storage, precision):
    # the second call on a smaller context (fewer individuals and markers)
    for n, m, seed in ((20_001, 900, 301), (3_001, 300, 303)):
        pr = make_problem(n, m, seed=seed, causal_frac=0.02)
        a, b = make(pr, storage, precision), make(pr, storage, precision)
        for it in range(2):
            order, u, z = inputs(pr.m, it)
            a.sweep(order, u, z)
            assert noalloc(b, order, u, z) == 0, brr.lib().brr_hip_last_error(b.ctx)
            sa, sb = a.get_state(), b.get_state()
            for k in ("eps", "beta", "comp", "acum"):
                assert np.array_equal(sa[k], sb[k]), (n, it, k)
            ca, cb = a.get_scalars(), b.get_scalars()
            for k in ("mu", "sigmaE", "epsilon_sum"):
                assert ca[k] == cb[k], (n, it, k)
            assert np.array_equal(ca["v"], cb["v"]) and np.array_equal(ca["betasqn"], cb["betasqn"])
            assert b.stats()["markers"] == pr.m and b.stats()["rounds"] == a.stats()["rounds"]
        a.close()
        b.close()


def test_noalloc_refuses_a_bad_order(gpu_lib):
    pr = make_problem(2_001, 200, seed=305)
    h = make(pr, brr.STORAGE_CODES)
    st0 = h.get_state()
    order, u, z = inputs(pr.m, 0)
    bad = order.copy()
    bad[7] = bad[8]
    assert noalloc(h, bad, u, z) == brr._abi.EINVAL
    assert b"permutation" in brr.lib().brr_hip_last_error(h.ctx)
    big = order.copy()
    big[3] = pr.m
    assert noalloc(h, big, u, z) == brr._abi.EINVAL
    st1 = h.get_state()
    for k in ("eps", "beta", "comp", "acum"):
        assert np.array_equal(st0[k], st1[k]), k
    assert noalloc(h, order, u, z) == 0         # and the next good call runs
    h.close()
