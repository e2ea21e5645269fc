"""The per-marker BayesRKernel interface on a host residual (HipBayesRKernel).

brr_hip_kernel_compute_num mirrors BayesRKernel::computeNum(const VectorXd &epsilon,
double beta_old) and brr_hip_kernel_epsilon_change mirrors calculateEpsilonChange
(bayesrkernel.h:11-15; dense densebayesrkernel.cpp:10-20; sparse
sparsebayesrkernel.cpp:10-34 with raggedbayesrkernel.cpp:10-35 or
eigenbayesrkernel.cpp). Both are checked through the C ABI against the oracle's
orc_compute_num / orc_apply_epsilon_change on the same host residual:
num within 1e-12 of its scale (the device sums in another order), the residual
change bit for bit (the same per-element operation on a zero vector), the
epsilonSum change within 1e-14 relative; the context's own residual is untouched.
"""
import numpy as np
import pytest

import bayesrrcmd_amd as brr
import _oracle as orc
from _data import make_problem, spec_dosage

pytestmark = pytest.mark.gpu

S = np.array([0.01, 0.001, 0.0001])
CASES = [("dense", orc.DENSE, brr.LAYOUT_DENSE_2BIT, brr.STORAGE_AUTO),
         ("ragged-lists", orc.RAGGED, brr.LAYOUT_RAGGED, brr.STORAGE_LISTS),
         ("ragged-codes", orc.RAGGED, brr.LAYOUT_RAGGED, brr.STORAGE_CODES),
         ("eigen", orc.EIGEN, brr.LAYOUT_EIGEN_SPARSE, brr.STORAGE_AUTO)]


def oracle_for(pr, olay, eigen):
    if olay == orc.DENSE:
        return orc.OracleProblem(pr.n, pr.m, orc.DENSE, 4, 1, S, bed=pr.bed, lut=pr.lut)
    return orc.OracleProblem(pr.n, pr.m, olay, 4, 1, S, ragged=eigen if olay == orc.EIGEN else pr.ragged)


@pytest.mark.parametrize("name,olay,layout,storage", CASES, ids=[c[0] for c in CASES])
def test_kernel_calls_on_a_host_residual(gpu_lib, name, olay, layout, storage):
    pr = make_problem(4001, 60, seed=401, miss=0.03)
    eigen = brr.build_eigen(pr.bed, pr.n) if olay == orc.EIGEN else None
    o = oracle_for(pr, olay, eigen)
    h = brr.HipSweep(pr.n, pr.m, 4, 1, layout, ragged_storage=storage)
    if layout == brr.LAYOUT_DENSE_2BIT:
        h.upload_dense(pr.bed, pr.lut)
    else:
        h.upload_ragged(eigen if eigen is not None else pr.ragged)
    h.set_model(S)
    state_eps = np.linspace(-1.0, 1.0, pr.n)
    h.set_state(state_eps, np.zeros(pr.m), np.zeros(pr.m, np.int32))
    rng = np.random.default_rng(402)
    for j in rng.choice(pr.m, 12, replace=False):
        j = int(j)
        eps = rng.normal(0.0, 0.8, pr.n)
        es = float(rng.normal(0.0, 30.0))
        b_old, b_new = float(rng.normal(0.0, 0.05)), float(rng.normal(0.0, 0.05))
        o.eps[:] = eps
        o.state.epsSum = es
        want = o.compute_num(j, b_old)
        got = h.kernel_compute_num(j, eps, es, b_old)
        assert abs(got - want) <= 1e-12 * max(1.0, abs(want)), (name, j, got, want)
        o.eps[:] = 0.0
        o.state.epsSum = 0.0
        o.apply_epsilon_change(j, b_old, b_new)
        d, des = h.kernel_epsilon_change(j, b_old, b_new)
        assert np.array_equal(d, o.eps), (name, j)
        want_es = o.state.epsSum if olay != orc.DENSE else 0.0
        assert abs(des - want_es) <= 1e-14 * max(1.0, abs(want_es)), (name, j, des, want_es)
    assert np.array_equal(h.get_state()["eps"], state_eps)     # the context's residual is untouched
    h.close()


def test_dense_kernel_num_is_the_standardised_dot(gpu_lib):
    """computeNum of a dense marker is Cx . epsilon + beta_old (N - 1) with Cx the
    DenseMarker column (densebayesrkernel.cpp:10-15), here recomputed in numpy."""
    pr = make_problem(3001, 20, seed=403)
    h = brr.HipSweep(pr.n, pr.m, 4, 1, brr.LAYOUT_DENSE_2BIT)
    h.upload_dense(pr.bed, pr.lut)
    h.set_model(S)
    d = spec_dosage(pr.bed, pr.n)
    x = np.take_along_axis(pr.lut, np.where(d < 0, 3, d), axis=1)
    eps = np.random.default_rng(404).normal(0.0, 1.0, pr.n)
    for j in range(pr.m):
        want = float(x[j] @ eps) + 0.02 * (pr.n - 1)
        assert abs(h.kernel_compute_num(j, eps, 0.0, 0.02) - want) <= 1e-11 * abs(want)
    h.close()
