"""Pins the counter-based stream of the oracle (oracle/rng.c).

Checks that do not retype the oracle: Random123's published Philox4x32-10
known-answer vectors, libm / scipy reference values for log, exp and the normal
quantile, distribution tests (KS, moments) for the uniform / normal / gamma /
inverse-scaled-chi-square / Dirichlet parameterisations of
distributions_boost.cpp:22-65, the libstdc++ random_shuffle rule driven by the
C library's own rand(), and bit-identity with the product's independent
implementation (bayesrrcmd_amd/csrc/brr_rng.h).
"""
import ctypes as C

import numpy as np
import pytest
from scipy import special, stats

import _oracle as orc
from _oracle import p


# Random123 kat_vectors: philox4x32 with 10 rounds (ctr[4], key[2]) -> out[4]
KAT = [
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000),
     (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff), (0xffffffff, 0xffffffff),
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,expect", KAT)
def test_philox_known_answers(ctr, key, expect):
    out = orc.philox(ctr, key)
    assert [int(x) for x in out] == list(expect)


def test_det_log_matches_libm():
    L = orc.lib()
    rng = np.random.default_rng(1)
    xs = np.concatenate([
        rng.uniform(0, 1, 2000), 10.0 ** rng.uniform(-300, 300, 2000),
        np.array([1.0, 2.0, 0.5, np.sqrt(2.0), 1 - 2 ** -53, 2 ** -1074, 5e-324 * 7, 1e-310]),
    ])
    got = np.array([L.orc_det_log(float(x)) for x in xs])
    ref = np.log(xs)
    err = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)
    # exact zero at 1, and within a few ulp elsewhere
    assert L.orc_det_log(1.0) == 0.0
    assert np.max(err[np.abs(ref) > 1e-3]) < 4e-16
    assert np.max(np.abs(got - ref)[np.abs(ref) <= 1e-3]) < 1e-18
    assert L.orc_det_log(0.0) == -np.inf


def test_det_exp_matches_libm():
    L = orc.lib()
    xs = np.concatenate([np.linspace(-745, 709, 4001), np.linspace(-1, 1, 2001)])
    got = np.array([L.orc_det_exp(float(x)) for x in xs])
    ref = np.exp(xs)
    mask = ref > 1e-300
    assert np.max(np.abs(got[mask] - ref[mask]) / ref[mask]) < 4e-16
    assert L.orc_det_exp(0.0) == 1.0
    assert L.orc_det_exp(-800.0) == 0.0


def test_inverse_normal_cdf_matches_scipy():
    L = orc.lib()
    ps = np.concatenate([np.linspace(1e-6, 1 - 1e-6, 20001), 10.0 ** -np.arange(1, 300, 7.0)])
    ps = np.concatenate([ps, 1 - ps[ps > 1e-15]])
    got = np.array([L.orc_inv_norm_cdf(float(x)) for x in ps])
    ref = special.ndtri(ps)
    assert np.max(np.abs(got - ref) / np.maximum(1.0, np.abs(ref))) < 1e-14
    # symmetry is exact (the algorithm folds on q = p - 1/2)
    assert L.orc_inv_norm_cdf(0.25) == -L.orc_inv_norm_cdf(0.75)


def _uz(seed, it, m, chain=0):
    u = np.zeros(m)
    z = np.zeros(m)
    orc.lib().orc_fill_marker_uz(seed, it, chain, m, p(u, C.c_double), p(z, C.c_double))
    return u, z


def test_marker_uz_distribution():
    u, z = _uz(20261202, 3, 200000)
    assert u.min() >= 0.0 and u.max() < 1.0
    assert stats.kstest(u, "uniform").pvalue > 1e-3
    assert stats.kstest(z, "norm").pvalue > 1e-3
    assert abs(z.mean()) < 0.01 and abs(z.var() - 1) < 0.01
    # u and z of one marker come from disjoint words of the block: uncorrelated
    assert abs(np.corrcoef(u, z)[0, 1]) < 0.01
    # streams of different iterations / chains / seeds differ
    u2, _ = _uz(20261202, 4, 1000)
    u3, _ = _uz(20261202, 3, 1000, chain=1)
    u4, _ = _uz(20261203, 3, 1000)
    assert not np.array_equal(u[:1000], u2) and not np.array_equal(u[:1000], u3)
    assert not np.array_equal(u[:1000], u4)


def test_uniform_from_block_words():
    # u = (x0 << 32 | x1) >> 11 * 2^-53 for the block (marker, iteration, 1, chain)
    seed, it, j = 0x123456789ABCDEF, 7, 11
    x = orc.philox((j, it, 1, 0), (seed & 0xFFFFFFFF, seed >> 32))
    expect = ((int(x[0]) << 32 | int(x[1])) >> 11) * 2.0 ** -53
    u = C.c_double()
    z = C.c_double()
    orc.lib().orc_marker_uz(seed, it, j, 0, C.byref(u), C.byref(z))
    assert u.value == expect
    pz = (((int(x[2]) << 32 | int(x[3])) >> 11) + 0.5) * 2.0 ** -53
    assert abs(z.value - special.ndtri(pz)) < 1e-13


def test_shuffle_libc_rule_matches_rand():
    """orc_shuffle_libc applies libstdc++'s rule to the C library's rand()."""
    libc = C.CDLL("libc.so.6")
    libc.rand.restype = C.c_int
    m = 1000
    libc.srand(5)
    r = [libc.rand() for _ in range(m - 1)]
    expect = list(range(m))
    for i in range(1, m):
        j = r[i - 1] % (i + 1)
        expect[i], expect[j] = expect[j], expect[i]
    order = np.arange(m, dtype=np.uint32)
    libc.srand(5)
    orc.lib().orc_shuffle_libc(m, p(order, C.c_uint32))
    assert order.tolist() == expect


def test_shuffle_philox_is_permutation_and_uniform():
    m = 5
    counts = {}
    for it in range(6000):
        order = np.arange(m, dtype=np.uint32)
        orc.lib().orc_shuffle_philox(99, it, 0, m, p(order, C.c_uint32))
        key = tuple(order.tolist())
        counts[key] = counts.get(key, 0) + 1
    assert len(counts) == 120            # all 5! permutations reachable
    obs = np.array(list(counts.values()))
    assert stats.chisquare(obs).pvalue > 1e-4


@pytest.mark.parametrize("shape", [0.05, 0.3, 1.0, 2.5, 50.0, 5000.0])
def test_gamma_distribution(shape):
    L = orc.lib()
    n = 20000
    g = np.array([L.orc_rgamma(7, 1, d, 0, shape, 1.0) for d in range(n)])
    assert np.all(g >= 0)
    assert stats.kstest(g, "gamma", args=(shape,)).pvalue > 1e-4
    # scale parameter multiplies
    assert L.orc_rgamma(7, 1, 3, 0, shape, 2.5) == 2.5 * L.orc_rgamma(7, 1, 3, 0, shape, 1.0)


def test_inv_scaled_chisq_parameterisation():
    # X = inv_scaled_chisq(dof, s): 1/X ~ Gamma(dof/2, scale 2/(dof s))
    L = orc.lib()
    dof, s = 30.0, 0.7
    x = np.array([L.orc_inv_scaled_chisq_rng(3, 2, d, 0, dof, s) for d in range(20000)])
    assert stats.kstest(1.0 / x, "gamma", args=(dof / 2, 0.0, 2 / (dof * s))).pvalue > 1e-4


def test_norm_rng_mean_and_variance_argument():
    L = orc.lib()
    x = np.array([L.orc_norm_rng(5, 0, d, 0, 3.0, 4.0) for d in range(20000)])
    assert abs(x.mean() - 3.0) < 0.05 and abs(x.var() - 4.0) < 0.15  # variance, not sd


def test_dirichlet_mean():
    L = orc.lib()
    alpha = np.array([50.0, 3.0, 2.0, 1.0])
    out = np.zeros(4)
    acc = np.zeros(4)
    n = 4000
    for it in range(n):
        L.orc_dirichlet_rng(11, it, 100, 0, 4, p(alpha, C.c_double), p(out, C.c_double))
        assert abs(out.sum() - 1) < 1e-12
        acc += out
    assert np.allclose(acc / n, alpha / alpha.sum(), atol=0.01)


# ---- product stream == oracle stream (independent implementations) ----

def test_product_stream_is_bit_identical_to_oracle():
    import bayesrrcmd_amd as brr
    seed = 20261202
    for it in (0, 5, 0xFFFFFFFE):
        u, z = brr.marker_uz(seed, it, 2, 3000)
        uo, zo = _uz(seed, it, 3000, chain=2)
        assert np.array_equal(u, uo) and np.array_equal(z, zo)
    for mode in (0,):
        a = np.arange(5000, dtype=np.uint32)
        brr.shuffle(seed, 9, 1, a, mode)
        b = np.arange(5000, dtype=np.uint32)
        orc.lib().orc_shuffle_philox(seed, 9, 1, 5000, p(b, C.c_uint32))
        assert np.array_equal(a, b)
