"""Synthetic BayesR problems for the tests (genotypes, phenotype, marker arenas)."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

import bayesrrcmd_amd as brr

DEFAULT_S = (0.01, 0.001, 0.0001)


def spec_dosage(bed: np.ndarray, n: int) -> np.ndarray:
    """(m, n) dosage of the counted allele from PLINK bytes (-1 = missing)."""
    bits = np.unpackbits(bed[..., None], axis=-1, bitorder="little").reshape(bed.shape[0], -1)
    code = (bits[:, 0::2] | (bits[:, 1::2] << 1))[:, :n]
    return np.array([2, -1, 1, 0], dtype=np.int64)[code]


@dataclass
class Problem:
    n: int
    m: int
    bed: np.ndarray
    lut: np.ndarray
    ragged: brr.RaggedArena
    y: np.ndarray
    causal: np.ndarray
    beta_true: np.ndarray
    maf: np.ndarray

    def x_dense(self) -> np.ndarray:
        """(n, m) standardised genotype matrix (DenseMarker columns)."""
        d = spec_dosage(self.bed, self.n)
        d = np.where(d < 0, 3, d)
        return np.take_along_axis(self.lut, d, axis=1).T.copy()


def make_problem(n, m, seed, h2=0.5, causal_frac=0.01, maf=(0.01, 0.5), miss=0.0) -> Problem:
    bed, mafs = brr.synth_bed(seed, n, m, maf[0], maf[1], miss)
    lut = brr.build_dense(bed, n)
    ragged = brr.build_ragged(bed, n)
    rng = np.random.default_rng(seed)
    mc = max(1, int(round(causal_frac * m)))
    causal = np.sort(rng.choice(m, mc, replace=False))
    b = rng.normal(0.0, np.sqrt(h2 / mc), mc)
    d = spec_dosage(bed[causal], n)
    d = np.where(d < 0, 3, d)
    x = np.take_along_axis(lut[causal], d, axis=1)      # (mc, n)
    g = b @ x
    e = rng.normal(0.0, np.sqrt(1.0 - h2), n)
    y = (g + e).astype(np.float32).astype(np.float64)    # the reference stores y as float
    beta_true = np.zeros(m)
    beta_true[causal] = b
    return Problem(n, m, bed, lut, ragged, y, causal, beta_true, mafs)


@dataclass
class BwProblem:
    n: int
    m: int
    bed: np.ndarray
    ragged: brr.RaggedArena
    y: np.ndarray          # log event / censoring times (rounded through float)
    fail: np.ndarray       # 1 = event observed, 0 = right-censored
    causal: np.ndarray
    beta_true: np.ndarray


def make_bw_problem(n, m, seed, h2=0.5, causal_frac=0.05, censor=0.1, alpha=3.0, mu=4.0,
                    maf=(0.05, 0.5), miss=0.0) -> BwProblem:
    """Weibull survival data: log T = mu + g + (log E + EulerGamma) / alpha,
    E ~ Exp(1); var(g) / (var(g) + pi^2 / (6 alpha^2)) = h2; a fraction of the
    individuals is right-censored at C ~ U(0, T) (fail = 0, time = C)."""
    bed, _ = brr.synth_bed(seed, n, m, maf[0], maf[1], miss)
    ragged = brr.build_ragged(bed, n)
    lut = brr.build_dense(bed, n)
    rng = np.random.default_rng(seed + 1)
    mc = max(1, int(round(causal_frac * m)))
    causal = np.sort(rng.choice(m, mc, replace=False))
    b = rng.normal(0.0, 1.0, mc)
    d = spec_dosage(bed[causal], n)
    x = np.take_along_axis(lut[causal], np.where(d < 0, 3, d), axis=1)
    g = b @ x
    ve = np.pi ** 2 / (6 * alpha ** 2)
    scale = np.sqrt(h2 / (1 - h2) * ve / max(g.var(), 1e-12))
    g *= scale
    b *= scale
    logt = mu + g + (np.log(rng.exponential(1.0, n)) + 0.5772156649015329) / alpha
    fail = np.ones(n)
    cens = rng.random(n) < censor
    fail[cens] = 0.0
    logt[cens] = np.log(rng.uniform(0.0, 1.0, cens.sum())) + logt[cens]
    y = logt.astype(np.float32).astype(np.float64)
    beta_true = np.zeros(m)
    beta_true[causal] = b
    return BwProblem(n, m, bed, ragged, y, fail, causal, beta_true)
