#!/usr/bin/env python3
"""Benchmark of the MI355X BayesR marker sweep on one GPU.

Workload (default, `--config c2`): plain BayesR ppbayes at scale -- N = 100,000
individuals, M = 100,000 SNPs on the ragged layout (ascending index lists of
genotype 1 and 2), MAF ~ U(0.005, 0.5), 0.5% causal SNPs, h2 = 0.4, synthetic
data from a fixed seed. A step is one Gibbs iteration of the chain (updateMu,
the sweep over all M markers, the sigmaE / sigmaG / pi draws), run by the
persistent device kernel with genotypes and chain state resident in HBM.
`--config c1` is the small dense case (N = 5,000, M = 10,000 .bed 2-bit).
`--config c3` is the batched case: T = 64 traits (one phenotype and chain each,
h2 ~ U(0.1, 0.8), causal fraction ~ U(0.1%, 2%)) over one dense 2-bit matrix of
N = 50,000 x M = 50,000, chain length 100, all traits in one persistent launch.
`--config c5` is BayesW (Weibull survival): N = 100,000, M = 50,000 ragged,
0.5% causal, log T = 4 + g + (log E + EulerGamma) / 3, 10% right-censored,
mixture variances 0.001, 0.01, 0.1, 7 Gauss-Hermite points, chain length 100.
`--config c5-dense` is the same BayesW workload on the dense 2-bit layout.

Prints one JSON line: marker-updates/s (whole job), the roofline of the sweep
kernel measured with HIP events on its stream, and the CPU reference sweep
(the FP64 oracle built with the reference's Release flags) timed on a bounded
sample of the same workload on this host. The default line also carries a
"batched" object (the c3 workload measured the same way: batched vs plain), a
"bayesw" object (the c5 workload) and a "decode" object (the c2 .bed file
decoded on the device into either layout, brr_hip_load_bed).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Gibbs marker-updates/sec at N=100k (batched vs plain GPU) + % HBM roofline"
HBM_PEAK_GBS = 8000.0
S_DEFAULT = np.array([0.01, 0.001, 0.0001])

CONFIGS = {
    "c2": dict(workload="plain BayesR ppbayes, ragged layout, N=100000 M=100000, single chain",
               n=100_000, m=100_000, layout="ragged", maf=(0.005, 0.5), causal=0.005, h2=0.4,
               chain=200, precision="fp32"),
    "c2-dense": dict(workload="plain BayesR ppbayes, dense .bed 2-bit layout, N=100000 M=100000, single chain",
                     n=100_000, m=100_000, layout="dense", maf=(0.005, 0.5), causal=0.005, h2=0.4,
                     chain=200, precision="fp32"),
    "c1": dict(workload="plain BayesR ppbayes, dense .bed 2-bit, N=5000 M=10000, single chain",
               n=5_000, m=10_000, layout="dense", maf=(0.01, 0.5), causal=0.01, h2=0.5, chain=10),
    "c3": dict(workload="batched traits: 64 BayesR chains (one phenotype each) over one dense .bed 2-bit "
                        "matrix, N=50000 M=50000, shared marker order",
               n=50_000, m=50_000, layout="dense", maf=(0.01, 0.5), traits=64, causal=(0.001, 0.02),
               h2=(0.1, 0.8), chain=100, precision="fp32"),
    "c3-n100k": dict(workload="batched traits at the plain workload's size: 64 BayesR chains over one dense "
                              ".bed 2-bit matrix, N=100000 M=100000, shared marker order",
                     n=100_000, m=100_000, layout="dense", maf=(0.005, 0.5), traits=64, causal=(0.001, 0.02),
                     h2=(0.1, 0.8), chain=20, precision="fp32"),
    "c5": dict(workload="BayesW Weibull survival, ragged layout, N=100000 M=50000, single chain",
               n=100_000, m=50_000, layout="ragged", maf=(0.005, 0.5), causal=0.005, h2=0.5, alpha=3.0,
               mu=4.0, censor=0.1, C=(0.001, 0.01, 0.1), quad_points=7, chain=100),
    "c4": dict(workload="uneven batch: 32 independent cohorts, N ~ logU(2,000, 200,000), M ~ U(10,000, 80,000), "
                        "dense 2-bit or ragged per cohort (whichever has fewer bytes, at least one of each), "
                        "one BayesR chain each, run cohort after cohort",
               cohorts=32, n=(2000, 200000), m=(10000, 80000), maf=(0.01, 0.5), h2=(0.2, 0.7), causal=0.005,
               chain=20, layout="mixed"),
    "c5-dense": dict(workload="BayesW Weibull survival, dense .bed 2-bit layout (DenseBayesW), N=100000 M=50000, "
                              "single chain",
                     n=100_000, m=50_000, layout="dense", maf=(0.005, 0.5), causal=0.005, h2=0.5, alpha=3.0,
                     mu=4.0, censor=0.1, C=(0.001, 0.01, 0.1), quad_points=7, chain=100),
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_data(cfg, seed, threads):
    import bayesrrcmd_amd as brr
    n, m = cfg["n"], cfg["m"]
    t0 = time.time()
    bed, maf = brr.synth_bed(seed, n, m, cfg["maf"][0], cfg["maf"][1], 0.0, threads=threads)
    log(f"genotypes {m} x {n} generated in {time.time() - t0:.1f}s")
    rng = np.random.default_rng(seed)
    mc = max(1, int(round(cfg["causal"] * m)))
    causal = np.sort(rng.choice(m, mc, replace=False))
    b = rng.normal(0.0, np.sqrt(cfg["h2"] / mc), mc)
    lut_c = brr.build_dense(bed[causal], n, threads=threads)
    bits = np.unpackbits(bed[causal][..., None], axis=-1, bitorder="little").reshape(mc, -1)
    code = (bits[:, 0::2] | (bits[:, 1::2] << 1))[:, :n]
    dos = np.array([2, 3, 1, 0])[code]
    g = b @ np.take_along_axis(lut_c, dos, axis=1)
    y = (g + rng.normal(0.0, np.sqrt(1.0 - cfg["h2"]), n)).astype(np.float32).astype(np.float64)
    t0 = time.time()
    if cfg["layout"] == "ragged":
        arena = brr.build_ragged(bed, n, threads=threads)
        dt = time.time() - t0
        log(f"ragged lists built in {dt:.1f}s ({arena.nbytes() / 1e9:.2f} GB, {threads} threads)")
        return dict(bed=bed, y=y, ragged=arena, lut=None, build_s=dt)
    lut = brr.build_dense(bed, n, threads=threads)
    dt = time.time() - t0
    log(f"dense tables built in {dt:.1f}s ({threads} threads)")
    return dict(bed=bed, y=y, ragged=None, lut=lut, build_s=dt)


def make_batch_data(cfg, seed, threads):
    """Genotypes plus T phenotypes: per trait h2 ~ U(h2), causal fraction ~ U(causal),
    causal set uniform, beta ~ N(0, h2 / Mc), y = X beta + e, e ~ N(0, 1 - h2)."""
    import bayesrrcmd_amd as brr
    n, m, T = cfg["n"], cfg["m"], cfg["traits"]
    t0 = time.time()
    bed, _ = brr.synth_bed(seed, n, m, cfg["maf"][0], cfg["maf"][1], 0.0, threads=threads)
    lut = brr.build_dense(bed, n, threads=threads)
    log(f"genotypes {m} x {n} and tables generated in {time.time() - t0:.1f}s")
    rng = np.random.default_rng(seed + 3)
    beta = np.zeros((m, T))
    h2 = rng.uniform(cfg["h2"][0], cfg["h2"][1], T)
    frac = rng.uniform(cfg["causal"][0], cfg["causal"][1], T)
    for t in range(T):
        mc = max(1, int(round(frac[t] * m)))
        causal = rng.choice(m, mc, replace=False)
        beta[causal, t] = rng.normal(0.0, np.sqrt(h2[t] / mc), mc)
    t0 = time.time()
    g = brr.synth_xb(bed, n, lut, beta, threads=threads)
    y = (g + rng.normal(0.0, 1.0, (n, T)) * np.sqrt(1.0 - h2)).astype(np.float32).astype(np.float64)
    log(f"{T} phenotypes in {time.time() - t0:.1f}s")
    return dict(bed=bed, lut=lut, y=y, h2=h2, frac=frac)


def algorithmic_bytes(cfg, data, storage, markers_per_step, changes, eps_elem=4):
    """Per step: sum_j B_j with B_j = G_j + 4 N (1 + delta_j): G_j the column
    bytes of the device copy (index lists: 4 (n1 + n2 + nmiss) + 32; 2-bit
    codes: ceil(N/4)), 4 N one read of an FP32 residual, delta_j = 1 when marker
    j's beta changed (one more residual pass, the update). The figure is fixed
    whatever precision the kernel computes in, so the FP32 and FP64 paths are
    priced alike. The residual terms are logical passes: the kernel keeps the
    residual in registers."""
    n, m = cfg["n"], cfg["m"]
    if cfg["layout"] == "ragged" and storage == 1:
        r = data["ragged"]
        g_total = 4.0 * (r.ones.size + r.twos.size + r.miss.size) + 32.0 * m
    else:
        g_total = float(m * ((n + 3) // 4))
    return g_total * (markers_per_step / m) + eps_elem * n * (markers_per_step + changes)


def chain_orders(seed, m, count):
    """The marker orders of iterations 0 .. count-1 of one chain (each iteration
    reshuffles the previous order, as the reference's runGibbs does)."""
    import bayesrrcmd_amd as brr
    order = np.arange(m, dtype=np.uint32)
    orders = np.empty((count, m), dtype=np.uint32)
    for it in range(count):
        brr.shuffle(seed, it, 0, order)
        orders[it] = order
    return orders


def errors_vs_fp64(h, brr, data, m, args, orders, hist32, fresh_chain):
    """Errors of the FP32-residual chain against the FP64 path of the same
    device code (whose parity with the FP64 oracle the GPU tests pin):
    - per marker, from one identical late-chain state (the end of the timed
      chain): one FP32 and one FP64 sweep over the same order and draws; num
      (computeNum, whose dot part is x_j . eps) and beta compared marker by marker;
    - the whole timed chain re-run in FP64 from the same initial state and
      random stream: sigmaG, sigmaE and h2 = sigmaG / (sigmaG + sigmaE) per
      iteration, the first iteration whose model size m0 differs (a component
      flipped and the chains parted), and both chains' posterior means of h2."""
    t0 = time.time()
    st = h.get_state()
    sc = h.get_scalars()
    order = orders[-1].copy()
    brr.shuffle(args.seed, len(orders), 0, order)
    u, z = brr.marker_uz(args.seed, len(orders), 0, m)

    def one_sweep(precision):
        h.set_precision(precision, args.refresh)
        h.set_state(st["eps"], st["beta"], st["comp"])
        h.set_scalars(sc["mu"], sc["sigmaE"], sc["epsilon_sum"], sc["sigmaG"], sc["pi"])
        h.record_num(True)
        h.sweep(order, u, z)
        out = h.get_state()
        out["num"] = h.get_num()
        h.record_num(False)
        return out

    a = one_sweep(brr.PRECISION_FP32_REFRESH)
    b = one_sweep(brr.PRECISION_FP64)
    scale = float(np.max(np.abs(b["num"])))
    dnum = np.abs(a["num"] - b["num"])
    same = a["comp"] == b["comp"]
    dbeta = np.abs(a["beta"] - b["beta"])[same]
    # the chain again in FP64 (its device rate over the timed iterations, after the warm-up ones)
    fresh_chain()
    rows, k_ms = [], 0.0
    for it in range(len(orders)):
        rows.append(h.run_chain(args.seed, 0, it, orders[it])[0])
        if it >= args.warmup:
            k_ms += h.stats()["kernel_ms"]
    hist64 = np.array(rows)
    h.set_precision(brr.PRECISION_FP32_REFRESH, args.refresh)
    h2a = hist32[:, 4] / (hist32[:, 4] + hist32[:, 1])
    h2b = hist64[:, 4] / (hist64[:, 4] + hist64[:, 1])
    parted = np.flatnonzero(hist32[:, 2] != hist64[:, 2])
    first = int(parted[0]) if parted.size else None
    upto = first if first is not None else len(orders)
    burn = min(len(orders) // 2, 50)
    sem = lambda x: float(np.std(x) / np.sqrt(max(1, len(x))))
    out = {
        "reference": "device FP64 path from the same state and random stream (pinned to the FP64 oracle "
                     "by the GPU parity tests)",
        "state": f"end of the timed chain (iteration {len(orders)}), one sweep of all {m} markers",
        "num_max_abs": float(dnum.max()), "num_max_rel_to_max_num": float(dnum.max() / scale),
        "num_rms": float(np.sqrt(np.mean(dnum ** 2))), "num_scale_max_abs": scale,
        "components_identical": bool(same.all()), "components_differing": int((~same).sum()),
        "beta_max_abs": float(dbeta.max()) if dbeta.size else 0.0,
        "eps_max_abs_after_sweep": float(np.max(np.abs(a["eps"] - b["eps"]))),
        "chain_iterations": len(orders),
        "chain_first_parting_iteration": first,
        "h2_max_abs_before_parting": float(np.max(np.abs(h2a[:upto] - h2b[:upto]))) if upto else None,
        "sigmaG_max_rel_before_parting": (float(np.max(np.abs(hist32[:upto, 4] - hist64[:upto, 4]) /
                                                       hist64[:upto, 4])) if upto else None),
        "h2_abs_err_at_iterations": {str(i): float(abs(h2a[i] - h2b[i]))
                                     for i in (0, 1, 2, 5, 10, 20, 50, 100, 150, len(orders) - 1) if i < len(orders)},
        "h2_final_abs": float(abs(h2a[-1] - h2b[-1])),
        "sigmaG_final_rel": float(abs(hist32[-1, 4] - hist64[-1, 4]) / hist64[-1, 4]),
        "h2_posterior_mean_fp32": float(np.mean(h2a[burn:])), "h2_posterior_mean_fp64": float(np.mean(h2b[burn:])),
        "h2_posterior_mean_mc_sem": sem(h2b[burn:]),
        "fp64_chain_device_marker_updates_per_s": (m * (len(orders) - args.warmup) / (k_ms * 1e-3)
                                                   if k_ms > 0 else None),
        "seconds": time.time() - t0,
    }
    log(f"errors vs fp64: num max {out['num_max_abs']:.3e} (rel {out['num_max_rel_to_max_num']:.2e}), "
        f"components identical {out['components_identical']}, chain parts at {first}, "
        f"h2 max err {out['h2_max_abs_before_parting']}")
    return out


def cpu_baseline(cfg, data, seed, budget_s, window=0):
    """Reference-style sweep on the host: the FP64 oracle built with -O3
    -ffast-math -funroll-loops, one thread, over the first markers of the same
    workload for a chain of a few iterations."""
    import ctypes as C
    sys.path.insert(0, str(ROOT / "tests"))
    import _oracle as orc
    L = orc.lib(fast=True)
    n = cfg["n"]
    m_s = min(cfg["m"], 2000 if cfg["layout"] == "ragged" else 4000)
    iters = 0
    t_total = 0.0
    mk = 0
    if cfg["layout"] == "ragged":
        r = data["ragged"]

        class Sub:
            pass
        sub = Sub()
        sub.ones_off = np.ascontiguousarray(r.ones_off[: m_s + 1])
        sub.twos_off = np.ascontiguousarray(r.twos_off[: m_s + 1])
        sub.miss_off = np.ascontiguousarray(r.miss_off[: m_s + 1])
        sub.ones, sub.twos, sub.miss = r.ones, r.twos, r.miss
        sub.stats = np.ascontiguousarray(r.stats[:m_s])
        prob = orc.OracleProblem(n, m_s, orc.RAGGED, 4, 1, S_DEFAULT, ragged=sub)
    else:
        prob = orc.OracleProblem(n, m_s, orc.DENSE, 4, 1, S_DEFAULT, bed=data["bed"][:m_s],
                                 lut=data["lut"][:m_s])
    y = np.ascontiguousarray(data["y"])
    order = np.arange(m_s, dtype=np.uint32)
    hist = np.zeros((1, 5))
    # one iteration at a time until the budget is spent (same chain, continued)
    L.orc_init(C.byref(prob.data), C.byref(prob.model), C.byref(prob.state), orc.p(y, C.c_double), seed, 0)
    u = np.zeros(m_s)
    z = np.zeros(m_s)
    while t_total < budget_s and iters < cfg["chain"]:
        t0 = time.perf_counter()
        L.orc_update_mu(C.byref(prob.data), C.byref(prob.state), seed, iters, 0)
        L.orc_fill_marker_uz(seed, iters, 0, m_s, orc.p(u, C.c_double), orc.p(z, C.c_double))
        L.orc_shuffle_philox(seed, iters, 0, m_s, orc.p(order, C.c_uint32))
        if window:
            L.orc_sweep_async(C.byref(prob.data), C.byref(prob.model), C.byref(prob.state),
                              orc.p(order, C.c_uint32), orc.p(u, C.c_double), orc.p(z, C.c_double), window)
        else:
            L.orc_sweep(C.byref(prob.data), C.byref(prob.model), C.byref(prob.state), orc.p(order, C.c_uint32),
                        orc.p(u, C.c_double), orc.p(z, C.c_double), None)
        L.orc_post_sweep(C.byref(prob.data), C.byref(prob.model), C.byref(prob.state), seed, iters, 0)
        t_total += time.perf_counter() - t0
        iters += 1
        mk += m_s
    return dict(value=mk / t_total, unit="marker-updates/s", cores=1, kind="port",
                sample=(f"oracle (FP64 restatement of the reference sweep"
                        f"{f', asynchronous windows of {window}' if window else ''}, -O3 -ffast-math "
                        f"-funroll-loops, 1 thread): first {m_s} markers of the same {cfg['layout']} "
                        f"workload at N={n}, {iters} chain iterations from init, {t_total:.1f}s"))


def cpu_baseline_batched(cfg, data, seed, budget_s, workers):
    """T independent reference chains on the host, as T brr processes would run:
    `workers` threads (ctypes releases the GIL) each drive one oracle chain
    (-O3 -ffast-math, FP64) over the first markers of the batched workload."""
    import ctypes as C
    import threading
    sys.path.insert(0, str(ROOT / "tests"))
    import _oracle as orc
    L = orc.lib(fast=True)
    n = cfg["n"]
    m_s = min(cfg["m"], 2000)
    bed = np.ascontiguousarray(data["bed"][:m_s])
    lut = np.ascontiguousarray(data["lut"][:m_s])
    done = [0] * workers
    spent = [0.0] * workers

    def worker(t):
        prob = orc.OracleProblem(n, m_s, orc.DENSE, 4, 1, S_DEFAULT, bed=bed, lut=lut)
        y = np.ascontiguousarray(data["y"][:, t])
        L.orc_init(C.byref(prob.data), C.byref(prob.model), C.byref(prob.state), orc.p(y, C.c_double), seed, t)
        order = np.arange(m_s, dtype=np.uint32)
        u, z = np.zeros(m_s), np.zeros(m_s)
        it = 0
        while spent[t] < budget_s and it < cfg["chain"]:
            t0 = time.perf_counter()
            L.orc_update_mu(C.byref(prob.data), C.byref(prob.state), seed, it, t)
            L.orc_fill_marker_uz(seed, it, t, m_s, orc.p(u, C.c_double), orc.p(z, C.c_double))
            L.orc_shuffle_philox(seed, it, 0, m_s, orc.p(order, C.c_uint32))
            L.orc_sweep(C.byref(prob.data), C.byref(prob.model), C.byref(prob.state), orc.p(order, C.c_uint32),
                        orc.p(u, C.c_double), orc.p(z, C.c_double), None)
            L.orc_post_sweep(C.byref(prob.data), C.byref(prob.model), C.byref(prob.state), seed, it, t)
            spent[t] += time.perf_counter() - t0
            it += 1
            done[t] += m_s

    t0 = time.perf_counter()
    ths = [threading.Thread(target=worker, args=(t,)) for t in range(workers)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    wall = time.perf_counter() - t0
    return dict(value=sum(done) / wall, unit="marker-updates/s", cores=workers, kind="port",
                sample=(f"oracle (FP64 restatement of the reference sweep, -O3 -ffast-math -funroll-loops): "
                        f"{workers} chains of the batched workload on {workers} host threads, one chain per "
                        f"thread, first {m_s} markers at N={n}, {sum(done) // m_s} chain iterations in total, "
                        f"{wall:.1f}s wall"))


def make_bw_data(cfg, seed, threads):
    """Genotypes + Weibull survival: log T = mu + g + (log E + EulerGamma) / alpha,
    E ~ Exp(1), var(g) / (var(g) + pi^2 / (6 alpha^2)) = h2; a uniform fraction
    `censor` of individuals is right-censored at C ~ U(0, T) (fail = 0)."""
    import bayesrrcmd_amd as brr
    n, m = cfg["n"], cfg["m"]
    t0 = time.time()
    bed, _ = brr.synth_bed(seed, n, m, cfg["maf"][0], cfg["maf"][1], 0.0, threads=threads)
    log(f"genotypes {m} x {n} generated in {time.time() - t0:.1f}s")
    rng = np.random.default_rng(seed + 5)
    mc = max(1, int(round(cfg["causal"] * m)))
    causal = np.sort(rng.choice(m, mc, replace=False))
    beta = np.zeros(m)
    beta[causal] = rng.normal(0.0, 1.0, mc)
    lut_c = brr.build_dense(bed[causal], n, threads=threads)
    g = brr.synth_xb(bed[causal], n, lut_c, beta[causal], threads=threads)
    ve = np.pi ** 2 / (6 * cfg["alpha"] ** 2)
    g *= np.sqrt(cfg["h2"] / (1 - cfg["h2"]) * ve / g.var())
    logt = cfg["mu"] + g + (np.log(rng.exponential(1.0, n)) + 0.5772156649015329) / cfg["alpha"]
    fail = np.ones(n)
    cens = rng.choice(n, int(round(cfg["censor"] * n)), replace=False)
    fail[cens] = 0.0
    logt[cens] += np.log(rng.uniform(0.0, 1.0, cens.size))
    y = logt.astype(np.float32).astype(np.float64)
    t0 = time.time()
    if cfg["layout"] == "dense":
        lut = brr.build_dense(bed, n, threads=threads)
        log(f"dense tables built in {time.time() - t0:.1f}s")
        return dict(bed=bed, y=y, fail=fail, ragged=None, lut=lut)
    arena = brr.build_ragged(bed, n, threads=threads)
    log(f"ragged lists built in {time.time() - t0:.1f}s ({arena.nbytes() / 1e9:.2f} GB)")
    return dict(bed=bed, y=y, fail=fail, ragged=arena, lut=None)


def cpu_baseline_bw(cfg, data, seed, budget_s):
    """Reference-style BayesW chain on the host: the FP64 oracle built with -O3
    -ffast-math -funroll-loops, one thread, over the first markers of the same
    workload (sampleMu, sweep, sampleAlpha, sigma_b / pi per iteration)."""
    import ctypes as C
    sys.path.insert(0, str(ROOT / "tests"))
    import _oracle as orc
    L = orc.lib(fast=True)
    n = cfg["n"]
    if cfg["layout"] == "dense":
        # the dense kernel's sums run over every individual per quadrature node and density value
        m_s = min(cfg["m"], 100)
        w = orc.BayesWProblem(n, m_s, None, data["fail"], K=4, C=tuple(cfg["C"]), quad_points=cfg["quad_points"],
                              bed=np.ascontiguousarray(data["bed"][:m_s]), lut=np.ascontiguousarray(data["lut"][:m_s]))
    else:
        m_s = min(cfg["m"], 1000)
        r = data["ragged"]

        class Sub:
            pass
        sub = Sub()
        sub.ones_off = np.ascontiguousarray(r.ones_off[: m_s + 1])
        sub.twos_off = np.ascontiguousarray(r.twos_off[: m_s + 1])
        sub.miss_off = np.ascontiguousarray(r.miss_off[: m_s + 1])
        sub.ones, sub.twos, sub.miss = r.ones, r.twos, r.miss
        sub.stats = np.ascontiguousarray(r.stats[:m_s])
        w = orc.BayesWProblem(n, m_s, sub, data["fail"], K=4, C=tuple(cfg["C"]), quad_points=cfg["quad_points"])
    y = np.ascontiguousarray(data["y"])
    fail = w.fail
    L.orc_bw_init(C.byref(w.data), C.byref(w.model), C.byref(w.state), orc.p(y, C.c_double),
                  orc.p(fail, C.c_double))
    order = np.arange(m_s, dtype=np.uint32)
    u, z = np.zeros(m_s), np.zeros(m_s)
    iters, t_total = 0, 0.0
    while t_total < budget_s and iters < cfg["chain"]:
        t0 = time.perf_counter()
        assert L.orc_bw_sample_mu(C.byref(w.data), C.byref(w.state), seed, iters, 0) == 0
        w.vi[:] = np.exp(w.state.alpha * w.eps - 0.577215664901532)
        L.orc_shuffle_philox(seed, iters, 0, m_s, orc.p(order, C.c_uint32))
        w.v[:] = 1.0
        L.orc_fill_marker_uz(seed, iters, 0, m_s, orc.p(u, C.c_double), orc.p(z, C.c_double))
        assert L.orc_bw_sweep(C.byref(w.data), C.byref(w.model), C.byref(w.state), orc.p(order, C.c_uint32),
                              orc.p(u, C.c_double), seed, iters, 0) == 0
        assert L.orc_bw_sample_alpha(C.byref(w.data), C.byref(w.state), orc.p(fail, C.c_double), seed, iters,
                                     0) == 0
        L.orc_bw_post_sweep(C.byref(w.data), C.byref(w.model), C.byref(w.state), seed, iters, 0)
        t_total += time.perf_counter() - t0
        iters += 1
    return dict(value=iters * m_s / t_total, unit="marker-updates/s", cores=1, kind="port",
                sample=(f"oracle (FP64 restatement of the reference BayesW chain, -O3 -ffast-math "
                        f"-funroll-loops, 1 thread): first {m_s} markers of the same {cfg['layout']} workload at "
                        f"N={n}, {iters} chain iterations from init, {t_total:.1f}s"))


def run_bayesw(cfg, args, with_cpu):
    """The BayesW workload on the device: returns the fields of its JSON line."""
    import bayesrrcmd_amd as brr
    n, m = cfg["n"], cfg["m"]
    steps = args.steps if (args.steps is not None and args.config in ("c5", "c5-dense")) else cfg["chain"]
    data = make_bw_data(cfg, args.seed, args.threads)
    K = len(cfg["C"]) + 1
    h = brr.HipBayesW(n, m, K, quad_points=cfg["quad_points"], workgroups=args.workgroups)
    t0 = time.time()
    if cfg["layout"] == "dense":
        h.upload_dense(data["bed"], data["lut"])
        nnz = 0
    else:
        h.upload_ragged(data["ragged"])
        nnz = data["ragged"].ones.size + data["ragged"].twos.size
    h.set_model(np.array(cfg["C"]), data["fail"])
    log(f"bayesw: uploaded in {time.time() - t0:.1f}s")

    def fresh_chain():
        eps, sc = brr.bayesw_init(data["y"], data["fail"], m, K)
        h.set_state(eps, np.zeros(m), np.zeros(m, np.int32))
        h.set_scalars(sc["mu"], sc["alpha"], sc["sigma_b"], sc["d"], sc["pi"], sc["v"])

    # one chain from init: its first `warmup` iterations untimed, the next `steps` timed
    fresh_chain()
    orders = chain_orders(args.seed, m, args.warmup + steps)
    for it in range(args.warmup):
        h.run_chain(args.seed, 0, it, orders[it])
    kernel_ms, changes, rounds, hist, phases = [], [], [], [], []
    t_start = time.perf_counter()
    for it in range(args.warmup, args.warmup + steps):
        hist.append(h.run_chain(args.seed, 0, it, orders[it])[0])
        s = h.stats()
        kernel_ms.append(s["kernel_ms"])
        changes.append(s["changes"])
        rounds.append(s["rounds"])
        phases.append(list(s["phase_ms"].values()))
        if it % 20 == 0 or it == args.warmup + steps - 1:
            ph = ", ".join(f"{k} {v:.1f}" for k, v in s["phase_ms"].items())
            log(f"bayesw iteration {it}: {s['kernel_ms']:.2f} ms, rounds {s['rounds']}, changes {s['changes']}, "
                f"mu {hist[-1][0]:.4f} alpha {hist[-1][1]:.4f} nonzero {hist[-1][3]:.0f}, "
                f"core clock {s['core_clock_mhz']:.0f} MHz | phases ms: {ph}")
    wall = time.perf_counter() - t_start
    stats = h.stats()
    hist = np.array(hist)
    avg_ms = float(np.mean(kernel_ms))
    # B_j = G_j + 4 N + 12 N delta_j: the 2-bit column, one FP32 vi read, and on
    # a change an eps read, an eps write and a vi write (FP32 words, whatever
    # precision the kernel computes in)
    g_col = float((n + 3) // 4)
    bytes_per_launch = float(np.mean([m * g_col + 4.0 * n * m + 12.0 * n * c for c in changes]))
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    burn = min(steps // 2, 30)
    out = {
        "value": float(steps) * m / wall,
        "unit": "marker-updates/s",
        "steps": steps,
        "ms_per_step": wall * 1e3 / steps,
        "dtype": "f64",
        "config": {
            "workload": cfg["workload"], "N": n, "M": m,
            "layout": ("dense 2-bit" if cfg["layout"] == "dense" else "ragged (held on the device as 2-bit codes)"),
            "maf": list(cfg["maf"]), "causal_fraction": cfg["causal"], "C": list(cfg["C"]),
            "quad_points": cfg["quad_points"], "censored_fraction": cfg["censor"], "timed_iterations": steps,
            "chain": f"one chain from init: {args.warmup} untimed warm-up iterations, then the timed ones",
            "precision": "fp64", "workgroups": stats["workgroups"], "block_markers": stats["block_markers"],
            "slice": stats["slice"], "ragged_list_bytes": 4 * nnz if nnz else None,
            "parallelism": "single GPU, persistent cooperative kernel",
            "alpha_posterior_mean": float(hist[burn:, 1].mean()) if steps > 1 else float("nan"),
            "mu_posterior_mean": float(hist[burn:, 0].mean()) if steps > 1 else float("nan"),
            "alpha_simulated": cfg["alpha"], "mu_simulated": cfg["mu"],
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None,
            "achieved_basis": ("logical: B_j = G_j + 4 N + 12 N delta_j per marker-update (the 2-bit column, "
                               "an FP32 vi pass, and on a change an eps read, an eps write and a vi write, "
                               "whatever precision the kernel computes in); eps and vi live in registers"),
            "limiter": "latency: exchange rounds, class likelihoods and ARMS draws",
            "kernel": "brr::bayesw_sweep_kernel (one launch per iteration)",
            "avg_launch_ms": avg_ms, "algorithmic_bytes_per_launch": bytes_per_launch,
            "rounds_per_launch": float(np.mean(rounds)), "changes_per_launch": float(np.mean(changes)),
            "phase_ms_per_launch": dict(zip(brr.HipBayesW.PHASES, np.mean(phases, axis=0).round(3).tolist())),
        },
    }
    h.close()
    if with_cpu:
        log("timing the CPU reference BayesW chain ...")
        out["cpu_baseline"] = cpu_baseline_bw(cfg, data, args.seed, args.cpu_budget)
    return out


def run_batched(cfg, args, with_cpu):
    """The batched workload on the device: returns the fields of its JSON line."""
    import bayesrrcmd_amd as brr
    n, m, T = cfg["n"], cfg["m"], cfg["traits"]
    steps = args.steps if (args.steps is not None and args.config.startswith("c3")) else cfg["chain"]
    precision = args.precision if args.config.startswith("c3") and args.precision else cfg.get("precision", "fp64")
    data = make_batch_data(cfg, args.seed, args.threads)
    h = brr.HipSweep(n, m, 4, 1, brr.LAYOUT_DENSE_2BIT, workgroups=args.workgroups, traits=T)
    t0 = time.time()
    h.upload_dense(data["bed"], data["lut"])
    h.set_model(S_DEFAULT)
    if precision == "fp32":
        h.set_precision(brr.PRECISION_FP32)
    log(f"batched N={n}: uploaded in {time.time() - t0:.1f}s; precision {precision}")

    def fresh_chain(seed):
        eps = np.empty((n, T))
        for t in range(T):
            e, sc = brr.bayesr_init(data["y"][:, t], 1, 4, seed, chain=t)
            eps[:, t] = e
            h.set_scalars(sc["mu"], sc["sigmaE"], sc["epsilon_sum"], sc["sigmaG"], sc["pi"], trait=t)
        h.set_state(eps, np.zeros((m, T)), np.zeros((m, T), np.int32))

    # one set of chains from init: the first `warmup` iterations untimed, the next `steps` timed
    fresh_chain(args.seed)
    orders = chain_orders(args.seed, m, args.warmup + steps)
    for it in range(args.warmup):
        h.run_chain(args.seed, 0, it, orders[it])
    kernel_ms, changes, rounds, phases, hist = [], [], [], [], []
    t_start = time.perf_counter()
    for it in range(args.warmup, args.warmup + steps):
        hist.append(h.run_chain(args.seed, 0, it, orders[it])[0])
        s = h.stats()
        kernel_ms.append(s["kernel_ms"])
        changes.append(s["changes"])
        rounds.append(s["rounds"])
        phases.append(list(s["phase_ms"].values()))
        if it % 20 == 0 or it == args.warmup + steps - 1:
            ph = ", ".join(f"{k} {v:.1f}" for k, v in s["phase_ms"].items())
            log(f"batched iteration {it}: {s['kernel_ms']:.2f} ms, rounds {s['rounds']}, changes {s['changes']}, "
                f"decision rounds {s['decision_rounds']} | phases ms: {ph}")
    wall = time.perf_counter() - t_start
    stats = h.stats()
    hist = np.array(hist)                                  # (steps, T, 5)
    avg_ms = float(np.mean(kernel_ms))
    # B_j = G_j / T + 4 N (1 + delta_j) per (marker, trait): the column once for
    # all traits, one FP32 residual read per trait, one more per change
    g_col = float((n + 3) // 4)
    bytes_per_launch = float(np.mean([m * g_col + 4.0 * n * (T * m + c) for c in changes]))
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    # matrix-core work per launch: dots 2 N T per marker, Gram 2 N per pair of a 32-marker block
    flops = 2.0 * n * T * m + 2.0 * n * (m / 32.0) * (32 * 31 / 2)
    burn = min(steps // 2, 30)
    sE, sG = hist[burn:, :, 1], hist[burn:, :, 4]
    h2_post = (sG / (sG + sE)).mean(axis=0) if steps > 1 else np.full(T, np.nan)
    out = {
        "value": float(steps) * m * T / wall,
        "unit": "marker-updates/s",
        "steps": steps,
        "ms_per_step": wall * 1e3 / steps,
        "dtype": "f32+f64" if precision == "fp32" else "f64",
        "config": {
            "workload": cfg["workload"], "N": n, "M": m, "traits": T, "layout": "dense 2-bit",
            "maf": list(cfg["maf"]), "causal_fraction_range": list(cfg["causal"]), "h2_range": list(cfg["h2"]),
            "S": S_DEFAULT.tolist(), "timed_iterations": steps,
            "chain": f"one set of chains from init: {args.warmup} untimed warm-up iterations, then the timed ones",
            "precision": ("fp32 residual tiles (f32 matrix cores, per-tile products summed in fp64), fp64 "
                          "decisions and draws, no refresh" if precision == "fp32" else "fp64"),
            "workgroups": stats["workgroups"], "block_markers": stats["block_markers"], "slice": stats["slice"],
            "parallelism": "single GPU, persistent cooperative kernel, all traits in one launch",
            "h2_abs_err_mean": float(np.mean(np.abs(h2_post - data["h2"]))),
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None,
            "achieved_basis": ("logical: B_j = G_j / T + 4 N (1 + delta_j) per (marker, trait) update (the "
                               "column once for all traits, an FP32 residual pass per trait, one more per change); "
                               "the residual tiles live in registers"),
            "limiter": "latency: decision rounds and exchange rounds",
            "kernel": "brr::bayesr_batch_kernel (one launch per iteration)",
            "avg_launch_ms": avg_ms, "algorithmic_bytes_per_launch": bytes_per_launch,
            "mfma_tflops": flops / (avg_ms * 1e-3) / 1e12,
            "mfma_peak_tflops": 157.3 if precision == "fp32" else 78.6,
            "rounds_per_launch": float(np.mean(rounds)), "changes_per_launch": float(np.mean(changes)),
            "phase_ms_per_launch": dict(zip(brr.HipSweep.BATCH_PHASES, np.mean(phases, axis=0).round(3).tolist())),
        },
    }
    h.close()
    if with_cpu:
        log("timing the CPU reference chains (batched) ...")
        workers = max(1, min(T, args.threads))
        out["cpu_baseline"] = cpu_baseline_batched(cfg, data, args.seed, args.cpu_budget, workers)
    return out


def run_cohorts(cfg, args, with_cpu):
    """The uneven batch: every cohort's .bed decoded on the device (brr_hip_load_bed) into a context sized to
    its N; then all cohorts' chains of `steps` iterations side by side (brr_hip_run_cohorts: cohorts packed
    onto disjoint workgroup sets of persistent launches, the launches captured in one hipGraph). The lane
    runs per slice (the one launch-shape choice the cohorts share) are picked by timing 2 iterations of the
    whole batch at each feasible value. value = all cohorts' marker-updates / the wall time of the timed
    call (orders upload, graph capture and launch included; data generation and upload excluded)."""
    import bayesrrcmd_amd as brr
    steps = args.steps if args.steps is not None else cfg["chain"]
    rng = np.random.default_rng(args.seed + 4)
    K = cfg["cohorts"]
    ns = np.exp(rng.uniform(np.log(cfg["n"][0]), np.log(cfg["n"][1]), K)).astype(int)
    ms = rng.integers(cfg["m"][0], cfg["m"][1] + 1, K)
    h2s = rng.uniform(cfg["h2"][0], cfg["h2"][1], K)
    # bytes of each layout at this MAF range: ragged 4 (n1 + n2) + 32 per marker with E[n1 + n2] / N =
    # E[2p - p^2] for p ~ U(lo, hi); dense ceil(N/4)
    lo, hi = cfg["maf"]
    nnz_frac = (hi ** 2 - lo ** 2) / (hi - lo) - (hi ** 3 - lo ** 3) / (3 * (hi - lo))
    layouts = ["ragged" if 4 * nnz_frac * n + 32 < (n + 3) // 4 else "dense" for n in ns]
    for want in ("dense", "ragged"):
        if want not in layouts:
            layouts[int(np.argmin(ns)) if want == "ragged" else int(np.argmax(ns))] = want
    precision = args.precision or "fp32"
    hs, ys, cpu = [], [], None
    t0 = time.time()
    for c in range(K):
        n, m, lay = int(ns[c]), int(ms[c]), layouts[c]
        bed, _ = brr.synth_bed(args.seed + 100 + c, n, m, lo, hi, 0.0, threads=args.threads)
        mc = max(1, int(round(cfg["causal"] * m)))
        causal = np.sort(rng.choice(m, mc, replace=False))
        b = rng.normal(0.0, np.sqrt(h2s[c] / mc), mc)
        g = brr.synth_xb(bed[causal], n, brr.build_dense(bed[causal], n, threads=args.threads), b,
                         threads=args.threads)
        y = (g + rng.normal(0.0, np.sqrt(1.0 - h2s[c]), n)).astype(np.float32).astype(np.float64)
        code = brr.LAYOUT_RAGGED if lay == "ragged" else brr.LAYOUT_DENSE_2BIT
        h = brr.HipSweep(n, m, 4, 1, code)
        h.load_bed(brr.bed_file_image(bed), code)
        h.set_model(S_DEFAULT)
        if precision == "fp32":
            h.set_precision(brr.PRECISION_FP32_REFRESH, args.refresh)
        hs.append(h)
        ys.append(y)
        if with_cpu and cpu is None and lay == "dense":
            k = min(m, 4000)                          # the markers the CPU sample uses (dense layout)
            data = dict(bed=bed[:k], lut=brr.build_dense(bed[:k], n, threads=args.threads), y=y, ragged=None)
            cpu = cpu_baseline(dict(n=n, m=m, layout="dense", chain=steps), data, args.seed, args.cpu_budget)
        del bed
    log(f"{K} cohorts generated and decoded on the device in {time.time() - t0:.1f}s "
        f"(N {int(ns.min())}..{int(ns.max())}, M {int(ms.min())}..{int(ms.max())}, "
        f"{layouts.count('dense')} dense / {layouts.count('ragged')} ragged)")

    def orders_for(steps_):
        out = []
        for c in range(K):
            o = np.empty((steps_, int(ms[c])), dtype=np.uint32)
            cur = np.arange(int(ms[c]), dtype=np.uint32)
            for it in range(steps_):
                brr.shuffle(args.seed, it, c, cur)
                o[it] = cur
            out.append(o)
        return out

    def fresh():
        for c, h in enumerate(hs):
            eps, sc = brr.bayesr_init(ys[c], 1, 4, args.seed, chain=c)
            h.set_state(eps, np.zeros(h.m), np.zeros(h.m, np.int32))
            h.set_scalars(sc["mu"], sc["sigmaE"], sc["epsilon_sum"], sc["sigmaG"], sc["pi"])

    # the autotuned shape: the least lane runs per slice a cohort may use (1 or 2; more runs, fewer
    # workgroups per cohort, so more cohorts per launch), each value timed on 2 iterations of the batch
    def shape(rmin):
        for h in hs:
            r = max(rmin, -(-h.n // (512 * 256)))
            h.set_launch(-(-h.n // (512 * r)), 0)

    tune, o2 = [], orders_for(2)
    for rmin in (1, 2):
        shape(rmin)
        fresh()
        t1 = time.perf_counter()
        _, rep = brr.run_cohorts(hs, args.seed, 0, o2, chains=list(range(K)))
        tune.append(dict(min_lane_runs=rmin, launches=rep["launches"], seconds=time.perf_counter() - t1))
        log(f"tune: >= {rmin} lane runs per slice -> {rep['launches']} launches, "
            f"{tune[-1]['seconds']:.2f} s for 2 iterations")
    best = min(tune, key=lambda e: e["seconds"])
    shape(best["min_lane_runs"])
    orders = orders_for(steps)
    fresh()
    t1 = time.perf_counter()
    hists, rep = brr.run_cohorts(hs, args.seed, 0, orders, chains=list(range(K)))
    wall = time.perf_counter() - t1
    per, bytes_total, updates = [], 0.0, 0
    for c, h in enumerate(hs):
        st = h.stats()
        cfg_c = dict(n=h.n, m=h.m, layout="dense")     # the device holds codes in either layout
        bytes_total += algorithmic_bytes(cfg_c, None, 2, steps * h.m, st["changes"])
        updates += h.m * steps
        per.append(dict(N=h.n, M=h.m, layout=layouts[c], workgroups=st["workgroups"], launch=rep["launch_of"][c],
                        rounds=st["rounds"], changes=st["changes"]))
    for h in hs:
        h.close()
    device_s = rep["device_ms"] * 1e-3
    achieved = bytes_total / device_s / 1e9
    log(f"c4: {updates / wall / 1e6:.2f} M marker-updates/s, {rep['launches']} launches in one graph, "
        f"device {rep['device_ms']:.0f} ms, wall {wall * 1e3:.0f} ms")
    out = dict(value=updates / wall, unit="marker-updates/s", steps=steps,
               ms_per_step=wall * 1e3 / steps, dtype="f32+f64" if precision == "fp32" else "f64",
               config={"workload": cfg["workload"].replace("run cohort after cohort",
                                                           "cohorts side by side in persistent launches"),
                       "cohorts": K, "timed_iterations_per_cohort": steps,
                       "precision": (f"fp32 residual, fp64 refresh every {args.refresh} markers"
                                     if precision == "fp32" else "fp64"),
                       "parallelism": "single GPU; cohorts packed onto disjoint workgroup sets of persistent "
                                      "cooperative launches (at most one workgroup per compute unit), the "
                                      "launches captured in one hipGraph",
                       "launches": rep["launches"], "min_lane_runs_per_slice": best["min_lane_runs"],
                       "autotune": tune, "per_cohort": per},
               roofline={"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                         "achieved_basis": "logical: B_j = ceil(N/4) + 4 N (1 + delta_j) per marker-update (an FP32 "
                                           "residual pass, whatever the precision); the residual passes are served "
                                           "from registers",
                         "limiter": "latency (exchange rounds)",
                         "kernel": "brr::bayesr_cohort_kernel (all cohorts, one graph)",
                         "device_ms": rep["device_ms"]})
    if cpu is not None:
        out["cpu_baseline"] = cpu
    return out


def decode_bytes(n, m, layout):
    """Algorithmic bytes of the unpack kernel: read the packed columns (M ceil(N/4)), write the
    16-byte-aligned code columns, 16 B of counts and (ragged) 32 B of statistics per marker."""
    cb = (n + 3) // 4
    col = (cb + 15) // 16 * 16
    return float(m) * (cb + col + 16 + (32 if layout == "ragged" else 0))


def run_decode(cfg, data, args, repeats=3):
    """The workload's .bed file decoded on the device (brr_hip_load_bed) in both layouts; the file
    image is in host memory, its copy to HBM precedes the timed kernels."""
    import bayesrrcmd_amd as brr
    n, m = cfg["n"], cfg["m"]
    f = brr.bed_file_image(data["bed"])
    out = {"workload": f"{cfg['workload']}: its .bed file ({f.size / 1e9:.2f} GB) decoded on the device",
           "file_bytes": int(f.size)}
    for name, layout in (("ragged", brr.LAYOUT_RAGGED), ("dense", brr.LAYOUT_DENSE_2BIT)):
        h = brr.HipSweep(n, m, 4, 1, layout)
        unpack, tables, wall = [], [], []
        for _ in range(repeats):
            t0 = time.perf_counter()
            t = h.load_bed(f, layout)
            wall.append(time.perf_counter() - t0)
            unpack.append(t["unpack"])
            tables.append(t["tables"])
        got = h.marker_tables()
        h.close()
        k = 2000
        ref = data["ragged"].stats if (name == "ragged" and data.get("ragged") is not None) else None
        if ref is None:
            try:
                ref = (brr.build_ragged(data["bed"][:k], n).stats if name == "ragged"
                       else brr.build_dense(data["bed"][:k], n))
            except brr.BrrHipError:            # a marker without genotype 1 / 2 calls: no host statistics
                ref = np.zeros((0, 4))
        same = bool(np.array_equal(got[: ref.shape[0]].view(np.uint64), ref.view(np.uint64)))
        avg = float(np.mean(unpack))
        nbytes = decode_bytes(n, m, name)
        achieved = nbytes / (avg * 1e-3) / 1e9
        out[name] = {
            "unpack_ms": avg, "tables_ms": float(np.mean(tables)) if name == "dense" else None,
            "load_bed_wall_s": float(np.mean(wall)),
            "markers_checked_bit_exact_vs_host_builder": int(ref.shape[0]), "bit_exact": same,
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "brr::bed_unpack_kernel",
                         "algorithmic_bytes_per_launch": nbytes},
        }
        log(f"decode {name}: unpack {avg:.3f} ms ({achieved:.0f} GB/s), tables "
            f"{np.mean(tables):.3f} ms, load_bed wall {np.mean(wall):.2f} s, bit-exact {same}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed Gibbs iterations (default: chain length)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--workgroups", type=int, default=0)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate", action="store_true",
                    help="run a 1 GiB device copy first (reference kernel for HBM counters)")
    ap.add_argument("--ragged-storage", default="auto", choices=["auto", "lists", "codes"],
                    help="device copy of ragged columns: smaller of lists / 2-bit codes, or forced")
    ap.add_argument("--threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-batched", action="store_true", help="skip the batched (c3) leg of the default line")
    ap.add_argument("--no-bayesw", action="store_true", help="skip the BayesW (c5) leg of the default line")
    ap.add_argument("--no-decode", action="store_true", help="skip the device .bed decode leg of the default line")
    ap.add_argument("--async", dest="async_window", type=int, default=0,
                    help="ParallelGraph's asynchronous sweep in windows of this many markers (1..64; "
                         "0 = the exact sequential sweep)")
    ap.add_argument("--precision", default=None, choices=["fp32", "fp64"],
                    help="plain sweep residual arithmetic: fp32 (FP32 residual and dots, compensated wave "
                         "reductions, FP64 refresh every --refresh markers; the c2 default) or fp64")
    ap.add_argument("--refresh", type=int, default=1024, help="FP64 residual refresh interval (markers)")
    ap.add_argument("--no-probe", action="store_true", help="skip the exchange-latency probe of the headline line")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shape of the headline line")
    ap.add_argument("--no-errors", action="store_true",
                    help="skip the FP64 comparison run of an fp32 line (errors_vs_fp64)")
    ap.add_argument("--decode-only", action="store_true",
                    help="only the decode leg: the c2 genotypes decoded on the device, its JSON object printed")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path "
              f"(--gpus {args.gpus}, WORLD_SIZE {world}). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import bayesrrcmd_amd as brr
    if args.calibrate:
        # a plain device copy of known size: FETCH_SIZE / WRITE_SIZE of this
        # kernel give the counter scale for the sweep's HBM traffic
        ms = brr.calibration_copy(1 << 30)
        log(f"calibration copy: 1 GiB read + 1 GiB written in {ms:.3f} ms ({2.0 * (1 << 30) / ms / 1e6:.0f} GB/s)")
    cfg = CONFIGS[args.config]
    if args.decode_only:
        t0 = time.time()
        bed, _ = brr.synth_bed(args.seed, cfg["n"], cfg["m"], cfg["maf"][0], cfg["maf"][1], 0.0,
                               threads=args.threads)
        log(f"genotypes {cfg['m']} x {cfg['n']} generated in {time.time() - t0:.1f}s")
        print(json.dumps({"decode": run_decode(cfg, {"bed": bed}, args)}), flush=True)
        return
    if args.config in ("c3", "c3-n100k", "c4", "c5", "c5-dense"):
        run = {"c3": run_batched, "c3-n100k": run_batched, "c4": run_cohorts}.get(args.config, run_bayesw)
        b = run(cfg, args, not args.no_cpu)
        line = {"metric": METRIC, "value": b["value"], "unit": b["unit"], "n_gpus": 1, "steps": b["steps"],
                "warmup": args.warmup, "ms_per_step": b["ms_per_step"], "higher_is_better": True,
                "scaling": "weak", "vs_baseline": None, "dtype": b["dtype"],
                "data": ("synthetic (Philox-seeded PLINK genotypes, HWE, 64 phenotypes y = X beta + e)"
                         if args.config.startswith("c3") else
                         "synthetic (Philox-seeded PLINK genotypes, HWE, one phenotype y = X beta + e per cohort)"
                         if args.config == "c4" else
                         "synthetic (Philox-seeded PLINK genotypes, HWE, Weibull log-times, 10% censored)"),
                "config": b["config"], "roofline": b["roofline"]}
        if "cpu_baseline" in b:
            line["cpu_baseline"] = b["cpu_baseline"]
        print(json.dumps(line), flush=True)
        return
    steps = args.steps if args.steps is not None else cfg["chain"]
    n, m = cfg["n"], cfg["m"]
    data = make_data(cfg, args.seed, args.threads)

    layout = brr.LAYOUT_RAGGED if cfg["layout"] == "ragged" else brr.LAYOUT_DENSE_2BIT
    storage = {"auto": brr.STORAGE_AUTO, "lists": brr.STORAGE_LISTS, "codes": brr.STORAGE_CODES}[args.ragged_storage]
    h = brr.HipSweep(n, m, 4, 1, layout, workgroups=args.workgroups, block_markers=args.block,
                     ragged_storage=storage)
    t0 = time.time()
    if data["ragged"] is not None:
        h.upload_ragged(data["ragged"])
    else:
        h.upload_dense(data["bed"], data["lut"])
    h.set_model(S_DEFAULT)
    if args.async_window:
        h.set_async(args.async_window)
    precision = args.precision or cfg.get("precision", "fp64")
    if args.async_window:
        precision = "fp64"
    if precision == "fp32":
        h.set_precision(brr.PRECISION_FP32_REFRESH, args.refresh)
    log(f"uploaded in {time.time() - t0:.1f}s; precision {precision}")

    def fresh_chain():
        eps, sc = brr.bayesr_init(data["y"], 1, 4, args.seed)
        h.set_state(eps, np.zeros(m), np.zeros(m, np.int32))
        h.set_scalars(sc["mu"], sc["sigmaE"], sc["epsilon_sum"], sc["sigmaG"], sc["pi"])

    # one chain from init: its first `warmup` iterations are untimed (the start of
    # the burn-in), the next `steps` are timed
    fresh_chain()
    orders = chain_orders(args.seed, m, args.warmup + steps)
    hist = []
    for it in range(args.warmup):
        hist.append(h.run_chain(args.seed, 0, it, orders[it])[0])
    # the launch shape: the autotuner times each candidate (workgroups x markers per
    # block) on one iteration from the chain's state after the warm-up, restores that
    # state, and keeps the fastest
    tune = None
    if not (args.workgroups or args.block or args.async_window or args.no_autotune):
        t0 = time.time()
        tune = h.autotune(args.seed + 2, 1)
        st = h.stats()
        log(f"autotune: {len(tune)} shapes in {time.time() - t0:.1f}s -> {st['workgroups']} workgroups, "
            f"{st['block_markers']}-marker blocks")

    kernel_ms, changes, rounds, phases = [], [], [], []
    t_start = time.perf_counter()
    for it in range(args.warmup, args.warmup + steps):
        hist.append(h.run_chain(args.seed, 0, it, orders[it])[0])
        s = h.stats()
        kernel_ms.append(s["kernel_ms"])
        changes.append(s["changes"])
        rounds.append(s["rounds"])
        phases.append(list(s["phase_ms"].values()))
        if it % 20 == 0 or it == args.warmup + steps - 1:
            ph = ", ".join(f"{k} {v:.1f}" for k, v in s["phase_ms"].items())
            log(f"iteration {it}: {s['kernel_ms']:.2f} ms, rounds {s['rounds']}, changes {s['changes']} "
                f"| phases ms: {ph}")
    wall = time.perf_counter() - t_start
    hist = np.array(hist)
    stats = h.stats()
    errors = None
    if precision == "fp32" and not args.no_errors:
        errors = errors_vs_fp64(h, brr, data, m, args, orders, hist, fresh_chain)
    # the latency floor of a round: exchange rounds of a full block with no other work
    probe_us = h.probe_exchange(2000) if not (args.async_window or args.no_probe) else None

    total_updates = float(steps) * m
    value = total_updates / wall
    kernel_ms = np.array(kernel_ms)
    avg_ms = float(kernel_ms.mean())
    bytes_per_launch = np.mean([algorithmic_bytes(cfg, data, stats["storage"], m, c) for c in changes])
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    # h2 over the whole chain (warm-up and timed iterations) after its burn-in
    sigmaE, sigmaG = hist[:, 1], hist[:, 4]
    burn = min(len(hist) // 2, 50)
    h2 = float(np.mean(sigmaG[burn:] / (sigmaG[burn:] + sigmaE[burn:]))) if len(hist) > 1 else float("nan")
    # the late part of the timed iterations (the chain near its stationary state), device time
    late = kernel_ms[-max(1, steps // 4):]

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "marker-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if precision == "fp32" else "f64",
        "data": "synthetic (Philox-seeded PLINK genotypes, HWE, y = X beta + e)",
        "config": {
            "workload": cfg["workload"], "N": n, "M": m, "layout": cfg["layout"],
            "device_storage": "2-bit codes" if stats["storage"] == 2 else "index lists",
            "maf": list(cfg["maf"]), "causal_fraction": cfg["causal"], "h2_simulated": cfg["h2"],
            "S": S_DEFAULT.tolist(), "timed_iterations": steps,
            "chain": (f"one chain from init: iterations 0..{args.warmup - 1} untimed (warm-up), "
                      f"{args.warmup}..{args.warmup + steps - 1} timed" if args.warmup else
                      f"one chain from init: iterations 0..{steps - 1} timed"),
            "precision": (f"fp32 residual and dots (compensated wave reductions), fp64 refresh every "
                          f"{args.refresh} markers, fp64 decisions and draws" if precision == "fp32" else "fp64"),
            "workgroups": stats["workgroups"], "block_markers": stats["block_markers"],
            "slice": stats["slice"], "parallelism": "single GPU, persistent cooperative kernel",
            "schedule": (f"asynchronous windows of {args.async_window} markers (ParallelGraph)"
                         if args.async_window else "sequential (exact)"),
            "h2_posterior_mean": h2,
            "autotune": ([{k: (round(v, 3) if isinstance(v, float) else v) for k, v in e.items()} for e in tune]
                         if tune else None),
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None,
            "traffic_note": ("HBM bytes per launch are measured by separate rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE "
                             "runs of this benchmark (profiles/bayesr_sweep_c2_ragged_N100k_M100k_hbm_counters.md); "
                             "a counter run cannot share this process"),
            "achieved_basis": ("logical: sum over the launch's markers of B_j = G_j + 4 N (1 + delta_j), G_j the "
                               "column's device bytes, 4 N one FP32 residual pass (whatever precision the kernel "
                               "computes in), delta_j = 1 when beta_j changed; the residual passes are served from "
                               "registers, so HBM moves only the columns and the exchange words"),
            "kernel": "brr::bayesr_sweep_kernel (one launch per iteration)",
            "avg_launch_ms": avg_ms, "algorithmic_bytes_per_launch": bytes_per_launch,
            "column_bytes_per_launch": float(m) * ((n + 3) // 4) if stats["storage"] == 2 else None,
            "column_GBps": (float(m) * ((n + 3) // 4) / (avg_ms * 1e-3) / 1e9) if stats["storage"] == 2 else None,
            "limiter": {
                "kind": "latency: exchange rounds (one per block of markers plus one per other beta change)",
                "rounds_per_launch": float(np.mean(rounds)),
                "exchange_us_per_round_probe": probe_us,
                "exchange_floor_ms": (float(np.mean(rounds)) * probe_us * 1e-3) if probe_us else None,
                "floor_fraction_of_launch": (float(np.mean(rounds)) * probe_us * 1e-3 / avg_ms) if probe_us else None,
            },
            "rounds_per_launch": float(np.mean(rounds)), "changes_per_launch": float(np.mean(changes)),
            "phase_ms_per_launch": dict(zip(brr.HipSweep.PHASES, np.mean(phases, axis=0).round(3).tolist())),
        },
    }
    line["late_iterations"] = {
        "iterations": f"the last {len(late)} timed ({args.warmup + steps - len(late)}..{args.warmup + steps - 1})",
        "kernel_ms_per_iteration": float(np.mean(late)),
        "device_marker_updates_per_s": float(m / (np.mean(late) * 1e-3)),
    }
    if errors is not None:
        line["errors_vs_fp64"] = errors
    h.close()
    if not args.no_cpu:
        log("timing the CPU reference sweep ...")
        line["cpu_baseline"] = cpu_baseline(cfg, data, args.seed, args.cpu_budget, args.async_window)
    if not args.no_decode:
        line["decode"] = run_decode(cfg, data, args)
        line["decode"]["host_builder_s"] = data.get("build_s")
    if args.config == "c2" and not args.no_batched:
        del data
        data = None
        line["batched"] = run_batched(CONFIGS["c3"], args, not args.no_cpu)
        line["batched_n100k"] = run_batched(CONFIGS["c3-n100k"], args, False)
    if args.config == "c2" and not args.no_bayesw:
        data = None
        line["bayesw"] = run_bayesw(CONFIGS["c5"], args, not args.no_cpu)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
