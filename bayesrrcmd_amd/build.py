"""Build libbrr_hip.so (gfx950) in-tree, and the oracle when requested.

The library is compiled with hipcc for gfx950 only. Floating-point contraction
is off so that host and device evaluate the counter-based stream bit-identically
and every FMA in a kernel is written explicitly; packed FP32 instructions are
disabled for all device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
LIB = PKG / "libbrr_hip.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    res = subprocess.run(cmd, cwd=str(ROOT), capture_output=True, text=True)
    out = "\n".join(line for line in (res.stdout + res.stderr).splitlines()
                    if "packed-fp32-ops" not in line)
    if out.strip():
        print(out, flush=True)
    if res.returncode != 0:
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd)}")


# device translation units: the host API with its small kernels, and one file per
# kernel family (compiled in parallel)
HIP_SOURCES = ("api.hip", "kern_sweep_seq.hip", "kern_sweep_async.hip", "kern_sweep_f32.hip", "kern_cohort.hip",
               "kern_batch.hip", "kern_batch_f32.hip", "kern_bayesw.hip")
HOST_SOURCES = ("host_markers.cpp", "host_driver.cpp", "synth.cpp")


def build_lib() -> Path:
    from concurrent.futures import ThreadPoolExecutor
    objdir = PKG / "_objs"
    objdir.mkdir(exist_ok=True)
    cmds, objs = [], []
    for src in HOST_SOURCES:
        obj = objdir / (Path(src).stem + ".o")
        cmds.append(["g++", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-pthread",
                     f"-I{INCLUDE}", f"-I{CSRC}", "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(str(obj))
    for src in HIP_SOURCES:
        obj = objdir / (Path(src).stem + ".o")
        cmds.append([HIPCC, *HIP_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(str(obj))
    workers = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(max_workers=workers) as pool:
        for f in [pool.submit(_run, c) for c in cmds]:
            f.result()
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-pthread", "-lz", "-o", str(tmp)])
    os.replace(tmp, LIB)
    shutil.rmtree(objdir, ignore_errors=True)
    return LIB


def build_oracle() -> None:
    _run(["make", "-s", "-C", str(ROOT / "oracle")])


def main(argv: list[str]) -> int:
    build_lib()
    if "--no-oracle" not in argv:
        build_oracle()
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
