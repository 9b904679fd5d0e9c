"""Build the gfx950 library (lib/libpct.so) and the CPU oracle (oracle/_build/libpct_oracle.so).

hipcc compiles every translation unit of csrc/ for gfx950 with -ffp-contract=off (the arithmetic
is specified operation by operation) and without packed FP32 instructions; units are compiled in
parallel into _build/ and linked into lib/libpct.so.  The oracle is plain C built with gcc.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "lib" / "libpct.so"
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "_build" / "libpct_oracle.so"

ARCH = os.environ.get("PCT_OFFLOAD_ARCH", "gfx950")
COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-result"]
HIP_ONLY = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-munsafe-fp-atomics"]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def build_library(verbose: bool = False) -> Path:
    hipcc = _hipcc()
    BUILD.mkdir(exist_ok=True)
    LIB.parent.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    objs = []
    jobs = []
    for src in sources:
        obj = BUILD / (src.name + ".o")
        objs.append(obj)
        cmd = [hipcc, *COMMON]
        if src.suffix == ".hip":
            cmd += ["-x", "hip", *HIP_ONLY]
        cmd += ["-c", str(src), "-o", str(obj)]
        jobs.append(cmd)
    workers = min(len(jobs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(workers) as ex:
        for cmd, fut in zip(jobs, [ex.submit(_run, c) for c in jobs]):
            fut.result()
            if verbose:
                print("compiled", cmd[-3], flush=True)
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, "-shared", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> Path:
    ORACLE_LIB.parent.mkdir(exist_ok=True)
    tmp = ORACLE_LIB.with_suffix(".so.tmp")
    # -O3 and AVX2/FMA code (x86-64-v3, which every host of the GPU boxes runs); with -ffp-contract=off and
    # no -ffast-math the floating-point results are those of -O0
    _run(["gcc", "-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared", "-Wall", "-Wextra",
          "-o", str(tmp), str(ORACLE_DIR / "pct_oracle.c"), str(ORACLE_DIR / "pct_oracle_fp64.c"),
          str(ORACLE_DIR / "pct_net_oracle.c"),
          str(ORACLE_DIR / "pct_tvs_oracle.c"), "-lm"])
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


def build_all(verbose: bool = False) -> None:
    build_library(verbose)
    build_oracle()


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print(LIB)
    print(ORACLE_LIB)
