"""Build libhclpp_hip.so (gfx950) in-tree with hipcc.

Every translation unit is compiled with:
  * --offload-arch=gfx950 only;
  * -ffp-contract=off: no a*b+c contraction anywhere, so the FP64 kernels
    round exactly like the reference's CPU build (bit-identical results);
  * -Xclang -target-feature -Xclang -packed-fp32-ops: no packed FP32 VALU ops.
With HCLPP_IEEE_DIVISION=1 in the environment, every device division is the full
IEEE sequence (-DHCLPP_IEEE_DIVISION; physics.hpp fdiv) and the FP64 kernels are
about a fifth slower: results then equal IEEE division bit for bit also where values
fall below about 1e-290.
The fused-step sources are also scheduled for instruction-level parallelism
(-amdgpu-sched-strategy=max-ilp): their kernels are bound by dependent FP64
VALU chains, and on MI355X that schedule gives the FP64 march kernel 2 % (2D)
to 7 % (3D) more throughput at the same occupancy.  Scheduling reorders
instructions only, so results are unchanged.

Usage: python -m heraclespp_amd.build [-j N] [--verbose]
"""

from __future__ import annotations

import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
OBJDIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libhclpp_hip.so")

SOURCES = ["ref_kernels.hip", "step_fused.hip", "step_fused_fx.hip", "adjoint.hip", "abi.cpp"]
EXTRA = {src: ["-mllvm", "-amdgpu-sched-strategy=max-ilp"] for src in ("step_fused.hip", "step_fused_fx.hip")}

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", CSRC,
    "-I", INCLUDE,
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
]


def _flags(src: str) -> list:
    extra = ["-DHCLPP_IEEE_DIVISION"] if os.environ.get("HCLPP_IEEE_DIVISION") == "1" else []
    return COMMON + extra + EXTRA.get(src, [])


def _compile(src: str, verbose: bool) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj] + _flags(src)
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    warn = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if warn and verbose:
        print(warn, file=sys.stderr)
    return obj


def build(jobs: int | None = None, verbose: bool = False) -> str:
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.j, a.verbose))


if __name__ == "__main__":
    main()
