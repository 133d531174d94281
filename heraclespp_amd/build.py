"""Build libhclpp_hip.so (gfx950) in-tree with hipcc.

Every translation unit is compiled with:
  * --offload-arch=gfx950 only;
  * -ffp-contract=off: no a*b+c contraction anywhere, so the FP64 kernels
    round exactly like the reference's CPU build (bit-identical results);
  * -Xclang -target-feature -Xclang -packed-fp32-ops: no packed FP32 VALU ops.
The same sources built with -DHCLPP_IEEE_DIVISION give libhclpp_hip_ieee.so
(`--ieee`, or HCLPP_IEEE_DIVISION=1 in the environment for the default name): every
device division is then the full IEEE sequence (physics.hpp fdiv) and the FP64
kernels are about a fifth slower, with results equal to IEEE division bit for bit
also where values fall below about 1e-290.  The package loads the library named by
HCLPP_LIB when that is set (heraclespp_amd/_lib.py).
The fused-step sources are also scheduled for instruction-level parallelism
(-amdgpu-sched-strategy=max-ilp): their kernels are bound by dependent FP64
VALU chains, and on MI355X that schedule gives the FP64 march kernel 2 % (2D)
to 7 % (3D) more throughput at the same occupancy.  Scheduling reorders
instructions only, so results are unchanged.

libhclpp_io.so (output and restart files in the reference's HDF5 layout,
csrc/io_hdf5.cpp) is host code: g++ against the HDF5 C library found under
HCLPP_HDF5_ROOT (default /opt/conda), with that library's directory as rpath.

Usage: python -m heraclespp_amd.build [-j N] [--verbose] [--ieee | --all | --io]
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
LIB_IEEE = os.path.join(HERE, "libhclpp_hip_ieee.so")
LIB_IO = os.path.join(HERE, "libhclpp_io.so")
HDF5_ROOT = os.environ.get("HCLPP_HDF5_ROOT", "/opt/conda")

SOURCES = ["ref_kernels.hip", "step_fused.hip", "step_fused_fx.hip", "step_fused_fast.hip", "pressure_fix_fused.hip", "adjoint.hip",
           "abi.cpp"]
EXTRA = {src: ["-mllvm", "-amdgpu-sched-strategy=max-ilp"] for src in ("step_fused.hip", "step_fused_fx.hip", "step_fused_fast.hip")}
# the fast FP64 arithmetic is not bound to the reference's rounding: let the compiler fuse
# multiplies and adds there
EXTRA["step_fused_fast.hip"] += ["-ffp-contract=fast"]

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


def _flags(src: str, ieee: bool) -> list:
    extra = ["-DHCLPP_IEEE_DIVISION"] if ieee else []
    return COMMON + extra + EXTRA.get(src, [])


def _compile(src: str, verbose: bool, ieee: bool) -> str:
    objdir = os.path.join(OBJDIR, "ieee") if ieee else OBJDIR
    os.makedirs(objdir, exist_ok=True)
    obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj] + _flags(src, ieee)
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    warn = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if warn and verbose:
        print(warn, file=sys.stderr)
    return obj


def build(jobs: int | None = None, verbose: bool = False, ieee: bool | None = None) -> str:
    """Compile and link one library: the default build, or with ieee=True libhclpp_hip_ieee.so
    (ieee=None: the default build, with IEEE division when HCLPP_IEEE_DIVISION=1)."""
    env_ieee = os.environ.get("HCLPP_IEEE_DIVISION") == "1"
    lib = LIB_IEEE if ieee else LIB
    ieee = bool(ieee) or env_ieee
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose, ieee), SOURCES))
    tmp = lib + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, lib)
    return lib


def build_io(verbose: bool = False) -> str:
    """libhclpp_io.so: g++ on csrc/io_hdf5.cpp against the HDF5 C library."""
    inc, libdir = os.path.join(HDF5_ROOT, "include"), os.path.join(HDF5_ROOT, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        raise RuntimeError(f"hdf5.h not found under {inc} (set HCLPP_HDF5_ROOT)")
    tmp = LIB_IO + ".tmp"
    cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", "-I", inc, "-I", INCLUDE, os.path.join(CSRC, "io_hdf5.cpp"),
           "-o", tmp, "-L", libdir, "-lhdf5", f"-Wl,-rpath,{libdir}"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed on io_hdf5.cpp:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB_IO)
    return LIB_IO


def build_all(jobs: int | None = None, verbose: bool = False) -> list:
    """The default and IEEE-division libraries (compiled concurrently) and the I/O library."""
    jobs = jobs or (os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=3) as ex:
        futs = [ex.submit(build, max(1, jobs // 2), verbose, flag) for flag in (None, True)]
        fio = ex.submit(build_io, verbose)
        return [f.result() for f in futs] + [fio.result()]


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--ieee", action="store_true", help="build libhclpp_hip_ieee.so (IEEE division everywhere)")
    ap.add_argument("--all", action="store_true", help="build every library")
    ap.add_argument("--io", action="store_true", help="build libhclpp_io.so only")
    a = ap.parse_args()
    if a.io:
        print(build_io(a.verbose))
    elif a.all:
        print("\n".join(build_all(a.j, a.verbose)))
    else:
        print(build(a.j, a.verbose, True if a.ieee else None))


if __name__ == "__main__":
    main()
