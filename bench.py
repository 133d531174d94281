"""Benchmark of the HERACLES++ hydro step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload kelvin_helmholtz]

A step is one fused MUSCL-Hancock + Riemann + Godunov pass over the whole
grid plus the ghost fill and the device-side dt reduction (hclpp_step).  The
default workload is the 2D Kelvin-Helmholtz configuration (2048 x 2048,
periodic, VanLeer + HLLC, gamma 1.4, FP64) on synthetic initial data.  The
state is resident in HBM before the timed region; W untimed steps run
first, then exactly K steps are timed between torch.cuda.synchronize()
calls.  Prints one JSON line:
  value      Gcell-updates/s = interior cells x K / wall seconds
  roofline   the fused kernel's algorithmic HBM bytes (read + write the
             conservative state once: 2 (ndim + 2) x 8 B per cell) divided by
             its mean duration, measured with HIP events on its stream,
             against the 8 TB/s HBM3E peak; `traffic` is the measured
             HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled
             by a copy-kernel calibration) read from profiles/, or null
             roofline.valu: the FP64 kernel is bound by its VALU issue rate, not
             by HBM; from the committed counter summary (SQ_ACTIVE_INST_VALU per
             cell) and the live kernel time, the fraction of the 1024 SIMDs'
             issue slots used and the VALU roof in Gcell/s
  cpu_baseline  the CPU oracle (a restatement of the reference's unfused
             kernels, OpenMP) on a bounded sample of the same workload, timed
             on this host's cores
  mixed_fp32 the same workload with FP32 storage (FP64 pressure recovery, dt and
             totals), timed the same way after the FP64 run: Gcell/s, kernel time,
             ratio to FP64, and the relative L1 distance of its density to the FP64
             density after the same number of steps (the FP64 path is bit-identical
             to the reference arithmetic)
  rayleigh_taylor_512  the 3D Rayleigh-Taylor 512^3 run (Minmod + HLLC, uniform
             gravity), 2 + 8 steps in FP64 and with FP32 storage: Gcell/s, kernel
             time and fraction of the HBM peak at the algorithmic bytes, the
             FP32/FP64 throughput ratio and the density L1 distance between the two
             (with the default workload only; --no-3d skips it)
  adjoint    differentiable-Euler 1024^2 (VanAlbada + HLLC, 400 steps,
             checkpoint every 20): forward, checkpointed backward and reverse-
             kernel cost per step, and central-difference checks of the gradient
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Gcell-updates/s & HBM roofline %, FP64 vs mixed FP32; adjoint/forward ratio"

WORKLOADS = {
    # name: (setup factory kwargs at full size, cpu-baseline sample kwargs, description)
    "kelvin_helmholtz": ({"n": 2048}, {"n": 2048}, "2D Kelvin-Helmholtz 2048x2048 periodic, VanLeer+HLLC"),
    "rayleigh_taylor": ({"n": 512}, {"n": 128}, "3D Rayleigh-Taylor 512^3, g_z=-0.1, Minmod+HLLC"),
    "sedov": ({"nr": 512, "nth": 128, "nph": 256}, {"nr": 128, "nth": 32, "nph": 64},
              "3D spherical Sedov 512x128x256, HLL+VanLeer"),
    "smooth_euler": ({"n": 1024}, {"n": 1024}, "2D smooth periodic Euler 1024x1024, VanAlbada+HLLC"),
    "sod": ({"n": 1000}, {"n": 1000}, "1D Sod 1000 cells, Minmod+HLL"),
}


def _setup(name, kw):
    from heraclespp_amd import setups

    return setups.CONFIGS[name](**kw)


def _traffic(workload: str, precision: str):
    """Measured HBM bytes per fused-kernel launch from the committed rocprofv3 summary, if any."""
    path = os.path.join(ROOT, "profiles", f"k_step_march_{workload}_{precision}_traffic.txt")
    if not os.path.exists(path):
        return None
    m = re.search(r"^traffic_bytes_per_launch\s*[:=]\s*([0-9.eE+]+)", open(path).read(), re.M)
    return float(m.group(1)) if m else None


def _valu_units(workload: str, precision: str):
    """VALU issue units per interior cell of the fused kernel (SQ_ACTIVE_INST_VALU x 64 / cells) from the
    committed rocprofv3 counter summary, if any."""
    path = os.path.join(ROOT, "profiles", f"k_step_march_{workload}_{precision}_valu.txt")
    if not os.path.exists(path):
        return None
    m = re.search(r"VALU issue units \(SQ_ACTIVE_INST_VALU\)\s+([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def cpu_baseline(name: str, seconds: float):
    """Time the oracle (reference kernel sequence on the CPU) on a bounded sample of the workload."""
    from oracle import oracle as orc

    kw = WORKLOADS[name][1]
    s = _setup(name, kw)
    o = orc.Oracle(s.ndim, s.n, geom=s.geom, limiter=s.reconstruction, riemann=s.riemann_solver, gamma=s.gamma,
                   bc=s.bc, xmin=s.xmin, xmax=s.xmax, g=s.gravity)
    o.init_from_prim(s.rho, s.u, s.P)
    o.step(s.cfl, s.t_end, 1 << 30)  # warm-up (page faults, thread pool)
    cells = 1
    for d in range(s.ndim):
        cells *= s.n[d]
    steps = 0
    t0 = time.perf_counter()
    while True:
        o.step(s.cfl, s.t_end, 1 << 30)
        steps += 1
        el = time.perf_counter() - t0
        if el >= seconds or steps >= 10000:
            break
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": cells * steps / el / 1e9,
        "unit": "Gcell-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{steps} steps of {name} at {'x'.join(str(v) for v in s.n[:s.ndim])} after 1 warm-up step "
                  f"({el:.1f} s, OpenMP {threads} threads, unfused reference kernel order)",
    }


def adjoint_block(n: int, steps: int, every: int, fd_dirs: int):
    """Differentiable-Euler workload: forward rollout, checkpointed backward and reverse-kernel cost per step,
    plus a central-difference check of the gradient on a few seeded unit directions."""
    import numpy as np
    import torch

    from heraclespp_amd.adjoint import diff_euler_problem

    prob = diff_euler_problem(n, steps, seed=0)
    prob.gradient(every=every)  # warm-up: allocations, first launches
    torch.cuda.synchronize()

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, out

    g = torch.tensor(prob.gamma, dtype=torch.float64)
    with torch.no_grad():
        t_fwd, _ = timed(lambda: prob.loss(prob.rho0, g, every))
    x = prob.rho0.clone().requires_grad_(True)
    gg = g.clone().requires_grad_(True)
    t_fwd_ckpt, L = timed(lambda: prob.loss(x, gg, every))
    t_bwd, _ = timed(lambda: L.backward())
    # reverse kernels alone: `steps` reverse steps on one state
    eng = prob.spec.engine(prob.gamma)
    U = eng.load(*prob.cons(prob.rho0, g))
    lam, lam2 = eng.new_state(), eng.new_state()
    lam[0].normal_()
    dg = torch.zeros(1, dtype=torch.float64, device="cuda")
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    eng.vjp(U, lam, lam2, dg, prob.dts[0])
    lams = [lam if k % 2 == 0 else lam2 for k in range(steps + 1)]
    e0.record()
    eng.vjps([U] * steps, lams, dg, prob.dts[:steps])  # one C call: no per-step host overhead
    e1.record()
    torch.cuda.synchronize()
    t_rev = e0.elapsed_time(e1) / 1e3
    rng = np.random.Generator(np.random.Philox(1))
    worst = 0.0
    for _ in range(fd_dirs):
        v = torch.from_numpy(rng.standard_normal((n, n))).cuda()
        v = v / torch.linalg.norm(v)
        fd = prob.fd(v, 0.0, 1e-6)
        ad = float(torch.sum(x.grad * v))
        worst = max(worst, abs(fd - ad) / max(abs(fd), 1e-300))
    # dL/dgamma of the density loss vanishes (uniform u and P: the density is a pure contact, which HLLC
    # carries independently of the wave speeds); check the gamma path on the loss with an energy term
    probE = diff_euler_problem(n, steps, seed=0, energy_weight=1.0, target_offset=0.01)
    _, _, gE = probE.gradient(every=every)
    fdg = probE.fd(torch.zeros_like(probE.rho0), 1.0, 1e-6)
    relg = abs(fdg - gE) / max(abs(fdg), 1e-300)
    f = t_fwd / steps * 1e3
    return {
        "workload": f"differentiable Euler {n}x{n} periodic, VanAlbada+HLLC, {steps} steps, checkpoint every {every}",
        "forward_ms_per_step": round(f, 5),
        "forward_with_checkpoints_ms_per_step": round(t_fwd_ckpt / steps * 1e3, 5),
        "backward_ms_per_step": round(t_bwd / steps * 1e3, 5),
        "reverse_kernels_ms_per_step": round(t_rev / steps * 1e3, 5),
        "ratio_reverse_to_forward": round(t_rev / t_fwd, 3),
        "ratio_backward_to_forward": round(t_bwd / t_fwd, 3),
        "fd_directions": fd_dirs,
        "fd_rel_err_max": float(f"{worst:.3e}"),
        "fd_rel_err_gamma": float(f"{relg:.3e}"),
        "fd_gamma_loss": "density + energy misfit, target mean density shifted by 0.01",
    }


def calibrate_copy(nbytes: int):
    """A plain device copy of known size, for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import torch

    a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "launches": 3}))


def _timed_run(workload: str, precision: str, steps: int, warmup: int):
    """W untimed then K timed fused steps of a workload, the state resident in HBM.  Returns a
    record: the setup (without its arrays), interior cells, wall seconds of the K steps (between
    synchronisations), Gcell-updates/s, the fused kernel's mean time in ms (HIP events on its
    stream), its launch plan, and the interior density after all steps."""
    from types import SimpleNamespace

    import torch

    from heraclespp_amd import _lib
    from heraclespp_amd.hydro import Grid, PerfectGas, UniformGravity
    from heraclespp_amd.solver import HydroSolver

    s = _setup(workload, WORKLOADS[workload][0])
    grid = Grid(s.ndim, s.n, s.ng, s.geom, s.xmin, s.xmax)
    solver = HydroSolver(grid, PerfectGas(s.gamma, 1.0), reconstruction=s.reconstruction,
                         riemann_solver=s.riemann_solver, gravity=UniformGravity(s.gravity), bc=s.bc, cfl=s.cfl,
                         precision=precision)
    solver.init_from_prim(s.rho, s.u, s.P)
    del s.rho, s.u, s.P
    cells = 1
    for d in range(s.ndim):
        cells *= s.n[d]
    solver.start_clock(1e300, warmup + steps + 1)
    for _ in range(warmup):
        solver.step_async()
    torch.cuda.synchronize()
    L = _lib.lib()
    _lib.check(L.hclpp_profile_begin(solver.ctx.handle, steps))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        solver.step_async()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    kms, nl = ctypes.c_double(), ctypes.c_int()
    _lib.check(L.hclpp_profile_end(solver.ctx.handle, ctypes.byref(kms), ctypes.byref(nl)))
    clk = solver.finish()
    if clk.iter != warmup + steps:
        raise RuntimeError(f"expected {warmup + steps} {precision} steps, clock counted {clk.iter}")
    seglen, nwg, occ = ctypes.c_int(), ctypes.c_longlong(), ctypes.c_int()
    _lib.check(L.hclpp_fused_plan(solver.ctx.handle, ctypes.byref(seglen), ctypes.byref(nwg), ctypes.byref(occ)))
    rho = solver.rho.cpu().numpy()[s.interior()]
    del solver
    torch.cuda.empty_cache()
    return SimpleNamespace(setup=s, cells=cells, wall=t1 - t0, value=cells * steps / (t1 - t0) / 1e9,
                           kernel_ms=kms.value / max(nl.value, 1), rho=rho,
                           plan={"segment_planes": seglen.value, "workgroups": nwg.value, "resident_per_cu": occ.value})


def _rel_l1(a, b):
    import numpy as np

    return float(np.sum(np.abs(a - b)) / np.sum(np.abs(b)))


def mixed_block(workload: str, steps: int, warmup: int, value64: float, rho64):
    """The workload again with FP32 storage: throughput (same timing as the FP64 line) and the
    density's relative L1 distance to the FP64 run after the same number of steps."""
    r = _timed_run(workload, "fp32", steps, warmup)
    return {"value": round(r.value, 4), "unit": "Gcell-updates/s", "kernel_ms": round(r.kernel_ms, 5),
            "ratio_to_fp64": round(r.value / value64, 3), "density_rel_l1_vs_fp64": float(f"{_rel_l1(r.rho, rho64):.3e}"),
            "steps_compared": warmup + steps}


def block_3d(steps: int, warmup: int):
    """The 3D Rayleigh-Taylor 512^3 run (Minmod + HLLC, uniform gravity) in FP64 and with FP32
    storage: throughput, kernel time and HBM fraction at the algorithmic bytes, the FP32/FP64
    ratio and the density's relative L1 distance between the two after the same steps."""
    out = {"workload": WORKLOADS["rayleigh_taylor"][2], "steps": steps, "warmup": warmup}
    rho64 = None
    for precision, word in (("fp64", 8), ("fp32", 4)):
        r = _timed_run("rayleigh_taylor", precision, steps, warmup)
        gbs = 2 * (r.setup.ndim + 2) * word * r.cells / (r.kernel_ms / 1e3) / 1e9
        out[precision] = {"value": round(r.value, 4), "unit": "Gcell-updates/s", "kernel_ms": round(r.kernel_ms, 4),
                          "hbm_achieved_gbs": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 5)}
        if precision == "fp64":
            rho64 = r.rho
        else:
            out["fp32"]["ratio_to_fp64"] = round(r.value / out["fp64"]["value"], 3)
            out["fp32"]["density_rel_l1_vs_fp64"] = float(f"{_rel_l1(r.rho, rho64):.3e}")
        del r
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="kelvin_helmholtz", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"],
                    help="fp32: FP32 storage and arithmetic with FP64 pressure recovery, dt and totals")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiable-Euler adjoint measurement")
    ap.add_argument("--no-mixed", action="store_true", help="skip the FP32-storage run of the same workload")
    ap.add_argument("--no-3d", action="store_true", help="skip the 3D Rayleigh-Taylor 512^3 FP64 / FP32 block")
    ap.add_argument("--adjoint-n", type=int, default=1024)
    ap.add_argument("--adjoint-steps", type=int, default=400)
    ap.add_argument("--adjoint-fd-dirs", type=int, default=4)
    a = ap.parse_args()

    if a.gpus != 1:
        print("bench.py: the hydro step runs on one GPU over the whole domain; "
              "multi-GPU runs are not supported (--gpus must be 1).", file=sys.stderr)
        sys.exit(2)
    if a.calibrate_copy:
        calibrate_copy(a.calibrate_copy)
        return

    import torch

    torch.cuda.set_device(0)
    r = _timed_run(a.workload, a.precision, a.steps, a.warmup)
    s, cells = r.setup, r.cells
    wall = r.wall
    value = r.value
    kernel_s = r.kernel_ms / 1e3
    word = 8 if a.precision == "fp64" else 4
    bytes_alg = 2 * (s.ndim + 2) * word * cells
    achieved = bytes_alg / kernel_s / 1e9
    traffic = _traffic(a.workload, a.precision)
    units = _valu_units(a.workload, a.precision)
    # FP64 VALU issue capacity: 1024 SIMDs, one wave64 instruction per 4 cycles each, at 2.4 GHz
    valu = None
    if units is not None:
        issued = units * cells / 64
        capacity = 1024 * kernel_s * 2.4e9 / 4
        valu = {"issue_units_per_cell": units, "busy_frac": round(issued / capacity, 4),
                "roof_gcells": round(39.3e12 / units / 1e9, 3)}
    line = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall / a.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f32 (FP64 pressure recovery, dt, totals)",
        "data": "synthetic",
        "config": {
            "workload": WORKLOADS[a.workload][2],
            "cells": cells,
            "shape": s.n[:s.ndim],
            "reconstruction": s.reconstruction,
            "riemann_solver": s.riemann_solver,
            "storage": a.precision,
            "path": "fused (hclpp_step: march kernel + ghost fill + device dt)",
            "launch": r.plan,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": traffic,
            "kernel": "k_step_march",
            "kernel_ms": round(kernel_s * 1e3, 5),
            "bytes_per_launch": bytes_alg,
            "valu": valu,
        },
        "cpu_baseline": None,
        "mixed_fp32": None,
        "rayleigh_taylor_512": None,
        "adjoint": None,
    }
    rho64 = r.rho
    del r
    if a.precision == "fp64" and not a.no_mixed:
        line["mixed_fp32"] = mixed_block(a.workload, a.steps, a.warmup, value, rho64)
        torch.cuda.empty_cache()
    if a.precision == "fp64" and a.workload == "kelvin_helmholtz" and not a.no_3d:
        line["rayleigh_taylor_512"] = block_3d(8, 2)
    if not a.no_adjoint:
        line["adjoint"] = adjoint_block(a.adjoint_n, a.adjoint_steps, 20, a.adjoint_fd_dirs)
        torch.cuda.empty_cache()
    if not a.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(a.workload, a.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
