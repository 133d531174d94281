"""Benchmark of the HERACLES++ hydro step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload rayleigh_taylor]

A step is one fused MUSCL-Hancock + Riemann + Godunov pass over the whole
grid plus the ghost fill and the device-side dt reduction (hclpp_step).  The
default workload is the 3D Rayleigh-Taylor configuration (512^3, uniform
gravity g_z = -0.1, periodic x/y, reflexive z walls, Minmod + HLLC, gamma
1.4, FP64) on synthetic initial data.  The state is resident in HBM before
the timed region; W untimed steps run first, then more untimed steps until at
least --warmup-seconds of warm-up have passed (the GPU clocks settle), then
exactly K steps are timed between torch.cuda.synchronize() calls.  Prints
one JSON line:
  value      Gcell-updates/s = interior cells x K / wall seconds
  roofline   the fused kernel's algorithmic HBM bytes (read + write the
             conservative state once: 2 (ndim + 2) x 8 B per cell) divided by
             its mean duration, measured with HIP events on its stream,
             against the 8 TB/s HBM3E peak (`frac`, also `hbm_frac`);
             `traffic` is the measured HBM bytes per launch (rocprofv3
             FETCH_SIZE + WRITE_SIZE, scaled by a copy-kernel calibration)
             read from profiles/, or null.  roofline.valu: from the committed
             counter summary (SQ_ACTIVE_INST_VALU per cell) and the live kernel
             time, the fraction of the 1024 SIMDs' FP64 issue slots used and the
             VALU roof in Gcell/s; `bound` names the roof the kernel is closer to
  cpu_baseline  the CPU oracle (a restatement of the reference's unfused
             kernels, OpenMP) on a bounded sample of the same workload (a
             smaller grid of the same configuration), timed on this host
  mixed_fp32 the same workload with FP32 storage (FP64 pressure recovery, dt and
             totals), timed the same way after the FP64 run: Gcell/s, kernel time,
             ratio to FP64, and the relative L1 distance of its density to the FP64
             density after the configured step count (500 for Rayleigh-Taylor;
             both runs continue untimed to it)
  fast_fp64  the same workload with the opt-in fast FP64 arithmetic (the reference's
             formulas without its rounding; hclpp_set_fast_arithmetic): Gcell/s, kernel
             time, ratio to the reference arithmetic, density L1 to the reference run
             after the configured step count, and relative-L1 divergence curves against
             the reference arithmetic (Rayleigh-Taylor 128^3 to 500 steps, Kelvin-Helmholtz
             2048^2 to 2000 steps)
  kelvin_helmholtz_2048  the 2D Kelvin-Helmholtz 2048^2 run (VanLeer + HLLC) in
             FP64 and with FP32 storage, timed the same way (default workload only)
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
    # name: (setup factory kwargs at full size, cpu-baseline sample kwargs, description, steps after which the
    # FP32-storage density is compared with the FP64 density: the configured run length, or a fixed window)
    "kelvin_helmholtz": ({"n": 2048}, {"n": 2048}, "2D Kelvin-Helmholtz 2048x2048 periodic, VanLeer+HLLC", 2000),
    "rayleigh_taylor": ({"n": 512}, {"n": 128},
                        "3D Rayleigh-Taylor 512^3, g_z=-0.1, periodic x/y, reflexive z, Minmod+HLLC", 500),
    "sedov": ({"nr": 512, "nth": 128, "nph": 256}, {"nr": 128, "nth": 32, "nph": 64},
              "3D spherical Sedov 512x128x256, HLL+VanLeer", 200),
    "smooth_euler": ({"n": 1024}, {"n": 1024}, "2D smooth periodic Euler 1024x1024, VanAlbada+HLLC", 400),
    "sod": ({"n": 1000}, {"n": 1000}, "1D Sod 1000 cells, Minmod+HLL", 500),
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
        "host_physical_cores": _physical_cores(),
        "kind": "port",
        "sample": f"{steps} steps of {name} at {'x'.join(str(v) for v in s.n[:s.ndim])} after 1 warm-up step "
                  f"({el:.1f} s, OpenMP {threads} threads -- this job's CPU share -- unfused reference kernel "
                  f"order; per-cell rate of the same configuration on a smaller grid)",
    }


def _physical_cores():
    """Physical cores of this host (distinct (core, socket) pairs of lscpu), or None."""
    import subprocess

    try:
        out = subprocess.run(["lscpu", "-p=core,socket"], capture_output=True, text=True, timeout=10).stdout
    except (OSError, subprocess.SubprocessError):
        return None
    pairs = {line.strip() for line in out.splitlines() if line.strip() and not line.startswith("#")}
    return len(pairs) or None


def adjoint_block(n: int, steps: int, every: int, fd_dirs: int):
    """Differentiable-Euler workload: forward rollout, backward under the default schedule (every state
    kept when the rollout fits in half the free device memory) and under periodic checkpoints, reverse-kernel
    cost per step, plus a central-difference check of the gradient on a few seeded unit directions."""
    import numpy as np
    import torch

    from heraclespp_amd.adjoint import diff_euler_problem, pick_schedule, state_bytes

    prob = diff_euler_problem(n, steps, seed=0)
    prob.gradient(every=every)  # warm-up: allocations, first launches
    prob.gradient(every=every, schedule="periodic")
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
    schedule = pick_schedule(prob.spec, steps)
    runs = {}
    for sched in dict.fromkeys([schedule, "periodic"]):
        x = prob.rho0.clone().requires_grad_(True)
        gg = g.clone().requires_grad_(True)
        t_f, L = timed(lambda: prob.loss(x, gg, every, schedule=sched))
        t_b, _ = timed(lambda: L.backward())
        runs[sched] = (t_f, t_b, x.grad, gg.grad)
    t_fwd_kept, t_bwd, xgrad, _ = runs[schedule]
    t_fwd_ckpt, t_bwd_ckpt, xgrad_ckpt, _ = runs["periodic"]
    same = bool(torch.equal(xgrad, xgrad_ckpt)) and bool(torch.equal(runs[schedule][3], runs["periodic"][3]))
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
        ad = float(torch.sum(xgrad * v))
        worst = max(worst, abs(fd - ad) / max(abs(fd), 1e-300))
    # dL/dgamma of the density loss vanishes (uniform u and P: the density is a pure contact, which HLLC
    # carries independently of the wave speeds); check the gamma path on the loss with an energy term
    probE = diff_euler_problem(n, steps, seed=0, energy_weight=1.0, target_offset=0.01)
    _, _, gE = probE.gradient(every=every)
    fdg = probE.fd(torch.zeros_like(probE.rho0), 1.0, 1e-6)
    relg = abs(fdg - gE) / max(abs(fdg), 1e-300)
    f = t_fwd / steps * 1e3
    kept = {"store_all": steps + 1, "periodic": len(range(0, steps, every)) + min(every, steps)}[schedule]
    return {
        "workload": f"differentiable Euler {n}x{n} periodic, VanAlbada+HLLC, {steps} steps",
        "backward_schedule": schedule if schedule == "store_all" else f"periodic, checkpoint every {every}",
        "backward_states_GB": round(kept * state_bytes(prob.spec) / 1e9, 2),
        "forward_ms_per_step": round(f, 5),
        "forward_keeping_states_ms_per_step": round(t_fwd_kept / steps * 1e3, 5),
        "backward_ms_per_step": round(t_bwd / steps * 1e3, 5),
        "reverse_kernels_ms_per_step": round(t_rev / steps * 1e3, 5),
        "ratio_reverse_to_forward": round(t_rev / t_fwd, 3),
        "ratio_backward_to_forward": round(t_bwd / t_fwd, 3),
        "checkpointed_every": every,
        "checkpointed_forward_ms_per_step": round(t_fwd_ckpt / steps * 1e3, 5),
        "checkpointed_backward_ms_per_step": round(t_bwd_ckpt / steps * 1e3, 5),
        "ratio_checkpointed_backward_to_forward": round(t_bwd_ckpt / t_fwd, 3),
        "gradients_bitwise_equal_across_schedules": same,
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


def _timed_run(workload: str, precision: str, steps: int, warmup: int, warmup_seconds: float, total: int,
               arithmetic: str = "reference", eps_pf: float | None = None):
    """Fused steps of a workload with the state resident in HBM: `warmup` untimed steps, more until
    `warmup_seconds` have passed (at most `total` - `steps` in all), then exactly `steps` timed steps
    between synchronisations, then untimed steps up to `total` (if more).  Returns a record: the setup
    (without its arrays), interior cells, wall seconds of the timed steps, Gcell-updates/s, the fused
    kernel's mean time in ms over the timed steps (HIP events on its stream), its launch plan, the
    warm-up step count, and the interior density after max(total, steps run) steps."""
    from types import SimpleNamespace

    import torch

    from heraclespp_amd import _lib
    from heraclespp_amd.hydro import Grid, PerfectGas, UniformGravity
    from heraclespp_amd.solver import HydroSolver

    s = _setup(workload, WORKLOADS[workload][0])
    grid = Grid(s.ndim, s.n, s.ng, s.geom, s.xmin, s.xmax)
    solver = HydroSolver(grid, PerfectGas(s.gamma, 1.0), reconstruction=s.reconstruction,
                         riemann_solver=s.riemann_solver, gravity=UniformGravity(s.gravity), bc=s.bc, cfl=s.cfl,
                         precision=precision, arithmetic=arithmetic,
                         pressure_fix="Off" if eps_pf is None else "On", eps_pf=1e-6 if eps_pf is None else eps_pf)
    solver.init_from_prim(s.rho, s.u, s.P)
    del s.rho, s.u, s.P
    cells = 1
    for d in range(s.ndim):
        cells *= s.n[d]
    open_ended = total <= 0  # no step count to reach: the run ends with the timed steps
    cap = (1 << 30) if open_ended else max(total, warmup + steps)
    solver.start_clock(1e300, cap)
    for _ in range(warmup):
        solver.step_async()
    torch.cuda.synchronize()
    done = warmup
    t_w = time.perf_counter()
    while time.perf_counter() - t_w < warmup_seconds and (open_ended or done + steps < cap):
        for _ in range(10 if open_ended else min(10, cap - steps - done)):
            solver.step_async()
            done += 1
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
    total = done + steps if open_ended else cap
    for _ in range(total - done - steps):
        solver.step_async()
    clk = solver.finish()
    if clk.iter != total:
        raise RuntimeError(f"expected {total} {precision} steps, clock counted {clk.iter}")
    seglen, nwg, occ = ctypes.c_int(), ctypes.c_longlong(), ctypes.c_int()
    _lib.check(L.hclpp_fused_plan(solver.ctx.handle, ctypes.byref(seglen), ctypes.byref(nwg), ctypes.byref(occ)))
    rho = solver.rho.cpu().numpy()[s.interior()]
    del solver
    torch.cuda.empty_cache()
    return SimpleNamespace(setup=s, cells=cells, wall=t1 - t0, value=cells * steps / (t1 - t0) / 1e9,
                           kernel_ms=kms.value / max(nl.value, 1), rho=rho, warmup_steps=done, total=total,
                           plan={"segment_planes": seglen.value, "workgroups": nwg.value, "resident_per_cu": occ.value})


def _rel_l1(a, b):
    import numpy as np

    return float(np.sum(np.abs(a - b)) / np.sum(np.abs(b)))


def _hbm(r, precision):
    word = 8 if precision == "fp64" else 4
    gbs = 2 * (r.setup.ndim + 2) * word * r.cells / (r.kernel_ms / 1e3) / 1e9
    return round(gbs, 1), round(gbs / HBM_PEAK_GBS, 5)


def mixed_block(workload: str, a, value64: float, rho64, total: int):
    """The workload again with FP32 storage: throughput (same timing as the FP64 line) and the
    density's relative L1 distance to the FP64 run after the same `total` steps."""
    r = _timed_run(workload, "fp32", a.steps, a.warmup, a.warmup_seconds, total)
    gbs, frac = _hbm(r, "fp32")
    return {"value": round(r.value, 4), "unit": "Gcell-updates/s", "kernel_ms": round(r.kernel_ms, 5),
            "hbm_achieved_gbs": gbs, "hbm_frac": frac, "ratio_to_fp64": round(r.value / value64, 3),
            "density_rel_l1_vs_fp64": float(f"{_rel_l1(r.rho, rho64):.3e}"), "steps_compared": r.total,
            "warmup_steps_run": r.warmup_steps}


def divergence_curve(workload: str, kw: dict, total: int, every: int):
    """The fast FP64 arithmetic against the reference arithmetic (bit-identical to the oracle) on the
    same initial state: relative L1 distance of rho, rhou and E every `every` steps up to `total`, and
    the first checkpoint at which each exceeds 1e-12."""
    import torch

    from heraclespp_amd.hydro import Grid, PerfectGas, UniformGravity
    from heraclespp_amd.solver import HydroSolver

    s = _setup(workload, kw)
    grid = Grid(s.ndim, s.n, s.ng, s.geom, s.xmin, s.xmax)
    sols = []
    for arith in ("reference", "fast"):
        sol = HydroSolver(grid, PerfectGas(s.gamma, 1.0), reconstruction=s.reconstruction, riemann_solver=s.riemann_solver,
                          gravity=UniformGravity(s.gravity), bc=s.bc, cfl=s.cfl, arithmetic=arith)
        sol.init_from_prim(s.rho, s.u, s.P)
        sol.start_clock(1e300, total)
        sols.append(sol)
    I = (Ellipsis,) + s.interior()
    curve, first = [], {}
    for n in range(every, total + 1, every):
        for sol in sols:
            for _ in range(every):
                sol.step_async()
        torch.cuda.synchronize()
        # the state after `launched` steps: buffer (cur0 + launched) % 2 (no early stop before `total`)
        cur = [sol._bufs()[(sol._cur0 + sol._launched) % 2] for sol in sols]
        row = {"n": n}
        for k, nm in enumerate(("rho", "rhou", "E")):
            ref, fast = cur[0][k][I], cur[1][k][I]
            d = float((fast - ref).abs().sum() / ref.abs().sum())
            row[nm] = float(f"{d:.3e}")
            if d > 1e-12 and nm not in first:
                first[nm] = n
        curve.append(row)
    for sol in sols:
        clk = sol.finish()
        if clk.iter != total:
            raise RuntimeError(f"expected {total} steps, clock counted {clk.iter}")
    del sols
    torch.cuda.empty_cache()
    return {"workload": f"{workload} {'x'.join(str(v) for v in s.n[:s.ndim])}", "rel_l1": curve,
            "first_n_above_1e-12": {nm: first.get(nm) for nm in ("rho", "rhou", "E")}}


def fast_block(workload: str, a, value_ref: float, rho_ref, total: int):
    """The workload with the opt-in fast FP64 arithmetic (hclpp_set_fast_arithmetic): throughput timed
    as the reference line, its density's relative L1 distance to the reference run after the same
    `total` steps, and divergence curves against the reference arithmetic on Rayleigh-Taylor 128^3
    (500 steps) and Kelvin-Helmholtz 2048^2 (2000 steps)."""
    r = _timed_run(workload, "fp64", a.steps, a.warmup, a.warmup_seconds, total, arithmetic="fast")
    gbs, frac = _hbm(r, "fp64")
    out = {"value": round(r.value, 4), "unit": "Gcell-updates/s", "kernel_ms": round(r.kernel_ms, 5),
           "hbm_achieved_gbs": gbs, "hbm_frac": frac, "ratio_to_reference": round(r.value / value_ref, 3),
           "density_rel_l1_vs_reference": float(f"{_rel_l1(r.rho, rho_ref):.3e}"), "steps_compared": r.total,
           "launch": r.plan, "warmup_steps_run": r.warmup_steps}
    units = _valu_units(workload, "fp64_fast")
    if units is not None:
        out["valu"] = {"issue_units_per_cell": units, "roof_gcells": round(39.3e12 / units / 1e9, 3),
                       "source": f"profiles/k_step_march_{workload}_fp64_fast_valu.txt"}
    del r
    out["divergence"] = [divergence_curve("rayleigh_taylor", {"n": 128}, 500, 50),
                         divergence_curve("kelvin_helmholtz", {"n": 2048}, 2000, 200)]
    return out


def kh_block(a):
    """The 2D Kelvin-Helmholtz 2048^2 run (VanLeer + HLLC) in FP64 and with FP32 storage: throughput,
    kernel time and HBM fraction at the algorithmic bytes, and the FP32/FP64 density L1 after its
    configured 2000 steps."""
    out = {"workload": WORKLOADS["kelvin_helmholtz"][2], "steps": a.steps}
    total = WORKLOADS["kelvin_helmholtz"][3]
    r = _timed_run("kelvin_helmholtz", "fp64", a.steps, a.warmup, a.warmup_seconds, total)
    gbs, frac = _hbm(r, "fp64")
    out["fp64"] = {"value": round(r.value, 4), "unit": "Gcell-updates/s", "kernel_ms": round(r.kernel_ms, 5),
                   "hbm_achieved_gbs": gbs, "hbm_frac": frac, "warmup_steps_run": r.warmup_steps}
    rho64, v64, total = r.rho, r.value, r.total
    del r
    out["fp32"] = mixed_block("kelvin_helmholtz", a, v64, rho64, total)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--warmup-seconds", type=float, default=1.0,
                    help="keep stepping untimed until this much warm-up time has passed")
    ap.add_argument("--workload", default="rayleigh_taylor", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"],
                    help="fp32: FP32 storage and arithmetic with FP64 pressure recovery, dt and totals")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiable-Euler adjoint measurement")
    ap.add_argument("--no-mixed", action="store_true", help="skip the FP32-storage run of the same workload")
    ap.add_argument("--no-kh", action="store_true", help="skip the 2D Kelvin-Helmholtz 2048^2 FP64 / FP32 block")
    ap.add_argument("--no-fast", action="store_true", help="skip the fast-FP64-arithmetic block")
    ap.add_argument("--pressure-fix", type=float, default=None, metavar="EPS",
                    help="run the headline workload with the pressure fix on (eps_pf = EPS; FP64, fused)")
    ap.add_argument("--arithmetic", default="reference", choices=["reference", "fast"],
                    help="fast: the headline run with the opt-in fast FP64 arithmetic (not bit-identical to the reference)")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="stop after the timed steps instead of continuing to the configured step count")
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
    if a.arithmetic == "fast" and a.precision != "fp64":
        ap.error("--arithmetic fast is an FP64 mode")
    if a.pressure_fix is not None and (a.precision != "fp64" or a.arithmetic != "reference"):
        ap.error("--pressure-fix runs with FP64 storage and the reference arithmetic")
    total = 0 if (a.no_accuracy or (a.no_mixed and a.no_fast) or a.precision != "fp64" or a.arithmetic != "reference"
                  or a.pressure_fix is not None) else WORKLOADS[a.workload][3]
    r = _timed_run(a.workload, a.precision, a.steps, a.warmup, a.warmup_seconds, total, arithmetic=a.arithmetic,
                   eps_pf=a.pressure_fix)
    s, cells = r.setup, r.cells
    wall = r.wall
    value = r.value
    kernel_s = r.kernel_ms / 1e3
    word = 8 if a.precision == "fp64" else 4
    bytes_alg = 2 * (s.ndim + 2) * word * cells
    achieved = bytes_alg / kernel_s / 1e9
    hbm_frac = achieved / HBM_PEAK_GBS
    tag = a.precision + ("_fast" if a.arithmetic == "fast" else "")
    traffic = _traffic(a.workload, tag)
    units = _valu_units(a.workload, tag)
    # FP64 VALU issue capacity: 1024 SIMDs, one wave64 FP64 instruction per 4 cycles each, at 2.4 GHz
    valu = None
    bound = "hbm"
    if units is not None:
        issued = units * cells / 64
        capacity = 1024 * kernel_s * 2.4e9 / 4
        roof = 39.3e12 / units / 1e9
        valu = {"issue_units_per_cell": units, "busy_frac": round(issued / capacity, 4),
                "roof_gcells": round(roof, 3), "kernel_gcells": round(cells / kernel_s / 1e9, 4),
                "frac_of_roof": round(cells / kernel_s / 1e9 / roof, 4),
                "source": f"profiles/k_step_march_{a.workload}_{tag}_valu.txt"}
        if issued / capacity > hbm_frac:
            bound = "valu"
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
            "arithmetic": a.arithmetic,
            "pressure_fix": None if a.pressure_fix is None else {"eps_pf": a.pressure_fix},
            "path": "fused (hclpp_step: march kernel + ghost fill + device dt)",
            "launch": r.plan,
            "warmup_steps_run": r.warmup_steps,
            "warmup_seconds_min": a.warmup_seconds,
        },
        "roofline": {
            "bound": bound,
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(hbm_frac, 5),
            "hbm_frac": round(hbm_frac, 5),
            "traffic": traffic,
            "kernel": "k_step_march",
            "kernel_ms": round(kernel_s * 1e3, 5),
            "bytes_per_launch": bytes_alg,
            "valu": valu,
        },
        "cpu_baseline": None,
        "mixed_fp32": None,
        "fast_fp64": None,
        "kelvin_helmholtz_2048": None,
        "adjoint": None,
    }
    rho64, total = r.rho, r.total
    del r
    reference_fp64 = a.precision == "fp64" and a.arithmetic == "reference" and a.pressure_fix is None
    if reference_fp64 and not a.no_mixed:
        line["mixed_fp32"] = mixed_block(a.workload, a, value, rho64, total)
        torch.cuda.empty_cache()
    if reference_fp64 and not a.no_fast:
        line["fast_fp64"] = fast_block(a.workload, a, value, rho64, total)
        torch.cuda.empty_cache()
    if reference_fp64 and a.workload == "rayleigh_taylor" and not a.no_kh:
        line["kelvin_helmholtz_2048"] = kh_block(a)
    if not a.no_adjoint:
        line["adjoint"] = adjoint_block(a.adjoint_n, a.adjoint_steps, 20, a.adjoint_fd_dirs)
        torch.cuda.empty_cache()
    if not a.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(a.workload, a.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
