"""Differentiable rollouts of the hydro step: a torch.autograd.Function over N fused steps
whose backward runs the hand-written reverse-mode step kernels (hclpp_step_adjoint).

* dt is held fixed: a rollout replays a dt schedule (by default the CFL schedule of the
  unperturbed forward, recorded on the device by hclpp_clock_record_dt), so the function
  being differentiated is U_N = S_{dt_N} o ... o S_{dt_1}(U_0; gamma).
* The schedule trades memory for recomputation.  `schedule="store_all"` keeps every state
  of the forward (N + 1 ghost-inclusive states, written in place by the fused steps) and the
  backward reverses all N steps in one call, with no recomputation.  `"periodic"` keeps the
  state every `checkpoint_every` steps; the backward walks the segments from last to first,
  recomputes each segment's states from its checkpoint with the same fused forward kernel
  (bit-identical to the original forward) and runs the reverse steps over it.  `"binomial"`
  keeps at most `snapshots` states alive and places them by the binomial (revolve) rule,
  trading more recomputation for O(snapshots) memory.  `"auto"` (the default) takes
  store_all when the N + 1 states fit in `memory_budget` bytes (by default half the device
  memory free at the call), else periodic.  Every schedule gives bit-identical gradients.
* A rollout whose inputs need no gradient (or that runs with grad disabled) keeps no state.
* Inputs and outputs are interior arrays (ghost cells are filled by the boundary
  conditions inside the rollout, so they carry no gradient).
* gamma is a differentiable input: its gradient is the sum over steps of the reverse
  kernels' FP64 grid reductions.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import torch

from ._lib import check, lib
from .hydro import Context, Grid, PerfectGas, UniformGravity


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _ptr_array(ts):
    return (ctypes.c_void_p * len(ts))(*[t.data_ptr() for t in ts])


@dataclass
class RolloutSpec:
    """Grid and scheme of a rollout (Cartesian or spherical geometry, HLL or HLLC, any limiter and
    boundaries)."""

    grid: Grid
    reconstruction: str
    riemann_solver: str
    bc: list
    gravity: UniformGravity | None = None
    mmw: float = 1.0
    _engines: dict = field(default_factory=dict, repr=False)
    _side_engines: dict = field(default_factory=dict, repr=False)

    def engine(self, gamma: float) -> "StepEngine":
        e = self._engines.get(gamma)
        if e is None:
            e = StepEngine(self, gamma)
            self._engines[gamma] = e
        return e

    def side_engine(self, gamma: float) -> "StepEngine":
        """An engine whose context runs on a stream of its own (default priority), for the segment
        recomputation that overlaps the reverse steps."""
        e = self._side_engines.get(gamma)
        if e is None:
            stream = torch.cuda.Stream()
            with torch.cuda.stream(stream):
                e = StepEngine(self, gamma, own_context=True)
            e.stream = stream
            self._side_engines[gamma] = e
        return e

    @property
    def interior(self):
        g = self.grid
        sl = [slice(g.Nghost[d], g.Nghost[d] + g.Nx_glob_ng[d]) for d in range(3)]
        return (sl[2], sl[1], sl[0])


class StepEngine:
    """One context: fixed-dt forward steps (hclpp_step_dt) and reverse steps (hclpp_step_adjoint)."""

    def __init__(self, spec: RolloutSpec, gamma: float, own_context: bool = False):
        """own_context: a context of its own (its own stream, clock and scratch) instead of the
        grid's shared one for this configuration."""
        self.spec = spec
        self.gamma = gamma
        args = (PerfectGas(gamma, spec.mmw), spec.gravity, spec.bc, None, 0, spec.reconstruction, spec.riemann_solver, "fp64")
        self.ctx = Context(spec.grid, *args) if own_context else spec.grid.context(*args)
        self.nd = spec.grid.ndim

    def new_state(self):
        g, nd = self.spec.grid, self.nd
        return (g.zeros(), g.zeros(nd), g.zeros())

    def new_states(self, n: int):
        """n states as views of one uninitialised allocation per variable, for step outputs: a step
        writes every cell of its output (the interior, then the whole ghost shell)."""
        shape, nd = self.spec.grid.shape3, self.nd
        rho = torch.empty((n,) + shape, dtype=torch.float64, device="cuda")
        rhou = torch.empty((n, nd) + shape, dtype=torch.float64, device="cuda")
        E = torch.empty((n,) + shape, dtype=torch.float64, device="cuda")
        return [(rho[k], rhou[k], E[k]) for k in range(n)]

    def load(self, rho, rhou, E, out=None):
        """Interior arrays -> ghost-inclusive state with its ghosts filled."""
        U = out if out is not None else self.new_state()
        it = self.spec.interior
        U[0][it] = rho
        U[1][(slice(None),) + it] = rhou
        U[2][it] = E
        c = self.ctx
        check(lib().hclpp_fill_ghosts(c.handle, _ptr(U[0]), _ptr(U[1]), _ptr(U[2]), ctypes.c_void_p(0)))
        return U

    def start(self, U, cfl: float = 1.0, max_iter: int = 2 ** 31 - 1):
        check(lib().hclpp_clock_start(self.ctx.handle, 0.0, 0, 1e300, cfl, max_iter, _ptr(U[0]), _ptr(U[1]), _ptr(U[2])))

    def step(self, src, dst, dt: float):
        check(lib().hclpp_step_dt(self.ctx.handle, float(dt), _ptr(src[0]), _ptr(src[1]), _ptr(src[2]), _ptr(dst[0]),
                                  _ptr(dst[1]), _ptr(dst[2])))

    def cfl_step(self, src, dst):
        check(lib().hclpp_step(self.ctx.handle, _ptr(src[0]), _ptr(src[1]), _ptr(src[2]), _ptr(dst[0]), _ptr(dst[1]),
                               _ptr(dst[2])))

    def steps(self, states, dts):
        """len(dts) steps, step k from states[k] to states[k + 1] (one C call)."""
        n = len(dts)
        if n == 0:
            return
        arr = (ctypes.c_double * n)(*dts)
        cols = [_ptr_array([U[v] for U in states]) for v in range(3)]
        check(lib().hclpp_steps_dt(self.ctx.handle, n, arr, *cols))

    def vjps(self, states, lams, dgamma, dts):
        """Reverse steps len(dts)-1 .. 0: lams[k] = VJP of step k (state states[k]) applied to lams[k + 1]."""
        n = len(dts)
        if n == 0:
            return
        arr = (ctypes.c_double * n)(*dts)
        cols = [_ptr_array([U[v] for U in states]) for v in range(3)]
        lcols = [_ptr_array([L[v] for L in lams]) for v in range(3)]
        check(lib().hclpp_steps_adjoint(self.ctx.handle, n, arr, *cols, *lcols, _ptr(dgamma)))

    def vjp(self, U, lam_new, lam, dgamma, dt: float):
        check(lib().hclpp_step_adjoint(self.ctx.handle, float(dt), _ptr(U[0]), _ptr(U[1]), _ptr(U[2]), _ptr(lam_new[0]),
                                       _ptr(lam_new[1]), _ptr(lam_new[2]), _ptr(lam[0]), _ptr(lam[1]), _ptr(lam[2]),
                                       _ptr(dgamma)))

    def clock(self):
        from . import _lib

        out = _lib.Clock()
        check(lib().hclpp_clock_read(self.ctx.handle, ctypes.byref(out)))
        return out

    def check_clock(self, steps: int):
        clk = self.clock()
        if clk.status:
            raise RuntimeError(f"iteration {clk.iter}: detected invalid volumic internal energy")
        if clk.iter != steps:
            raise RuntimeError(f"expected {steps} steps, the clock counted {clk.iter}")


def cfl_schedule(spec: RolloutSpec, rho, rhou, E, gamma: float, nsteps: int, cfl: float) -> list[float]:
    """dt of each of nsteps CFL-limited fused steps from (rho, rhou, E), recorded on the device."""
    eng = spec.engine(float(gamma))
    a = eng.load(rho, rhou, E)
    b = eng.new_state()
    log = torch.zeros(nsteps, dtype=torch.float64, device="cuda")
    eng.start(a, cfl=cfl, max_iter=nsteps)
    check(lib().hclpp_clock_record_dt(eng.ctx.handle, _ptr(log), nsteps))
    for _ in range(nsteps):
        eng.cfl_step(a, b)
        a, b = b, a
    eng.check_clock(nsteps)
    return [float(v) for v in log.cpu()]


def forward_states(eng: StepEngine, U0, dts, keep):
    """Run the steps dts from state U0; returns {step index: state clone} for the indices in keep, and the final state."""
    N = len(dts)
    bufs = (tuple(t.clone() for t in U0), eng.new_state())
    kept = {0: tuple(t.clone() for t in U0)} if 0 in keep else {}
    eng.start(bufs[0])
    marks = sorted(k for k in keep if 0 < k < N) + [N]
    at, cur = 0, 0
    for m in marks:
        n = m - at
        eng.steps([bufs[(cur + j) % 2] for j in range(n + 1)], dts[at:m])
        cur = (cur + n) % 2
        at = m
        if m in keep:
            kept[m] = tuple(t.clone() for t in bufs[cur])
    eng.check_clock(N)
    return kept, bufs[cur]


def binomial_split(steps: int, snaps: int) -> int:
    """Steps to advance before placing the next snapshot (revolve's binomial rule).

    With s snapshots and t recomputation sweeps, beta(s, t) = C(s + t, s) steps can be
    reversed; the split keeps the remaining problem on both sides within one sweep less."""
    if steps <= 1 or snaps <= 0:
        return 0
    t = 0
    while math.comb(snaps + t, snaps) < steps:
        t += 1
    # smallest advance m that leaves the right part (steps - m) reversible with snaps - 1 snapshots
    # in t sweeps; the left part m then fits snaps snapshots in t - 1 sweeps (beta(s,t) = beta(s,t-1) + beta(s-1,t))
    left_cap = math.comb(snaps + t - 1, snaps) if t > 0 else 1
    m = steps - math.comb(snaps - 1 + t, snaps - 1)
    return max(1, min(m, left_cap, steps - 1))


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rho, rhou, E, gamma, spec, dts, every, schedule, snapshots, overlap):
        eng = spec.engine(float(gamma))
        N = len(dts)
        if schedule == "store_all":
            states = [eng.load(rho, rhou, E)] + eng.new_states(N)
            eng.start(states[0])
            eng.steps(states, dts)
            eng.check_clock(N)
            kept, last = dict(enumerate(states)), states[N]
        else:
            keep = {"periodic": set(range(0, N, every)), "binomial": {0}, "none": set()}[schedule]
            kept, last = forward_states(eng, eng.load(rho, rhou, E), dts, keep)
        ctx.spec, ctx.dts, ctx.every, ctx.schedule, ctx.snapshots = spec, list(dts), every, schedule, snapshots
        ctx.overlap = overlap
        ctx.kept = kept
        ctx.gamma_shape, ctx.gamma_device = gamma.shape, gamma.device
        ctx.gamma_value = float(gamma)
        it = spec.interior
        return last[0][it].clone(), last[1][(slice(None),) + it].clone(), last[2][it].clone()

    @staticmethod
    def backward(ctx, g_rho, g_rhou, g_E):
        spec = ctx.spec
        eng = spec.engine(ctx.gamma_value)
        it = spec.interior
        lam = eng.new_state()
        lam2 = eng.new_state()
        for dst, g in zip(lam, (g_rho, g_rhou, g_E)):
            if g is not None:
                dst[(Ellipsis,) + it] = g
        dgam = torch.zeros(1, dtype=torch.float64, device="cuda")
        state = {"lam": lam, "lam2": lam2, "reverse_steps": 0, "forward_steps": 0}

        def reverse(U, n):
            eng.vjp(U, state["lam"], state["lam2"], dgam, ctx.dts[n])
            state["lam"], state["lam2"] = state["lam2"], state["lam"]
            state["reverse_steps"] += 1

        def batched(seg, s, e):
            _reverse_segment(eng, seg, s, e, ctx.dts, dgam, state)

        N = len(ctx.dts)
        if ctx.schedule == "none":
            raise RuntimeError("the rollout ran without grad and kept no state to reverse")
        eng.start(ctx.kept[0])  # the clock only gates the recomputation steps here; dt comes from the schedule
        if ctx.schedule == "store_all":
            _reverse_segment(eng, [ctx.kept[k] for k in range(N)], 0, N, ctx.dts, dgam, state)
        elif ctx.schedule == "periodic" and ctx.overlap and len(ctx.kept) > 1:
            side = spec.side_engine(ctx.gamma_value)
            side.start(ctx.kept[0])
            _reverse_periodic_overlapped(eng, side, ctx.kept, ctx.dts, ctx.every, dgam, state)
            side.check_clock(state["forward_steps"])
            state["forward_steps"] = 0  # the main engine's clock counted none
        elif ctx.schedule == "periodic":
            _reverse_periodic(eng, ctx.kept, ctx.dts, ctx.every, reverse, state, batched)
        else:
            _reverse_binomial(eng, ctx.kept[0], 0, N, ctx.dts, ctx.snapshots, reverse, state)
        ctx.kept = None
        eng.check_clock(state["forward_steps"])
        lam = state["lam"]
        out = (lam[0][it].clone(), lam[1][(slice(None),) + it].clone(), lam[2][it].clone())
        _Rollout.last_backward = {"reverse_steps": state["reverse_steps"], "recomputed_steps": state.get("recomputed", state["forward_steps"])}
        return out + (dgam.reshape(ctx.gamma_shape).to(ctx.gamma_device), None, None, None, None, None, None)


def _reverse_segment(eng, seg, s, e, dts, dgam, state):
    """Reverse steps e-1 .. s in one call; seg[k - s] is the state before step k.  The adjoints
    alternate between state["lam"] and state["lam2"] so that adjoint e is the current one."""
    n = e - s
    lams = [state["lam"] if (e - k) % 2 == 0 else state["lam2"] for k in range(s, e + 1)]
    eng.vjps(seg, lams, dgam, dts[s:e])
    if n % 2:
        state["lam"], state["lam2"] = state["lam2"], state["lam"]
    state["reverse_steps"] += n


def _reverse_periodic(eng, kept, dts, every, reverse, state, batched=None):
    """Segments last to first: recompute the segment's states from its checkpoint, then reverse it.

    `batched(states, s, e)` (optional) reverses steps e-1 .. s in one call; otherwise `reverse`
    is called per step."""
    N = len(dts)
    pool = [eng.new_state() for _ in range(min(every, N))]
    for s in sorted(kept, reverse=True):
        e = min(s + every, N)
        seg = [kept[s]] + pool[:e - 1 - s]
        if hasattr(eng, "steps"):
            eng.steps(seg, dts[s:e - 1])
        else:
            for n in range(s, e - 1):
                eng.step(seg[n - s], seg[n - s + 1], dts[n])
        state["forward_steps"] += e - 1 - s
        if batched is not None:
            batched(seg, s, e)
        else:
            for n in range(e - 1, s - 1, -1):
                reverse(seg[n - s], n)


def _reverse_periodic_overlapped(eng, side, kept, dts, every, dgam, state):
    """_reverse_periodic with the recomputation of the next segment (on the side engine's stream)
    running while the current one is reversed (on the main engine's stream, the caller's current
    stream).  Two pools of segment states alternate; events order each recomputation after the
    reversal that last read its pool, and each reversal after its segment's recomputation.  The
    kernels and their order on each stream are those of _reverse_periodic, so the gradients are
    bit-identical to it."""
    N = len(dts)
    starts = sorted(kept, reverse=True)
    pools = [[eng.new_state() for _ in range(min(every, N))] for _ in range(2)]
    ready = [torch.cuda.Event(), torch.cuda.Event()]
    free = [None, None]
    main = torch.cuda.current_stream()
    segs = [None, None]

    def recompute(i, s):
        e = min(s + every, N)
        seg = [kept[s]] + pools[i][:e - 1 - s]
        with torch.cuda.stream(side.stream):
            side.stream.wait_stream(main) if free[i] is None else side.stream.wait_event(free[i])
            side.steps(seg, dts[s:e - 1])
            ready[i].record(side.stream)
        state["recomputed"] = state.get("recomputed", 0) + e - 1 - s
        segs[i] = (seg, s, e)

    recompute(0, starts[0])
    for idx in range(len(starts)):
        i = idx % 2
        if idx + 1 < len(starts):
            recompute(1 - i, starts[idx + 1])
        seg, s, e = segs[i]
        main.wait_event(ready[i])
        _reverse_segment(eng, seg, s, e, dts, dgam, state)
        ev = torch.cuda.Event()
        ev.record(main)
        free[i] = ev
    main.wait_stream(side.stream)
    state["forward_steps"] = state["recomputed"]


def _reverse_binomial(eng, U, first, last, dts, snaps, reverse, state):
    """Reverse steps [first, last) starting from state U at step `first`, holding at most `snaps` extra states."""
    if last - first == 1:
        reverse(U, first)
        return
    if snaps <= 0:  # no snapshot left: recompute from U for every step, last to first
        tmp = [eng.new_state(), eng.new_state()]
        for n in range(last - 1, first - 1, -1):
            cur = U
            for m in range(first, n):
                eng.step(cur, tmp[(m - first) % 2], dts[m])
                cur = tmp[(m - first) % 2]
                state["forward_steps"] += 1
            reverse(cur, n)
        return
    m = binomial_split(last - first, snaps)
    cur = tuple(t.clone() for t in U)
    nxt = eng.new_state()
    for k in range(first, first + m):
        eng.step(cur, nxt, dts[k])
        cur, nxt = nxt, cur
        state["forward_steps"] += 1
    _reverse_binomial(eng, cur, first + m, last, dts, snaps - 1, reverse, state)
    del cur, nxt
    _reverse_binomial(eng, U, first, first + m, dts, snaps, reverse, state)


def state_bytes(spec: RolloutSpec) -> int:
    """Bytes of one ghost-inclusive FP64 state (rho, rhou, E) of the spec's grid."""
    g = spec.grid
    cells = g.Nx_local_wg[0] * g.Nx_local_wg[1] * g.Nx_local_wg[2]
    return cells * (g.ndim + 2) * 8


def pick_schedule(spec: RolloutSpec, nsteps: int, memory_budget: int | None = None) -> str:
    """"store_all" when the nsteps + 1 states of a rollout fit in memory_budget bytes (default: half the
    device memory free now), else "periodic"."""
    if memory_budget is None:
        memory_budget = torch.cuda.mem_get_info()[0] // 2
    return "store_all" if (nsteps + 1) * state_bytes(spec) <= memory_budget else "periodic"


def rollout(rho, rhou, E, gamma, spec: RolloutSpec, dts, checkpoint_every: int = 20, schedule: str = "auto",
            snapshots: int = 8, overlap: bool = True, memory_budget: int | None = None):
    """Differentiable N-step rollout of interior arrays (rho, rhou, E) with the fixed dt schedule `dts`.

    gamma: a 0-d float64 tensor (may require grad).  schedule: "store_all" (every state kept, no
    recomputation), "periodic" (a checkpoint every `checkpoint_every` steps, segment
    recomputation), "binomial" (at most `snapshots` live states, revolve placement) or "auto"
    (store_all if the N + 1 states fit in `memory_budget` bytes, default half the free device
    memory; else periodic).  overlap (periodic schedule): recompute the next segment on a second
    stream while the current one is reversed (same results, bit for bit).  With grad disabled, or
    no input that requires grad, no state is kept."""
    if schedule not in ("auto", "store_all", "periodic", "binomial"):
        raise ValueError("schedule must be 'auto', 'store_all', 'periodic' or 'binomial'")
    if checkpoint_every < 1:
        raise ValueError("checkpoint_every must be >= 1")
    if not torch.is_tensor(gamma):
        gamma = torch.tensor(float(gamma), dtype=torch.float64)
    dts = list(dts)
    if not (torch.is_grad_enabled() and any(t.requires_grad for t in (rho, rhou, E, gamma))):
        schedule = "none"
    elif schedule == "auto":
        schedule = pick_schedule(spec, len(dts), memory_budget)
    return _Rollout.apply(rho, rhou, E, gamma, spec, dts, checkpoint_every, schedule, snapshots, overlap)


class DiffEulerProblem:
    """The differentiable-Euler configuration: smooth 2D periodic flow (setups.smooth_euler), a fixed
    CFL dt schedule from the unperturbed state, loss L = 1/2 sum (rho_T - rho_tgt)^2 dV with rho_tgt
    the rollout of a seeded perturbation of rho0.  Inputs of the loss: rho0 (interior) and gamma;
    u and P are held fixed and the conservative state is formed in torch (differentiable).

    With uniform u and P the scheme carries the density as a pure contact, which HLLC resolves
    without any dependence on the wave speeds: rho_T, hence this loss, does not depend on gamma.
    `energy_weight` > 0 adds w/2 sum (E_T - E_tgt)^2 dV, which does (E = P/(gamma-1) + rho u^2/2);
    its gamma derivative is -P/(gamma-1)^2 sum (E_T - E_tgt) dV, which vanishes for a zero-mean
    perturbation (energy is conserved), so `target_offset` shifts the perturbation's mean."""

    def __init__(self, n: int, steps: int, seed: int = 0, gamma: float = 1.4, target_amp: float = 0.02,
                 energy_weight: float = 0.0, target_offset: float = 0.0):
        import numpy as np

        from . import setups

        s = setups.smooth_euler(n=n, steps=steps, seed=seed)
        self.setup = s
        self.grid = Grid(2, s.n, 2)
        self.spec = RolloutSpec(self.grid, s.reconstruction, s.riemann_solver, s.bc)
        it = s.interior()
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
        self.rho0 = dev(s.rho[it])
        self.u = dev(s.u[(slice(None),) + it])
        self.P = dev(s.P[it])
        self.gamma = float(gamma)
        self.energy_weight = float(energy_weight)
        self.dV = 1.0 / (n * n)
        r, m, E = self.cons(self.rho0, torch.tensor(self.gamma, dtype=torch.float64))
        self.dts = cfl_schedule(self.spec, r, m, E, self.gamma, steps, s.cfl)
        xc = (np.arange(n) + 0.5) / n
        X, Y = np.meshgrid(xc, xc)
        pert = setups.smooth_density(X, Y, seed + 1, modes=8, kmax=4, amp=target_amp) - 1.0 + target_offset
        with torch.no_grad():
            rt, mt, et = self.cons(self.rho0 + dev(pert), torch.tensor(self.gamma, dtype=torch.float64))
            out = rollout(rt, mt, et, self.gamma, self.spec, self.dts)
            self.target, self.target_E = out[0], out[2]

    def cons(self, rho0, gamma):
        """E = P / (gamma - 1) + rho |u|^2 / 2, m = rho u."""
        return rho0, rho0.unsqueeze(0) * self.u, self.P / (gamma - 1) + 0.5 * rho0 * (self.u * self.u).sum(0)

    def loss(self, rho0, gamma, every=20, schedule="auto", snapshots=8):
        r, m, E = self.cons(rho0, gamma)
        rT, _, ET = rollout(r, m, E, gamma, self.spec, self.dts, checkpoint_every=every, schedule=schedule,
                            snapshots=snapshots)
        L = 0.5 * torch.sum((rT - self.target) ** 2) * self.dV
        if self.energy_weight:
            L = L + 0.5 * self.energy_weight * torch.sum((ET - self.target_E) ** 2) * self.dV
        return L

    def gradient(self, every=20, schedule="auto", snapshots=8):
        x = self.rho0.clone().requires_grad_(True)
        g = torch.tensor(self.gamma, dtype=torch.float64, requires_grad=True)
        L = self.loss(x, g, every, schedule, snapshots)
        L.backward()
        return L.detach(), x.grad, float(g.grad)

    def fd(self, v, vg: float, h: float) -> float:
        """Central difference of L along (v, vg); the loss difference is formed term by term."""
        with torch.no_grad():
            out = []
            for sgn in (1.0, -1.0):
                gam = self.gamma + sgn * h * vg
                r, m, E = self.cons(self.rho0 + sgn * h * v, torch.tensor(gam, dtype=torch.float64))
                out.append(rollout(r, m, E, gam, self.spec, self.dts))
            (Rp, _, Ep), (Rm, _, Em) = out
            d = torch.sum((Rp - Rm) * (Rp + Rm - 2 * self.target))
            if self.energy_weight:
                d = d + self.energy_weight * torch.sum((Ep - Em) * (Ep + Em - 2 * self.target_E))
            return float(d * 0.5 * self.dV / (2 * h))


def diff_euler_problem(n: int = 1024, steps: int = 400, seed: int = 0, energy_weight: float = 0.0,
                       target_offset: float = 0.0) -> DiffEulerProblem:
    return DiffEulerProblem(n, steps, seed, energy_weight=energy_weight, target_offset=target_offset)
