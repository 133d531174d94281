"""Time-loop driver: the reference main.cpp:305-311 (initialisation) and
:354-531 (loop body) on one GPU.

Two paths, both HIP:
  * fused (default): one hclpp_step (hclpp_step_fx with passive scalars)
    per iteration -- the fused reconstruct/predict/Riemann/update kernel, the
    pressure fix when it is on, the ghost fill and the device clock (dt, t_end
    clipping, max_iter, the internal-energy check) -- with no host
    synchronisation inside the loop;
  * operators: the reference operator sequence (time_step,
    minimum_internal_energy, MUSCL-Hancock reconstruction, Godunov update,
    pressure fix, boundaries, cons->prim), one kernel per reference kernel,
    host-synchronised on dt exactly like the reference.  Used to bisect
    parity and for user-supplied strategies.
With an output schedule (simulate(), io.OutputSchedule) dt is clipped to the
output times and outputs are written in the reference's HDF5 layout, as
main.cpp:313-345 and :357-375, :532-557 do; restart() reads such a file back
(main.cpp:223-238, then :305-311).
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check, lib
from .hydro import (Context, Grid, InternalMassGravity, PerfectGas, PointMassGravity, Range, UniformGravity)


def _dev(a, shape, device="cuda"):
    if isinstance(a, np.ndarray):
        a = torch.from_numpy(np.ascontiguousarray(a))
    a = a.to(device=device, dtype=torch.float64).contiguous()
    if tuple(a.shape) != tuple(shape):
        raise ValueError(f"expected shape {tuple(shape)}, got {tuple(a.shape)}")
    return a


class HydroSolver:
    def __init__(self, grid: Grid, eos: PerfectGas, *, reconstruction: str = "VanLeer", riemann_solver: str = "HLL",
                 gravity=None, bc=None, bc_order=None, cfl: float = 0.4, nfx: int = 0, fused: bool = True,
                 precision: str = "fp64", pressure_fix: str = "Off", eps_pf: float = 1e-6, arithmetic: str = "reference"):
        if precision not in ("fp64", "fp32"):
            raise ValueError("precision must be 'fp64' or 'fp32' (FP32 storage with FP64 pressure recovery)")
        if pressure_fix not in ("On", "Off"):
            raise ValueError("pressure_fix must be 'On' or 'Off'")
        self.pressure_fix = pressure_fix == "On"  # main.cpp:459
        self.eps_pf = float(eps_pf)
        if precision == "fp32" and (not fused or self.pressure_fix):
            raise ValueError("FP32 storage is a fused-step mode (no pressure fix)")
        if arithmetic not in ("reference", "fast"):
            raise ValueError("arithmetic must be 'reference' or 'fast'")
        if arithmetic == "fast" and (not fused or precision != "fp64" or nfx > 0 or self.pressure_fix):
            raise ValueError("the fast FP64 arithmetic is a fused FP64 mode (no passive scalars, no pressure fix)")
        self.arithmetic = arithmetic
        self.grid = grid
        self.eos = eos
        self.gravity = gravity
        self.cfl = float(cfl)
        self.nfx = nfx
        self.fused = fused
        self.precision = precision
        self.ctx = Context(grid, eos, gravity, bc, bc_order, nfx, reconstruction, riemann_solver, precision)
        if self.fused and self.pressure_fix:  # the fused step recomputes the face states the fix reads
            check(lib().hclpp_set_pressure_fix(self.ctx.handle, 1, self.eps_pf))
        if arithmetic == "fast":
            check(lib().hclpp_set_fast_arithmetic(self.ctx.handle, 1))
        nd = grid.ndim
        z = grid.zeros
        self._cons = [(z(), z(nd), z()), (z(), z(nd), z())]
        if precision == "fp32":
            f = lambda *lead: grid.zeros(*lead, dtype=torch.float32)  # noqa: E731
            self._cons32 = [(f(), f(nd), f()), (f(), f(nd), f())]
            self._fxs32 = [f(nfx), f(nfx)] if nfx else None
        self._cur = 0
        self._fxs = [z(nfx), z(nfx)] if nfx else None  # passive scalars (mass fractions), with the state
        self.u = z(nd)
        self.P = z()
        self.t = 0.0
        self.iter = 0
        self._clock_started = False
        # output schedule of main.cpp:362-375 (frequencies 0: none) and the output counters
        self._sched = (0, 0.0, 0.0)
        self.output_id = self.iter_output_id = self.time_output_id = -1
        self.outputs_record = []  # (iter, t) of the outputs of this run, for the XDMF side file
        self.restarted = False
        if not self.fused:
            self.rho_rec = z(nd, 2)
            self.rhou_rec = z(nd, nd, 2)
            self.E_rec = z(nd, 2)
            self.u_rec = z(nd, nd, 2)
            self.P_rec = z(nd, 2)
            self.fx_rec = z(nfx, nd, 2) if nfx else None

    # ---- state access
    @property
    def rho(self):
        return self._cons[self._cur][0]

    @property
    def rhou(self):
        return self._cons[self._cur][1]

    @property
    def E(self):
        return self._cons[self._cur][2]

    @property
    def fx(self):
        """Passive scalars (mass fractions) of the current state, or None when nfx = 0."""
        return self._fxs[self._cur] if self.nfx else None

    @property
    def fx_new(self):
        return self._fxs[1 - self._cur] if self.nfx else None

    def _r(self, ngeff) -> ctypes.Structure:
        return self.grid.range.with_ghosts(ngeff)._c()

    def update_primitives(self):
        """conv_cons_to_prim over all ghosts (main.cpp:311, :526)."""
        c = self.ctx
        check(lib().hclpp_conv_cons_to_prim(c.handle, ctypes.byref(self._r(self.grid.Ng)), c._chk(self.rho),
                                            c._chk(self.rhou, (self.grid.ndim,)), c._chk(self.E),
                                            c._chk(self.u, (self.grid.ndim,)), c._chk(self.P)))

    # ---- main.cpp:305-311
    def init_from_prim(self, rho, u, P, fx=None, t0: float = 0.0, iter0: int = 0):
        g = self.grid
        nd = g.ndim
        rho_d = _dev(rho, g.shape3)
        self.u.copy_(_dev(u, (nd,) + g.shape3))
        self.P.copy_(_dev(P, g.shape3))
        self._cur = 0
        r, m, E = self._cons[0]
        r.copy_(rho_d)
        if self.nfx:
            self.fx.copy_(_dev(fx, (self.nfx,) + g.shape3))
        c = self.ctx
        check(lib().hclpp_conv_prim_to_cons(c.handle, ctypes.byref(self._r(0)), c._chk(r), c._chk(self.u, (nd,)),
                                            c._chk(self.P), c._chk(m, (nd,)), c._chk(E)))
        check(lib().hclpp_fill_ghosts(c.handle, c._chk(r), c._chk(m, (nd,)), c._chk(E),
                                      c._fx(self.fx, (self.nfx,), "fx")))
        self._refresh_gravity()  # main.cpp:309
        self.update_primitives()
        self.t = float(t0)
        self.iter = int(iter0)
        self._clock_started = False

    # ---- main.cpp:223-238 (read_pdi), then :305-311
    def restart(self, filename: str):
        """Continue from an output file: read t, iter, the output counters and the primitive
        state (rho, u, P and the passive scalars) of its interior, then initialise as
        main.cpp:305-311 does (prim -> cons, boundaries, cons -> prim): E is rebuilt from P."""
        from .io import read_state

        hdr, rho, u, P, fx, _ = read_state(self.grid, filename, self.nfx)
        self.init_from_prim(rho, u, P, fx, t0=hdr.current_time, iter0=hdr.iter)
        self.output_id, self.iter_output_id, self.time_output_id = hdr.output_id, hdr.iter_output_id, hdr.time_output_id
        self.outputs_record = []  # main.cpp:210: a restarted run indexes its own outputs only
        self.restarted = True
        return hdr

    def set_outputs(self, iter_output_frequency: int = 0, time_first_output: float = 0.0, time_output_frequency: float = 0.0):
        """Output schedule of main.cpp:362-375 for the following steps (frequencies 0: none)."""
        self._sched = (int(iter_output_frequency), float(time_first_output), float(time_output_frequency))

    def write_output(self, schedule, record: bool = True) -> str:
        """write_pdi (io.cpp:185-324) of the current state with the current output counters; with
        record, the output also joins the run's XDMF side file (main.cpp:320-322, :536-538; the
        output written before the invalid-internal-energy error, main.cpp:399-425, does not)."""
        from .io import Header, write_state, write_xdmf

        hdr = Header(self.output_id, self.iter_output_id, self.time_output_id, self.iter, self.t, self.eos.adiabatic_index())
        path = write_state(self.grid, self.eos, schedule.directory, schedule.prefix, hdr, self.rho, self.u, self.P, self.E,
                           self.fx if self.nfx else None)
        if record:
            self.outputs_record.append((self.iter, self.t))
            write_xdmf(schedule.directory, schedule.prefix, self.grid.Nx_glob_ng, self.grid.ndim, self.nfx, self.output_id,
                       self.outputs_record)
        return path

    def simulate(self, t_end: float, max_iter: int, schedule=None, t_ini: float = 0.0) -> list:
        """The time loop of main.cpp:313-557 with outputs: an initial output when a frequency is set and
        the run is not a restart, dt clipped to the output times, an output after every step that ends
        at one (the final step always does), and an output before the invalid-internal-energy error.
        Returns the paths written."""
        from .io import OutputSchedule

        schedule = schedule or OutputSchedule()
        written = []
        if not self.restarted and schedule.enabled:  # main.cpp:313-345
            self.iter_output_id += 1
            self.output_id += 1
            written.append(self.write_output(schedule))
        self.set_outputs(schedule.iter_frequency, schedule.first_output(t_ini), schedule.time_frequency)
        if self.fused:
            self._simulate_fused(t_end, max_iter, schedule, written)
        else:
            while True:
                try:
                    done, make_output = self._step_operators(t_end, max_iter, outputs=True)
                except _InvalidInternalEnergy:
                    self.output_id += 1  # main.cpp:399-425: output, then the error
                    written.append(self.write_output(schedule, record=False))
                    raise
                if make_output:
                    self.output_id += 1
                    written.append(self.write_output(schedule))
                if done:
                    break
        return written

    def _simulate_fused(self, t_end, max_iter, schedule, written, check_every: int = 64):
        c = self.ctx
        it, tf, ff = self._sched
        check(lib().hclpp_clock_set_outputs(c.handle, it, tf, ff, self.iter_output_id, self.time_output_id))
        self._clock_start(t_end, max_iter)
        while True:
            it0, cur0 = self.iter, self._cur
            bufs = self._bufs()
            launched = 0
            while True:
                for _ in range(check_every):
                    self._launch_step(bufs[(cur0 + launched) % 2], bufs[(cur0 + launched + 1) % 2])
                    launched += 1
                clk = self.clock()
                if clk.done or clk.status:
                    break
            steps = clk.iter - it0
            self._cur = (cur0 + steps) % 2
            self.t, self.iter = clk.t, clk.iter
            self.iter_output_id, self.time_output_id = clk.iter_output_id, clk.time_output_id
            self._sync_back()
            self.update_primitives()
            if clk.status == _lib.HCLPP_ERR_NEGATIVE_DT:
                raise RuntimeError("Error: the time step is negative")
            if clk.status:
                self.output_id += 1  # main.cpp:399-425
                written.append(self.write_output(schedule, record=False))
                raise RuntimeError(f"Time = {clk.t}, iteration = {clk.iter}: detected invalid volumic internal energy")
            if clk.output_due:
                self.output_id += 1
                written.append(self.write_output(schedule))
            if clk.exiting:
                return
            r, m, E, _ = self._bufs()[self._cur]
            resume = lib().hclpp_clock_resume_f32 if self.precision == "fp32" else lib().hclpp_clock_resume
            check(resume(c.handle, _ptr(r), _ptr(m), _ptr(E)))

    # ---- main.cpp:354-531
    def run(self, t_end: float, max_iter: int, check_every: int = 64) -> int:
        """Advance until t_end or max_iter (as the reference loop); returns the number of steps taken."""
        if self.fused:
            return self._run_fused(t_end, max_iter, check_every)
        n = 0
        while True:
            done, _ = self._step_operators(t_end, max_iter)
            n += 1
            if done:
                return n

    def _bufs(self):
        fp32 = self.precision == "fp32"
        cons = self._cons32 if fp32 else self._cons
        fxs = (self._fxs32 if fp32 else self._fxs) if self.nfx else [None, None]
        return [cons[i] + (fxs[i],) for i in range(2)]

    def _cast(self, src, dst):
        n = src.numel()
        fn = lib().hclpp_cast_f64_to_f32 if src.dtype == torch.float64 else lib().hclpp_cast_f32_to_f64
        check(fn(self.ctx.handle, _ptr(src), _ptr(dst), n))

    def _clock_start(self, t_end, max_iter):
        c = self.ctx
        if self.precision == "fp32":
            for src, dst in zip(self._cons[self._cur], self._cons32[self._cur]):
                self._cast(src, dst)
            if self.nfx:
                self._cast(self._fxs[self._cur], self._fxs32[self._cur])
            r, m, E = self._cons32[self._cur]
            check(lib().hclpp_clock_start_f32(c.handle, self.t, self.iter, float(t_end), self.cfl, int(max_iter), _ptr(r), _ptr(m),
                                              _ptr(E)))
        else:
            r, m, E = self._cons[self._cur]
            check(lib().hclpp_clock_start(c.handle, self.t, self.iter, float(t_end), self.cfl, int(max_iter), _ptr(r), _ptr(m),
                                          _ptr(E)))

    def _launch_step(self, a, b):
        fp32 = self.precision == "fp32"
        if self.nfx:
            step = lib().hclpp_step_fx_f32 if fp32 else lib().hclpp_step_fx
            check(step(self.ctx.handle, _ptr(a[0]), _ptr(a[1]), _ptr(a[2]), _ptr(a[3]), _ptr(b[0]), _ptr(b[1]), _ptr(b[2]),
                       _ptr(b[3])))
        else:
            step = lib().hclpp_step_f32 if fp32 else lib().hclpp_step
            check(step(self.ctx.handle, _ptr(a[0]), _ptr(a[1]), _ptr(a[2]), _ptr(b[0]), _ptr(b[1]), _ptr(b[2])))

    def _sync_back(self):
        if self.precision == "fp32":
            for src, dst in zip(self._cons32[self._cur], self._cons[self._cur]):
                self._cast(src, dst)
            if self.nfx:
                self._cast(self._fxs32[self._cur], self._fxs[self._cur])

    def _run_fused(self, t_end, max_iter, check_every):
        it0 = self.iter
        cur0 = self._cur
        check(lib().hclpp_clock_set_outputs(self.ctx.handle, 0, 0.0, 0.0, -1, -1))  # no output pauses
        self._clock_start(t_end, max_iter)
        bufs = self._bufs()
        launched = 0
        while True:
            batch = check_every
            if max_iter - it0 - launched > 0:
                batch = min(batch, max_iter - it0 - launched)
            for _ in range(batch):
                self._launch_step(bufs[(cur0 + launched) % 2], bufs[(cur0 + launched + 1) % 2])
                launched += 1
            clk = self.clock()
            if clk.done or clk.status:
                break
        steps = clk.iter - it0
        self._cur = (cur0 + steps) % 2
        self.t = clk.t
        self.iter = clk.iter
        self._sync_back()
        self.update_primitives()
        if clk.status:
            raise RuntimeError(f"Time = {clk.t}, iteration = {clk.iter}: detected invalid volumic internal energy")
        return steps

    def start_clock(self, t_end: float, max_iter: int):
        """Start the device clock on the current state (for callers that drive step_async themselves)."""
        check(lib().hclpp_clock_set_outputs(self.ctx.handle, 0, 0.0, 0.0, -1, -1))
        self._clock_start(t_end, max_iter)
        self._launched = 0
        self._cur0 = self._cur
        self._it0 = self.iter

    def step_async(self):
        """Enqueue one fused step (no host synchronisation)."""
        bufs = self._bufs()
        self._launch_step(bufs[(self._cur0 + self._launched) % 2], bufs[(self._cur0 + self._launched + 1) % 2])
        self._launched += 1

    def finish(self):
        """Synchronise with the device clock after step_async calls."""
        clk = self.clock()
        steps = clk.iter - self._it0
        self._cur = (self._cur0 + steps) % 2
        self.t = clk.t
        self.iter = clk.iter
        self._sync_back()
        if clk.status:
            raise RuntimeError(f"Time = {clk.t}, iteration = {clk.iter}: detected invalid volumic internal energy")
        return clk

    def clock(self) -> _lib.Clock:
        out = _lib.Clock()
        check(lib().hclpp_clock_read(self.ctx.handle, ctypes.byref(out)))
        return out

    def _refresh_gravity(self):
        """make_gravity(param, grid, rho) of main.cpp:309/:528: only internal-mass gravity depends on rho."""
        if isinstance(self.gravity, InternalMassGravity):
            self.gravity.refresh(self.ctx, self.rho)

    def _step_operators(self, t_end, max_iter, outputs: bool = False):
        """One iteration of main.cpp:354-531; returns (exit, make_output).  With outputs, dt is
        clipped to the schedule's output times and the output counters advance (:362-375)."""
        g = self.grid
        c = self.ctx
        nd = g.ndim
        L = lib()
        h = c.handle
        inv = ctypes.c_double()
        check(L.hclpp_time_step(h, ctypes.byref(self._r(g.Ng)), c._chk(self.rho), c._chk(self.u, (nd,)),
                                c._chk(self.P), ctypes.byref(inv)))
        dt = self.cfl * inv.value  # main.cpp:355
        exit_ = False
        make_output = False
        it_freq, t_first, t_freq = self._sched if outputs else (0, 0.0, 0.0)
        if it_freq > 0 and (self.iter + 1) >= (self.iter_output_id + 1) * it_freq:  # :362-368
            make_output = True
            self.iter_output_id += 1
        if t_freq > 0:  # :369-375
            next_output = t_first + ((self.time_output_id + 1) * t_freq)
            if (self.t + dt) >= next_output:
                dt = next_output - self.t
                if dt <= 0:
                    raise RuntimeError("Error: the time step is negative")
                make_output = True
                self.time_output_id += 1
        if (self.t + dt) >= t_end:  # :377-381
            dt = t_end - self.t
            exit_ = True
            make_output = True
        if (self.iter + 1) >= max_iter:  # :382-385
            exit_ = True
            make_output = True
        me = ctypes.c_double()
        check(L.hclpp_min_internal_energy(h, ctypes.byref(self._r(0)), c._chk(self.rho), c._chk(self.rhou, (nd,)),
                                          c._chk(self.E), ctypes.byref(me)))
        if np.isnan(me.value) or me.value < 0:  # :397-425
            raise _InvalidInternalEnergy(f"Time = {self.t}, iteration = {self.iter}: detected invalid volumic internal energy")
        nfx = self.nfx
        fx = c._fx(self.fx, (nfx,), "fx")
        check(L.hclpp_hydro_reconstruction(h, _lib.LIMITER[c.limiter], ctypes.byref(self._r(1)), dt / 2,
                                           c._chk(self.rho), c._chk(self.u, (nd,)), c._chk(self.P), fx,
                                           c._chk(self.rho_rec, (nd, 2)), c._chk(self.rhou_rec, (nd, nd, 2)),
                                           c._chk(self.E_rec, (nd, 2)), c._fx(self.fx_rec, (nfx, nd, 2), "fx_rec"),
                                           c._chk(self.P_rec, (nd, 2)), c._chk(self.u_rec, (nd, nd, 2))))
        nxt = self._cons[1 - self._cur]
        check(L.hclpp_godunov(h, _lib.RIEMANN[c.riemann], ctypes.byref(self._r(0)), dt, c._chk(self.rho),
                              c._chk(self.rhou, (nd,)), c._chk(self.E), fx, c._chk(self.rho_rec, (nd, 2)),
                              c._chk(self.rhou_rec, (nd, nd, 2)), c._chk(self.E_rec, (nd, 2)),
                              c._fx(self.fx_rec, (nfx, nd, 2), "fx_rec"), c._chk(nxt[0]), c._chk(nxt[1], (nd,)),
                              c._chk(nxt[2]), c._fx(self.fx_new, (nfx,), "fx_new")))
        if self.pressure_fix:  # :459-475
            check(L.hclpp_pressure_fix(h, ctypes.byref(self._r(0)), dt, self.eps_pf, c._chk(self.rho),
                                       c._chk(self.rhou, (nd,)), c._chk(self.E), c._chk(self.rho_rec, (nd, 2)),
                                       c._chk(self.rhou_rec, (nd, nd, 2)), c._chk(self.E_rec, (nd, 2)), c._chk(nxt[0]),
                                       c._chk(nxt[1], (nd,)), c._chk(nxt[2])))
        check(L.hclpp_fill_ghosts(h, c._chk(nxt[0]), c._chk(nxt[1], (nd,)), c._chk(nxt[2]),
                                  c._fx(self.fx_new, (nfx,), "fx_new")))
        self._cur = 1 - self._cur
        self.update_primitives()
        self._refresh_gravity()  # :528
        self.t += dt
        self.iter += 1
        return exit_, make_output


class _InvalidInternalEnergy(RuntimeError):
    """main.cpp:420-424: the run stops on a NaN or negative internal energy."""


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())
