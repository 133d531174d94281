"""Time-loop driver: the reference main.cpp:305-311 (initialisation) and
:354-531 (loop body) on one GPU.

Two paths, both HIP:
  * fused (default): one hclpp_step (hclpp_step_fx with one passive scalar)
    per iteration -- the fused reconstruct/predict/Riemann/update kernel, the
    ghost fill and the device clock (dt, t_end clipping, max_iter, the
    internal-energy check) -- with no host synchronisation inside the loop;
  * operators: the reference operator sequence (time_step,
    minimum_internal_energy, MUSCL-Hancock reconstruction, Godunov update,
    pressure fix, boundaries, cons->prim), one kernel per reference kernel,
    host-synchronised on dt exactly like the reference.  Needed for more than
    one passive scalar and for the pressure fix, and used to bisect parity.
Time outputs are not modelled (the reference clips dt to output times only
when an output frequency is configured).
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
                 precision: str = "fp64", pressure_fix: str = "Off", eps_pf: float = 1e-6):
        if precision not in ("fp64", "fp32"):
            raise ValueError("precision must be 'fp64' or 'fp32' (FP32 storage with FP64 pressure recovery)")
        if pressure_fix not in ("On", "Off"):
            raise ValueError("pressure_fix must be 'On' or 'Off'")
        self.pressure_fix = pressure_fix == "On"  # main.cpp:459 (operator path: it needs the face states)
        self.eps_pf = float(eps_pf)
        if precision == "fp32" and (not fused or nfx > 1 or self.pressure_fix):
            raise ValueError("FP32 storage is a fused-step mode (at most one passive scalar, no pressure fix)")
        self.grid = grid
        self.eos = eos
        self.gravity = gravity
        self.cfl = float(cfl)
        self.nfx = nfx
        # the fused step carries at most one passive scalar; the operator path takes any nfx
        self.fused = fused and nfx <= 1 and not self.pressure_fix
        self.precision = precision
        self.ctx = Context(grid, eos, gravity, bc, bc_order, nfx, reconstruction, riemann_solver, precision)
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
    def init_from_prim(self, rho, u, P, fx=None, t0: float = 0.0):
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
        self.iter = 0
        self._clock_started = False

    # ---- main.cpp:354-531
    def run(self, t_end: float, max_iter: int, check_every: int = 64) -> int:
        """Advance until t_end or max_iter (as the reference loop); returns the number of steps taken."""
        if self.fused:
            return self._run_fused(t_end, max_iter, check_every)
        n = 0
        while True:
            done = self._step_operators(t_end, max_iter)
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

    def _step_operators(self, t_end, max_iter) -> bool:
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
        if (self.t + dt) >= t_end:  # :377-381
            dt = t_end - self.t
            exit_ = True
        if (self.iter + 1) >= max_iter:  # :382-385
            exit_ = True
        me = ctypes.c_double()
        check(L.hclpp_min_internal_energy(h, ctypes.byref(self._r(0)), c._chk(self.rho), c._chk(self.rhou, (nd,)),
                                          c._chk(self.E), ctypes.byref(me)))
        if np.isnan(me.value) or me.value < 0:  # :397-425
            raise RuntimeError(f"Time = {self.t}, iteration = {self.iter}: detected invalid volumic internal energy")
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
        return exit_


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())
