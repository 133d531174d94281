"""Shared helpers: build the oracle and the GPU solver from one setup description."""

from __future__ import annotations

import numpy as np

from heraclespp_amd import setups
from heraclespp_amd.hydro import Grid, PerfectGas, UniformGravity
from oracle import oracle as orc


def make_oracle(s: setups.Setup, **kw) -> orc.Oracle:
    o = orc.Oracle(s.ndim, s.n, geom=s.geom, limiter=kw.get("limiter", s.reconstruction),
                   riemann=kw.get("riemann", s.riemann_solver), gamma=s.gamma, bc=s.bc, xmin=s.xmin, xmax=s.xmax,
                   g=s.gravity, gravity=kw.get("gravity", "Uniform"), central_mass=kw.get("central_mass", 1.0),
                   nfx=kw.get("nfx", 0), bc_order=kw.get("bc_order"))
    return o


def make_solver(s: setups.Setup, fused=True, precision="fp64", **kw):
    from heraclespp_amd.hydro import PointMassGravity
    from heraclespp_amd.solver import HydroSolver

    grid = Grid(s.ndim, s.n, s.ng, s.geom, s.xmin, s.xmax)
    if kw.get("gravity") == "Point_mass":
        g = PointMassGravity(kw.get("central_mass", 1.0), grid)
    else:
        g = UniformGravity(s.gravity)
    return HydroSolver(grid, PerfectGas(s.gamma, 1.0), reconstruction=kw.get("limiter", s.reconstruction),
                       riemann_solver=kw.get("riemann", s.riemann_solver), gravity=g, bc=s.bc, cfl=s.cfl, fused=fused,
                       nfx=kw.get("nfx", 0), bc_order=kw.get("bc_order"), precision=precision)


def run_oracle(o: orc.Oracle, s: setups.Setup, nsteps: int, t_end=None):
    t_end = s.t_end if t_end is None else t_end
    o.init_from_prim(s.rho, s.u, s.P)
    for _ in range(nsteps):
        _, ex = o.step(s.cfl, t_end, nsteps)
        if ex:
            break
    return o


def rel_l1(a: np.ndarray, b: np.ndarray) -> float:
    den = np.sum(np.abs(b))
    return float(np.sum(np.abs(a - b)) / den) if den > 0 else float(np.sum(np.abs(a - b)))


def interior(s: setups.Setup, a: np.ndarray) -> np.ndarray:
    return a[(Ellipsis,) + s.interior()]
