"""Host-side mirror of the HERACLES++ hydro operator interfaces, over the HIP C-ABI.

The classes and functions keep the reference names, argument meaning and
error behaviour (factories raise on unknown names, the EoS rejects invalid
parameters, ``Range.with_ghosts`` rejects NgEff > Ng), so code written
against the reference operators reads the same here.  Arrays are
``torch.float64`` tensors on the GPU in the reference memory layout
(Kokkos::LayoutLeft): the reference view ``a(i, j, k, c...)`` is the
C-contiguous tensor ``A[c..., k, j, i]``, components first:

  rho, P, E         (nz, ny, nx)
  u, rhou           (ndim, nz, ny, nx)
  fx                (nfx, nz, ny, nx)
  var_rec, P_rec    (ndim, 2, nz, ny, nx)          # [idim, side]
  rho_rec, E_rec    (ndim, 2, nz, ny, nx)
  u_rec, rhou_rec   (ndim, ndim, 2, nz, ny, nx)    # [component, idim, side]
  fx_rec            (nfx, ndim, 2, nz, ny, nx)

with (nx, ny, nz) the ghost-inclusive extents.  Every operator runs a
hand-written HIP kernel; nothing here computes on the CPU.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import check, lib


class DomainError(ValueError):
    """std::domain_error of the reference (invalid EoS parameters)."""


# ------------------------------------------------------------ thermodynamics
class PerfectGas:
    """thermodynamics/perfect_gas.hpp:17-90, perfect_gas.cpp:13-22."""

    def __init__(self, gamma: float, mmw: float):
        if not math.isfinite(gamma) or gamma <= 1:
            raise DomainError("Invalid gamma")
        if not math.isfinite(mmw) or mmw <= 0:
            raise DomainError("Invalid mmw")
        self._gamma = float(gamma)
        self._mmw = float(mmw)

    def adiabatic_index(self) -> float:
        return self._gamma

    def mean_molecular_weight(self) -> float:
        return self._mmw

    @staticmethod
    def is_valid(rho: float, P: float) -> bool:
        return math.isfinite(rho) and rho > 0 and math.isfinite(P) and P > 0


# ------------------------------------------------------------ mesh
class Range:
    """mesh/range.hpp:19-55, range.cpp:23-141 (single rank: Corner_min = 0)."""

    def __init__(self, cmin, cmax, Ng: int, ndim: int, NgEff: int = 0):
        for d in range(3):
            if cmin[d] < 0 or cmax[d] < 0:
                raise RuntimeError("Cmin < 0 || Cmax < 0")
            if cmin[d] > cmax[d]:
                raise RuntimeError("Cmin > Cmax")
        self.Corner_min = list(cmin)
        self.Corner_max = list(cmax)
        self.Ng = Ng
        self.NgEff = NgEff
        self.ndim = ndim
        self.Nghost = [Ng if d < ndim else 0 for d in range(3)]
        # cell (Nc) and face (Nf) bounds with 0, 1, 2 ghost layers (range.cpp:44-81)
        for name in ("Nc_min_0g", "Nc_max_0g", "Nf_min_0g", "Nf_max_0g", "Nc_min_1g", "Nc_max_1g", "Nf_min_1g",
                     "Nf_max_1g", "Nc_min_2g", "Nc_max_2g", "Nf_min_2g", "Nf_max_2g"):
            setattr(self, name, [0, 0, 0])
        for d in range(3):
            e = 1 if d < ndim else 0
            lo = self.Nghost[d]
            hi = self.Nghost[d] + cmax[d] - cmin[d]
            for g in range(3):
                self.__dict__[f"Nc_min_{g}g"][d] = lo - g * e
                self.__dict__[f"Nc_max_{g}g"][d] = hi + g * e
                self.__dict__[f"Nf_min_{g}g"][d] = lo - g * e
                self.__dict__[f"Nf_max_{g}g"][d] = hi + g * e + 1

    def with_ghosts(self, NgEff: int) -> "Range":
        if NgEff > self.Ng:
            raise RuntimeError("NgEff > Ng")
        return Range(self.Corner_min, self.Corner_max, self.Ng, self.ndim, NgEff)

    def no_ghosts(self) -> "Range":
        return self.with_ghosts(0)

    def all_ghosts(self) -> "Range":
        return self.with_ghosts(self.Ng)

    def cell_range(self):
        """range.cpp:120-135: [Ng-NgEff, Ng+N+NgEff) on active dims, [Ng, Ng+N) on inactive dims."""
        b, e = [0, 0, 0], [0, 0, 0]
        for d in range(3):
            eff = self.NgEff if d < self.ndim else 0
            b[d] = self.Nghost[d] - eff
            e[d] = self.Nghost[d] + self.Corner_max[d] - self.Corner_min[d] + eff
        return b, e

    def _c(self) -> _lib.Range:
        b, e = self.cell_range()
        r = _lib.Range()
        for d in range(3):
            r.begin[d] = b[d]
            r.end[d] = e[d]
        return r


def regular_nodes(n: int, ng: int, xmin: float, xmax: float) -> np.ndarray:
    """mesh/grid_type.cpp:23-30 (Regular): x(i) = xmin + (i - ng) * ((xmax - xmin) / n)."""
    d = (xmax - xmin) / n
    return np.array([xmin + ((i - ng) * d) for i in range(n + 2 * ng + 1)], dtype=np.float64)


class Grid:
    """mesh/grid.hpp:23-75 on one rank: extents, node coordinates, cell sizes and centres."""

    def __init__(self, ndim: int, Nx_glob_ng, Ng: int = 2, geom: str = "Cartesian", xmin=(0.0, 0.0, 0.0),
                 xmax=(1.0, 1.0, 1.0), nodes=None):
        if ndim not in (1, 2, 3):
            raise ValueError("ndim must be 1, 2 or 3")
        if geom not in _lib.GEOM:
            raise ValueError(f"unknown geometry {geom}")
        self.ndim = ndim
        self.geom = geom
        self.Ng = Ng
        n = [int(v) for v in list(Nx_glob_ng) + [1] * (3 - len(Nx_glob_ng))]
        for d in range(ndim, 3):
            n[d] = 1
        self.Nx_glob_ng = n
        self.Nghost = [Ng if d < ndim else 0 for d in range(3)]
        self.Nx_local_wg = [n[d] + 2 * self.Nghost[d] for d in range(3)]
        if nodes is None:
            nodes = [regular_nodes(n[d], self.Nghost[d], xmin[d], xmax[d]) for d in range(3)]
        self.x = [np.ascontiguousarray(v, dtype=np.float64) for v in nodes]
        for d in range(3):
            if self.x[d].shape != (self.Nx_local_wg[d] + 1,):
                raise ValueError(f"dim {d}: expected {self.Nx_local_wg[d] + 1} node coordinates")
        self.dx = [self.x[d][1:] - self.x[d][:-1] for d in range(3)]            # grid.cpp:31-35
        self.x_center = [(self.x[d][:-1] + self.x[d][1:]) / 2 for d in range(3)]  # grid.cpp:37-44
        self.range = Range([0, 0, 0], n, Ng, ndim)
        self.shape3 = (self.Nx_local_wg[2], self.Nx_local_wg[1], self.Nx_local_wg[0])
        self._contexts: dict = {}

    @property
    def x0(self):
        return self.x[0]

    @property
    def x1(self):
        return self.x[1]

    @property
    def x2(self):
        return self.x[2]

    def zeros(self, *lead, dtype=torch.float64, device="cuda"):
        return torch.zeros(tuple(lead) + self.shape3, dtype=dtype, device=device)

    def geometry_context(self, bc=None, bc_order=None, nfx: int = 0) -> "Context":
        """The device copy of this grid's tables for the operators that read only the grid and the arrays
        they are given (face reconstruction, integrate, minimum_internal_energy, the ghost fill): a context
        whose equation of state, gravity and scheme fields those operators never read (they hold valid
        neutral values, as the context checks them)."""
        return self.context(PerfectGas(1.4, 1.0), bc=bc, bc_order=bc_order, nfx=nfx)

    def context(self, eos: PerfectGas, gravity=None, bc=None, bc_order=None, nfx: int = 0, limiter="VanLeer",
                riemann="HLL", precision="fp64") -> "Context":
        key = (eos.adiabatic_index(), eos.mean_molecular_weight(), _gravity_key(gravity), tuple(bc or ()),
               tuple(bc_order or ()), nfx, limiter, riemann, precision)
        ctx = self._contexts.get(key)
        if ctx is None:
            ctx = Context(self, eos, gravity, bc, bc_order, nfx, limiter, riemann, precision)
            self._contexts[key] = ctx
        return ctx


# ------------------------------------------------------------ gravity
class UniformGravity:
    """gravity/gravity.hpp:18-41."""

    def __init__(self, g):
        self.g = [float(v) for v in g]

    def __call__(self, i, j, k, d):
        return self.g[d]


class PointMassGravity:
    """gravity/gravity.hpp:45-71, gravity.cpp:44-55 (g(i) = -G M / xc(i)^2 along dim 0)."""

    def __init__(self, central_mass: float, grid: Grid):
        self.central_mass = float(central_mass)
        self.g = [-6.6743015e-11 * self.central_mass / (xc * xc) for xc in grid.x_center[0]]

    def __call__(self, i, j, k, d):
        return self.g[i] if d == 0 else 0.0


class InternalMassGravity:
    """gravity/gravity.hpp:75-103, gravity.cpp:59-128: g(i) = -G M(<xc(i)) / xc(i)^2 along dim 0, with
    M the central mass plus the mass of the shell-averaged density inside xc(i).  The table lives on
    the device and is rebuilt from rho every step (main.cpp:309, :528); `refresh` does it on demand."""

    def __init__(self, central_mass: float, grid: Grid):
        self.central_mass = float(central_mass)
        self.grid = grid

    def refresh(self, ctx: "Context", rho):
        """make_internal_mass_gravity(central_mass, grid, rho): rebuild ctx's table from rho (ghost-filled)."""
        check(lib().hclpp_internal_mass_gravity(ctx.handle, ctx._chk(rho, (), "rho")))

    def table(self, ctx: "Context") -> np.ndarray:
        out = np.zeros(self.grid.Nx_local_wg[0])
        check(lib().hclpp_gravity_table(ctx.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out


def make_uniform_gravity(gx0: float, gx1: float, gx2: float) -> UniformGravity:
    return UniformGravity((gx0, gx1, gx2))


def make_point_mass_gravity(central_mass: float, grid: Grid) -> PointMassGravity:
    return PointMassGravity(central_mass, grid)


def make_internal_mass_gravity(central_mass: float, grid: Grid) -> InternalMassGravity:
    return InternalMassGravity(central_mass, grid)


def _gravity_key(g):
    if g is None:
        return ("Uniform", 0.0, 0.0, 0.0)
    if isinstance(g, PointMassGravity):
        return ("Point_mass", g.central_mass)
    if isinstance(g, InternalMassGravity):
        return ("Internal_mass", g.central_mass)
    return ("Uniform",) + tuple(g.g)


# ------------------------------------------------------------ device context
def _ptr(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


class Context:
    """An hclpp_ctx: grid tables, EoS, gravity, boundaries and scheme choice on the device."""

    def __init__(self, grid: Grid, eos: PerfectGas, gravity=None, bc=None, bc_order=None, nfx=0, limiter="VanLeer",
                 riemann="HLL", precision="fp64"):
        nd = grid.ndim
        bc = list(bc) if bc else ["Periodic"] * (2 * nd)
        if len(bc) != 2 * nd:
            raise ValueError("one boundary condition per face")
        for b in bc:
            if b not in _lib.BC:
                raise RuntimeError(f"Unknown boundary condition : {b}.")
        bc_order = list(bc_order) if bc_order is not None else list(range(2 * nd))
        if limiter not in _lib.LIMITER:
            raise RuntimeError(f"Unknown face reconstruction algorithm: {limiter}.")
        if riemann not in _lib.RIEMANN:
            raise RuntimeError(f"Invalid riemann solver: {riemann}.")
        self.grid, self.eos, self.nfx = grid, eos, nfx
        self.limiter, self.riemann, self.precision = limiter, riemann, precision
        self.bc, self.bc_order = bc, bc_order
        cfg = _lib.Config()
        cfg.ndim = nd
        cfg.geom = _lib.GEOM[grid.geom]
        cfg.limiter = _lib.LIMITER[limiter]
        cfg.riemann = _lib.RIEMANN[riemann]
        if isinstance(gravity, PointMassGravity):
            cfg.gravity = _lib.GRAVITY["Point_mass"]
            cfg.central_mass = gravity.central_mass
        elif isinstance(gravity, InternalMassGravity):
            cfg.gravity = _lib.GRAVITY["Internal_mass"]
            cfg.central_mass = gravity.central_mass
        else:
            cfg.gravity = _lib.GRAVITY["Uniform"]
            g = gravity.g if gravity is not None else (0.0, 0.0, 0.0)
            for d in range(3):
                cfg.g[d] = g[d]
        cfg.nfx = nfx
        cfg.ng = grid.Ng
        for d in range(3):
            cfg.n[d] = grid.Nx_glob_ng[d]
            cfg.x[d] = grid.x[d].ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        for f in range(6):
            cfg.bc[f] = _lib.BC[bc[f]] if f < 2 * nd else 0
            cfg.bc_order[f] = bc_order[f] if f < 2 * nd else 0
        cfg.gamma = eos.adiabatic_index()
        cfg.mmw = eos.mean_molecular_weight()
        cfg.precision = _lib.PRECISION[precision]
        h = ctypes.c_void_p()
        check(lib().hclpp_ctx_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().hclpp_ctx_destroy(h)
            self._h = None

    @property
    def handle(self):
        lib().hclpp_set_stream(self._h, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        return self._h

    # shape checks (the reference asserts equal extents)
    def _chk(self, t, lead=(), name="array", dtype=torch.float64):
        if t is None:
            raise ValueError(f"{name} is None")
        want = tuple(lead) + self.grid.shape3
        if tuple(t.shape) != want or t.dtype != dtype or not t.is_cuda or not t.is_contiguous():
            raise ValueError(f"{name}: expected a contiguous {dtype} cuda tensor of shape {want}, got "
                             f"{tuple(t.shape)} {t.dtype} {t.device}")
        return _ptr(t)

    def _fx(self, t, lead, name):
        if self.nfx == 0 and (t is None or t.numel() == 0):
            return ctypes.c_void_p(0)
        return self._chk(t, lead, name)


# ------------------------------------------------------------ array conversions
def conv_prim_to_cons(range_: Range, eos: PerfectGas, rho, u, P, rhou, E, grid: Grid | None = None):
    """hydro/array_conversion.hpp:32-97."""
    grid = grid or _grid_of(range_)
    c = grid.context(eos)
    nd = grid.ndim
    check(lib().hclpp_conv_prim_to_cons(c.handle, ctypes.byref(range_._c()), c._chk(rho, (), "rho"),
                                        c._chk(u, (nd,), "u"), c._chk(P, (), "P"), c._chk(rhou, (nd,), "rhou"),
                                        c._chk(E, (), "E")))


def conv_cons_to_prim(range_: Range, eos: PerfectGas, rho, rhou, E, u, P, grid: Grid | None = None):
    """hydro/array_conversion.hpp:107-172."""
    grid = grid or _grid_of(range_)
    c = grid.context(eos)
    nd = grid.ndim
    check(lib().hclpp_conv_cons_to_prim(c.handle, ctypes.byref(range_._c()), c._chk(rho, (), "rho"),
                                        c._chk(rhou, (nd,), "rhou"), c._chk(E, (), "E"), c._chk(u, (nd,), "u"),
                                        c._chk(P, (), "P")))


_GRIDS: dict = {}


def _grid_of(range_: Range) -> Grid:
    """The conversions need only extents: a Cartesian grid sized by the range's local domain."""
    n = [range_.Corner_max[d] - range_.Corner_min[d] for d in range(3)]
    key = (range_.ndim, tuple(n), range_.Ng)
    if key not in _GRIDS:
        _GRIDS[key] = Grid(range_.ndim, n, range_.Ng)
    return _GRIDS[key]


# ------------------------------------------------------------ face reconstruction
class IFaceReconstruction:
    """hydro/face_reconstruction.hpp:21-43."""

    def execute(self, range_: Range, grid: Grid, var, var_rec):
        raise NotImplementedError


class LimitedLinearReconstruction(IFaceReconstruction):
    """hydro/limited_linear_reconstruction.hpp:25-66 with one of the slope limiters of slope_limiters.hpp."""

    def __init__(self, slope_limiter: str):
        self.slope_limiter = slope_limiter
        self._lim = _lib.LIMITER[slope_limiter]

    def execute(self, range_: Range, grid: Grid, var, var_rec):
        c = grid.geometry_context()
        check(lib().hclpp_face_reconstruction(c.handle, self._lim, ctypes.byref(range_._c()), c._chk(var, (), "var"),
                                              c._chk(var_rec, (grid.ndim, 2), "var_rec")))


def factory_face_reconstruction(slope: str) -> IFaceReconstruction:
    """hydro/face_reconstruction.cpp:31-50."""
    if slope in _lib.LIMITER:
        return LimitedLinearReconstruction(slope)
    raise RuntimeError(f"Unknown face reconstruction algorithm: {slope}.")


# ------------------------------------------------------------ Hancock predictor
class IExtrapolationReconstruction:
    """hydro/extrapolation_reconstruction.hpp:27-54."""

    def execute(self, range_, grid, gravity, dt_reconstruction, u_rec, P_rec, rho_rec, rhou_rec, E_rec, fx_rec):
        raise NotImplementedError


class HancockExtrapolationReconstruction(IExtrapolationReconstruction):
    """hydro/extrapolation_reconstruction.hpp:56-339."""

    def __init__(self, eos: PerfectGas, bc=None, bc_order=None):
        self.eos = eos

    def execute(self, range_, grid, gravity, dt, u_rec, P_rec, rho_rec, rhou_rec, E_rec, fx_rec):
        nd = grid.ndim
        nfx = 0 if fx_rec is None else fx_rec.shape[0]
        c = grid.context(self.eos, gravity, nfx=nfx)
        check(lib().hclpp_hancock_extrapolation(
            c.handle, ctypes.byref(range_._c()), float(dt), c._chk(u_rec, (nd, nd, 2), "u_rec"),
            c._chk(P_rec, (nd, 2), "P_rec"), c._chk(rho_rec, (nd, 2), "rho_rec"),
            c._chk(rhou_rec, (nd, nd, 2), "rhou_rec"), c._chk(E_rec, (nd, 2), "E_rec"),
            c._fx(fx_rec, (nfx, nd, 2), "fx_rec")))


class IHydroReconstruction:
    """hydro/hydro_reconstruction.hpp:30-59."""

    def execute(self, range_, grid, gravity, dt, rho, u, P, fx, rho_rec, rhou_rec, E_rec, fx_rec):
        raise NotImplementedError


class MUSCLHancockHydroReconstruction(IHydroReconstruction):
    """hydro/hydro_reconstruction.hpp:61-130: P_rec and u_rec are the operator's own scratch views.

    execute composes the injected strategies as the reference does: the face reconstruction of rho, each
    velocity component, P and each passive scalar, the prim -> cons conversion of every face state, then the
    extrapolation step.  When both strategies are the library's own (LimitedLinearReconstruction and
    HancockExtrapolationReconstruction with this operator's EoS), the same sequence runs as one C-ABI call
    (hclpp_hydro_reconstruction), which launches the same kernels."""

    def __init__(self, face_reconstruction: IFaceReconstruction, hancock_reconstruction: IExtrapolationReconstruction,
                 eos: PerfectGas, P_rec, u_rec):
        self.face_reconstruction = face_reconstruction
        self.hancock_reconstruction = hancock_reconstruction
        self.eos = eos
        self.P_rec = P_rec
        self.u_rec = u_rec

    def _library_strategies(self) -> bool:
        h = self.hancock_reconstruction
        return (type(self.face_reconstruction) is LimitedLinearReconstruction
                and type(h) is HancockExtrapolationReconstruction
                and h.eos.adiabatic_index() == self.eos.adiabatic_index()
                and h.eos.mean_molecular_weight() == self.eos.mean_molecular_weight())

    def execute(self, range_, grid, gravity, dt, rho, u, P, fx, rho_rec, rhou_rec, E_rec, fx_rec):
        nd = grid.ndim
        nfx = 0 if fx is None else fx.shape[0]
        c = grid.context(self.eos, gravity, nfx=nfx)
        if not self._library_strategies():
            face = self.face_reconstruction
            face.execute(range_, grid, rho, rho_rec)  # hydro_reconstruction.hpp:99-109
            for d in range(nd):
                face.execute(range_, grid, u[d], self.u_rec[d])
            face.execute(range_, grid, P, self.P_rec)
            for f in range(nfx):
                face.execute(range_, grid, fx[f], fx_rec[f])
            # :111-126, conv_prim_to_cons of every face state
            check(lib().hclpp_face_prim_to_cons(
                c.handle, ctypes.byref(range_._c()), c._chk(rho_rec, (nd, 2), "rho_rec"),
                c._chk(self.u_rec, (nd, nd, 2), "u_rec"), c._chk(self.P_rec, (nd, 2), "P_rec"),
                c._chk(rhou_rec, (nd, nd, 2), "rhou_rec"), c._chk(E_rec, (nd, 2), "E_rec")))
            # :128
            self.hancock_reconstruction.execute(range_, grid, gravity, dt, self.u_rec, self.P_rec, rho_rec, rhou_rec,
                                                E_rec, fx_rec)
            return
        check(lib().hclpp_hydro_reconstruction(
            c.handle, self.face_reconstruction._lim, ctypes.byref(range_._c()), float(dt), c._chk(rho, (), "rho"),
            c._chk(u, (nd,), "u"), c._chk(P, (), "P"), c._fx(fx, (nfx,), "fx"), c._chk(rho_rec, (nd, 2), "rho_rec"),
            c._chk(rhou_rec, (nd, nd, 2), "rhou_rec"), c._chk(E_rec, (nd, 2), "E_rec"),
            c._fx(fx_rec, (nfx, nd, 2), "fx_rec"), c._chk(self.P_rec, (nd, 2), "P_rec"),
            c._chk(self.u_rec, (nd, nd, 2), "u_rec")))


# ------------------------------------------------------------ Godunov update
class IGodunovScheme:
    """hydro/godunov_scheme.hpp:31-64."""

    def execute(self, range_, grid, gravity, dt, rho, rhou, E, fx, rho_rec, rhou_rec, E_rec, fx_rec, rho_new, rhou_new,
                E_new, fx_new):
        raise NotImplementedError


class RiemannBasedGodunovScheme(IGodunovScheme):
    """hydro/godunov_scheme.hpp:66-303 with the HLL, HLLC or "Low Mach" solver of riemann_solver.hpp."""

    def __init__(self, riemann_solver: str, eos: PerfectGas):
        self.riemann_solver = riemann_solver
        self._rs = _lib.RIEMANN[riemann_solver]
        self.eos = eos

    def execute(self, range_, grid, gravity, dt, rho, rhou, E, fx, rho_rec, rhou_rec, E_rec, fx_rec, rho_new, rhou_new,
                E_new, fx_new):
        nd = grid.ndim
        nfx = 0 if fx is None else fx.shape[0]
        c = grid.context(self.eos, gravity, nfx=nfx)
        check(lib().hclpp_godunov(
            c.handle, self._rs, ctypes.byref(range_._c()), float(dt), c._chk(rho, (), "rho"),
            c._chk(rhou, (nd,), "rhou"), c._chk(E, (), "E"), c._fx(fx, (nfx,), "fx"),
            c._chk(rho_rec, (nd, 2), "rho_rec"), c._chk(rhou_rec, (nd, nd, 2), "rhou_rec"),
            c._chk(E_rec, (nd, 2), "E_rec"), c._fx(fx_rec, (nfx, nd, 2), "fx_rec"), c._chk(rho_new, (), "rho_new"),
            c._chk(rhou_new, (nd,), "rhou_new"), c._chk(E_new, (), "E_new"), c._fx(fx_new, (nfx,), "fx_new")))


def factory_godunov_scheme(riemann_solver: str, eos: PerfectGas) -> IGodunovScheme:
    """hydro/godunov_scheme.hpp:305-321."""
    if riemann_solver in _lib.RIEMANN:
        return RiemannBasedGodunovScheme(riemann_solver, eos)
    raise RuntimeError(f"Invalid riemann solver: {riemann_solver}.")


# ------------------------------------------------------------ reductions
def time_step(range_: Range, eos: PerfectGas, grid: Grid, rho, u, P) -> float:
    """hydro/time_step.hpp:35-82: 1 / max over the range of sum_d (|u_d| + c) / dx_geom_d."""
    c = grid.context(eos)
    out = ctypes.c_double()
    check(lib().hclpp_time_step(c.handle, ctypes.byref(range_._c()), c._chk(rho, (), "rho"),
                                c._chk(u, (grid.ndim,), "u"), c._chk(P, (), "P"), ctypes.byref(out)))
    return out.value


def minimum_internal_energy(range_: Range, grid: Grid, rho, rhou, E) -> float:
    """analysis/internal_energy.cpp:20-45."""
    c = grid.geometry_context()
    out = ctypes.c_double()
    check(lib().hclpp_min_internal_energy(c.handle, ctypes.byref(range_._c()), c._chk(rho, (), "rho"),
                                          c._chk(rhou, (grid.ndim,), "rhou"), c._chk(E, (), "E"), ctypes.byref(out)))
    return out.value


def integrate(range_: Range, grid: Grid, var) -> float:
    """analysis/integration.cpp:22-37."""
    c = grid.geometry_context()
    out = ctypes.c_double()
    check(lib().hclpp_integrate(c.handle, ctypes.byref(range_._c()), c._chk(var, (), "var"), ctypes.byref(out)))
    return out.value


# ------------------------------------------------------------ boundaries
@dataclass
class DistributedBoundaryCondition:
    """boundary_distribute.hpp:28-77 on one rank: periodic self-exchange, then border BCs in priority order."""

    grid: Grid
    bc: list
    bc_order: list | None = None

    def __call__(self, rho, rhou, E, fx=None):
        nfx = 0 if fx is None else fx.shape[0]
        c = self.grid.geometry_context(bc=self.bc, bc_order=self.bc_order, nfx=nfx)
        nd = self.grid.ndim
        check(lib().hclpp_fill_ghosts(c.handle, c._chk(rho, (), "rho"), c._chk(rhou, (nd,), "rhou"),
                                      c._chk(E, (), "E"), c._fx(fx, (nfx,), "fx")))


# ------------------------------------------------------------ pressure fix
def pressure_fix(range_: Range, eos: PerfectGas, grid: Grid, dt: float, eps: float, rho, rhou, E, rho_rec, rhou_rec, E_rec,
                 rho_new, rhou_new, E_new, gravity=None, bc=None):
    """hydro/pressure_fix.hpp:19-172: E_new is updated in place (range no_ghosts in main.cpp:459-475)."""
    c = grid.context(eos, gravity, bc)
    nd = grid.ndim
    check(lib().hclpp_pressure_fix(c.handle, ctypes.byref(range_._c()), float(dt), float(eps), c._chk(rho, (), "rho"),
                                   c._chk(rhou, (nd,), "rhou"), c._chk(E, (), "E"), c._chk(rho_rec, (nd, 2), "rho_rec"),
                                   c._chk(rhou_rec, (nd, nd, 2), "rhou_rec"), c._chk(E_rec, (nd, 2), "E_rec"),
                                   c._chk(rho_new, (), "rho_new"), c._chk(rhou_new, (nd,), "rhou_new"),
                                   c._chk(E_new, (), "E_new")))
