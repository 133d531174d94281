"""ctypes front-end of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP product.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
module.  The product (heraclespp_amd) never imports it and has no CPU
fallback.

Array convention: numpy arrays in C order with reversed axes, i.e. a
Kokkos::LayoutLeft view ``a(i, j, k, c)`` of the reference is the numpy array
``A[c, k, j, i]``; component axes come first, x is the last (fastest) axis.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

LIMITERS = {"Constant": 0, "VanLeer": 1, "Minmod": 2, "VanAlbada": 3}
SOLVERS = {"HLL": 0, "HLLC": 1, "Low Mach": 2}
GEOMS = {"Cartesian": 0, "Spherical": 1}
GRAVITIES = {"Uniform": 0, "Point_mass": 1, "Internal_mass": 2}
BCS = {"Periodic": 0, "NullGradient": 1, "Reflexive": 2}


class _Cfg(ctypes.Structure):
    _fields_ = [
        ("ndim", ctypes.c_int),
        ("geom", ctypes.c_int),
        ("limiter", ctypes.c_int),
        ("riemann", ctypes.c_int),
        ("gravity", ctypes.c_int),
        ("nfx", ctypes.c_int),
        ("ng", ctypes.c_int),
        ("n", ctypes.c_int * 3),
        ("bc", ctypes.c_int * 6),
        ("bc_order", ctypes.c_int * 6),
        ("gamma", ctypes.c_double),
        ("mmw", ctypes.c_double),
        ("g", ctypes.c_double * 3),
        ("central_mass", ctypes.c_double),
        ("x", ctypes.POINTER(ctypes.c_double) * 3),
        ("pressure_fix", ctypes.c_int),
        ("eps_pf", ctypes.c_double),
    ]


def build() -> str:
    """Compile liboracle.so with the committed Makefile (gcc only)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        vp = ctypes.c_void_p
        L.or_create.restype = vp
        L.or_create.argtypes = [ctypes.POINTER(_Cfg)]
        L.or_destroy.argtypes = [vp]
        L.or_last_error.restype = ctypes.c_char_p
        for name in ("or_evol_from_pres", "or_pres_from_evol"):
            f = getattr(L, name)
            f.restype = ctypes.c_double
            f.argtypes = [ctypes.c_double] * 4
        L.or_speed_of_sound.restype = ctypes.c_double
        L.or_speed_of_sound.argtypes = [ctypes.c_double] * 3
        for name in ("or_ek_cons", "or_ek_prim"):
            f = getattr(L, name)
            f.restype = ctypes.c_double
            f.argtypes = [ctypes.c_int, dp]
        for name in ("or_to_prim", "or_to_cons"):
            getattr(L, name).argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, dp, dp]
        for name in ("or_flux_prim", "or_flux_cons"):
            getattr(L, name).argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, dp, ctypes.c_int, dp]
        L.or_riemann.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, dp, dp, ctypes.c_int, dp]
        L.or_limiter.restype = ctypes.c_double
        L.or_limiter.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.or_ncells_wg.restype = ctypes.c_int64
        L.or_ncells_wg.argtypes = [vp]
        L.or_extents.argtypes = [vp, ip]
        L.or_geometry.restype = dp
        L.or_geometry.argtypes = [vp, ctypes.c_int]
        L.or_range.argtypes = [vp, ctypes.c_int, ip]
        L.or_conv_prim_to_cons.argtypes = [vp, ip, dp, dp, dp, dp, dp]
        L.or_conv_cons_to_prim.argtypes = [vp, ip, dp, dp, dp, dp, dp]
        L.or_face_reconstruction.argtypes = [vp, ip, dp, dp]
        L.or_hancock.argtypes = [vp, ip, ctypes.c_double, dp, dp, dp, dp, dp, dp]
        L.or_hydro_reconstruction.argtypes = [vp, ip, ctypes.c_double] + [dp] * 10
        L.or_godunov.argtypes = [vp, ip, ctypes.c_double] + [dp] * 12
        for name in ("or_time_step", "or_integrate"):
            getattr(L, name).restype = ctypes.c_double
        L.or_time_step.argtypes = [vp, ip, dp, dp, dp]
        L.or_integrate.argtypes = [vp, ip, dp]
        L.or_min_internal_energy.restype = ctypes.c_double
        L.or_min_internal_energy.argtypes = [vp, ip, dp, dp, dp]
        L.or_boundaries.argtypes = [vp, dp, dp, dp, dp]
        L.or_pressure_fix.argtypes = [vp, ip, ctypes.c_double, ctypes.c_double] + [dp] * 9
        L.or_internal_mass_gravity.argtypes = [vp, dp]
        L.or_gravity_table.restype = dp
        L.or_gravity_table.argtypes = [vp]
        L.or_init_from_prim.restype = ctypes.c_int
        L.or_init_from_prim.argtypes = [vp, dp, dp, dp, dp]
        L.or_step.restype = ctypes.c_int
        L.or_step.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_int, dp, ip]
        L.or_step_outputs.restype = ctypes.c_int
        L.or_step_outputs.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_int, dp, ip, ip]
        L.or_set_outputs.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.or_output_ids.argtypes = [vp, ip, ip]
        L.or_state.restype = dp
        L.or_state.argtypes = [vp, ctypes.c_int]
        L.or_time.restype = ctypes.c_double
        L.or_time.argtypes = [vp]
        L.or_iter.restype = ctypes.c_int
        L.or_iter.argtypes = [vp]
        L.or_set_time.argtypes = [vp, ctypes.c_double, ctypes.c_int]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _pv(v):
    a = np.ascontiguousarray(v, dtype=np.float64)
    return a, _p(a)


# ---------------------------------------------------------------- scalars
def evol_from_pres(gamma, mmw, rho, P):
    return lib().or_evol_from_pres(gamma, mmw, rho, P)


def pres_from_evol(gamma, mmw, rho, evol):
    return lib().or_pres_from_evol(gamma, mmw, rho, evol)


def speed_of_sound(gamma, rho, P):
    return lib().or_speed_of_sound(gamma, rho, P)


def ek_cons(cons):
    a, p = _pv(cons)
    return lib().or_ek_cons(len(a) - 2, p)


def ek_prim(prim):
    a, p = _pv(prim)
    return lib().or_ek_prim(len(a) - 2, p)


def _vec1(fn, v, *args):
    a, p = _pv(v)
    out = np.zeros_like(a)
    fn(len(a) - 2, *args[:2], p, *args[2:], _p(out))
    return out


def to_prim(cons, gamma, mmw=1.0):
    return _vec1(lib().or_to_prim, cons, gamma, mmw)


def to_cons(prim, gamma, mmw=1.0):
    return _vec1(lib().or_to_cons, prim, gamma, mmw)


def flux_prim(prim, locdim, gamma, mmw=1.0):
    return _vec1(lib().or_flux_prim, prim, gamma, mmw, locdim)


def flux_cons(cons, locdim, gamma, mmw=1.0):
    return _vec1(lib().or_flux_cons, cons, gamma, mmw, locdim)


def riemann(solver: str, consL, consR, locdim, gamma, mmw=1.0):
    a, pa = _pv(consL)
    b, pb = _pv(consR)
    out = np.zeros_like(a)
    lib().or_riemann(SOLVERS[solver], len(a) - 2, gamma, mmw, pa, pb, locdim, _p(out))
    return out


def limiter(name: str, diffR, diffL):
    return lib().or_limiter(LIMITERS[name], float(diffR), float(diffL))


# ---------------------------------------------------------------- grid
def regular_nodes(n: int, ng: int, xmin: float, xmax: float) -> np.ndarray:
    """mesh/grid_type.cpp:23-30: x(i) = xmin + (i - ng) * ((xmax - xmin) / n), n + 2 ng + 1 nodes."""
    d = (xmax - xmin) / n
    return np.array([xmin + ((i - ng) * d) for i in range(n + 2 * ng + 1)], dtype=np.float64)


class Oracle:
    """One oracle context: grid geometry plus the reference driver's arrays."""

    def __init__(self, ndim, n, *, geom="Cartesian", limiter="Minmod", riemann="HLL", gamma=1.4, mmw=1.0,
                 gravity="Uniform", g=(0.0, 0.0, 0.0), central_mass=1.0, nfx=0, ng=2,
                 xmin=(0.0, 0.0, 0.0), xmax=(1.0, 1.0, 1.0), nodes=None, bc=None, bc_order=None,
                 pressure_fix=False, eps_pf=1e-6):
        self.ndim = ndim
        self.n = [int(v) for v in (list(n) + [1, 1, 1])[:3]]
        for d in range(ndim, 3):
            self.n[d] = 1
        self.ng = ng
        self.nfx = nfx
        if nodes is None:
            nodes = [regular_nodes(self.n[d], ng if d < ndim else 0, xmin[d], xmax[d]) for d in range(3)]
        self._nodes = [np.ascontiguousarray(x, dtype=np.float64) for x in nodes]
        cfg = _Cfg()
        cfg.ndim = ndim
        cfg.geom = GEOMS[geom]
        cfg.limiter = LIMITERS[limiter]
        cfg.riemann = SOLVERS[riemann]
        cfg.gravity = GRAVITIES[gravity]
        cfg.nfx = nfx
        cfg.ng = ng
        for d in range(3):
            cfg.n[d] = self.n[d]
            cfg.g[d] = g[d]
            cfg.x[d] = _p(self._nodes[d])
        bc = bc or ["Periodic"] * (2 * ndim)
        for f in range(6):
            cfg.bc[f] = BCS[bc[f]] if f < len(bc) else 0
            cfg.bc_order[f] = (bc_order[f] if bc_order is not None else f) if f < 2 * ndim else 0
        cfg.gamma = gamma
        cfg.mmw = mmw
        cfg.central_mass = central_mass
        cfg.pressure_fix = 1 if pressure_fix else 0
        cfg.eps_pf = eps_pf
        self._cfg = cfg
        self._L = lib()
        h = self._L.or_create(ctypes.byref(cfg))
        if not h:
            raise RuntimeError(lib().or_last_error().decode())
        self._h = ctypes.c_void_p(h)
        ext = (ctypes.c_int * 3)()
        lib().or_extents(self._h, ext)
        self.nwg = [ext[0], ext[1], ext[2]]
        self.shape3 = (self.nwg[2], self.nwg[1], self.nwg[0])
        self.plane = self.nwg[0] * self.nwg[1] * self.nwg[2]

    def __del__(self):
        h = getattr(self, "_h", None)
        L = getattr(self, "_L", None)  # held by the object: module globals may be gone at exit
        if h and L is not None:
            L.or_destroy(h)
            self._h = None

    # ---- ranges (mesh/range.cpp)
    def range(self, ngeff):
        r = (ctypes.c_int * 6)()
        lib().or_range(self._h, ngeff, r)
        return r

    def zeros(self, *lead):
        return np.zeros(tuple(lead) + self.shape3, dtype=np.float64)

    def geometry(self, which):
        """0..2 dx, 3..5 xc (1D); 6 ds (3, nz, ny, nx); 7 dv (nz, ny, nx)."""
        p = lib().or_geometry(self._h, which)
        if which < 6:
            nn = self.nwg[which % 3]
            return np.ctypeslib.as_array(p, shape=(nn,)).copy()
        if which == 6:
            return np.ctypeslib.as_array(p, shape=(3,) + self.shape3).copy()
        return np.ctypeslib.as_array(p, shape=self.shape3).copy()

    # ---- array kernels on caller arrays
    def conv_prim_to_cons(self, ngeff, rho, u, P, rhou, E):
        lib().or_conv_prim_to_cons(self._h, self.range(ngeff), _p(rho), _p(u), _p(P), _p(rhou), _p(E))

    def conv_cons_to_prim(self, ngeff, rho, rhou, E, u, P):
        lib().or_conv_cons_to_prim(self._h, self.range(ngeff), _p(rho), _p(rhou), _p(E), _p(u), _p(P))

    def face_reconstruction(self, ngeff, var, var_rec):
        lib().or_face_reconstruction(self._h, self.range(ngeff), _p(var), _p(var_rec))

    def hancock(self, ngeff, dt, u_rec, P_rec, rho_rec, rhou_rec, E_rec, fx_rec):
        lib().or_hancock(self._h, self.range(ngeff), dt, _p(u_rec), _p(P_rec), _p(rho_rec), _p(rhou_rec),
                         _p(E_rec), _p(fx_rec))

    def hydro_reconstruction(self, ngeff, dt, rho, u, P, fx, rho_rec, rhou_rec, E_rec, fx_rec, P_rec, u_rec):
        lib().or_hydro_reconstruction(self._h, self.range(ngeff), dt, _p(rho), _p(u), _p(P), _p(fx), _p(rho_rec),
                                      _p(rhou_rec), _p(E_rec), _p(fx_rec), _p(P_rec), _p(u_rec))

    def godunov(self, ngeff, dt, rho, rhou, E, fx, rho_rec, rhou_rec, E_rec, fx_rec, rho_new, rhou_new, E_new, fx_new):
        lib().or_godunov(self._h, self.range(ngeff), dt, _p(rho), _p(rhou), _p(E), _p(fx), _p(rho_rec),
                         _p(rhou_rec), _p(E_rec), _p(fx_rec), _p(rho_new), _p(rhou_new), _p(E_new), _p(fx_new))

    def time_step(self, ngeff, rho, u, P):
        return lib().or_time_step(self._h, self.range(ngeff), _p(rho), _p(u), _p(P))

    def min_internal_energy(self, ngeff, rho, rhou, E):
        return lib().or_min_internal_energy(self._h, self.range(ngeff), _p(rho), _p(rhou), _p(E))

    def integrate(self, ngeff, var):
        return lib().or_integrate(self._h, self.range(ngeff), _p(var))

    def boundaries(self, rho, rhou, E, fx):
        lib().or_boundaries(self._h, _p(rho), _p(rhou), _p(E), _p(fx))

    def pressure_fix(self, ngeff, dt, eps, rho, rhou, E, rho_rec, rhou_rec, E_rec, rho_new, rhou_new, E_new):
        lib().or_pressure_fix(self._h, self.range(ngeff), dt, eps, _p(rho), _p(rhou), _p(E), _p(rho_rec), _p(rhou_rec),
                              _p(E_rec), _p(rho_new), _p(rhou_new), _p(E_new))

    def internal_mass_gravity(self, rho):
        """Rebuild the radial gravity table from rho (ghost-filled); returns a copy of the table."""
        lib().or_internal_mass_gravity(self._h, _p(rho))
        return self.gravity_table()

    def gravity_table(self):
        return np.ctypeslib.as_array(lib().or_gravity_table(self._h), shape=(self.nwg[0],)).copy()

    # ---- driver
    def init_from_prim(self, rho, u, P, fx=None):
        fxa = fx if fx is not None else self.zeros(max(self.nfx, 1))
        lib().or_init_from_prim(self._h, _p(rho), _p(u), _p(P), _p(fxa))

    def step(self, cfl, t_end, max_iter=1 << 30):
        dt = ctypes.c_double()
        ex = ctypes.c_int()
        rc = lib().or_step(self._h, cfl, t_end, max_iter, ctypes.byref(dt), ctypes.byref(ex))
        if rc != 0:
            raise RuntimeError(lib().or_last_error().decode())
        return dt.value, bool(ex.value)

    def set_outputs(self, iter_output_frequency=0, time_first_output=0.0, time_output_frequency=0.0, iter_output_id=-1,
                    time_output_id=-1):
        """Output schedule of main.cpp:362-375 for the following steps."""
        lib().or_set_outputs(self._h, int(iter_output_frequency), float(time_first_output), float(time_output_frequency),
                             int(iter_output_id), int(time_output_id))

    def step_outputs(self, cfl, t_end, max_iter=1 << 30):
        """One step under the output schedule: (dt, exit, make_output)."""
        dt, ex, out = ctypes.c_double(), ctypes.c_int(), ctypes.c_int()
        rc = lib().or_step_outputs(self._h, cfl, t_end, max_iter, ctypes.byref(dt), ctypes.byref(ex), ctypes.byref(out))
        if rc != 0:
            raise RuntimeError(lib().or_last_error().decode())
        return dt.value, bool(ex.value), bool(out.value)

    def output_ids(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        lib().or_output_ids(self._h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def state(self, which):
        """Live view of the driver array: rho, rhou, E, fx, u, P."""
        idx = {"rho": 0, "rhou": 1, "E": 2, "fx": 3, "u": 4, "P": 5}[which]
        ncomp = {"rho": 0, "rhou": self.ndim, "E": 0, "fx": self.nfx, "u": self.ndim, "P": 0}[which]
        p = lib().or_state(self._h, idx)
        shape = ((ncomp,) if ncomp else ()) + self.shape3
        return np.ctypeslib.as_array(p, shape=shape)

    @property
    def t(self):
        return lib().or_time(self._h)

    @property
    def iter(self):
        return lib().or_iter(self._h)

    def set_time(self, t, it):
        lib().or_set_time(self._h, t, it)
