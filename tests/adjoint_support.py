"""Test support for the reverse-mode step: a fixed-dt forward step through the oracle's
operators (the reference kernel sequence of main.cpp:427-479) and ctypes access to the
host build of the adjoint bodies (tests/cpp/libadjoint_host.so)."""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from oracle import oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))
HOST_LIB = os.path.join(HERE, "cpp", "libadjoint_host.so")

_host = None


def build_host() -> str:
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "cpp"), HOST_LIB], check=True)
    return HOST_LIB


def host_lib():
    global _host
    if _host is None:
        if not os.path.exists(HOST_LIB):
            build_host()
        L = ctypes.CDLL(HOST_LIB)
        dp = ctypes.POINTER(ctypes.c_double)
        L.adjh_step.restype = ctypes.c_int
        L.adjh_step.argtypes = ([ctypes.c_int] * 4 + [ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.POINTER(dp),
                                ctypes.POINTER(ctypes.c_int), ctypes.c_double, dp, ctypes.c_double] + [dp] * 9 + [dp])
        _host = L
    return _host


def _p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Problem:
    """A small problem (Cartesian, or spherical on [xmin, xmax]): oracle context, nodes, boundary
    kinds, gravity."""

    def __init__(self, ndim, n, *, limiter, riemann, bc, g=(0.0, 0.0, 0.0), gamma=1.4, xmax=(1.0, 1.0, 1.0),
                 geom="Cartesian", xmin=(0.0, 0.0, 0.0)):
        self.ndim = ndim
        self.n = [int(v) for v in (list(n) + [1, 1, 1])[:3]]
        self.limiter, self.riemann, self.bc, self.g, self.gamma = limiter, riemann, list(bc), tuple(g), gamma
        self.geom, self.xmin, self.xmax = geom, xmin, xmax
        self.o = self.oracle(gamma)
        self.ng = 2
        self.nodes = [orc.regular_nodes(self.n[d], 2 if d < ndim else 0, xmin[d], xmax[d]) for d in range(3)]

    def oracle(self, gamma):
        return orc.Oracle(self.ndim, self.n, geom=self.geom, limiter=self.limiter, riemann=self.riemann, gamma=gamma,
                          bc=self.bc, g=self.g, xmin=self.xmin, xmax=self.xmax)

    def interior(self):
        sl = [slice(None)] * 3
        for d in range(self.ndim):
            sl[2 - d] = slice(2, 2 + self.n[d])
        return tuple(sl)

    def random_state(self, seed, noise=0.05, u0=0.0):
        """Smooth positive state plus seeded noise and a uniform x-velocity u0, ghosts filled by the
        boundary conditions."""
        o = self.o
        rng = np.random.default_rng(seed)
        rho, E, rhou = o.zeros(), o.zeros(), o.zeros(self.ndim)
        zc = np.meshgrid(*[np.arange(o.nwg[2 - a]) for a in range(3)], indexing="ij")
        phase = sum(0.37 * (a + 1) * zc[a] for a in range(3))
        rho[...] = 1.0 + 0.3 * np.sin(phase) + noise * rng.standard_normal(rho.shape)
        u = 0.4 * np.cos(phase)[None] * np.ones((self.ndim,) + rho.shape) + 0.2 * rng.standard_normal((self.ndim,) + rho.shape)
        u[0] += u0
        P = 1.0 + 0.2 * np.cos(1.3 * phase) + noise * rng.standard_normal(rho.shape)
        fx = o.zeros(1)
        o.conv_prim_to_cons(o.ng, rho, np.ascontiguousarray(u), P, rhou, E)
        o.boundaries(rho, rhou, E, fx)
        return rho, rhou, E

    def stable_dt(self, rho, rhou, E, cfl=0.3):
        o = self.o
        u, P = o.zeros(self.ndim), o.zeros()
        o.conv_cons_to_prim(o.ng, rho, rhou, E, u, P)
        return cfl * o.time_step(o.ng, rho, u, P)  # or_time_step returns 1 / max rate

    def step(self, rho, rhou, E, dt, gamma=None):
        """One step with dt given: boundaries, cons->prim, MUSCL-Hancock (dt/2), Godunov (dt)."""
        o = self.o if gamma is None else self.oracle(gamma)
        nd = self.ndim
        rho, rhou, E = rho.copy(), rhou.copy(), E.copy()
        fx = o.zeros(1)
        o.boundaries(rho, rhou, E, fx)
        u, P = o.zeros(nd), o.zeros()
        o.conv_cons_to_prim(o.ng, rho, rhou, E, u, P)
        rho_rec, E_rec, P_rec = o.zeros(nd, 2), o.zeros(nd, 2), o.zeros(nd, 2)
        rhou_rec, u_rec, fx_rec = o.zeros(nd, nd, 2), o.zeros(nd, nd, 2), o.zeros(1, nd, 2)
        o.hydro_reconstruction(1, dt / 2, rho, u, P, fx, rho_rec, rhou_rec, E_rec, fx_rec, P_rec, u_rec)
        rn, mn, en, fn = o.zeros(), o.zeros(nd), o.zeros(), o.zeros(1)
        o.godunov(0, dt, rho, rhou, E, fx, rho_rec, rhou_rec, E_rec, fx_rec, rn, mn, en, fn)
        return rn, mn, en

    def host_adjoint(self, rho, rhou, E, lrho_new, lrhou_new, lE_new, dt):
        """Reverse step on the host build of the GPU bodies; returns (lrho, lrhou, lE, dgamma)."""
        L = host_lib()
        o = self.o
        lr, lm, le = o.zeros(), o.zeros(self.ndim), o.zeros()
        n = (ctypes.c_int * 3)(*self.n)
        xs = (ctypes.POINTER(ctypes.c_double) * 3)(*[_p(x) for x in self.nodes])
        bcs = (ctypes.c_int * 6)(*[orc.BCS[b] for b in self.bc] + [0] * (6 - len(self.bc)))
        g = np.ascontiguousarray(self.g, dtype=np.float64)
        dg = np.zeros(1)
        rc = L.adjh_step(self.ndim, orc.GEOMS[self.geom], orc.SOLVERS[self.riemann], orc.LIMITERS[self.limiter], n, 2, xs, bcs,
                         self.gamma, _p(g),
                         dt, _p(rho), _p(rhou), _p(E), _p(lrho_new), _p(lrhou_new), _p(lE_new), _p(lr), _p(lm), _p(le),
                         _p(dg))
        assert rc == 0
        return lr, lm, le, float(dg[0])
