"""Initial conditions of the five benchmark configurations, at any resolution.

Each function returns a ``Setup``: the grid and scheme parameters and the
primitive state (rho, u, P) as numpy arrays over the ghost-inclusive extents
(ghost values are zero; the driver fills them with the boundary conditions).
Random perturbations use numpy's counter-based Philox generator with an
explicit seed, so the same arrays can be fed to any implementation.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .hydro import regular_nodes


@dataclass
class Setup:
    name: str
    ndim: int
    n: list
    geom: str
    xmin: tuple
    xmax: tuple
    reconstruction: str
    riemann_solver: str
    bc: list
    gamma: float
    cfl: float
    t_end: float
    max_iter: int
    gravity: tuple = (0.0, 0.0, 0.0)
    ng: int = 2
    rho: np.ndarray = field(default=None, repr=False)
    u: np.ndarray = field(default=None, repr=False)
    P: np.ndarray = field(default=None, repr=False)

    def nodes(self):
        return [regular_nodes(self.n[d] if d < self.ndim else 1, self.ng if d < self.ndim else 0, self.xmin[d],
                              self.xmax[d]) for d in range(3)]

    @property
    def shape3(self):
        nwg = [(self.n[d] + 2 * self.ng) if d < self.ndim else 1 for d in range(3)]
        return (nwg[2], nwg[1], nwg[0])

    def interior(self):
        ng = self.ng
        sl = [slice(ng, ng + self.n[d]) if d < self.ndim else slice(0, 1) for d in range(3)]
        return (sl[2], sl[1], sl[0])


def _centers(s: Setup):
    x = s.nodes()
    return [(v[:-1] + v[1:]) / 2 for v in x]


def _alloc(s: Setup):
    s.rho = np.zeros(s.shape3)
    s.u = np.zeros((s.ndim,) + s.shape3)
    s.P = np.zeros(s.shape3)


def sod(n: int = 1000) -> Setup:
    """Sod shock tube: 1D, [0,1], (1,0,1) | (0.125,0,0.1) at x = 0.5, Minmod + HLL, CFL 0.4, t_end 0.2."""
    s = Setup("sod", 1, [n, 1, 1], "Cartesian", (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), "Minmod", "HLL",
              ["NullGradient", "NullGradient"], 1.4, 0.4, 0.2, 1 << 30)
    _alloc(s)
    xc = _centers(s)[0]
    left = xc <= 0.5  # shock_tube/setup.hpp:78
    s.rho[0, 0, :] = np.where(left, 1.0, 0.125)
    s.P[0, 0, :] = np.where(left, 1.0, 0.1)
    _zero_ghosts(s)
    return s


def kelvin_helmholtz(n: int = 2048, seed: int = 0, steps: int = 2000) -> Setup:
    """2D periodic [0,1]^2: rho 2 in |y-0.5| < 0.25 else 1, u_x = +-0.5, p = 2.5,
    u_y = 0.01 U(-1,1) sin(4 pi x); VanLeer + HLLC, CFL 0.4."""
    s = Setup("kelvin_helmholtz", 2, [n, n, 1], "Cartesian", (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), "VanLeer", "HLLC",
              ["Periodic"] * 4, 1.4, 0.4, 1e30, steps)
    _alloc(s)
    xc, yc, _ = _centers(s)
    X, Y = np.meshgrid(xc, yc)  # (ny, nx)
    band = np.abs(Y - 0.5) < 0.25
    rng = np.random.Generator(np.random.Philox(seed))
    noise = rng.uniform(-1.0, 1.0, size=X.shape)
    s.rho[0] = np.where(band, 2.0, 1.0)
    s.u[0, 0] = np.where(band, 0.5, -0.5)
    s.u[1, 0] = 0.01 * noise * np.sin(4 * np.pi * X)
    s.P[0] = 2.5
    _zero_ghosts(s)
    return s


def rayleigh_taylor(n: int = 512, seed: int = 0, steps: int = 500, g: float = -0.1) -> Setup:
    """3D [0,1]^3: rho 2 above z = 0.5, 1 below, hydrostatic P with P(0.5) = 2.5 under g_z = -0.1,
    u_z = 0.01 U(-1,1) within two cells of the interface; periodic x, y, reflexive z; Minmod + HLLC."""
    s = Setup("rayleigh_taylor", 3, [n, n, n], "Cartesian", (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), "Minmod", "HLLC",
              ["Periodic"] * 4 + ["Reflexive"] * 2, 1.4, 0.4, 1e30, steps, gravity=(0.0, 0.0, g))
    _alloc(s)
    xc, yc, zc = _centers(s)
    dz = 1.0 / n
    heavy = zc > 0.5
    rho_z = np.where(heavy, 2.0, 1.0)
    P_z = 2.5 + rho_z * g * (zc - 0.5)  # dP/dz = rho g on each side of the interface
    s.rho[:] = rho_z[:, None, None]
    s.P[:] = P_z[:, None, None]
    band = np.abs(zc - 0.5) < 2 * dz
    rng = np.random.Generator(np.random.Philox(seed))
    kb = np.nonzero(band)[0]
    s.u[2][kb] = 0.01 * rng.uniform(-1.0, 1.0, size=(len(kb), s.shape3[1], s.shape3[2]))
    _zero_ghosts(s)
    return s


def sedov(nr: int = 512, nth: int = 128, nph: int = 256, t_end: float = 1.0, E0: float = 1.0,
          gamma: float = 1.4) -> Setup:
    """3D spherical blast: r in [0.01, 1.2], theta in [0, pi], phi in [0, 2 pi); rho 1, p 1e-5; energy E0
    spread over the two innermost radial shells; HLL + VanLeer; reflexive r_min and poles, free r_max,
    periodic phi."""
    s = Setup("sedov", 3, [nr, nth, nph], "Spherical", (0.01, 0.0, 0.0), (1.2, np.pi, 2 * np.pi), "VanLeer", "HLL",
              ["Reflexive", "NullGradient", "Reflexive", "Reflexive", "Periodic", "Periodic"], gamma, 0.4, t_end,
              1 << 30)
    _alloc(s)
    x = s.nodes()
    r = x[0]
    ng = s.ng
    # volume of the two innermost interior shells over the whole sphere
    vol = 4.0 / 3.0 * np.pi * (r[ng + 2] ** 3 - r[ng] ** 3)
    s.rho[:] = 1.0
    s.P[:] = 1e-5
    s.P[:, :, ng:ng + 2] = (gamma - 1) * E0 / vol
    _zero_ghosts(s)
    return s


def smooth_euler(n: int = 1024, seed: int = 0, modes: int = 16, kmax: int = 8, amp: float = 0.2,
                 steps: int = 400, ux: float = 1.0, uy: float = 0.5, p: float = 1.0) -> Setup:
    """2D periodic [0,1]^2 smooth flow: rho = 1 + amp sum_m a_m cos(2 pi k_m.x + phi_m) with seeded integer
    wave vectors |k| <= kmax, sum |a_m| = 1; u = (1, 0.5), p = 1; VanAlbada + HLLC."""
    s = Setup("smooth_euler", 2, [n, n, 1], "Cartesian", (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), "VanAlbada", "HLLC",
              ["Periodic"] * 4, 1.4, 0.4, 1e30, steps)
    _alloc(s)
    xc, yc, _ = _centers(s)
    X, Y = np.meshgrid(xc, yc)
    s.rho[0] = smooth_density(X, Y, seed, modes, kmax, amp)
    s.u[0, 0] = ux
    s.u[1, 0] = uy
    s.P[0] = p
    _zero_ghosts(s)
    return s


def smooth_density(X, Y, seed, modes=16, kmax=8, amp=0.2):
    rng = np.random.Generator(np.random.Philox(seed))
    ks = []
    while len(ks) < modes:
        k = rng.integers(-kmax, kmax + 1, size=2)
        if 0 < k[0] ** 2 + k[1] ** 2 <= kmax ** 2:
            ks.append(k)
    a = rng.uniform(0.0, 1.0, size=modes)
    a = a / np.sum(np.abs(a))
    ph = rng.uniform(0.0, 2 * np.pi, size=modes)
    rho = np.ones_like(X)
    for m in range(modes):
        rho = rho + amp * a[m] * np.cos(2 * np.pi * (ks[m][0] * X + ks[m][1] * Y) + ph[m])
    return rho


def _zero_ghosts(s: Setup):
    mask = np.ones(s.shape3, dtype=bool)
    mask[s.interior()] = False
    s.rho[mask] = 0.0
    s.P[mask] = 0.0
    for d in range(s.ndim):
        s.u[d][mask] = 0.0


CONFIGS = {
    "sod": sod,
    "kelvin_helmholtz": kelvin_helmholtz,
    "rayleigh_taylor": rayleigh_taylor,
    "sedov": sedov,
    "smooth_euler": smooth_euler,
}
