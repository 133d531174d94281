"""Pins of the oracle's pressure fix (hydro/pressure_fix.hpp) and internal-mass gravity
(gravity/gravity.cpp:59-128) against closed forms.

Pressure fix: for a cell flagged (e < eps ek) the new total energy is
e + alpha dE_godunov + (1 - alpha) dt (-P div u - div(u e)) + ek, alpha = (e / (eps ek))^2
(0 below eps / 10).  With face velocities sampled from a linear field u = (a x + c0, b y + c1)
the discrete divergence is exactly a + b; with a uniform flow U and a varying internal energy
the energy divergence is the first-order upwind difference.  Unflagged cells keep E_new.

Internal-mass gravity: for a density that is constant on each radial cell layer, the enclosed
mass at a cell centre is M + 4 pi / 3 * sum(rho_layer (r_hi^3 - r_lo^3)), so
g = -G M(r) / r^2 has a closed form (uniform and two-layer densities, 1D and 3D spherical
grids, in 3D with a density modulated in phi whose shell average is the layer value)."""

import numpy as np
import pytest

from oracle import oracle as orc

G_NEWTON = 6.6743015e-11


def _cart(nd, n, xmax=(1.0, 1.0, 1.0)):
    return orc.Oracle(nd, n, xmax=xmax, bc=["Periodic"] * (2 * nd))


def _centers(o, d):
    x = orc.regular_nodes(o.n[d], o.ng if d < o.ndim else 0, 0.0, [1.0, 1.0, 1.0][d])
    return (x[:-1] + x[1:]) / 2


def _linear_problem(o, a, b, c0, c1, e0, rho0=1.0):
    """Cell state and MUSCL face states sampled from u = (a x + c0, b y + c1), uniform rho0 and internal energy e0."""
    nd = o.ndim
    xc = _centers(o, 0)
    X = xc[None, None, :] * np.ones(o.shape3)
    dx = 1.0 / o.n[0]
    if nd == 2:
        yc = _centers(o, 1)
        Y = yc[None, :, None] * np.ones(o.shape3)
    rho = np.full(o.shape3, rho0)
    u = np.zeros((nd,) + o.shape3)
    u[0] = a * X + c0
    if nd == 2:
        u[1] = b * Y + c1
    ek = 0.5 * rho * np.sum(u * u, axis=0)
    E = e0 + ek
    rhou = rho[None] * u
    rho_rec, E_rec, rhou_rec = o.zeros(nd, 2), o.zeros(nd, 2), o.zeros(nd, nd, 2)
    for idim in range(nd):
        for s in (0, 1):
            uf = u.copy()
            if idim == 0:
                uf[0] = a * (X + (s - 0.5) * dx) + c0
            else:
                dy = 1.0 / o.n[1]
                uf[1] = b * (Y + (s - 0.5) * dy) + c1
            rho_rec[idim, s] = rho0
            for v in range(nd):
                rhou_rec[v, idim, s] = rho0 * uf[v]
            E_rec[idim, s] = e0 + 0.5 * rho0 * np.sum(uf * uf, axis=0)
    return rho, rhou, E, ek, rho_rec, rhou_rec, E_rec


def _interior(o):
    sl = [slice(None)] * 3
    for d in range(o.ndim):
        sl[2 - d] = slice(o.ng, o.ng + o.n[d])
    return tuple(sl)


@pytest.mark.parametrize("nd", [1, 2])
def test_pressure_fix_linear_velocity_gives_exact_divergence(nd):
    o = _cart(nd, [8, 6][:nd])
    a, b, c0, c1, e0, dt, eps = 0.3, -0.2, 5.0, 4.0, 0.05, 0.01, 0.1
    b = b if nd == 2 else 0.0
    rho, rhou, E, ek, rho_rec, rhou_rec, E_rec = _linear_problem(o, a, b, c0, c1, e0)
    E_new = E.copy()
    o.pressure_fix(0, dt, eps, rho, rhou, E, rho_rec, rhou_rec, E_rec, rho.copy(), rhou.copy(), E_new)
    it = _interior(o)
    assert np.all(e0 < eps / 10 * ek[it])  # alpha = 0 everywhere
    P = (1.4 - 1) * e0
    want = e0 - dt * (a + b) * (P + e0) + ek
    np.testing.assert_allclose(E_new[it], want[it], rtol=0, atol=2e-14)


@pytest.mark.parametrize("U", [3.0, -3.0])
def test_pressure_fix_upwinds_the_internal_energy(U):
    o = _cart(1, [10])
    rho = np.ones(o.shape3)
    e = 0.01 * (1 + 0.3 * np.sin(np.arange(o.nwg[0]) * 0.7))[None, None, :] * np.ones(o.shape3)
    rhou = np.full((1,) + o.shape3, U)
    ek = 0.5 * U * U * np.ones(o.shape3)
    E = e + ek
    rho_rec, E_rec, rhou_rec = o.zeros(1, 2), o.zeros(1, 2), o.zeros(1, 1, 2)
    rho_rec[...] = 1.0
    rhou_rec[...] = U
    E_rec[...] = 0.01 + ek  # face energies only enter through the face velocity, which is U
    dt, eps, dx = 0.02, 0.1, 1.0 / 10
    E_new = E.copy()
    o.pressure_fix(0, dt, eps, rho, rhou, E, rho_rec, rhou_rec, E_rec, rho.copy(), rhou.copy(), E_new)
    it = _interior(o)
    e1 = e[0, 0]
    if U > 0:
        div = U * (e1[2:12] - e1[1:11]) / dx
    else:
        div = U * (e1[3:13] - e1[2:12]) / dx
    P = 0.4 * e1[2:12]
    want = e1[2:12] + dt * (-P * 0.0 - div) + ek[0, 0, 2:12]
    np.testing.assert_allclose(E_new[it][0, 0], want, rtol=0, atol=1e-14)


def test_pressure_fix_blends_with_the_godunov_change():
    o = _cart(2, [6, 5])
    a, b, c0, c1, dt, eps = 0.3, -0.2, 5.0, 4.0, 0.01, 0.1
    rho, rhou, E, ek, rho_rec, rhou_rec, E_rec = _linear_problem(o, a, b, c0, c1, 0.0)
    it = _interior(o)
    e0 = 0.5 * eps * float(np.min(ek[it]))  # between eps/10 and eps times ek in every cell
    E = E + e0
    E_rec = E_rec + e0
    delta = 1e-3
    E_new = E + delta
    o.pressure_fix(0, dt, eps, rho, rhou, E, rho_rec, rhou_rec, E_rec, rho.copy(), rhou.copy(), E_new)
    alpha = (e0 * e0) / (ek * ek * eps * eps)
    assert np.all((alpha[it] > 0) & (alpha[it] < 1))
    de_pf = dt * (-(0.4 * e0) * (a + b) - e0 * (a + b))
    want = e0 + alpha * delta + (1 - alpha) * de_pf + ek
    np.testing.assert_allclose(E_new[it], want[it], rtol=0, atol=2e-14)


def test_pressure_fix_leaves_unflagged_cells():
    o = _cart(2, [6, 5])
    rho, rhou, E, ek, rho_rec, rhou_rec, E_rec = _linear_problem(o, 0.3, -0.2, 1.0, 1.0, 10.0)
    E_new = E + np.random.default_rng(0).random(E.shape)
    before = E_new.copy()
    o.pressure_fix(0, 0.01, 1e-6, rho, rhou, E, rho_rec, rhou_rec, E_rec, rho.copy(), rhou.copy(), E_new)
    np.testing.assert_array_equal(E_new, before)


def _sph(nd, n, rmin=0.5, rmax=1.5, M=3.0e10):
    return orc.Oracle(nd, n, geom="Spherical", gravity="Internal_mass", central_mass=M,
                      xmin=(rmin, 0.0, 0.0), xmax=(rmax, np.pi, 2 * np.pi),
                      bc=["Reflexive", "NullGradient"] + ["Reflexive", "Reflexive", "Periodic", "Periodic"][:2 * nd - 2])


def _enclosed(o, layer_rho, M):
    x = orc.regular_nodes(o.n[0], o.ng, 0.5, 1.5)
    xc = (x[:-1] + x[1:]) / 2
    ng = o.ng
    out = np.empty(o.nwg[0])
    for i in range(o.nwg[0]):
        if i < ng:
            out[i] = M
            continue
        m = M
        for s in range(ng, i):  # full layers below
            m += 4 * np.pi / 3 * layer_rho[s] * (x[s + 1] ** 3 - x[s] ** 3)
        m += 4 * np.pi / 3 * layer_rho[i] * (xc[i] ** 3 - x[i] ** 3)  # half of layer i
        out[i] = m
    return -G_NEWTON * out / (xc * xc)


@pytest.mark.parametrize("nd,n", [(1, [12]), (3, [10, 4, 8])])
@pytest.mark.parametrize("profile", ["uniform", "two_layer"])
def test_internal_mass_gravity_closed_form(nd, n, profile):
    M = 3.0e10
    o = _sph(nd, n, M=M)
    layer = np.full(o.nwg[0], 2.0)
    if profile == "two_layer":
        layer[o.ng + n[0] // 2:] = 0.5
    rho = np.ones(o.shape3) * layer[None, None, :]
    if nd == 3:  # phi modulation with zero shell average
        phi = (np.arange(o.nwg[2]) - o.ng + 0.5) * 2 * np.pi / n[2]
        rho = rho * (1 + 0.5 * np.cos(phi))[:, None, None]
    g = o.internal_mass_gravity(rho)
    np.testing.assert_allclose(g, _enclosed(o, layer, M), rtol=1e-12, atol=0)


def test_internal_mass_gravity_follows_the_density_every_step():
    """Driver: the table is rebuilt from the new density after each step (main.cpp:528)."""
    o = _sph(1, [16])
    rho = np.ones(o.shape3)
    rho[..., 2:6] = 3.0
    u = np.zeros((1,) + o.shape3)
    P = np.ones(o.shape3)
    o.init_from_prim(rho, u, P)
    g0 = o.gravity_table()
    o.step(0.4, 1e9, 1 << 30)
    g1 = o.gravity_table()
    rho_now = o.state("rho").copy()
    want = o.internal_mass_gravity(rho_now)
    np.testing.assert_array_equal(g1, want)
    assert not np.array_equal(g0, g1)


def test_pressure_fix_uses_face_areas_of_the_shifted_cell():
    """1D spherical, uniform radial flow U and internal energy e: div u = U (A(r+) - A(r-)) / V with
    A = 4 pi r^2 and V = 4 pi (r+^3 - r-^3) / 3 (the finite-volume divergence)."""
    o = orc.Oracle(1, [10], geom="Spherical", xmin=(0.5, 0, 0), xmax=(1.5, 1, 1), bc=["Reflexive", "NullGradient"])
    U, e0, dt, eps = 4.0, 0.05, 0.01, 0.1
    rho = np.ones(o.shape3)
    rhou = np.full((1,) + o.shape3, U)
    ek = 0.5 * U * U * np.ones(o.shape3)
    E = e0 + ek
    rho_rec, E_rec, rhou_rec = o.zeros(1, 2), o.zeros(1, 2), o.zeros(1, 1, 2)
    rho_rec[...] = 1.0
    rhou_rec[...] = U
    E_rec[...] = e0 + ek[None]
    E_new = E.copy()
    o.pressure_fix(0, dt, eps, rho, rhou, E, rho_rec, rhou_rec, E_rec, rho.copy(), rhou.copy(), E_new)
    r = orc.regular_nodes(10, 2, 0.5, 1.5)[2:13]
    A = 4 * np.pi * r * r
    V = 4 * np.pi * (r[1:] ** 3 - r[:-1] ** 3) / 3
    div = U * (A[1:] - A[:-1]) / V
    want = e0 - dt * (0.4 * e0 + e0) * div + ek[0, 0, 2:12]
    np.testing.assert_allclose(E_new[0, 0, 2:12], want, rtol=1e-13, atol=0)


# ---------------------------------------------------------------- passive scalars
# Closed forms for the scalar path of the predictor (extrapolation_reconstruction.hpp:87-108,
# 315-337) and of the Godunov update (godunov_scheme.hpp:122, 263-301).  With a uniform density,
# velocity U and pressure, the mass flux of every face is rho U, and the scalar obeys the
# textbook schemes for linear advection with Courant number nu = U dt / dx (Toro, "Riemann
# Solvers and Numerical Methods for Fluid Dynamics", ch. 13-14):
#   Constant limiter: first-order upwind,  y_i' = y_i - nu (y_i - y_{i-1})            (U > 0)
#   Minmod: MUSCL-Hancock, face y_R = y_i + (1 - nu) D_i / 2, D_i = minmod(y_{i+1} - y_i, y_i - y_{i-1}),
#           y_i' = y_i - nu (y_R(i) - y_R(i-1))                                          (U > 0)
# and the mirrored forms for U < 0.

def _uniform_flow_1d(n, U, limiter, y):
    o = orc.Oracle(1, [n], limiter=limiter, riemann="HLL", nfx=1, bc=["Periodic", "Periodic"])
    rho = np.ones(o.shape3)
    u = np.full((1,) + o.shape3, U)
    P = np.ones(o.shape3)
    fx = np.zeros((1,) + o.shape3)
    fx[0, 0, 0, o.ng:o.ng + n] = y
    fx[0, 0, 0, :o.ng] = y[-o.ng:]
    fx[0, 0, 0, o.ng + n:] = y[:o.ng]
    o.init_from_prim(rho, u, P, fx)
    dt, _ = o.step(0.4, 1e30)
    return o, dt


def _minmod(a, b):
    return np.where(a * b > 0, np.sign(a) * np.minimum(np.abs(a), np.abs(b)), 0.0)


@pytest.mark.parametrize("U", [0.7, -0.7])
def test_passive_scalar_first_order_upwind(U):
    n = 32
    y = np.where((np.arange(n) >= 8) & (np.arange(n) < 20), 0.7, 0.2)
    o, dt = _uniform_flow_1d(n, U, "Constant", y)
    nu = U * dt * n
    if U > 0:
        want = y - nu * (y - np.roll(y, 1))
    else:
        want = y - nu * (np.roll(y, -1) - y)
    np.testing.assert_allclose(o.state("fx")[0, 0, 0, 2:2 + n], want, rtol=0, atol=2e-15)


@pytest.mark.parametrize("U", [0.7, -0.7])
def test_passive_scalar_muscl_hancock(U):
    n = 40
    x = (np.arange(n) + 0.5) / n
    y = 0.5 + 0.3 * np.sin(2 * np.pi * x) + 0.1 * (x > 0.6)  # smooth part, a jump and extrema
    o, dt = _uniform_flow_1d(n, U, "Minmod", y)
    nu = abs(U) * dt * n
    D = _minmod(np.roll(y, -1) - y, y - np.roll(y, 1))
    if U > 0:
        yr = y + (1 - nu) * D / 2  # upwind face value at i+1/2
        want = y - nu * (yr - np.roll(yr, 1))
    else:
        yl = y - (1 - nu) * D / 2  # upwind face value at i-1/2
        want = y - nu * (yl - np.roll(yl, -1))
    np.testing.assert_allclose(o.state("fx")[0, 0, 0, 2:2 + n], want, rtol=0, atol=2e-15)


def _random_flow_with_scalar(nd, n, geom, seed):
    lo = SPH_LO if geom == "Spherical" else (0.0, 0.0, 0.0)
    hi = SPH_HI if geom == "Spherical" else (1.0, 1.0, 1.0)
    bc = (["Reflexive", "Reflexive", "Reflexive", "Reflexive", "Periodic", "Periodic"] if geom == "Spherical"
          else ["Periodic"] * 6)[:2 * nd]
    o = orc.Oracle(nd, n, geom=geom, xmin=lo, xmax=hi, limiter="VanLeer", riemann="HLLC", nfx=1, bc=bc)
    rng = np.random.default_rng(seed)
    rho = 1.0 + 0.3 * rng.random(o.shape3)
    u = 0.3 * rng.standard_normal((nd,) + o.shape3)
    P = 1.0 + 0.2 * rng.random(o.shape3)
    return o, rho, u, P, rng


SPH_LO = (0.3, 0.2, 0.0)
SPH_HI = (1.3, np.pi - 0.2, 2 * np.pi)


@pytest.mark.parametrize("nd,n,geom", [(1, [24], "Cartesian"), (2, [16, 12], "Cartesian"), (3, [10, 8, 6], "Cartesian"),
                                       (2, [12, 8], "Spherical"), (3, [8, 6, 8], "Spherical")])
def test_passive_scalar_of_one_follows_the_density(nd, n, geom):
    """y = 1 everywhere: the partial density y rho takes exactly the density's updates (same
    fluxes, same order), so y stays exactly 1 in every cell after several steps."""
    o, rho, u, P, _ = _random_flow_with_scalar(nd, n, geom, seed=nd)
    o.init_from_prim(rho, u, P, np.ones((1,) + o.shape3))
    for _ in range(4):
        o.step(0.4, 1e30)
    np.testing.assert_array_equal(o.state("fx"), 1.0)


@pytest.mark.parametrize("nd,n", [(1, [48]), (2, [20, 16]), (3, [10, 8, 6])])
def test_passive_scalar_mass_is_conserved(nd, n):
    """Periodic box, random flow and scalar in (0.2, 0.8) (no clamping): the scalar mass
    sum(y rho dV) is conserved to rounding, which needs the same upwinded flux on both sides
    of every face."""
    o, rho, u, P, rng = _random_flow_with_scalar(nd, n, "Cartesian", seed=10 + nd)
    y = 0.2 + 0.6 * rng.random((1,) + o.shape3)
    o.init_from_prim(rho, u, P, y)
    dv = o.geometry(7)
    it = _interior(o)

    def scalar_mass():
        return float(np.sum((o.state("fx")[0] * o.state("rho") * dv)[it]))

    m0 = scalar_mass()
    for _ in range(5):
        o.step(0.4, 1e30)
        fx = o.state("fx")[0][it]
        assert fx.min() > 0 and fx.max() < 1
    assert abs(scalar_mass() - m0) <= 1e-14 * abs(m0)


def test_passive_scalar_is_clamped():
    """A scalar outside [0, 1] on input comes out clamped (godunov_scheme.hpp:294-301)."""
    n = 16
    y = np.linspace(-0.5, 1.5, n)
    o, _ = _uniform_flow_1d(n, 0.3, "Constant", y)
    fx = o.state("fx")[0, 0, 0, 2:2 + n]
    assert fx.min() == 0.0 and fx.max() == 1.0
