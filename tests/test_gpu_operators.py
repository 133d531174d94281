"""Operator-level parity: each C-ABI operator (through the reference-shaped host
API of heraclespp_amd.hydro) against the oracle's kernel on the same random
state.  FP64 results must be bit-identical; reductions whose summation
order differs (integrate) are held to 1e-12 relative."""

import numpy as np
import pytest
import torch

import heraclespp_amd as H
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

GEOMS = [
    (1, "Cartesian", (12, 1, 1)),
    (2, "Cartesian", (10, 9, 1)),
    (3, "Cartesian", (7, 6, 5)),
    (1, "Spherical", (12, 1, 1)),
    (2, "Spherical", (10, 8, 1)),
    (3, "Spherical", (8, 6, 7)),
]
SPH_LO = (0.05, 0.0, 0.0)
SPH_HI = (1.0, np.pi, 2 * np.pi)


def _grids(nd, geom, n, **kw):
    lo = SPH_LO if geom == "Spherical" else (0.0, 0.0, 0.0)
    hi = SPH_HI if geom == "Spherical" else (1.0, 1.0, 1.0)
    g = H.Grid(nd, n, 2, geom, lo, hi)
    o = orc.Oracle(nd, n, geom=geom, xmin=lo, xmax=hi, **kw)
    return g, o


def _random_prim(o, seed, nfx=0):
    rng = np.random.default_rng(seed)
    rho = 1.0 + 0.5 * rng.random(o.shape3)
    u = rng.uniform(-0.7, 0.7, size=(o.ndim,) + o.shape3)
    P = 0.5 + rng.random(o.shape3)
    fx = rng.random((nfx,) + o.shape3) if nfx else None
    return rho, u, P, fx


def _cons(o, rho, u, P):
    rhou = np.zeros_like(u)
    E = np.zeros_like(rho)
    o.conv_prim_to_cons(o.ng, rho, u, P, rhou, E)
    return rhou, E


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _n(t):
    return t.cpu().numpy()


def test_conversions_reference_case(cuda):
    """tests/conversions.cpp:17-103: constant state on a 10^3 grid, gamma = 2, against scalar to_cons/to_prim."""
    n = 10
    eos = H.PerfectGas(2.0, 1.0)
    for nd in (1, 2, 3):
        shape = tuple(n if 2 - d < nd else 1 for d in range(3))
        rng = H.Range([0, 0, 0], [n if d < nd else 1 for d in range(3)], 0, nd)
        rho = torch.full(shape, 2.0, dtype=torch.float64, device="cuda")
        u = torch.full((nd,) + shape, -1.0, dtype=torch.float64, device="cuda")
        P = torch.full(shape, 10.0, dtype=torch.float64, device="cuda")
        rhou = torch.zeros_like(u)
        E = torch.zeros_like(rho)
        H.conv_prim_to_cons(rng.all_ghosts(), eos, rho, u, P, rhou, E)
        want = orc.to_cons([2.0] + [-1.0] * nd + [10.0], 2.0, 1.0)
        assert torch.all(rhou == want[1]) and torch.all(E == want[-1])
        rho2 = torch.full(shape, 2.0, dtype=torch.float64, device="cuda")
        rhou2 = torch.full((nd,) + shape, -2.0, dtype=torch.float64, device="cuda")
        E2 = torch.full(shape, 10.0, dtype=torch.float64, device="cuda")
        u2 = torch.zeros_like(rhou2)
        P2 = torch.zeros_like(rho2)
        H.conv_cons_to_prim(rng.all_ghosts(), eos, rho2, rhou2, E2, u2, P2)
        want = orc.to_prim([2.0] + [-2.0] * nd + [10.0], 2.0, 1.0)
        assert torch.all(u2 == want[1]) and torch.all(P2 == want[-1])


@pytest.mark.parametrize("nd,geom,n", GEOMS)
def test_conversions_random(cuda, nd, geom, n):
    g, o = _grids(nd, geom, n)
    rho, u, P, _ = _random_prim(o, 1)
    eos = H.PerfectGas(1.4, 1.0)
    rhou, E = _cons(o, rho, u, P)
    rhou_d = torch.zeros_like(_t(u))
    E_d = torch.zeros_like(_t(rho))
    H.conv_prim_to_cons(g.range.all_ghosts(), eos, _t(rho), _t(u), _t(P), rhou_d, E_d, grid=g)
    assert np.array_equal(_n(rhou_d), rhou) and np.array_equal(_n(E_d), E)
    u_o, P_o = np.zeros_like(u), np.zeros_like(P)
    o.conv_cons_to_prim(o.ng, rho, rhou, E, u_o, P_o)
    u_d, P_d = torch.zeros_like(rhou_d), torch.zeros_like(E_d)
    H.conv_cons_to_prim(g.range.all_ghosts(), eos, _t(rho), _t(rhou), _t(E), u_d, P_d, grid=g)
    assert np.array_equal(_n(u_d), u_o) and np.array_equal(_n(P_d), P_o)


@pytest.mark.parametrize("limiter", ["Constant", "VanLeer", "Minmod", "VanAlbada"])
@pytest.mark.parametrize("nd,geom,n", GEOMS[:3])
def test_face_reconstruction(cuda, nd, geom, n, limiter):
    g, o = _grids(nd, geom, n, limiter=limiter)
    rng = np.random.default_rng(2)
    var = rng.standard_normal(o.shape3)
    rec = np.zeros((nd, 2) + o.shape3)
    o.face_reconstruction(1, var, rec)
    rec_d = torch.zeros((nd, 2) + g.shape3, dtype=torch.float64, device="cuda")
    H.factory_face_reconstruction(limiter).execute(g.range.with_ghosts(1), g, _t(var), rec_d)
    assert np.array_equal(_n(rec_d), rec)


def test_factories_reject_unknown_names():
    with pytest.raises(RuntimeError, match="Unknown face reconstruction algorithm: Superbee."):
        H.factory_face_reconstruction("Superbee")
    with pytest.raises(RuntimeError, match="Invalid riemann solver: Roe."):
        H.factory_godunov_scheme("Roe", H.PerfectGas(1.4, 1.0))


class _CountingFace(H.IFaceReconstruction):
    """A user's face reconstruction (forwards to the library's and counts the calls)."""

    def __init__(self, inner):
        self.inner, self.calls = inner, 0

    def execute(self, range_, grid, var, var_rec):
        self.calls += 1
        self.inner.execute(range_, grid, var, var_rec)


class _CountingHancock(H.IExtrapolationReconstruction):
    """A user's extrapolation step (forwards to the library's and counts the calls)."""

    def __init__(self, inner):
        self.inner, self.calls = inner, 0

    def execute(self, *args):
        self.calls += 1
        self.inner.execute(*args)


@pytest.mark.parametrize("strategies", ["library", "user"])
@pytest.mark.parametrize("gravity", ["Uniform", "Point_mass"])
@pytest.mark.parametrize("nfx", [0, 2])
@pytest.mark.parametrize("nd,geom,n", GEOMS)
def test_hydro_reconstruction(cuda, nd, geom, n, nfx, gravity, strategies):
    """MUSCLHancockHydroReconstruction with the library's strategies (one C-ABI call) or with user
    strategies (composed through the interfaces, as hydro_reconstruction.hpp:85-129 does): both
    bit-identical to the oracle."""
    kw = dict(limiter="VanLeer", nfx=nfx, gravity=gravity, g=(0.3, -0.2, 0.1), central_mass=2e10)
    g, o = _grids(nd, geom, n, **kw)
    rho, u, P, fx = _random_prim(o, 3, nfx)
    fxa = fx if nfx else np.zeros((1,) + o.shape3)
    dt = 1e-3
    outs = [np.zeros((nd, 2) + o.shape3), np.zeros((nd, nd, 2) + o.shape3), np.zeros((nd, 2) + o.shape3),
            np.zeros((max(nfx, 1), nd, 2) + o.shape3), np.zeros((nd, 2) + o.shape3), np.zeros((nd, nd, 2) + o.shape3)]
    o.hydro_reconstruction(1, dt / 2, rho, u, P, fxa, *outs)
    grav = H.PointMassGravity(2e10, g) if gravity == "Point_mass" else H.UniformGravity((0.3, -0.2, 0.1))
    eos = H.PerfectGas(1.4, 1.0)
    d = [torch.zeros(x.shape, dtype=torch.float64, device="cuda") for x in outs]
    if not nfx:
        d[3] = None
    face, hancock = H.factory_face_reconstruction("VanLeer"), H.HancockExtrapolationReconstruction(eos)
    if strategies == "user":
        face, hancock = _CountingFace(face), _CountingHancock(hancock)
    rec = H.MUSCLHancockHydroReconstruction(face, hancock, eos, d[4], d[5])
    rec.execute(g.range.with_ghosts(1), g, grav, dt / 2, _t(rho), _t(u), _t(P), _t(fx) if nfx else None, d[0], d[1],
                d[2], d[3])
    if strategies == "user":
        assert face.calls == 1 + nd + 1 + nfx and hancock.calls == 1
    for k, name in enumerate(["rho_rec", "rhou_rec", "E_rec", "fx_rec", "P_rec", "u_rec"]):
        if d[k] is None:
            continue
        assert np.array_equal(_n(d[k]), outs[k]), name


@pytest.mark.parametrize("solver", ["HLL", "HLLC", "Low Mach"])
@pytest.mark.parametrize("nd,geom,n", GEOMS)
def test_godunov(cuda, nd, geom, n, solver):
    nfx = 1
    kw = dict(limiter="Minmod", riemann=solver, nfx=nfx, g=(0.2, 0.1, -0.3))
    g, o = _grids(nd, geom, n, **kw)
    rho, u, P, fx = _random_prim(o, 4, nfx)
    rhou, E = _cons(o, rho, u, P)
    dt = 1e-3
    rec = [np.zeros((nd, 2) + o.shape3), np.zeros((nd, nd, 2) + o.shape3), np.zeros((nd, 2) + o.shape3),
           np.zeros((nfx, nd, 2) + o.shape3), np.zeros((nd, 2) + o.shape3), np.zeros((nd, nd, 2) + o.shape3)]
    o.hydro_reconstruction(1, dt / 2, rho, u, P, fx, *rec)
    new = [np.zeros(o.shape3), np.zeros((nd,) + o.shape3), np.zeros(o.shape3), np.zeros((nfx,) + o.shape3)]
    o.godunov(0, dt, rho, rhou, E, fx, rec[0], rec[1], rec[2], rec[3], *new)
    eos = H.PerfectGas(1.4, 1.0)
    dn = [torch.zeros(x.shape, dtype=torch.float64, device="cuda") for x in new]
    H.factory_godunov_scheme(solver, eos).execute(
        g.range.no_ghosts(), g, H.UniformGravity((0.2, 0.1, -0.3)), dt, _t(rho), _t(rhou), _t(E), _t(fx), _t(rec[0]),
        _t(rec[1]), _t(rec[2]), _t(rec[3]), *dn)
    inner = (Ellipsis,) + tuple(slice(2, -2) if 2 - a < nd else slice(None) for a in range(3))
    for k, name in enumerate(["rho_new", "rhou_new", "E_new", "fx_new"]):
        assert np.array_equal(_n(dn[k])[inner], new[k][inner]), name


@pytest.mark.parametrize("nd,geom,n", GEOMS)
def test_reductions(cuda, nd, geom, n):
    g, o = _grids(nd, geom, n)
    rho, u, P, _ = _random_prim(o, 5)
    rhou, E = _cons(o, rho, u, P)
    eos = H.PerfectGas(1.4, 1.0)
    dt_o = o.time_step(2, rho, u, P)
    dt_g = H.time_step(g.range.all_ghosts(), eos, g, _t(rho), _t(u), _t(P))
    assert dt_g == dt_o
    me_o = o.min_internal_energy(0, rho, rhou, E)
    assert H.minimum_internal_energy(g.range.no_ghosts(), g, _t(rho), _t(rhou), _t(E)) == me_o
    m_o = o.integrate(0, rho)
    m_g = H.integrate(g.range.no_ghosts(), g, _t(rho))
    assert abs(m_g - m_o) <= 1e-12 * abs(m_o)


BC_SETS = [
    ["Periodic", "Periodic", "NullGradient", "Reflexive", "Reflexive", "NullGradient"],
    ["Reflexive", "NullGradient", "Periodic", "Periodic", "Reflexive", "Reflexive"],
    ["NullGradient", "NullGradient", "Reflexive", "Reflexive", "Periodic", "Periodic"],
]


@pytest.mark.parametrize("order", [None, [5, 3, 1, 0, 2, 4]])
@pytest.mark.parametrize("bcs", BC_SETS)
@pytest.mark.parametrize("nd", [1, 2, 3])
def test_boundaries(cuda, nd, bcs, order):
    bc = bcs[:2 * nd]
    bo = [b for b in (order or range(6)) if b < 2 * nd] if order else None
    n = (6, 5, 4)
    nfx = 1
    g, o = _grids(nd, "Cartesian", n, bc=bc, bc_order=bo, nfx=nfx)
    rho, u, P, fx = _random_prim(o, 6, nfx)
    rhou, E = _cons(o, rho, u, P)
    # the reference fills ghosts of arrays whose ghosts hold stale values: randomise them too
    ref = [rho.copy(), rhou.copy(), E.copy(), fx.copy()]
    o.boundaries(*ref)
    dev = [_t(rho), _t(rhou), _t(E), _t(fx)]
    H.DistributedBoundaryCondition(g, bc, bo)(*dev)
    for k in range(4):
        assert np.array_equal(_n(dev[k]), ref[k]), k


def test_fused_step_passive_scalar_rules(cuda):
    """The fused step carries the passive scalars through hclpp_step_fx (any nfx); the C-ABI
    refuses, before touching any array, hclpp_step on a context with nfx > 0."""
    import ctypes

    from heraclespp_amd import _lib, setups
    from tests.helpers import make_solver

    s = setups.kelvin_helmholtz(16)
    fake = [ctypes.c_void_p(64 * (k + 1)) for k in range(8)]
    for nfx in (1, 2):
        sol = make_solver(s, fused=True, nfx=nfx)
        assert sol.fused
        with pytest.raises(_lib.HclppError, match="hclpp_step_fx"):
            _lib.check(_lib.lib().hclpp_step(sol.ctx.handle, *fake[:6]))


@pytest.mark.parametrize("fused", [True, False])
def test_invalid_internal_energy_is_detected(cuda, fused):
    """main.cpp:397-425: a negative internal energy stops the run with an error (forced from the host)."""
    from heraclespp_amd import setups
    from tests.helpers import make_solver

    s = setups.kelvin_helmholtz(16)
    sol = make_solver(s, fused=fused)
    sol.init_from_prim(s.rho, s.u, s.P)
    sol.E[0, 5, 7] = -1.0  # interior cell: E < kinetic energy
    with pytest.raises(RuntimeError, match="invalid volumic internal energy"):
        sol.run(1e30, 5)
    assert sol.iter == 0
