"""The opt-in fast FP64 arithmetic of the fused step (hclpp_set_fast_arithmetic).

It keeps the FP64 state and every formula of the step but not the reference's rounding, so it is
checked against the oracle with tolerances, not bit for bit:
* one step from random states, every limiter x Riemann solver, 1D/2D/3D Cartesian and 3D
  spherical: relative L1 <= 1e-12 per conserved variable.  A wrong closed-form limiter, a dropped
  term or a wrong constant in the direct pressure formula shows as an O(1e-3) or larger difference;
* a short run of each configured workload shape (KH, RT, Sedov, Sod, smooth Euler) stays within
  1e-9 of the oracle, and mass is conserved to round-off on a periodic grid;
* Sod to t = 0.2 against the exact Riemann solution (test_gpu_analytic.py, fp64_fast);
* the C-ABI rules: FP64 storage only, no passive scalars, no pressure fix; switching back to the
  reference arithmetic gives the oracle's bits again.
"""

import ctypes

import numpy as np
import pytest

import heraclespp_amd as H
from heraclespp_amd import _lib, setups
from heraclespp_amd.solver import HydroSolver
from oracle import oracle as orc
from tests.helpers import make_oracle, make_solver, rel_l1

pytestmark = pytest.mark.gpu

VARS = ("rho", "rhou", "E")


def _fast_solver(s, **kw):
    sol = make_solver(s, **kw)
    _lib.check(_lib.lib().hclpp_set_fast_arithmetic(sol.ctx.handle, 1))
    return sol


def _random_state(o, nd, rng):
    rho = 1.0 + 0.5 * rng.random(o.shape3)
    u = 0.3 * rng.standard_normal((nd,) + o.shape3)
    P = 1.0 + 0.5 * rng.random(o.shape3)
    return rho, u, P


@pytest.mark.parametrize("riemann", ["HLL", "HLLC", "Low Mach"])
@pytest.mark.parametrize("limiter", ["Constant", "VanLeer", "Minmod", "VanAlbada"])
@pytest.mark.parametrize("nd,n,geom", [(1, [96], "Cartesian"), (2, [40, 24], "Cartesian"), (3, [14, 12, 10], "Cartesian"),
                                        (3, [16, 8, 12], "Spherical")], ids=["1d", "2d", "3d", "3d_spherical"])
def test_one_step_matches_oracle_to_rounding(cuda, nd, n, geom, limiter, riemann):
    if geom == "Spherical":
        lo, hi = (0.5, 0.3, 0.0), (1.5, np.pi - 0.3, 2 * np.pi)
        bc = ["Reflexive", "NullGradient", "Reflexive", "Reflexive", "Periodic", "Periodic"]
    else:
        lo, hi = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
        bc = ["Periodic"] * 6
    bc = bc[:2 * nd]
    o = orc.Oracle(nd, n, geom=geom, xmin=lo, xmax=hi, bc=bc, limiter=limiter, riemann=riemann, g=(0.0, -0.1, 0.0))
    rho, u, P = _random_state(o, nd, np.random.default_rng(nd + 10 * len(limiter)))
    o.init_from_prim(rho, u, P)
    o.step(0.4, 1e30, 1)
    g = H.Grid(nd, n, 2, geom, lo, hi)
    sol = HydroSolver(g, H.PerfectGas(1.4, 1.0), reconstruction=limiter, riemann_solver=riemann,
                      gravity=H.UniformGravity((0.0, -0.1, 0.0)), bc=bc, cfl=0.4, arithmetic="fast")
    sol.init_from_prim(rho, u, P)
    assert sol.run(1e30, 1) == 1
    # dt comes from the initial state, computed by the clock start with the reference arithmetic
    assert sol.t == o.t
    for nm in VARS:
        err = rel_l1(getattr(sol, nm).cpu().numpy(), o.state(nm))
        assert err <= 1e-12, (nm, err)


RUNS = {
    "kh2d": (lambda: setups.kelvin_helmholtz(96, seed=3), 200),
    "rt3d": (lambda: setups.rayleigh_taylor(24, seed=2), 60),
    "sedov3d_spherical": (lambda: setups.sedov(24, 12, 16, t_end=1.0), 30),
    "sod1d": (lambda: setups.sod(1000), 300),
    "smooth2d": (lambda: setups.smooth_euler(64, seed=1), 100),
}


@pytest.mark.parametrize("name", list(RUNS))
def test_short_runs_stay_close_to_oracle(cuda, name):
    make, steps = RUNS[name]
    s = make()
    o = make_oracle(s)
    o.init_from_prim(s.rho, s.u, s.P)
    for _ in range(steps):
        _, ex = o.step(s.cfl, s.t_end, steps)
        if ex:
            break
    sol = _fast_solver(s)
    sol.init_from_prim(s.rho, s.u, s.P)
    assert sol.run(s.t_end, steps) == o.iter
    assert abs(sol.t - o.t) <= 1e-9 * o.t
    I = (Ellipsis,) + s.interior()
    for nm in VARS:
        err = rel_l1(getattr(sol, nm).cpu().numpy()[I], o.state(nm)[I])
        print(f"[fast] {name} after {o.iter} steps: {nm} rel L1 {err:.3e}", flush=True)
        assert err <= 1e-9, (nm, err)


def test_mass_conserved_periodic(cuda):
    s = setups.kelvin_helmholtz(128, seed=4)
    sol = _fast_solver(s)
    sol.init_from_prim(s.rho, s.u, s.P)
    I = s.interior()
    m0 = float(sol.rho.cpu().numpy()[I].sum())
    sol.run(s.t_end, 300)
    m1 = float(sol.rho.cpu().numpy()[I].sum())
    assert abs(m1 - m0) <= 1e-13 * m0


def test_fast_arithmetic_rules(cuda):
    L = _lib.lib()
    s = setups.kelvin_helmholtz(16)
    with pytest.raises(_lib.HclppError, match="passive scalars"):
        _lib.check(L.hclpp_set_fast_arithmetic(make_solver(s, nfx=1).ctx.handle, 1))
    with pytest.raises(_lib.HclppError, match="FP64-storage"):
        _lib.check(L.hclpp_set_fast_arithmetic(make_solver(s, precision="fp32").ctx.handle, 1))
    sol = make_solver(s)
    _lib.check(L.hclpp_set_fast_arithmetic(sol.ctx.handle, 1))
    with pytest.raises(_lib.HclppError, match="reference arithmetic"):
        _lib.check(L.hclpp_set_pressure_fix(sol.ctx.handle, 1, ctypes.c_double(1e-5)))
    with pytest.raises(ValueError, match="fused FP64 mode"):
        HydroSolver(sol.grid, H.PerfectGas(1.4, 1.0), fused=False, arithmetic="fast")


def test_switching_back_restores_reference_bits(cuda):
    s = setups.kelvin_helmholtz(48, seed=2)
    o = make_oracle(s)
    o.init_from_prim(s.rho, s.u, s.P)
    for _ in range(20):
        o.step(s.cfl, s.t_end, 20)
    sol = make_solver(s)
    L = _lib.lib()
    _lib.check(L.hclpp_set_fast_arithmetic(sol.ctx.handle, 1))
    _lib.check(L.hclpp_set_fast_arithmetic(sol.ctx.handle, 0))
    sol.init_from_prim(s.rho, s.u, s.P)
    assert sol.run(s.t_end, 20) == 20
    for nm in VARS:
        np.testing.assert_array_equal(getattr(sol, nm).cpu().numpy(), o.state(nm), err_msg=nm)
