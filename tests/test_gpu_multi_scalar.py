"""Several passive scalars on the fused step (hclpp_step_fx with nfx > 1), against the oracle.

The fused kernel carries one scalar per launch; with nfx scalars the step launches it nfx times,
the first launch also writing the hydro update and the CFL/totals partials.  Each scalar is
reconstructed, extrapolated as a partial density and upwinded with the same mass flux
(extrapolation_reconstruction.hpp:87-88, 105-108, 315-337; godunov_scheme.hpp:263-301), so every
array must equal the oracle's, bit for bit, the clock included.

* Cartesian 1D/2D/3D with 2 to 4 scalars, including scalars that differ only by their index (the
  same field in every plane must give the same planes);
* the shape of the reference input inputs/v1d_3d.ini: 3D spherical wedge theta, phi in
  [pi/4, 3pi/4], five scalars, Minmod + HLLC, point-mass gravity, mmw 0.62, Reflexive /
  NullGradient radial walls and the pressure fix on (eps_pf 1e-5), on a small grid;
* FP32 storage with three scalars: within the FP32 tolerance of the FP64 oracle.
"""

import numpy as np
import pytest

import heraclespp_amd as H
from heraclespp_amd import setups
from heraclespp_amd.hydro import PointMassGravity
from heraclespp_amd.solver import HydroSolver
from oracle import oracle as orc
from tests.helpers import make_oracle, make_solver, rel_l1

pytestmark = pytest.mark.gpu


def _scalars(rng, shape3, nfx):
    return np.stack([0.05 + 0.9 * rng.random(shape3) for _ in range(nfx)])


def _compare(sol, o, tag):
    assert sol.t == o.t and sol.iter == o.iter, tag
    for nm, gpu in (("rho", sol.rho), ("rhou", sol.rhou), ("E", sol.E), ("fx", sol.fx)):
        got = gpu.cpu().numpy()
        ref = o.state(nm)
        print(f"[multi-fx] {tag} {nm}: maxabs={float(np.max(np.abs(got - ref))):.3e}", flush=True)
        np.testing.assert_array_equal(got, ref, err_msg=f"{tag} {nm}")


CASES = {
    "sod1d_nfx4": (lambda: setups.sod(400), 4, 40),
    "kh2d_nfx3": (lambda: setups.kelvin_helmholtz(40, seed=5), 3, 12),
    "rt3d_nfx2": (lambda: setups.rayleigh_taylor(16, seed=4), 2, 6),
}


@pytest.mark.parametrize("name", list(CASES))
def test_fused_step_several_scalars_match_oracle(cuda, name):
    make, nfx, steps = CASES[name]
    s = make()
    fx = _scalars(np.random.default_rng(11), s.rho.shape, nfx)
    print(f"\n[multi-fx] case={name} n={s.n} nfx={nfx} steps={steps}", flush=True)
    o = make_oracle(s, nfx=nfx)
    o.init_from_prim(s.rho, s.u, s.P, fx)
    for _ in range(steps):
        _, ex = o.step(s.cfl, s.t_end, steps)
        if ex:
            break
    sol = make_solver(s, fused=True, nfx=nfx)
    assert sol.fused
    sol.init_from_prim(s.rho, s.u, s.P, fx=fx)
    assert sol.run(s.t_end, steps) == o.iter
    _compare(sol, o, name)


def test_identical_scalars_stay_identical(cuda):
    """Three copies of one field: each launch computes the same scalar update, so the three planes
    are equal, and equal to the one-scalar step's plane."""
    s = setups.kelvin_helmholtz(32, seed=2)
    y = (s.rho > 1.5).astype(np.float64)
    sol3 = make_solver(s, nfx=3)
    sol3.init_from_prim(s.rho, s.u, s.P, fx=np.stack([y, y, y]))
    sol1 = make_solver(s, nfx=1)
    sol1.init_from_prim(s.rho, s.u, s.P, fx=y[None])
    assert sol3.run(s.t_end, 10) == sol1.run(s.t_end, 10) == 10
    f3, f1 = sol3.fx.cpu().numpy(), sol1.fx.cpu().numpy()
    for f in range(3):
        assert np.array_equal(f3[f], f1[0]), f
    for nm in ("rho", "rhou", "E"):
        assert np.array_equal(getattr(sol3, nm).cpu().numpy(), getattr(sol1, nm).cpu().numpy()), nm


def test_v1d_3d_shape_spherical_five_scalars_pressure_fix(cuda):
    nd, n = 3, [28, 6, 6]
    lo, hi = (0.5, np.pi / 4, np.pi / 4), (1.5, 3 * np.pi / 4, 3 * np.pi / 4)
    bc = ["Reflexive", "NullGradient", "Reflexive", "Reflexive", "Reflexive", "Reflexive"]
    mass, mmw, nfx, steps = 7.0e8, 0.62, 5, 12
    kw = dict(geom="Spherical", xmin=lo, xmax=hi, bc=bc, limiter="Minmod", riemann="HLLC", mmw=mmw, nfx=nfx,
              gravity="Point_mass", central_mass=mass)
    o = orc.Oracle(nd, n, pressure_fix=True, eps_pf=1e-5, **kw)
    rng = np.random.default_rng(5)
    # a cold radial wind (e / ek ~ 1e-6: rewritten by the fix) over a nearly uniform density, warm
    # (P = 1) in the inner 30 % of the radial cells
    rho = 1.0 + 1e-6 * rng.random(o.shape3)
    u = np.zeros((nd,) + o.shape3)
    u[0] = 3.0
    P = 1e-5 * (1.0 + 0.5 * rng.random(o.shape3))
    P[..., : int(0.3 * P.shape[-1])] = 1.0
    fx = _scalars(rng, o.shape3, nfx)
    o.init_from_prim(rho, u, P, fx)
    for _ in range(steps):
        o.step(0.4, 1e30, steps)
    ref = orc.Oracle(nd, n, **kw)
    ref.init_from_prim(rho, u, P, fx)
    for _ in range(steps):
        ref.step(0.4, 1e30, steps)
    assert np.sum(o.state("E") != ref.state("E")) > 0  # the fix acted
    print(f"\n[multi-fx] v1d_3d shape n={n} nfx={nfx} steps={steps}", flush=True)
    g = H.Grid(nd, n, 2, "Spherical", lo, hi)
    sol = HydroSolver(g, H.PerfectGas(1.4, mmw), reconstruction="Minmod", riemann_solver="HLLC",
                      gravity=PointMassGravity(mass, g), bc=bc, cfl=0.4, nfx=nfx, pressure_fix="On", eps_pf=1e-5)
    assert sol.fused
    sol.init_from_prim(rho, u, P, fx=fx)
    assert sol.run(1e30, steps) == steps
    _compare(sol, o, "v1d_3d_shape")


def test_fp32_storage_three_scalars(cuda):
    s = setups.kelvin_helmholtz(48, seed=3)
    nfx, steps = 3, 30
    fx = _scalars(np.random.default_rng(2), s.rho.shape, nfx)
    o = make_oracle(s, nfx=nfx)
    o.init_from_prim(s.rho, s.u, s.P, fx)
    for _ in range(steps):
        o.step(s.cfl, s.t_end, steps)
    sol = make_solver(s, precision="fp32", nfx=nfx)
    sol.init_from_prim(s.rho, s.u, s.P, fx=fx)
    assert sol.run(s.t_end, steps) == steps
    I = (Ellipsis,) + s.interior()
    for nm, gpu in (("rho", sol.rho), ("E", sol.E), ("fx", sol.fx)):
        err = rel_l1(gpu.cpu().numpy()[I], o.state(nm)[I])
        print(f"[multi-fx] fp32 {nm}: rel L1 {err:.3e}", flush=True)
        assert err <= 1e-5, nm
