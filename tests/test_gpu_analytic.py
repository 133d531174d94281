"""The GPU step against the analytic solutions of the reference's own test problems
(test/exact_shock_tube.py, test/exact_sedov.py), independent of the oracle:

* advection of a sine wave once around the periodic box (test/convergence_advection_sinus, 1D in
  the reference's CI; here 1D-3D): the L1 error falls as h^2;
* Sod (1000 cells, Minmod + HLL, CFL 0.4, t = 0.2): the fused step, FP64 (reference and fast
  arithmetic) and FP32 storage,
  within L1(rho) < 2e-3 of the exact Riemann solution (Toro ch. 4; star state pinned against
  Toro's table in the oracle tests);
* 3D spherical Sedov (48 x 8 x 16, HLL + VanLeer, E = 1 in the two innermost shells, rho = 1):
  the radius of the shell-averaged density peak follows the Sedov-Taylor law
  R = (E t^2 / (alpha rho))^(1/5), alpha(gamma = 1.4) = 0.851072 (Kamm & Timmes 2007), within
  one radial cell, after 20000 steps, and at the configuration's t_end = 1.0 on 96 x 8 x 16
  (about 148 000 steps: the CFL step is set by the narrowest cells around the inner radius, ghost
  cells included): the shock front (the outermost cell above half the peak
  density) within one radial cell of R(1) = 1.033, and the density maximum at most two cells
  behind it.  On the configured 512 x 128 x 256 grid the cells beside the poles at the inner
  radius are some 250 times narrower and t_end would take about 2e7 steps."""

import numpy as np
import pytest

from heraclespp_amd import setups
from oracle import exact_riemann as X
from oracle import oracle as orc
from tests.helpers import make_solver

pytestmark = pytest.mark.gpu


def _advection(nd, n):
    """rho = 1 + 0.1 sin(2 pi (x + y + z)) over the unit cube, u = 1 along every dim, P = 0.1,
    periodic, VanLeer + HLLC, CFL 0.4, t_end 1: the profile is advected once around the box."""
    s = setups.Setup(f"advection{nd}d", nd, [n if d < nd else 1 for d in range(3)], "Cartesian", (0.0, 0.0, 0.0),
                     (1.0, 1.0, 1.0), "VanLeer", "HLLC", ["Periodic"] * (2 * nd), 1.4, 0.4, 1.0, 1 << 30)
    s.rho = np.zeros(s.shape3)
    s.u = np.zeros((nd,) + s.shape3)
    s.P = np.full(s.shape3, 0.1)
    xc = [(v[:-1] + v[1:]) / 2 for v in s.nodes()]
    Z, Y, X = np.meshgrid(xc[2], xc[1], xc[0], indexing="ij")
    phase = sum((X, Y, Z)[:nd])
    s.rho[...] = 1.0 + 0.1 * np.sin(2 * np.pi * phase)
    s.u[...] = 1.0
    return s, phase


@pytest.mark.parametrize("nd,sizes,precision", [(1, (50, 100, 200, 400), "fp64"), (1, (50, 100, 200, 400), "fp64_fast"),
                                                (1, (50, 100, 200), "fp32"), (2, (32, 64, 128, 256), "fp64"),
                                                (3, (32, 64, 128), "fp64"), (3, (32, 64, 128), "fp32")])
def test_advection_is_second_order(cuda, nd, sizes, precision):
    """The reference's convergence test (test/convergence_advection_sinus, run by its CI in 1D),
    on the fused GPU step in 1D-3D: the L1 distance of rho to the exact cell averages after one
    period, sin(2 pi (x + y + z)) (sin(pi h) / (pi h))^nd, falls as h^2 (slope within 5 %)."""
    from heraclespp_amd import _lib

    errs, hs = [], []
    for n in sizes:
        s, phase = _advection(nd, n)
        sol = make_solver(s, precision=precision.replace("_fast", ""))
        if precision == "fp64_fast":
            _lib.check(_lib.lib().hclpp_set_fast_arithmetic(sol.ctx.handle, 1))
        sol.init_from_prim(s.rho, s.u, s.P)
        sol.run(s.t_end, 1 << 30)
        assert abs(sol.t - 1.0) < 1e-14
        h = 1.0 / n
        exact = 1.0 + 0.1 * np.sin(2 * np.pi * phase) * (np.sin(np.pi * h) / (np.pi * h)) ** nd
        it = s.interior()
        errs.append(float(np.mean(np.abs(sol.rho.cpu().numpy()[it] - exact[it]))))
        hs.append(h)
    slope = np.polyfit(np.log(hs), np.log(errs), 1)[0]
    print(f"[advection] {nd}D {precision} L1 {['%.3e' % e for e in errs]} slope {slope:.3f}", flush=True)
    assert abs(slope - 2) / 2 < 0.05, (slope, errs)


@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp64_fast"])
def test_sod_against_exact_solution(cuda, precision):
    from heraclespp_amd import _lib

    s = setups.sod(1000)
    sol = make_solver(s, precision=precision.replace("_fast", ""))
    if precision == "fp64_fast":  # the opt-in fast FP64 arithmetic (hclpp_set_fast_arithmetic)
        _lib.check(_lib.lib().hclpp_set_fast_arithmetic(sol.ctx.handle, 1))
    sol.init_from_prim(s.rho, s.u, s.P)
    sol.run(s.t_end, 1 << 30)
    assert abs(sol.t - 0.2) < 1e-15
    x = orc.regular_nodes(1000, 2, 0.0, 1.0)
    xc = ((x[:-1] + x[1:]) / 2)[2:-2]
    re, _, _ = X.solve(xc, 0.5, 0.2, (1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 1.4)
    l1 = float(np.mean(np.abs(sol.rho.cpu().numpy()[0, 0, 2:-2] - re)))
    print(f"\n[analytic] sod {precision}: L1(rho) = {l1:.3e}", flush=True)
    assert l1 < 2e-3


def test_sedov_shock_radius(cuda):
    s = setups.sedov(48, 8, 16, t_end=1.0)
    sol = make_solver(s)
    sol.init_from_prim(s.rho, s.u, s.P)
    steps = sol.run(s.t_end, 20000)
    assert steps == 20000
    rho = sol.rho.cpu().numpy()
    profile = rho[2:-2, 2:-2, 2:-2].mean(axis=(0, 1))
    x = orc.regular_nodes(48, 2, s.xmin[0], s.xmax[0])
    xc = ((x[:-1] + x[1:]) / 2)[2:-2]
    r_peak = xc[int(np.argmax(profile))]
    r_st = (1.0 * sol.t ** 2 / (0.851072 * 1.0)) ** 0.2
    dr = x[3] - x[2]
    print(f"\n[analytic] sedov t={sol.t:.4f}: peak radius {r_peak:.4f}, Sedov-Taylor {r_st:.4f}, cell {dr:.4f}", flush=True)
    assert abs(r_peak - r_st) <= dr


def test_sedov_configured_t_end(cuda):
    s = setups.sedov(96, 8, 16, t_end=1.0)
    sol = make_solver(s)
    sol.init_from_prim(s.rho, s.u, s.P)
    steps = sol.run(s.t_end, 1 << 30)
    assert sol.t == 1.0
    rho = sol.rho.cpu().numpy()
    profile = rho[2:-2, 2:-2, 2:-2].mean(axis=(0, 1))
    x = orc.regular_nodes(96, 2, s.xmin[0], s.xmax[0])
    xc = ((x[:-1] + x[1:]) / 2)[2:-2]
    k = int(np.argmax(profile))
    r_peak = xc[k]
    # the front: the outermost cell whose density is above halfway between the ambient value and the peak
    r_front = xc[k + int(np.nonzero(profile[k:] > 1.0 + 0.5 * (profile[k] - 1.0))[0][-1])]
    r_st = (1.0 * 1.0 ** 2 / (0.851072 * 1.0)) ** 0.2
    dr = x[3] - x[2]
    print(f"\n[analytic] sedov 96x8x16 t=1.0 after {steps} steps: front {r_front:.4f}, peak {r_peak:.4f}, "
          f"Sedov-Taylor {r_st:.4f}, cell {dr:.4f}", flush=True)
    assert abs(r_front - r_st) <= dr
    assert 0 <= r_st - r_peak <= 2 * dr  # a captured shock's density maximum trails the front by a cell or two
