"""FP32-storage / FP64-recovery fused step against the FP64 path (bit-identical to the oracle),
per configuration: the density relative L1 after N steps must stay below the bound in the table.
These are accuracy bounds, not parity: the mixed path rounds every stored value to FP32."""

import numpy as np
import pytest

from heraclespp_amd import setups
from tests.helpers import make_solver, rel_l1

pytestmark = pytest.mark.gpu

CASES = {
    # name: (setup, steps, bound on rel L1 of rho)
    "sod1d": (lambda: setups.sod(1000), None, 1e-5),
    "rayleigh_taylor3d": (lambda: setups.rayleigh_taylor(32, seed=1), 60, 1e-5),
    "sedov3d": (lambda: setups.sedov(32, 16, 32), 40, 1e-5),
    "kelvin_helmholtz2d": (lambda: setups.kelvin_helmholtz(128, seed=2), 100, 1e-5),
    "smooth2d": (lambda: setups.smooth_euler(128, seed=3), 100, 1e-5),
}


@pytest.mark.parametrize("name", list(CASES))
def test_fp32_storage_accuracy(cuda, name):
    make, nsteps, bound = CASES[name]
    s = make()
    n = nsteps if nsteps is not None else 1 << 30
    out = {}
    for prec in ("fp64", "fp32"):
        sol = make_solver(s, precision=prec)
        sol.init_from_prim(s.rho, s.u, s.P)
        steps = sol.run(s.t_end, n)
        out[prec] = (steps, sol.t, sol.rho.cpu().numpy()[s.interior()], sol.E.cpu().numpy()[s.interior()])
    l1 = rel_l1(out["fp32"][2], out["fp64"][2])
    l1E = rel_l1(out["fp32"][3], out["fp64"][3])
    print(f"\n[mixed] {name}: steps fp64={out['fp64'][0]} fp32={out['fp32'][0]} "
          f"t={out['fp64'][1]:.6g}/{out['fp32'][1]:.6g} relL1(rho)={l1:.3e} relL1(E)={l1E:.3e}")
    if nsteps is not None:
        assert out["fp32"][0] == out["fp64"][0] == nsteps
    assert l1 <= bound, l1
