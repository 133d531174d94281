"""Full-length runs of the configured workloads against the CPU oracle (opt-in: the oracle needs
minutes of host time, so these run only with HCLPP_LONG_RUNS=1).

* 2D Kelvin-Helmholtz 2048^2, VanLeer + HLLC, CFL 0.4, all 2000 configured steps: the fused HIP
  step's conservative state is bit-identical to the oracle's after the whole run, and the clock
  (t, step count) agrees exactly.
"""

import os
import time

import numpy as np
import pytest

from heraclespp_amd import setups
from tests.helpers import make_oracle, make_solver

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skipif(os.environ.get("HCLPP_LONG_RUNS") != "1", reason="full-length run: set HCLPP_LONG_RUNS=1"),
]


def test_kelvin_helmholtz_2048_full_run_is_bit_identical(cuda):
    s = setups.kelvin_helmholtz(2048, seed=0)
    assert s.max_iter == 2000
    t0 = time.perf_counter()
    sol = make_solver(s)
    sol.init_from_prim(s.rho, s.u, s.P)
    steps = sol.run(s.t_end, s.max_iter)
    t1 = time.perf_counter()
    print(f"\n[long-run] KH 2048^2: {steps} GPU steps in {t1 - t0:.1f} s, t = {sol.t:.9g}", flush=True)
    o = make_oracle(s)
    o.init_from_prim(s.rho, s.u, s.P)
    while o.iter < s.max_iter:
        _, ex = o.step(s.cfl, s.t_end, s.max_iter)
        if o.iter % 100 == 0:
            print(f"[long-run] oracle step {o.iter} ({time.perf_counter() - t1:.0f} s)", flush=True)
        if ex:
            break
    assert steps == o.iter == 2000
    assert sol.t == o.t
    for name, gpu in (("rho", sol.rho), ("rhou", sol.rhou), ("E", sol.E)):
        np.testing.assert_array_equal(gpu.cpu().numpy(), o.state(name), err_msg=name)
