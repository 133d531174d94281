"""Child process of tests/test_gpu_long_runs.py: the 3D spherical Sedov run (128 x 32 x 64, HLL +
VanLeer) through the library named by HCLPP_LIB (the IEEE-division build), writing the state
every EVERY steps as OUT/state_<step>.npz (written under a temporary name, then renamed).

Usage: HCLPP_LIB=.../libhclpp_hip_ieee.so python -m tests.ieee_division_run OUT STEPS EVERY
"""

import os
import sys

import numpy as np


def main() -> None:
    out, steps, every = sys.argv[1], int(sys.argv[2]), int(sys.argv[3])
    import torch

    from heraclespp_amd import _lib, setups
    from tests.helpers import make_solver

    if not _lib.LIB_PATH.endswith("libhclpp_hip_ieee.so"):
        raise SystemExit(f"expected the IEEE-division library, HCLPP_LIB gives {_lib.LIB_PATH}")
    torch.cuda.set_device(0)
    s = setups.sedov(128, 32, 64, t_end=1.0)
    sol = make_solver(s)
    sol.init_from_prim(s.rho, s.u, s.P)
    for step in range(every, steps + 1, every):
        sol.run(s.t_end, step)
        tmp = os.path.join(out, f"state_{step}.tmp.npz")
        np.savez(tmp, rho=sol.rho.cpu().numpy(), rhou=sol.rhou.cpu().numpy(), E=sol.E.cpu().numpy(), iter=sol.iter,
                 t=sol.t)
        os.replace(tmp, os.path.join(out, f"state_{step}.npz"))
    print(f"IEEE-division run: {sol.iter} steps, t = {sol.t!r}", flush=True)


if __name__ == "__main__":
    main()
