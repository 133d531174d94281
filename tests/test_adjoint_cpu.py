"""Reverse-mode step, checked on the CPU: the host build of the GPU kernel bodies
(adjoint_step.hpp) against central finite differences of the oracle's fixed-dt step
(the reference kernel sequence), in Cartesian and spherical geometry (the spherical cases
differentiate the geometric sources of the predictor and of the update,
source_terms.hpp:18-77).  The adjoint holds dt fixed, so the finite differences
do too.  Non-smooth points (limiter sign tests, wave-speed min/max, HLLC branch choices)
are differentiated on the branch the state takes; a branch decision flipping inside the
FD step would show as a mismatch far above the 1e-8 tolerance, and the seeded states
used here have none."""

import numpy as np
import pytest

from tests.adjoint_support import Problem

SPH = dict(geom="Spherical", xmin=(0.5, 0.3, 0.0), xmax=(1.5, np.pi - 0.3, 2 * np.pi))
SPH_BC = ["Reflexive", "NullGradient", "Reflexive", "Reflexive", "Periodic", "Periodic"]

CASES = [
    # id, ndim, n, limiter, riemann, bc, g[, x-velocity offset[, geometry]]
    ("1d_minmod_hll_ng_refl_g", 1, [12], "Minmod", "HLL", ["NullGradient", "Reflexive"], (-0.3, 0, 0)),
    ("1d_vanleer_hllc_periodic", 1, [10], "VanLeer", "HLLC", ["Periodic", "Periodic"], (0, 0, 0)),
    ("2d_vanalbada_hllc_mixed_g", 2, [9, 7], "VanAlbada", "HLLC", ["Periodic", "Periodic", "Reflexive", "NullGradient"],
     (0.0, -0.2, 0)),
    ("2d_vanleer_hll_periodic", 2, [8, 8], "VanLeer", "HLL", ["Periodic"] * 4, (0.1, 0.05, 0)),
    ("2d_constant_hllc_nullgrad", 2, [6, 5], "Constant", "HLLC", ["NullGradient"] * 4, (0, 0, 0)),
    ("2d_minmod_hllc_reflexive", 2, [7, 6], "Minmod", "HLLC", ["Reflexive", "NullGradient", "Reflexive", "Reflexive"],
     (0, 0, 0)),
    ("3d_minmod_hllc_mixed_g", 3, [5, 4, 4], "Minmod", "HLLC",
     ["Periodic", "Periodic", "Reflexive", "NullGradient", "Reflexive", "Reflexive"], (0, 0, -0.1)),
    ("3d_vanalbada_hll_periodic", 3, [4, 4, 5], "VanAlbada", "HLL", ["Periodic"] * 6, (0, 0, 0)),
    # supersonic along x (u ~ 2.5 > c): HLL's clipped wave speeds and HLLC's upwind-flux branch
    ("2d_vanleer_hllc_supersonic", 2, [8, 6], "VanLeer", "HLLC", ["Periodic"] * 4, (0, 0, 0), 2.5),
    ("1d_minmod_hll_supersonic", 1, [10], "Minmod", "HLL", ["Periodic"] * 2, (0, 0, 0), -2.5),
    # spherical (r, theta, phi): reflexive r_min and poles, free r_max, periodic phi
    ("1d_sph_vanleer_hllc", 1, [12], "VanLeer", "HLLC", SPH_BC[:2], (0, 0, 0), 0.0, SPH),
    ("2d_sph_minmod_hll_g", 2, [8, 6], "Minmod", "HLL", SPH_BC[:4], (-0.2, 0, 0), 0.0, SPH),
    ("3d_sph_vanalbada_hllc", 3, [5, 4, 4], "VanAlbada", "HLLC", SPH_BC, (0, 0, 0), 0.0, SPH),
    ("3d_sph_vanleer_hll", 3, [4, 5, 4], "VanLeer", "HLL", SPH_BC, (0, 0, 0), 0.0, SPH),
]


def _dot(p, a, b):
    it = p.interior()
    return float(sum(np.sum(x[(Ellipsis,) + it] * y[(Ellipsis,) + it]) for x, y in zip(a, b)))


def _setup(case, seed):
    _, nd, n, lim, rs, bc, g = case[:7]
    p = Problem(nd, n, limiter=lim, riemann=rs, bc=bc, g=g, **(case[8] if len(case) > 8 else {}))
    U = p.random_state(seed, u0=case[7] if len(case) > 7 else 0.0)
    dt = p.stable_dt(*U)
    rng = np.random.default_rng(seed + 100)
    lam_new = [rng.standard_normal(u.shape) for u in U]
    it = p.interior()
    for l in lam_new:  # only interior values of the output enter the loss
        mask = np.zeros(l.shape[-3:], dtype=bool)
        mask[it] = True
        l[..., ~mask] = 0.0
    return p, U, dt, lam_new


def _interior_direction(p, U, rng):
    V = []
    it = p.interior()
    for u in U:
        v = np.zeros_like(u)
        v[(Ellipsis,) + it] = rng.standard_normal(v[(Ellipsis,) + it].shape)
        V.append(v)
    return V


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_host_adjoint_matches_fd_of_oracle_step(case):
    p, U, dt, lam_new = _setup(case, seed=3)
    lam = p.host_adjoint(*U, *lam_new, dt)
    rng = np.random.default_rng(7)
    h = 1e-6
    for _ in range(3):
        V = _interior_direction(p, U, rng)
        up = p.step(*[u + h * v for u, v in zip(U, V)], dt)
        um = p.step(*[u - h * v for u, v in zip(U, V)], dt)
        fd = (_dot(p, lam_new, up) - _dot(p, lam_new, um)) / (2 * h)
        ad = _dot(p, lam[:3], V)
        assert abs(fd - ad) <= 1e-8 * max(1.0, abs(fd)), (fd, ad)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_host_adjoint_gamma_matches_fd(case):
    p, U, dt, lam_new = _setup(case, seed=5)
    *_, dg = p.host_adjoint(*U, *lam_new, dt)
    h = 1e-6
    up = p.step(*U, dt, gamma=p.gamma + h)
    um = p.step(*U, dt, gamma=p.gamma - h)
    fd = (_dot(p, lam_new, up) - _dot(p, lam_new, um)) / (2 * h)
    assert abs(fd - dg) <= 2e-8 * max(1.0, abs(fd)), (fd, dg)


@pytest.mark.parametrize("case", [CASES[0], CASES[2], CASES[5], CASES[8], CASES[11], CASES[12]],
                         ids=[CASES[i][0] for i in (0, 2, 5, 8, 11, 12)])
def test_host_adjoint_every_cell_and_component(case):
    """Full gradient, one unit direction per interior value: pins each boundary cell's ghost preimages."""
    p, U, dt, lam_new = _setup(case, seed=11)
    lam = p.host_adjoint(*U, *lam_new, dt)
    h = 1e-6
    it = p.interior()
    for a, u in enumerate(U):
        sub = u[(Ellipsis,) + it]
        for idx in np.ndindex(sub.shape):
            V = [np.zeros_like(w) for w in U]
            view = V[a][(Ellipsis,) + it]
            view[idx] = 1.0
            up = p.step(*[w + h * v for w, v in zip(U, V)], dt)
            um = p.step(*[w - h * v for w, v in zip(U, V)], dt)
            fd = (_dot(p, lam_new, up) - _dot(p, lam_new, um)) / (2 * h)
            ad = lam[a][(Ellipsis,) + it][idx]
            assert abs(fd - ad) <= 1e-7 * max(1.0, abs(fd)), (a, idx, fd, ad)


def test_host_adjoint_is_linear_in_the_output_adjoint():
    p, U, dt, lam_new = _setup(CASES[2], seed=13)
    a = p.host_adjoint(*U, *lam_new, dt)
    b = p.host_adjoint(*U, *[2.5 * l for l in lam_new], dt)
    for x, y in zip(a[:3], b[:3]):
        np.testing.assert_allclose(2.5 * x, y, rtol=1e-13, atol=1e-13)
    assert abs(2.5 * a[3] - b[3]) <= 1e-12 * max(1.0, abs(b[3]))


# ---------------------------------------------------------------- checkpoint schedules (host logic)
class _CountingEngine:
    """Stands in for a StepEngine: a state is its step index, a step adds one."""

    def __init__(self):
        import torch

        self.torch = torch
        self.count = 0

    def new_state(self):
        return (self.torch.zeros(1, dtype=self.torch.float64),)

    def step(self, src, dst, dt):
        dst[0].copy_(src[0] + 1)
        self.count += 1


def _reverse_order(run):
    seen = []

    def reverse(U, n):
        seen.append((int(U[0].item()), n))

    state = {"forward_steps": 0}
    run(reverse, state)
    return seen, state["forward_steps"]


@pytest.mark.parametrize("N,every", [(1, 20), (7, 3), (40, 20), (41, 20), (23, 23)])
def test_periodic_schedule_reverses_every_step_from_its_state(N, every):
    from heraclespp_amd.adjoint import _reverse_periodic

    eng = _CountingEngine()
    kept = {s: (eng.torch.full((1,), float(s), dtype=eng.torch.float64),) for s in range(0, N, every)}
    seen, fwd = _reverse_order(lambda rev, st: _reverse_periodic(eng, kept, [0.1] * N, every, rev, st))
    assert seen == [(n, n) for n in range(N - 1, -1, -1)]
    assert fwd == N - len(kept)  # each segment recomputed once, its last state never needed


@pytest.mark.parametrize("N,snaps", [(1, 2), (10, 0), (10, 1), (23, 3), (55, 3), (100, 4), (400, 8)])
def test_binomial_schedule_reverses_every_step_within_the_revolve_bound(N, snaps):
    import math

    from heraclespp_amd.adjoint import _reverse_binomial

    eng = _CountingEngine()
    U0 = (eng.torch.zeros(1, dtype=eng.torch.float64),)
    seen, fwd = _reverse_order(lambda rev, st: _reverse_binomial(eng, U0, 0, N, [0.1] * N, snaps, rev, st))
    assert seen == [(n, n) for n in range(N - 1, -1, -1)]
    if snaps > 0:
        t = 0
        while math.comb(snaps + t, snaps) < N:
            t += 1
        assert fwd <= t * N  # at most t sweeps of recomputation
    else:
        assert fwd == N * (N - 1) // 2


class _BatchedEngine(_CountingEngine):
    """Batched forms: steps() over a list of states, and a reverse map lam_k = 2 lam_{k+1} + state_k."""

    def steps(self, states, dts):
        for k in range(len(dts)):
            self.step(states[k], states[k + 1], dts[k])

    def vjps(self, states, lams, dgam, dts):
        for k in range(len(dts) - 1, -1, -1):
            lams[k][0].copy_(2 * lams[k + 1][0] + states[k][0])
            dgam += 1


@pytest.mark.parametrize("N,every", [(1, 20), (7, 3), (40, 20), (41, 20), (23, 23), (9, 2)])
def test_batched_segments_chain_the_adjoint_buffers(N, every):
    """The batched reverse pass gives the same result as the per-step recursion lam_k = 2 lam_{k+1} + k."""
    from heraclespp_amd.adjoint import _reverse_periodic, _reverse_segment

    torch = pytest.importorskip("torch")
    eng = _BatchedEngine()
    kept = {s: (torch.full((1,), float(s), dtype=torch.float64),) for s in range(0, N, every)}
    dgam = torch.zeros(1, dtype=torch.float64)
    state = {"lam": (torch.full((1,), 1.0, dtype=torch.float64),), "lam2": (torch.zeros(1, dtype=torch.float64),),
             "reverse_steps": 0, "forward_steps": 0}
    _reverse_periodic(eng, kept, [0.1] * N, every, None, state,
                      lambda seg, s, e: _reverse_segment(eng, seg, s, e, [0.1] * N, dgam, state))
    want = 1.0
    for k in range(N - 1, -1, -1):
        want = 2 * want + k
    assert float(state["lam"][0]) == want
    assert state["reverse_steps"] == N and float(dgam) == N
    assert state["forward_steps"] == N - len(kept)


@pytest.mark.parametrize("N", [1, 2, 7, 40])
def test_store_all_reverse_is_one_segment_over_every_kept_state(N):
    """The store-all backward: all N states kept, one reverse call over them, no recomputation."""
    from heraclespp_amd.adjoint import _reverse_segment

    torch = pytest.importorskip("torch")
    eng = _BatchedEngine()
    states = [(torch.zeros(1, dtype=torch.float64),)] + [eng.new_state() for _ in range(N)]
    eng.steps(states, [0.1] * N)
    dgam = torch.zeros(1, dtype=torch.float64)
    state = {"lam": (torch.full((1,), 1.0, dtype=torch.float64),), "lam2": (torch.zeros(1, dtype=torch.float64),),
             "reverse_steps": 0, "forward_steps": 0}
    _reverse_segment(eng, states[:N], 0, N, [0.1] * N, dgam, state)
    want = 1.0
    for k in range(N - 1, -1, -1):
        want = 2 * want + k
    assert float(state["lam"][0]) == want
    assert state["reverse_steps"] == N and float(dgam) == N and state["forward_steps"] == 0


def test_schedule_pick_follows_the_memory_budget():
    from heraclespp_amd import hydro as H
    from heraclespp_amd.adjoint import RolloutSpec, pick_schedule, state_bytes

    spec = RolloutSpec(H.Grid(2, [32, 16, 1], 2), "VanLeer", "HLLC", ["Periodic"] * 4)
    one = state_bytes(spec)
    assert one == 36 * 20 * 4 * 8  # ghost-inclusive cells x (rho, 2 momenta, E) x 8 B
    assert pick_schedule(spec, 10, 11 * one) == "store_all"
    assert pick_schedule(spec, 10, 11 * one - 1) == "periodic"
    spec3 = RolloutSpec(H.Grid(3, [8, 6, 4], 2), "Minmod", "HLL", ["Periodic"] * 6)
    assert state_bytes(spec3) == 12 * 10 * 8 * 5 * 8
