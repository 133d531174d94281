"""Reverse-mode step on the GPU (hclpp_step_adjoint through the C-ABI) and the differentiable
rollout built on it.

* The GPU reverse step equals the host build of the same kernel bodies (which the CPU suite
  pins to finite differences of the oracle's step) on the same inputs: bit-identical for the
  state adjoints, 1e-12 relative for dL/dgamma (a different summation order).
* hclpp_step_dt equals the oracle's fixed-dt step bit for bit, and replaying a recorded CFL
  dt schedule reproduces the CFL-driven run bit for bit.
* Rollout gradients (dL/drho0 along seeded unit directions, dL/dgamma) agree with central
  finite differences (h = 1e-6) of the same fixed-dt rollout; the periodic and binomial
  checkpoint schedules give bit-identical gradients.
"""

import ctypes

import numpy as np
import pytest
import torch

import heraclespp_amd as H
from heraclespp_amd import _lib, setups
from heraclespp_amd.adjoint import RolloutSpec, cfl_schedule, rollout
from tests.adjoint_support import Problem
from tests.test_adjoint_cpu import CASES, _setup

pytestmark = pytest.mark.gpu


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _gpu_ctx(p: Problem):
    grid = H.Grid(p.ndim, p.n, 2, p.geom, p.xmin, p.xmax)
    ctx = grid.context(H.PerfectGas(p.gamma, 1.0), H.UniformGravity(p.g), p.bc, None, 0, p.limiter, p.riemann, "fp64")
    return grid, ctx


def _gpu_adjoint(p, U, lam_new, dt):
    grid, ctx = _gpu_ctx(p)
    Ud = [_t(u) for u in U]
    Ld = [_t(l) for l in lam_new]
    out = [torch.zeros_like(u) for u in Ud]
    dg = torch.zeros(1, dtype=torch.float64, device="cuda")
    args = [_ptr(t) for t in Ud + Ld + out]
    _lib.check(_lib.lib().hclpp_step_adjoint(ctx.handle, dt, *args, _ptr(dg)))
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in out] + [float(dg.cpu()[0])]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_gpu_reverse_step_matches_host_build(cuda, case):
    print(f"[adjoint] case={case[0]}", flush=True)
    p, U, dt, lam_new = _setup(case, seed=3)
    host = p.host_adjoint(*U, *lam_new, dt)
    gpu = _gpu_adjoint(p, U, lam_new, dt)
    it = p.interior()
    for a, b in zip(gpu[:3], host[:3]):
        np.testing.assert_array_equal(a[(Ellipsis,) + it], b[(Ellipsis,) + it])
    assert abs(gpu[3] - host[3]) <= 1e-12 * max(1.0, abs(host[3]))


# Grids wide enough for the gather pass's workgroups (64 x 4 cells) to lie where no ghost copies
# from, which take the gather without the preimage search; the host build searches every cell.
LARGE_CASES = [
    ("2d_large_vanalbada_hllc_mixed_g", 2, [160, 24], "VanAlbada", "HLLC", ["Periodic", "Periodic", "Reflexive", "NullGradient"],
     (0.0, -0.2, 0)),
    ("3d_large_minmod_hll_periodic_g", 3, [136, 12, 10], "Minmod", "HLL", ["Periodic"] * 6, (0, 0, -0.1)),
    ("3d_large_vanleer_hllc_mixed", 3, [140, 13, 11], "VanLeer", "HLLC",
     ["Periodic", "Periodic", "Reflexive", "NullGradient", "Reflexive", "Reflexive"], (0, 0, 0)),
]


@pytest.mark.parametrize("case", LARGE_CASES, ids=[c[0] for c in LARGE_CASES])
def test_gpu_reverse_step_matches_host_build_on_wide_grids(cuda, case):
    print(f"[adjoint] case={case[0]}", flush=True)
    p, U, dt, lam_new = _setup(case, seed=6)
    host = p.host_adjoint(*U, *lam_new, dt)
    gpu = _gpu_adjoint(p, U, lam_new, dt)
    it = p.interior()
    for a, b in zip(gpu[:3], host[:3]):
        np.testing.assert_array_equal(a[(Ellipsis,) + it], b[(Ellipsis,) + it])
    assert abs(gpu[3] - host[3]) <= 1e-12 * max(1.0, abs(host[3]))


ADJ_COMBOS = [(nd, geom, lim, rs) for nd in (1, 2, 3) for geom in ("Cartesian", "Spherical")
              for lim in ("Constant", "VanLeer", "Minmod", "VanAlbada") for rs in ("HLL", "HLLC")]


@pytest.mark.parametrize("nd,geom,lim,rs", ADJ_COMBOS, ids=[f"{c[0]}d-{c[1]}-{c[2]}-{c[3]}" for c in ADJ_COMBOS])
def test_gpu_reverse_step_matches_host_build_every_scheme(cuda, nd, geom, lim, rs):
    """Every scheme combination the reverse step supports (48 kernel instances), each with a seeded
    mix of boundary kinds and gravity: GPU equal to the host build bit for bit."""
    seed = 100 * nd + 10 * (geom == "Spherical") + 2 * ("Constant", "VanLeer", "Minmod", "VanAlbada").index(lim) + (rs == "HLLC")
    rng = np.random.default_rng(seed)
    if geom == "Cartesian":
        n = {1: [11], 2: [9, 7], 3: [6, 5, 4]}[nd]
        kinds = [["Periodic", "Periodic"], ["NullGradient", "NullGradient"], ["Reflexive", "Reflexive"], ["NullGradient", "Reflexive"]]
        bc = sum((kinds[int(rng.integers(len(kinds)))] for _ in range(nd)), [])
        g = tuple(list(rng.uniform(-0.3, 0.3, nd)) + [0.0] * (3 - nd))
        p = Problem(nd, n, limiter=lim, riemann=rs, bc=bc, g=g)
    else:
        n = {1: [11], 2: [8, 6], 3: [5, 4, 4]}[nd]
        g = (float(rng.uniform(-0.3, 0.0)), 0.0, 0.0)  # radial
        p = Problem(nd, n, limiter=lim, riemann=rs, bc=["Reflexive", "NullGradient", "Reflexive", "Reflexive", "Periodic", "Periodic"][:2 * nd],
                    g=g, geom="Spherical", xmin=(0.5, 0.3, 0.0), xmax=(1.5, np.pi - 0.3, 2 * np.pi))
    U = p.random_state(seed)
    dt = p.stable_dt(*U)
    lam_new = [rng.standard_normal(u.shape) for u in U]
    it = p.interior()
    for lam in lam_new:
        mask = np.zeros(lam.shape[-3:], dtype=bool)
        mask[it] = True
        lam[..., ~mask] = 0.0
    host = p.host_adjoint(*U, *lam_new, dt)
    gpu = _gpu_adjoint(p, U, lam_new, dt)
    for a, b in zip(gpu[:3], host[:3]):
        np.testing.assert_array_equal(a[(Ellipsis,) + it], b[(Ellipsis,) + it])
    assert abs(gpu[3] - host[3]) <= 1e-12 * max(1.0, abs(host[3]))


def test_gpu_reverse_step_is_deterministic(cuda):
    p, U, dt, lam_new = _setup(CASES[6], seed=4)
    a = _gpu_adjoint(p, U, lam_new, dt)
    b = _gpu_adjoint(p, U, lam_new, dt)
    for x, y in zip(a[:3], b[:3]):
        np.testing.assert_array_equal(x, y)
    assert a[3] == b[3]


@pytest.mark.parametrize("case", [CASES[0], CASES[2], CASES[6]], ids=[CASES[i][0] for i in (0, 2, 6)])
def test_step_dt_matches_oracle_fixed_step(cuda, case):
    p, U, dt, _ = _setup(case, seed=8)
    want = p.step(*U, dt)
    grid, ctx = _gpu_ctx(p)
    Ud = [_t(u) for u in U]
    out = [torch.zeros_like(u) for u in Ud]
    L = _lib.lib()
    _lib.check(L.hclpp_clock_start(ctx.handle, 0.0, 0, 1e300, 0.4, 1 << 30, *[_ptr(t) for t in Ud]))
    _lib.check(L.hclpp_step_dt(ctx.handle, dt, *[_ptr(t) for t in Ud + out]))
    clk = _lib.Clock()
    _lib.check(L.hclpp_clock_read(ctx.handle, ctypes.byref(clk)))
    assert clk.iter == 1 and clk.status == 0 and clk.t == dt
    it = p.interior()
    for a, b in zip(out, want):
        np.testing.assert_array_equal(a.cpu().numpy()[(Ellipsis,) + it], b[(Ellipsis,) + it])


def test_reverse_step_rejects_unsupported_configs(cuda):
    L = _lib.lib()
    z = torch.zeros(1, dtype=torch.float64, device="cuda")
    args = [_ptr(z)] * 10
    g = H.Grid(2, [8, 8], 2)
    ctx = g.context(H.PerfectGas(1.4, 1.0), None, None, None, 1, "Minmod", "HLL", "fp64")  # a passive scalar
    assert L.hclpp_step_adjoint(ctx.handle, 0.1, *args) == _lib.HCLPP_ERR_UNSUPPORTED
    ctx = g.context(H.PerfectGas(1.4, 1.0), None, None, None, 0, "Minmod", "Low Mach", "fp64")
    assert L.hclpp_step_adjoint(ctx.handle, 0.1, *args) == _lib.HCLPP_ERR_UNSUPPORTED
    ctx = g.context(H.PerfectGas(1.4, 1.0), None, None, None, 0, "Minmod", "HLL", "fp64")
    assert L.hclpp_step_adjoint(ctx.handle, -1.0, *args) == _lib.HCLPP_ERR_ARG


# ---------------------------------------------------------------- rollouts
def _smooth(n, steps):
    s = setups.smooth_euler(n=n, steps=steps)
    grid = H.Grid(2, s.n, 2)
    spec = RolloutSpec(grid, s.reconstruction, s.riemann_solver, s.bc)
    it = s.interior()
    rho0 = _t(s.rho[it])
    u = _t(s.u[(slice(None),) + it])
    P = _t(s.P[it])
    return s, spec, rho0, u, P


def _cons0(rho0, u, P, gamma):
    """prim -> cons in torch (differentiable in rho0 and gamma): E = P / (gamma - 1) + rho |u|^2 / 2."""
    rhou = rho0.unsqueeze(0) * u
    E = P / (gamma - 1) + 0.5 * rho0 * (u * u).sum(0)
    return rho0, rhou, E


def _loss(spec, rho0, u, P, gamma, dts, target, every=8, schedule="periodic", snapshots=4, target_E=None, overlap=True):
    r, m, E = _cons0(rho0, u, P, gamma)
    rT, _, ET = rollout(r, m, E, gamma, spec, dts, checkpoint_every=every, schedule=schedule, snapshots=snapshots,
                        overlap=overlap)
    dV = 1.0 / rho0.numel()
    L = 0.5 * torch.sum((rT - target) ** 2) * dV
    if target_E is not None:
        L = L + 0.5 * torch.sum((ET - target_E) ** 2) * dV
    return L, rT


def _fd_loss_diff(spec, rho0, u, P, gamma, dts, target, v, h, target_E=None):
    """[L(x + h v) - L(x - h v)] / 2h, with the loss difference formed term-wise (no cancellation of L itself)."""
    with torch.no_grad():
        rp, mp, ep = _cons0(rho0 + h * v[0], u, P, gamma + h * v[1])
        rm, mm, em = _cons0(rho0 - h * v[0], u, P, gamma - h * v[1])
        Rp, _, Ep = rollout(rp, mp, ep, gamma + h * v[1], spec, dts)
        Rm, _, Em = rollout(rm, mm, em, gamma - h * v[1], spec, dts)
        dV = 1.0 / rho0.numel()
        d = torch.sum((Rp - Rm) * (Rp + Rm - 2 * target))
        if target_E is not None:
            d = d + torch.sum((Ep - Em) * (Ep + Em - 2 * target_E))
        return float(d * 0.5 * dV / (2 * h))


def test_cfl_schedule_replay_is_bit_identical(cuda):
    s, spec, rho0, u, P = _smooth(48, 30)
    gamma = torch.tensor(1.4, dtype=torch.float64)
    r, m, E = _cons0(rho0, u, P, gamma)
    dts = cfl_schedule(spec, r, m, E, 1.4, 30, s.cfl)
    assert len(dts) == 30 and all(d > 0 for d in dts)
    # the CFL-driven solver from the same conservative state
    solver = H.HydroSolver(spec.grid, H.PerfectGas(1.4, 1.0), reconstruction=s.reconstruction,
                           riemann_solver=s.riemann_solver, bc=s.bc, cfl=s.cfl)
    solver.init_from_prim(_t(s.rho), _t(s.u), _t(s.P))
    it = spec.interior
    solver.rho[it] = r
    solver.rhou[(slice(None),) + it] = m
    solver.E[it] = E
    lib = _lib.lib()
    c = solver.ctx
    _lib.check(lib.hclpp_fill_ghosts(c.handle, _ptr(solver.rho), _ptr(solver.rhou), _ptr(solver.E), ctypes.c_void_p(0)))
    solver.run(1e30, 30)
    rT = rollout(r, m, E, gamma, spec, dts)
    torch.testing.assert_close(rT[0], solver.rho[it], rtol=0, atol=0)
    torch.testing.assert_close(rT[2], solver.E[it], rtol=0, atol=0)


def test_rollout_gradient_matches_finite_differences(cuda):
    """Smooth flow with a pressure wave (so gamma enters the dynamics through the sound speed),
    loss on density and energy."""
    s, spec, rho0, u, P = _smooth(48, 40)
    X, _ = np.meshgrid(np.arange(48) / 48, np.arange(48) / 48)
    P = P + 0.1 * _t(np.cos(2 * np.pi * X))
    gamma = torch.tensor(1.4, dtype=torch.float64, requires_grad=True)
    r, m, E = _cons0(rho0, u, P, gamma.detach())
    dts = cfl_schedule(spec, r, m, E, 1.4, 40, s.cfl)
    rng = np.random.default_rng(0)
    X, Y = np.meshgrid(np.arange(48) / 48, np.arange(48) / 48)
    tgt0 = rho0 + 0.02 * _t(np.cos(2 * np.pi * (2 * X + Y)))
    with torch.no_grad():
        rt, mt, et = _cons0(tgt0, u, P, gamma.detach())
        target, _, target_E = rollout(rt, mt, et, gamma.detach(), spec, dts)
    x = rho0.clone().requires_grad_(True)
    L, _ = _loss(spec, x, u, P, gamma, dts, target, target_E=target_E)
    L.backward()
    g_rho, g_gam = x.grad, float(gamma.grad)
    h = 1e-6
    worst = 0.0
    for k in range(6):
        v = _t(rng.standard_normal((48, 48)))
        v = v / torch.linalg.norm(v)
        fd = _fd_loss_diff(spec, rho0, u, P, gamma.detach(), dts, target, (v, 0.0), h, target_E)
        ad = float(torch.sum(g_rho * v))
        rel = abs(fd - ad) / max(abs(fd), 1e-30)
        worst = max(worst, rel)
        print(f"[adjoint-fd] dir={k} fd={fd:.10e} adjoint={ad:.10e} rel={rel:.2e}", flush=True)
    fdg = _fd_loss_diff(spec, rho0, u, P, gamma.detach(), dts, target, (torch.zeros_like(rho0), 1.0), h, target_E)
    relg = abs(fdg - g_gam) / max(abs(fdg), 1e-30)
    print(f"[adjoint-fd] gamma fd={fdg:.10e} adjoint={g_gam:.10e} rel={relg:.2e}", flush=True)
    assert worst <= 1e-4
    assert relg <= 1e-4


def test_spherical_rollout_gradient_matches_finite_differences(cuda):
    """2D spherical (r, theta) rollout: the geometric sources of the predictor and of the update are
    differentiated; dL/drho0 along seeded unit directions and dL/dgamma against central differences.
    A random unit direction sees about |grad| / sqrt(N) of the gradient (N = 288 cells here), so each
    difference is measured against the larger of |fd| and |grad| / 10: a direction whose derivative
    happens to be small does not turn the finite difference's own error into a large ratio."""
    nr, nth, steps = 24, 12, 20
    lo, hi = (0.5, 0.3, 0.0), (1.5, np.pi - 0.3, 1.0)
    bc = ["Reflexive", "NullGradient", "Reflexive", "Reflexive"]
    grid = H.Grid(2, [nr, nth], 2, "Spherical", lo, hi)
    spec = RolloutSpec(grid, "VanLeer", "HLLC", bc)
    R, TH = np.meshgrid(np.linspace(0, 1, nr), np.linspace(0, 1, nth))
    R, TH = R[None], TH[None]  # interior arrays are [z, y, x], z of extent 1
    rho0 = _t(1.0 + 0.2 * np.sin(2 * np.pi * R) * np.cos(np.pi * TH))
    u = _t(np.stack([0.1 + 0.05 * np.cos(2 * np.pi * TH), 0.05 * np.sin(2 * np.pi * R)]))
    P = _t(1.0 + 0.1 * np.cos(2 * np.pi * (R + TH)))
    gamma = torch.tensor(1.4, dtype=torch.float64, requires_grad=True)
    r, m, E = _cons0(rho0, u, P, gamma.detach())
    dts = cfl_schedule(spec, r, m, E, 1.4, steps, 0.4)
    with torch.no_grad():
        rt, mt, et = _cons0(rho0 + 0.02 * _t(np.cos(np.pi * (R + 2 * TH))), u, P, gamma.detach())
        target, _, target_E = rollout(rt, mt, et, gamma.detach(), spec, dts)
    x = rho0.clone().requires_grad_(True)
    L, _ = _loss(spec, x, u, P, gamma, dts, target, every=5, target_E=target_E)
    L.backward()
    g_rho, g_gam = x.grad, float(gamma.grad)
    gnorm = float(torch.linalg.norm(g_rho))
    rng = np.random.default_rng(1)
    worst = 0.0
    for k in range(4):
        v = _t(rng.standard_normal((1, nth, nr)))
        v = v / torch.linalg.norm(v)
        fd = _fd_loss_diff(spec, rho0, u, P, gamma.detach(), dts, target, (v, 0.0), 1e-6, target_E)
        ad = float(torch.sum(g_rho * v))
        rel = abs(fd - ad) / max(abs(fd), 0.1 * gnorm, 1e-30)
        worst = max(worst, rel)
        print(f"[adjoint-fd spherical] dir={k} fd={fd:.10e} adjoint={ad:.10e} rel={rel:.2e}", flush=True)
    fdg = _fd_loss_diff(spec, rho0, u, P, gamma.detach(), dts, target, (torch.zeros_like(rho0), 1.0), 1e-6, target_E)
    relg = abs(fdg - g_gam) / max(abs(fdg), 1e-30)
    print(f"[adjoint-fd spherical] gamma fd={fdg:.10e} adjoint={g_gam:.10e} rel={relg:.2e}", flush=True)
    assert worst <= 1e-4
    assert relg <= 1e-4


def test_checkpoint_schedules_agree_bitwise(cuda):
    s, spec, rho0, u, P = _smooth(32, 23)
    gamma0 = torch.tensor(1.4, dtype=torch.float64)
    r, m, E = _cons0(rho0, u, P, gamma0)
    dts = cfl_schedule(spec, r, m, E, 1.4, 23, s.cfl)
    target = torch.ones_like(rho0)
    grads = []
    # periodic with the next segment recomputed on a second stream during each reversal (overlap),
    # and without; single-segment periodic; binomial with 3 and 1 snapshots; every state kept (one
    # reverse call over all steps), chosen explicitly and by the memory budget
    for schedule, every, snaps, overlap in (("periodic", 5, 0, True), ("periodic", 5, 0, False), ("periodic", 4, 0, True),
                                            ("periodic", 23, 0, True), ("binomial", 1, 3, False), ("binomial", 1, 1, False),
                                            ("store_all", 5, 0, False), ("auto", 5, 0, True)):
        x = rho0.clone().requires_grad_(True)
        gamma = gamma0.clone().requires_grad_(True)
        L, _ = _loss(spec, x, u, P, gamma, dts, target, every=every, schedule=schedule, snapshots=snaps, overlap=overlap)
        L.backward()
        grads.append((x.grad.clone(), gamma.grad.clone()))
    for g in grads[1:]:
        assert torch.equal(g[0], grads[0][0])
        assert torch.equal(g[1], grads[0][1])


def test_step_writes_every_cell_of_its_output(cuda):
    """The store-all schedule keeps step outputs in uninitialised memory: a step must write the whole
    ghost-inclusive output.  Outputs poisoned with NaN equal outputs that start at zero, bit for bit."""
    from heraclespp_amd.adjoint import StepEngine

    for bc in (None, ["Periodic", "Periodic", "NullGradient", "Reflexive"]):
        s, spec, rho0, u, P = _smooth(24, 3)
        if bc is not None:
            spec = RolloutSpec(spec.grid, s.reconstruction, s.riemann_solver, bc)
        gamma = torch.tensor(1.4, dtype=torch.float64)
        r, m, E = _cons0(rho0, u, P, gamma)
        eng = StepEngine(spec, 1.4)
        dts = [1e-3] * 3
        outs = []
        for poison in (0.0, float("nan")):
            states = [eng.load(r, m, E)] + eng.new_states(3)
            for U in states[1:]:
                for t in U:
                    t.fill_(poison)
            eng.start(states[0])
            eng.steps(states, dts)
            eng.check_clock(3)
            outs.append([t.clone() for U in states[1:] for t in U])
        for a, b in zip(*outs):
            assert torch.equal(a, b)


def test_auto_schedule_follows_the_memory_budget(cuda):
    """schedule="auto": every state kept (no recomputation) when the N + 1 states fit in the budget,
    periodic checkpoints (each segment recomputed once) when they do not; the same gradient."""
    from heraclespp_amd.adjoint import _Rollout, state_bytes

    s, spec, rho0, u, P = _smooth(32, 12)
    gamma0 = torch.tensor(1.4, dtype=torch.float64)
    r, m, E = _cons0(rho0, u, P, gamma0)
    dts = cfl_schedule(spec, r, m, E, 1.4, 12, s.cfl)
    target = torch.ones_like(rho0)
    need = 13 * state_bytes(spec)
    grads = []
    for budget, recomputed in ((need, 0), (need - 1, 12 - 3)):
        x = rho0.clone().requires_grad_(True)
        r, m, E = _cons0(x, u, P, gamma0)
        rT, _, _ = rollout(r, m, E, gamma0, spec, dts, checkpoint_every=5, memory_budget=budget)
        (0.5 * torch.sum((rT - target) ** 2)).backward()
        assert _Rollout.last_backward == {"reverse_steps": 12, "recomputed_steps": recomputed}
        grads.append(x.grad.clone())
    assert torch.equal(grads[0], grads[1])


def test_batched_reverse_steps_equal_single_steps(cuda):
    """hclpp_steps_adjoint over 6 steps (after the first, each step's primitives come from the
    gather pass of the step before, and its dgamma partials are folded by the faces pass of the
    step after) equals 6 hclpp_step_adjoint calls bit for bit, dgamma included; a batch that stops
    on a null state keeps the steps it ran, dgamma included, like a batch of just those steps."""
    from heraclespp_amd.adjoint import StepEngine, _ptr_array

    s, spec, rho0, u, P = _smooth(32, 6)
    gamma = torch.tensor(1.4, dtype=torch.float64)
    r, m, E = _cons0(rho0, u, P, gamma)
    dts = cfl_schedule(spec, r, m, E, 1.4, 6, s.cfl)
    n = len(dts)
    eng = StepEngine(spec, 1.4)
    states = [eng.load(r, m, E)] + [eng.new_state() for _ in range(n)]
    eng.start(states[0])
    eng.steps(states, dts)
    rng = np.random.default_rng(5)
    seed = eng.new_state()
    for t in seed:
        t.copy_(_t(rng.standard_normal(tuple(t.shape))))

    def batched(lo, hi, broken=None):
        """Reverse steps hi-1 .. lo in one call (state `broken`, if given, passed as null)."""
        lams = [eng.new_state() for _ in range(lo, hi)] + [tuple(t.clone() for t in seed)]
        dg = torch.zeros(1, dtype=torch.float64, device="cuda")
        cols = []
        for v in range(3):
            arr = _ptr_array([U[v] for U in states[lo:hi]])
            if broken is not None:
                arr[broken - lo] = None
            cols.append(arr)
        lcols = [_ptr_array([L[v] for L in lams]) for v in range(3)]
        rc = _lib.lib().hclpp_steps_adjoint(eng.ctx.handle, hi - lo, (ctypes.c_double * (hi - lo))(*dts[lo:hi]), *cols, *lcols,
                                            _ptr(dg))
        torch.cuda.synchronize()
        return rc, lams, dg

    rc, lams, dg = batched(0, n)
    assert rc == _lib.HCLPP_OK
    lam = tuple(t.clone() for t in seed)
    dg1 = torch.zeros(1, dtype=torch.float64, device="cuda")
    for k in range(n - 1, -1, -1):
        out = eng.new_state()
        eng.vjp(states[k], lam, out, dg1, dts[k])
        lam = out
    torch.cuda.synchronize()
    for a, b in zip(lams[0], lam):
        assert torch.equal(a, b)
    assert torch.equal(dg, dg1)

    # state 2 missing: steps 5, 4, 3 need it only as the next state of step 3, so steps 5 and 4 run
    rc, lams_b, dg_b = batched(0, n, broken=2)
    assert rc == _lib.HCLPP_ERR_ARG
    rc, lams_p, dg_p = batched(4, n)
    assert rc == _lib.HCLPP_OK
    for a, b in zip(lams_b[4], lams_p[0]):
        assert torch.equal(a, b)
    assert torch.equal(dg_b, dg_p)


def test_diff_euler_1024_gradient_64_directions(cuda):
    """The differentiable-Euler configuration at full size: 1024^2 periodic, VanAlbada + HLLC,
    400 steps under the default schedule (every state kept; the periodic schedule gives the same
    gradient bit for bit, test_checkpoint_schedules_agree_bitwise); loss 1/2 sum (rho_T - rho_tgt)^2 dV with the target the
    forward of a seeded perturbation of rho0; dL/drho0 along 64 seeded unit directions against
    central differences with h = 1e-6 (relative agreement <= 1e-4).

    dL/dgamma: this loss is independent of gamma (uniform u and P make the density a pure contact,
    carried by HLLC independently of the wave speeds), so the adjoint and the finite difference must
    both vanish to rounding; the gamma path is checked against central differences on the same
    problem with an energy term added to the loss and a target whose mean density differs."""
    from heraclespp_amd.adjoint import diff_euler_problem

    prob = diff_euler_problem(1024, 400, seed=0)
    L, g_rho, g_gam = prob.gradient(every=20)
    print(f"[diff-euler] loss={float(L):.10e} |dL/drho0|={float(torch.linalg.norm(g_rho)):.6e} dL/dgamma={g_gam:.3e}",
          flush=True)
    rng = np.random.Generator(np.random.Philox(1))
    worst = 0.0
    for k in range(64):
        v = _t(rng.standard_normal((1024, 1024)))
        v = v / torch.linalg.norm(v)
        fd = prob.fd(v, 0.0, 1e-6)
        ad = float(torch.sum(g_rho * v))
        rel = abs(fd - ad) / max(abs(fd), 1e-300)
        worst = max(worst, rel)
        if k < 4 or rel > 1e-5:
            print(f"[diff-euler] dir={k} fd={fd:.10e} adjoint={ad:.10e} rel={rel:.2e}", flush=True)
    fdg0 = prob.fd(torch.zeros_like(prob.rho0), 1.0, 1e-6)
    print(f"[diff-euler] worst rel over 64 directions={worst:.2e}; density-loss dL/dgamma: fd={fdg0:.3e} "
          f"adjoint={g_gam:.3e}", flush=True)
    assert worst <= 1e-4
    # both vanish: |dL/dgamma| is below 1e-6 of the typical directional derivative
    scale = float(torch.linalg.norm(g_rho))
    assert abs(g_gam) <= 1e-6 * scale and abs(fdg0) <= 1e-6 * scale

    probE = diff_euler_problem(1024, 400, seed=0, energy_weight=1.0, target_offset=0.01)
    _, gE_rho, gE_gam = probE.gradient(every=20)
    fdg = probE.fd(torch.zeros_like(probE.rho0), 1.0, 1e-6)
    relg = abs(fdg - gE_gam) / max(abs(fdg), 1e-300)
    print(f"[diff-euler] density+energy loss: dL/dgamma fd={fdg:.10e} adjoint={gE_gam:.10e} rel={relg:.2e}", flush=True)
    assert relg <= 1e-4
