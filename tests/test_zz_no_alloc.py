"""hclpp_time_step_noalloc (the time step with its reduction scratch in a static device array,
for callers that cannot allocate) against hclpp_time_step through the C-ABI: the same double,
bit for bit, on two calls in a row, the second on a smaller input, plus the oracle's value; and a
range too large for the scratch is refused before anything is launched."""

import ctypes

import numpy as np
import pytest
import torch

import heraclespp_amd as H
from heraclespp_amd import _lib
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _state(o, seed):
    rng = np.random.default_rng(seed)
    rho = 1.0 + 0.5 * rng.random(o.shape3)
    u = rng.uniform(-0.7, 0.7, size=(o.ndim,) + o.shape3)
    P = 0.5 + rng.random(o.shape3)
    return rho, u, P


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _both(g, eos, rng_, rho, u, P):
    """(hclpp_time_step, hclpp_time_step_noalloc) on the same context and arrays."""
    L = _lib.lib()
    c = g.context(eos)
    r = rng_._c()
    args = (c._chk(rho, (), "rho"), c._chk(u, (g.ndim,), "u"), c._chk(P, (), "P"))
    ref = ctypes.c_double()
    _lib.check(L.hclpp_time_step(c.handle, ctypes.byref(r), *args, ctypes.byref(ref)))
    out = torch.full((1,), np.nan, dtype=torch.float64, device="cuda")
    _lib.check(L.hclpp_time_step_noalloc(c.handle, ctypes.byref(r), *args, ctypes.c_void_p(out.data_ptr())))
    torch.cuda.synchronize()
    return ref.value, float(out.item())


@pytest.mark.parametrize("geom", ["Cartesian", "Spherical"])
def test_time_step_noalloc_equals_time_step(cuda, geom):
    lo = (0.05, 0.0, 0.0) if geom == "Spherical" else (0.0, 0.0, 0.0)
    hi = (1.0, np.pi, 2 * np.pi) if geom == "Spherical" else (1.0, 1.0, 1.0)
    eos = H.PerfectGas(1.4, 1.0)
    # first call: a 3D grid over all ghosts (1904 workgroup partials); second: a smaller grid
    for k, n in enumerate([(70, 61, 52), (9, 7, 5)]):
        print(f"[no-alloc] {geom} grid {n}, call {k + 1}", flush=True)
        g = H.Grid(3, n, 2, geom, lo, hi)
        o = orc.Oracle(3, n, geom=geom, xmin=lo, xmax=hi)
        rho, u, P = _state(o, 10 + k)
        want = o.time_step(2, rho, u, P)
        ref, got = _both(g, eos, g.range.all_ghosts(), _t(rho), _t(u), _t(P))
        assert ref == want
        assert got == ref, (got, ref)
        # a sub-range of the same arrays (fewer partials than the call before left in the scratch)
        ref, got = _both(g, eos, g.range.no_ghosts(), _t(rho), _t(u), _t(P))
        assert got == ref == o.time_step(0, rho, u, P)


def test_time_step_noalloc_refuses_a_range_beyond_its_scratch(cuda):
    # 1024 x 1024 x 300 interior cells need 17 x 257 x 304 = 1328176 > 2^20 workgroup partials.  The size
    # check comes before the argument check, so null arrays make this a pure host-side test: if the
    # size check failed to fire, the null check would refuse the call instead (nothing is launched).
    g = H.Grid(3, (1024, 1024, 300), 2, "Cartesian", (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    c = g.context(H.PerfectGas(1.4, 1.0))
    r = g.range.all_ghosts()._c()
    rc = _lib.lib().hclpp_time_step_noalloc(c.handle, ctypes.byref(r), None, None, None, None)
    assert rc == _lib.HCLPP_ERR_ARG
    msg = _lib.lib().hclpp_last_error().decode()
    assert "no-allocation scratch holds 1048576" in msg, msg
