"""The reference's own unit tests (tests/*.cpp), restated against the oracle
and the host mirror: known-answer values, Riemann consistency/symmetry and
limiter properties.  EXPECT_DOUBLE_EQ means "within 4 ulps" (gtest)."""

import json
import math
import os
import sys

import numpy as np
import pytest

from oracle import oracle as orc

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_kat.json")
KAT = json.load(open(GOLDEN))


def ulps(a, b):
    a, b = float(a), float(b)
    if a == b:
        return 0
    ia = np.array([a]).view(np.int64)[0]
    ib = np.array([b]).view(np.int64)[0]
    if (ia < 0) != (ib < 0):
        return abs(ia) + abs(ib)
    return abs(int(ia) - int(ib))


def double_eq(a, b):
    return ulps(a, b) <= 4


# ---------------------------------------------------------------- tests/euler_equations.cpp
def test_euler_known_values():
    k = KAT["euler_equations"]
    assert double_eq(orc.ek_prim(k["ek_prim"]["prim"]), k["ek_prim"]["expected"])
    assert double_eq(orc.ek_cons(k["ek_cons"]["cons"]), k["ek_cons"]["expected"])
    f = orc.flux_prim(k["flux_prim"]["prim"], 0, k["gamma"], k["mmw"])
    for got, want in zip(f, k["flux_prim"]["expected"]):
        assert double_eq(got, want)
    f = orc.flux_cons(k["flux_cons"]["cons"], 0, k["gamma"], k["mmw"])
    for got, want in zip(f, k["flux_cons"]["expected"]):
        assert double_eq(got, want)
    prim = k["roundtrip"]["prim"]
    back = orc.to_prim(orc.to_cons(prim, k["gamma"], k["mmw"]), k["gamma"], k["mmw"])
    for got, want in zip(back, prim):
        assert double_eq(got, want)


# ---------------------------------------------------------------- tests/riemann_solvers.cpp
@pytest.mark.parametrize("solver", ["HLL", "HLLC", "Low Mach"])
def test_riemann_consistency(solver):
    k = KAT["riemann_solvers"]["consistency"]
    cons = orc.to_cons(k["prim"], k["gamma"], k["mmw"])
    num = orc.riemann(solver, cons, cons, 0, k["gamma"], k["mmw"])
    phys = orc.flux_cons(cons, 0, k["gamma"], k["mmw"])
    for a, b in zip(num, phys):
        assert double_eq(a, b), (solver, num, phys)


@pytest.mark.parametrize("solver", ["HLL", "HLLC", "Low Mach"])
def test_riemann_symmetry(solver):
    k = KAT["riemann_solvers"]["symmetry"]
    g, m = k["gamma"], k["mmw"]
    cl = orc.to_cons(k["left"], g, m)
    cr = orc.to_cons(k["right"], g, m)
    f1 = orc.riemann(solver, cl, cr, 0, g, m)
    cl2, cr2 = cr.copy(), cl.copy()
    cl2[1], cr2[1] = -cl2[1], -cr2[1]
    f2 = orc.riemann(solver, cl2, cr2, 0, g, m)
    assert double_eq(f1[0], -f2[0])
    assert double_eq(f1[1], f2[1])
    assert double_eq(f1[2], -f2[2])


# ---------------------------------------------------------------- tests/slopes.cpp
def test_constant_limiter_is_zero():
    k = KAT["slopes"]
    assert orc.limiter("Constant", k["diffR"], k["diffL"]) == 0.0


@pytest.mark.parametrize("lim", ["Minmod", "VanAlbada", "VanLeer"])
def test_limiter_properties(lim):
    k = KAT["slopes"]
    dR, dL = k["diffR"], k["diffL"]
    assert double_eq(orc.limiter(lim, +dR, -dL), 0.0)
    assert double_eq(orc.limiter(lim, -dR, +dL), 0.0)
    d = k["identity"]
    assert double_eq(orc.limiter(lim, +d, +d), +d)
    assert double_eq(orc.limiter(lim, -d, -d), -d)
    a, b = k["symmetry"]
    assert double_eq(orc.limiter(lim, +a, +b), orc.limiter(lim, +b, +a))
    assert double_eq(orc.limiter(lim, -a, -b), orc.limiter(lim, -b, -a))


# ---------------------------------------------------------------- tests/conversions.cpp (arrays vs scalar)
@pytest.mark.parametrize("nd", [1, 2, 3])
def test_array_conversions_match_scalar(nd):
    k = KAT["conversions"]
    n = k["n"]
    o = orc.Oracle(nd, [n, n, n], gamma=k["gamma"], ng=2)
    prim = [k["prim"]["rho"]] + [k["prim"]["u"]] * nd + [k["prim"]["P"]]
    rho = np.full(o.shape3, prim[0])
    u = np.full((nd,) + o.shape3, prim[1])
    P = np.full(o.shape3, prim[-1])
    rhou, E = np.zeros_like(u), np.zeros_like(rho)
    o.conv_prim_to_cons(2, rho, u, P, rhou, E)
    want = orc.to_cons(prim, k["gamma"], 1.0)
    assert np.all(rhou == want[1]) and np.all(E == want[-1])
    cons = [k["cons"]["rho"]] + [k["cons"]["rhou"]] * nd + [k["cons"]["E"]]
    rho = np.full(o.shape3, cons[0])
    rhou = np.full((nd,) + o.shape3, cons[1])
    E = np.full(o.shape3, cons[-1])
    u, P = np.zeros_like(rhou), np.zeros_like(rho)
    o.conv_cons_to_prim(2, rho, rhou, E, u, P)
    want = orc.to_prim(cons, k["gamma"], 1.0)
    assert np.all(u == want[1]) and np.all(P == want[-1])


# ---------------------------------------------------------------- tests/thermodynamics.cpp
def _invalid_values(extra):
    return [-math.inf, -1.0, math.inf, math.nan] + extra


def test_eos_validity_range():
    from heraclespp_amd.hydro import DomainError, PerfectGas

    for g in _invalid_values([1.0]):
        with pytest.raises(DomainError, match="Invalid gamma"):
            PerfectGas(g, 1.0)
        with pytest.raises(RuntimeError, match="Invalid gamma"):
            orc.Oracle(1, [4], gamma=g)
    for g in [np.nextafter(1.0, math.inf), sys.float_info.max]:
        PerfectGas(g, 1.0)
        orc.Oracle(1, [4], gamma=g)
    for m in _invalid_values([0.0]):
        with pytest.raises(DomainError, match="Invalid mmw"):
            PerfectGas(1.4, m)
        with pytest.raises(RuntimeError, match="Invalid mmw"):
            orc.Oracle(1, [4], mmw=m)
    for m in [5e-324, 1.0, sys.float_info.max]:
        PerfectGas(1.4, m)
    eos = PerfectGas(1.4, 1.0)
    assert eos.adiabatic_index() == 1.4 and eos.mean_molecular_weight() == 1.0
    valid = [5e-324, sys.float_info.min, 1.0, sys.float_info.max]
    invalid = [-math.inf, -1.0, 0.0, math.inf, math.nan]
    for a in valid:
        for b in valid:
            assert PerfectGas.is_valid(a, b)
    for a in invalid:
        for b in invalid:
            assert not PerfectGas.is_valid(a, b) and not PerfectGas.is_valid(b, a)


# ---------------------------------------------------------------- tests/gravity.cpp (uniform)
def test_uniform_gravity():
    from heraclespp_amd.hydro import make_uniform_gravity

    k = KAT["gravity"]["uniform"]
    g = make_uniform_gravity(*k["g"])
    for i in range(5):
        for j in range(4):
            for kk in range(3):
                assert [g(i, j, kk, d) for d in range(3)] == k["g"]


# ---------------------------------------------------------------- tests/range.cpp
def test_range_reference_cases():
    from heraclespp_amd.hydro import Range

    for cmin, cmax in [([1, 0, 0], [0, 0, 0]), ([0, 1, 0], [0, 0, 0]), ([0, 0, 1], [0, 0, 0]),
                       ([-1, 0, 0], [0, 0, 0]), ([0, -1, 0], [0, 0, 0]), ([0, 0, -1], [0, 0, 0]),
                       ([0, 0, 0], [-1, 0, 0]), ([0, 0, 0], [0, -1, 0]), ([0, 0, 0], [0, 0, -1])]:
        with pytest.raises(RuntimeError):
            Range(cmin, cmax, 2, 3)
    for nd in (1, 2, 3):
        def mask(a, b):
            return [a[d] if d < nd else b[d] for d in range(3)]

        r = Range([1, 1, 1], [2, 2, 2], 2, nd)
        assert r.Corner_min == [1, 1, 1] and r.Corner_max == [2, 2, 2]
        r = Range([2, 3, 4], [2, 3, 4], 3, nd)
        assert r.Nc_min_0g == mask([3, 3, 3], [0, 0, 0]) and r.Nc_max_0g == mask([3, 3, 3], [0, 0, 0])
        r = Range([2, 3, 4], [4, 5, 6], 3, nd)
        assert r.Nc_min_0g == mask([3, 3, 3], [0, 0, 0]) and r.Nc_max_0g == mask([5, 5, 5], [2, 2, 2])
        assert r.Nc_min_1g == mask([2, 2, 2], [0, 0, 0]) and r.Nc_max_1g == mask([6, 6, 6], [2, 2, 2])
        assert r.Nc_min_2g == mask([1, 1, 1], [0, 0, 0]) and r.Nc_max_2g == mask([7, 7, 7], [2, 2, 2])
        assert r.Nf_min_0g == mask([3, 3, 3], [0, 0, 0]) and r.Nf_max_0g == mask([6, 6, 6], [3, 3, 3])
        assert r.Nf_min_1g == mask([2, 2, 2], [0, 0, 0]) and r.Nf_max_1g == mask([7, 7, 7], [3, 3, 3])
        assert r.Nf_min_2g == mask([1, 1, 1], [0, 0, 0]) and r.Nf_max_2g == mask([8, 8, 8], [3, 3, 3])
        with pytest.raises(RuntimeError, match="NgEff > Ng"):
            r.with_ghosts(4)
