"""Exact solution of the 1D Riemann problem for a perfect gas.

TEST INFRASTRUCTURE ONLY (analytic oracle for the Sod configuration).

Written from the textbook method (E. F. Toro, "Riemann Solvers and Numerical
Methods for Fluid Dynamics", 3rd ed., ch. 4): the star pressure is the root
of f_L(p) + f_R(p) + (u_R - u_L) = 0, with f_K the shock (Rankine-Hugoniot)
or rarefaction (isentropic) branch; Newton iterations run to convergence
(relative change < 1e-15), unlike the reference's test/exact_shock_tube.py,
whose loop never updates its iterate.  The solution is then sampled on
x/t rays.
"""

from __future__ import annotations

import math

import numpy as np


def _fk(p, rho, P, c, g):
    """Pressure function f_K and its derivative (Toro eqs. 4.6-4.7, 4.37)."""
    if p > P:  # shock
        A = 2.0 / ((g + 1.0) * rho)
        B = (g - 1.0) / (g + 1.0) * P
        q = math.sqrt(A / (p + B))
        return (p - P) * q, q * (1.0 - 0.5 * (p - P) / (B + p))
    # rarefaction
    pr = p / P
    f = 2.0 * c / (g - 1.0) * (pr ** ((g - 1.0) / (2.0 * g)) - 1.0)
    fd = 1.0 / (rho * c) * pr ** (-(g + 1.0) / (2.0 * g))
    return f, fd


def star_state(rhoL, uL, PL, rhoR, uR, PR, gamma):
    """Return (p_star, u_star) iterated to machine precision."""
    g = gamma
    cL = math.sqrt(g * PL / rhoL)
    cR = math.sqrt(g * PR / rhoR)
    if 2.0 * (cL + cR) / (g - 1.0) <= uR - uL:
        raise ValueError("vacuum is generated")
    p = max(1e-12, 0.5 * (PL + PR) - 0.125 * (uR - uL) * (rhoL + rhoR) * (cL + cR))  # PVRS guess
    for _ in range(200):
        fL, dL = _fk(p, rhoL, PL, cL, g)
        fR, dR = _fk(p, rhoR, PR, cR, g)
        pn = p - (fL + fR + uR - uL) / (dL + dR)
        if pn <= 0:
            pn = 0.5 * p
        if abs(pn - p) <= 1e-15 * 0.5 * (pn + p):
            p = pn
            break
        p = pn
    fL, _ = _fk(p, rhoL, PL, cL, g)
    fR, _ = _fk(p, rhoR, PR, cR, g)
    u = 0.5 * (uL + uR) + 0.5 * (fR - fL)
    return p, u


def sample(S, rhoL, uL, PL, rhoR, uR, PR, gamma, pstar, ustar):
    """State on the ray x/t = S (Toro section 4.5)."""
    g = gamma
    cL = math.sqrt(g * PL / rhoL)
    cR = math.sqrt(g * PR / rhoR)
    if S <= ustar:
        if pstar > PL:  # left shock
            SL = uL - cL * math.sqrt((g + 1) / (2 * g) * pstar / PL + (g - 1) / (2 * g))
            if S <= SL:
                return rhoL, uL, PL
            r = pstar / PL
            gg = (g - 1) / (g + 1)
            return rhoL * (r + gg) / (gg * r + 1), ustar, pstar
        SHL = uL - cL
        if S <= SHL:
            return rhoL, uL, PL
        cs = cL * (pstar / PL) ** ((g - 1) / (2 * g))
        STL = ustar - cs
        if S > STL:
            return rhoL * (pstar / PL) ** (1 / g), ustar, pstar
        w = 2 / (g + 1) + (g - 1) / ((g + 1) * cL) * (uL - S)
        return rhoL * w ** (2 / (g - 1)), 2 / (g + 1) * (cL + (g - 1) / 2 * uL + S), PL * w ** (2 * g / (g - 1))
    if pstar > PR:  # right shock
        SR = uR + cR * math.sqrt((g + 1) / (2 * g) * pstar / PR + (g - 1) / (2 * g))
        if S >= SR:
            return rhoR, uR, PR
        r = pstar / PR
        gg = (g - 1) / (g + 1)
        return rhoR * (r + gg) / (gg * r + 1), ustar, pstar
    SHR = uR + cR
    if S >= SHR:
        return rhoR, uR, PR
    cs = cR * (pstar / PR) ** ((g - 1) / (2 * g))
    STR = ustar + cs
    if S <= STR:
        return rhoR * (pstar / PR) ** (1 / g), ustar, pstar
    w = 2 / (g + 1) - (g - 1) / ((g + 1) * cR) * (uR - S)
    return rhoR * w ** (2 / (g - 1)), 2 / (g + 1) * (-cR + (g - 1) / 2 * uR + S), PR * w ** (2 * g / (g - 1))


def solve(x, x0, t, left, right, gamma):
    """Exact (rho, u, P) at positions x for a discontinuity at x0 released at t = 0."""
    rhoL, uL, PL = left
    rhoR, uR, PR = right
    ps, us = star_state(rhoL, uL, PL, rhoR, uR, PR, gamma)
    out = np.array([sample((xi - x0) / t, rhoL, uL, PL, rhoR, uR, PR, gamma, ps, us) for xi in np.asarray(x)])
    return out[:, 0], out[:, 1], out[:, 2]


def cell_averages(nodes, x0, t, left, right, gamma, sub=64):
    """Exact solution averaged over cells [nodes[i], nodes[i+1]] with `sub` midpoint samples."""
    nodes = np.asarray(nodes)
    res = []
    for a, b in zip(nodes[:-1], nodes[1:]):
        xs = a + (np.arange(sub) + 0.5) * (b - a) / sub
        r, u, p = solve(xs, x0, t, left, right, gamma)
        res.append((r.mean(), u.mean(), p.mean()))
    res = np.array(res)
    return res[:, 0], res[:, 1], res[:, 2]
