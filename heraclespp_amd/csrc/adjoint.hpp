// adjoint.hpp -- hand-written reverse-mode derivatives (vector-Jacobian products) of the
// pieces of one hydro step, for the adjoint step kernels.
//
// Each *_vjp function takes the forward inputs and the adjoint of the outputs and
// accumulates (+=) the adjoint of the inputs, plus the derivative with respect to
// gamma into `lg`.  They differentiate the mathematical functions of the forward
// path: the kb/mp temperature round trip of the reference EoS is the identity
// P = (gamma-1) e, so it is differentiated as such.  Non-smooth operations are
// differentiated piecewise on the branch the forward pass takes: the limiter
// sign test and Minmod selection, the min/max wave speeds, HLL's clipping of
// the wave speeds at zero, and HLLC's choice of star state (sign of u*) and of
// the supersonic branch (wsL * wsR > 0).
//
// The forward values that decide a branch (slopes, wave speeds, u*) are recomputed with the
// forward's own expressions.  The derivative formulas themselves are free of that constraint:
// they divide once per divisor and multiply by the reciprocal (each division is an 8-instruction
// sequence around a quarter-rate reciprocal, physics.hpp fdiv), which moves results by rounding
// only.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "physics.hpp"

namespace hclpp::adj {

template <int ND>
using S = State<ND, double>;

template <int ND>
__host__ __device__ __forceinline__ S<ND> zero()
{
    S<ND> z;
#pragma unroll
    for (int v = 0; v < ND + 2; ++v) z[v] = 0;
    return z;
}

// P = (gamma-1)(E - |m|^2 / 2 rho), u = m / rho: cons -> prim
template <int ND>
__host__ __device__ __forceinline__ void prim_vjp(const S<ND>& c, double gamma, const S<ND>& lp, S<ND>& lc, double& lg)
{
    double const rho = c[0];
    double const ir = fdiv(1, rho);
    double const gm1 = gamma - 1;
    double u2 = 0, ulu = 0;
    double u[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        u[d] = c[1 + d] * ir;
        u2 += u[d] * u[d];
        ulu += lp[1 + d] * u[d];
    }
    double const lP = lp[ND + 1];
    lc[0] += lp[0] - ulu * ir + lP * gm1 * 0.5 * u2;
#pragma unroll
    for (int d = 0; d < ND; ++d) lc[1 + d] += lp[1 + d] * ir - lP * gm1 * u[d];
    lc[ND + 1] += lP * gm1;
    lg += lP * (c[ND + 1] - 0.5 * rho * u2);
}

// m = rho u, E = P / (gamma-1) + rho |u|^2 / 2: prim -> cons
template <int ND>
__host__ __device__ __forceinline__ void cons_vjp(const S<ND>& p, double gamma, const S<ND>& lc, S<ND>& lp, double& lg)
{
    double const rho = p[0];
    double const gm1 = gamma - 1;
    double u2 = 0, lmu = 0;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        u2 += p[1 + d] * p[1 + d];
        lmu += lc[1 + d] * p[1 + d];
    }
    double const igm1 = fdiv(1, gm1);
    double const lE = lc[ND + 1];
    lp[0] += lc[0] + lmu + lE * 0.5 * u2;
#pragma unroll
    for (int d = 0; d < ND; ++d) lp[1 + d] += lc[1 + d] * rho + lE * rho * p[1 + d];
    lp[ND + 1] += lE * igm1;
    lg += -lE * p[ND + 1] * igm1 * igm1;
}

// physical flux along n from a primitive state
template <int ND>
__host__ __device__ __forceinline__ void flux_vjp(const S<ND>& p, int n, double gamma, const S<ND>& lf, S<ND>& lp, double& lg)
{
    double const rho = p[0], P = p[ND + 1];
    double const gm1 = gamma - 1;
    double const un = p[1 + n];
    double u2 = 0, au = 0;
#pragma unroll
    for (int k = 0; k < ND; ++k) {
        u2 += p[1 + k] * p[1 + k];
        au += lf[1 + k] * p[1 + k];
    }
    double const igm1 = fdiv(1, gm1);
    double const E = P * igm1 + 0.5 * rho * u2;
    double const aE = lf[ND + 1];
    lp[0] += lf[0] * un + un * au + aE * un * 0.5 * u2;
#pragma unroll
    for (int j = 0; j < ND; ++j) {
        double g = lf[1 + j] * rho * un + aE * un * rho * p[1 + j];
        if (j == n) g += lf[0] * rho + rho * au + aE * (E + P);
        lp[1 + j] += g;
    }
    lp[ND + 1] += lf[1 + n] + aE * un * (igm1 + 1);
    lg += -aE * un * P * igm1 * igm1;
}

// c = sqrt(gamma P / rho).  A zero adjoint contributes nothing, also where c is not a number: the
// forward's wave-speed min/max (fmin/fmax) pass over a NaN sound speed (a state with P < 0, as the
// Hancock state of a ghost cell whose volume is negative can be), and the branch it takes sends
// no adjoint to that speed.
__host__ __device__ __forceinline__ void sound_vjp(double rho, double P, double gamma, double c, double lc_, double& lrho, double& lP, double& lg)
{
    if (lc_ == 0) return;
    double const i2rc = 1 / (2 * rho * c);  // c / (2 rho) = c^2 / (2 rho c)
    lrho += -lc_ * c * c * i2rc;
    lP += lc_ * gamma * i2rc;
    lg += lc_ * P * i2rc;
}

// Which operand fmin(a, b) / fmax(a, b) returns: the first unless it is the larger / smaller one
// or a NaN (fmin and fmax return the other operand when one is a NaN).
__host__ __device__ __forceinline__ bool fmin_takes_first(double a, double b) { return std::isnan(b) || (!std::isnan(a) && a <= b); }
__host__ __device__ __forceinline__ bool fmax_takes_first(double a, double b) { return std::isnan(b) || (!std::isnan(a) && a >= b); }

// slope limiter: derivative of the limited slope w.r.t. (diffR, diffL) on the forward branch
template <int LIM>
__host__ __device__ __forceinline__ void limiter_grad(double dR, double dL, double& gR, double& gL)
{
    gR = 0;
    gL = 0;
    if constexpr (LIM == kConstant) {
        return;
    } else {
        if (!(dL * dR > 0)) return;
        if constexpr (LIM == kVanLeer) {  // 2 dR dL / (dR + dL)
            double const s = dR + dL;
            double const is2 = fdiv(1, s * s);
            gR = 2 * dL * dL * is2;
            gL = 2 * dR * dR * is2;
        } else if constexpr (LIM == kMinmod) {  // dR if r <= 1 else dL  (fmin(1, r) branch)
            double const r = fdiv(dR, dL);
            if (r < 1.) gR = 1;
            else if (r > 1.) gL = 1;
            else { gR = 0.5; gL = 0.5; }
        } else {  // VanAlbada: dR dL (dR + dL) / (dR^2 + dL^2)
            double const N = dR * dL * (dR + dL);
            double const D = dR * dR + dL * dL;
            double const iD2 = fdiv(1, D * D);
            gR = ((2 * dR * dL + dL * dL) * D - N * 2 * dR) * iD2;
            gL = ((2 * dR * dL + dR * dR) * D - N * 2 * dL) * iD2;
        }
    }
}

// HLL flux (riemann_solver.hpp:30-64)
template <int ND>
__host__ __device__ __forceinline__ void hll_vjp(const S<ND>& cL, const S<ND>& cR, int n, const PerfectGas& eos, const S<ND>& lF, S<ND>& lcL,
                                                 S<ND>& lcR, double& lg)
{
    double const g = eos.gamma;
    S<ND> const pL = cons_to_prim(cL, eos), pR = cons_to_prim(cR, eos);
    double const sL = eos.sound(pL[0], pL[ND + 1]), sR = eos.sound(pR[0], pR[ND + 1]);
    double const uL = pL[1 + n], uR = pR[1 + n];
    double const aL = uL - sL, bL = uR - sR, aR = uL + sL, bR = uR + sR;
    double const wsL = fmin(aL, bL), wsR = fmax(aR, bR);
    double const nL = fmin(wsL, 0.), pRw = fmax(wsR, 0.);
    S<ND> const fL = flux_of_prim(pL, n, eos), fR = flux_of_prim(pR, n, eos);
    double const D = pRw - nL;
    double const iD = fdiv(1, D);
    S<ND> lfL = zero<ND>(), lfR = zero<ND>();
    double lpR = 0, lnL = 0, lD = 0;
#pragma unroll
    for (int v = 0; v < ND + 2; ++v) {
        double const dU = cR[v] - cL[v];
        double const F = ((pRw * fL[v]) - (nL * fR[v]) + (nL * pRw * dU)) * iD;
        double const lN = lF[v] * iD;
        lD -= lF[v] * F * iD;
        lfL[v] += lN * pRw;
        lfR[v] -= lN * nL;
        lcR[v] += lN * nL * pRw;
        lcL[v] -= lN * nL * pRw;
        lpR += lN * (fL[v] + nL * dU);
        lnL += lN * (-fR[v] + pRw * dU);
    }
    lpR += lD;
    lnL -= lD;
    double const lwsL = wsL < 0 ? lnL : 0.;
    double const lwsR = wsR > 0 ? lpR : 0.;
    S<ND> lpL = zero<ND>(), lpRs = zero<ND>();
    double lsL = 0, lsR = 0;
    if (fmin_takes_first(aL, bL)) { lpL[1 + n] += lwsL; lsL -= lwsL; } else { lpRs[1 + n] += lwsL; lsR -= lwsL; }
    if (fmax_takes_first(aR, bR)) { lpL[1 + n] += lwsR; lsL += lwsR; } else { lpRs[1 + n] += lwsR; lsR += lwsR; }
    flux_vjp(pL, n, g, lfL, lpL, lg);
    flux_vjp(pR, n, g, lfR, lpRs, lg);
    sound_vjp(pL[0], pL[ND + 1], g, sL, lsL, lpL[0], lpL[ND + 1], lg);
    sound_vjp(pR[0], pR[ND + 1], g, sR, lsR, lpRs[0], lpRs[ND + 1], lg);
    prim_vjp(cL, g, lpL, lcL, lg);
    prim_vjp(cR, g, lpRs, lcR, lg);
}

// HLLC flux (riemann_solver.hpp:70-134)
template <int ND>
__host__ __device__ __forceinline__ void hllc_vjp(const S<ND>& cL, const S<ND>& cR, int n, const PerfectGas& eos, const S<ND>& lF, S<ND>& lcL,
                                                  S<ND>& lcR, double& lg)
{
    double const g = eos.gamma;
    S<ND> const pL = cons_to_prim(cL, eos), pR = cons_to_prim(cR, eos);
    double const sL = eos.sound(pL[0], pL[ND + 1]), sR = eos.sound(pR[0], pR[ND + 1]);
    double const uL = pL[1 + n], uR = pR[1 + n];
    double const PL = pL[ND + 1], PR = pR[ND + 1];
    double const aL = uL - sL, bL = uR - sR, aR = uL + sL, bR = uR + sR;
    double const wsL = fmin(aL, bL), wsR = fmax(aR, bR);
    double const rcL = pL[0] * (wsL - uL), rcR = pR[0] * (wsR - uR);
    double const num = PR - PL + (rcL * uL) - (rcR * uR), den = rcL - rcR;
    double const ustar = fdiv(num, den);
    bool const left = ustar > 0;
    // upwind (star-side) state, selected element-wise
    S<ND> cs, ps;
#pragma unroll
    for (int v = 0; v < ND + 2; ++v) {
        cs[v] = left ? cL[v] : cR[v];
        ps[v] = left ? pL[v] : pR[v];
    }
    double const Sw = left ? wsL : wsR;
    double const un = ps[1 + n];

    S<ND> lcs = zero<ND>(), lps = zero<ND>(), lfs = zero<ND>();
    double lS = 0, lustar = 0, lpstar = 0;
    if (wsL * wsR > 0) {
        lfs = lF;  // F = flux of the upwind state
    } else {
        double const A = Sw - un, B = Sw - ustar;
        double const iB = fdiv(1, B);
        double const rs = cs[0] * A * iB;
        double const C = (ustar * cs[0] * A) + ps[ND + 1];
        double const Es = ((A * iB) * cs[ND + 1]) + (((ustar - un) * iB) * C);
        // f0 = fs0 + S (rho* - rho), fE = fsE + S (E* - E)
        lfs[0] = lF[0];
        lfs[ND + 1] = lF[ND + 1];
        lS += lF[0] * (rs - cs[0]) + lF[ND + 1] * (Es - cs[ND + 1]);
        double lrs = lF[0] * Sw;
        lcs[0] -= lF[0] * Sw;
        // transverse momenta rho* u* v_t, normal momentum rho* u*^2 + p*
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            if (d == n) continue;
            lrs += lF[1 + d] * ustar * ps[1 + d];
            lustar += lF[1 + d] * rs * ps[1 + d];
            lps[1 + d] += lF[1 + d] * rs * ustar;
        }
        lrs += lF[1 + n] * ustar * ustar;
        lustar += lF[1 + n] * 2 * rs * ustar;
        lpstar += lF[1 + n];
        double const lEs = lF[ND + 1] * Sw;
        lcs[ND + 1] -= lF[ND + 1] * Sw;
        // E* = (A/B) E + ((u* - un)/B) C, C = u* rho A + P, A = S - un, B = S - u*
        double lA = lEs * cs[ND + 1] * iB;
        double lB = -lEs * Es * iB;
        lcs[ND + 1] += lEs * A * iB;
        lustar += lEs * C * iB;
        double lun = -lEs * C * iB;
        double const lC = lEs * (ustar - un) * iB;
        lustar += lC * cs[0] * A;
        lcs[0] += lC * ustar * A;
        lA += lC * ustar * cs[0];
        lps[ND + 1] += lC;
        // rho* = rho A / B
        lcs[0] += lrs * A * iB;
        lA += lrs * cs[0] * iB;
        lB -= lrs * rs * iB;
        lS += lA + lB;
        lun -= lA;
        lustar -= lB;
        lps[1 + n] += lun;
    }
    double lwsL = left ? lS : 0., lwsR = left ? 0. : lS;
    S<ND> lpL = zero<ND>(), lpR = zero<ND>();
    // p* = (PL + PR + rcL (u* - uL) + rcR (u* - uR)) / 2
    double lrcL = 0.5 * lpstar * (ustar - uL), lrcR = 0.5 * lpstar * (ustar - uR);
    lpL[ND + 1] += 0.5 * lpstar;
    lpR[ND + 1] += 0.5 * lpstar;
    lustar += 0.5 * lpstar * (rcL + rcR);
    lpL[1 + n] -= 0.5 * lpstar * rcL;
    lpR[1 + n] -= 0.5 * lpstar * rcR;
    // u* = num / den
    double const iden = fdiv(1, den);
    double const lnum = lustar * iden;
    double const lden = -lustar * ustar * iden;
    lpR[ND + 1] += lnum;
    lpL[ND + 1] -= lnum;
    lrcL += lnum * uL + lden;
    lpL[1 + n] += lnum * rcL;
    lrcR -= lnum * uR + lden;
    lpR[1 + n] -= lnum * rcR;
    // rcK = rhoK (wsK - uK)
    lpL[0] += lrcL * (wsL - uL);
    lwsL += lrcL * pL[0];
    lpL[1 + n] -= lrcL * pL[0];
    lpR[0] += lrcR * (wsR - uR);
    lwsR += lrcR * pR[0];
    lpR[1 + n] -= lrcR * pR[0];
    // flux of the upwind state, then scatter the upwind adjoints to their side
    flux_vjp(ps, n, g, lfs, lps, lg);
#pragma unroll
    for (int v = 0; v < ND + 2; ++v) {
        lpL[v] += left ? lps[v] : 0.;
        lpR[v] += left ? 0. : lps[v];
        lcL[v] += left ? lcs[v] : 0.;
        lcR[v] += left ? 0. : lcs[v];
    }
    // wave speeds
    double lsL = 0, lsR = 0;
    if (fmin_takes_first(aL, bL)) { lpL[1 + n] += lwsL; lsL -= lwsL; } else { lpR[1 + n] += lwsL; lsR -= lwsL; }
    if (fmax_takes_first(aR, bR)) { lpL[1 + n] += lwsR; lsL += lwsR; } else { lpR[1 + n] += lwsR; lsR += lwsR; }
    sound_vjp(pL[0], PL, g, sL, lsL, lpL[0], lpL[ND + 1], lg);
    sound_vjp(pR[0], PR, g, sR, lsR, lpR[0], lpR[ND + 1], lg);
    prim_vjp(cL, g, lpL, lcL, lg);
    prim_vjp(cR, g, lpR, lcR, lg);
}

template <int RS, int ND>
__host__ __device__ __forceinline__ void riemann_vjp(const S<ND>& cL, const S<ND>& cR, int n, const PerfectGas& eos, const S<ND>& lF, S<ND>& lcL,
                                                     S<ND>& lcR, double& lg)
{
    if constexpr (RS == kHLL) hll_vjp(cL, cR, n, eos, lF, lcL, lcR, lg);
    else hllc_vjp(cL, cR, n, eos, lF, lcL, lcR, lg);
}

}  // namespace hclpp::adj
