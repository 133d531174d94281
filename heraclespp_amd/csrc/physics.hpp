// physics.hpp -- device-side hydro algebra for the HERACLES++ hydro step on gfx950.
//
// A state is a small register array q[NV] with NV = NDIM + 2: q[0] density,
// q[1..NDIM] momentum (or velocity), q[NDIM+1] total energy (or pressure).
// Every expression keeps the operation order of the reference so the FP64
// path is bit-identical to it when compiled with -ffp-contract=off; the
// comments cite the reference file:line each function follows.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace hclpp {

// units/units.hpp:60-63 (SI code units, kg = m = s = K = 1)
constexpr double kKb = 1.380649E-23;
constexpr double kMp = 1.672649E-27;
constexpr double kGrav = 6.6743015E-11;
constexpr double kPi = 3.141592653589793238462643383279502884;

enum Limiter { kConstant = 0, kVanLeer = 1, kMinmod = 2, kVanAlbada = 3 };
enum Solver { kHLL = 0, kHLLC = 1, kLowMach = 2 };
enum Geom { kCartesian = 0, kSpherical = 1 };

// thermodynamics/perfect_gas.hpp:17-90 (gamma, gamma-1, mean molecular weight)
struct PerfectGas {
    double gamma;
    double gamma_m1;
    double mmw;

    // perfect_gas.hpp:69-72
    __host__ __device__ __forceinline__ double temp_from_pres(double rho, double P) const
    {
        return P * mmw * kMp / (rho * kKb);
    }
    // perfect_gas.hpp:52-56
    __host__ __device__ __forceinline__ double evol_from_temp(double rho, double T) const
    {
        return rho * kKb * T / (mmw * kMp * gamma_m1);
    }
    // perfect_gas.hpp:46-50
    __host__ __device__ __forceinline__ double evol_from_pres(double rho, double P) const
    {
        return evol_from_temp(rho, temp_from_pres(rho, P));
    }
    // perfect_gas.hpp:74-78
    __host__ __device__ __forceinline__ double temp_from_evol(double rho, double evol) const
    {
        return evol * mmw * kMp * gamma_m1 / (rho * kKb);
    }
    // perfect_gas.hpp:64-67
    __host__ __device__ __forceinline__ double pres_from_temp(double rho, double T) const
    {
        return rho * kKb * T / (mmw * kMp);
    }
    // perfect_gas.hpp:58-62
    __host__ __device__ __forceinline__ double pres_from_evol(double rho, double evol) const
    {
        return pres_from_temp(rho, temp_from_evol(rho, evol));
    }
    // perfect_gas.hpp:80-83
    __host__ __device__ __forceinline__ double sound(double rho, double P) const { return sqrt(gamma * P / rho); }
};

template <int ND>
struct State {
    static constexpr int NV = ND + 2;
    double q[NV];
    __host__ __device__ __forceinline__ double& operator[](int i) { return q[i]; }
    __host__ __device__ __forceinline__ double operator[](int i) const { return q[i]; }
};

// euler_equations.hpp:41-49: kinetic energy of a conservative state
template <int ND>
__host__ __device__ __forceinline__ double kinetic_cons(const State<ND>& c)
{
    double n = 0;
#pragma unroll
    for (int d = 0; d < ND; ++d) n += c[1 + d] * c[1 + d];
    return n / c[0] / 2;
}

// euler_equations.hpp:51-59: kinetic energy of a primitive state
template <int ND>
__host__ __device__ __forceinline__ double kinetic_prim(const State<ND>& p)
{
    double n = 0;
#pragma unroll
    for (int d = 0; d < ND; ++d) n += p[1 + d] * p[1 + d];
    return p[0] * n / 2;
}

// euler_equations.hpp:61-65
template <int ND>
__host__ __device__ __forceinline__ double internal_energy(const State<ND>& c)
{
    return c[ND + 1] - kinetic_cons(c);
}

// euler_equations.hpp:111-122
template <int ND>
__host__ __device__ __forceinline__ State<ND> cons_to_prim(const State<ND>& c, const PerfectGas& eos)
{
    State<ND> p;
    double const evol = internal_energy(c);
    p[0] = c[0];
#pragma unroll
    for (int d = 0; d < ND; ++d) p[1 + d] = c[1 + d] / c[0];
    p[ND + 1] = eos.pres_from_evol(c[0], evol);
    return p;
}

// euler_equations.hpp:124-134
template <int ND>
__host__ __device__ __forceinline__ State<ND> prim_to_cons(const State<ND>& p, const PerfectGas& eos)
{
    State<ND> c;
    c[0] = p[0];
#pragma unroll
    for (int d = 0; d < ND; ++d) c[1 + d] = p[0] * p[1 + d];
    c[ND + 1] = eos.evol_from_pres(p[0], p[ND + 1]) + kinetic_prim(p);
    return c;
}

// euler_equations.hpp:72-86: physical flux along `dir` from a primitive state
template <int ND>
__host__ __device__ __forceinline__ State<ND> flux_of_prim(const State<ND>& p, int dir, const PerfectGas& eos)
{
    State<ND> f;
    double const E = kinetic_prim(p) + eos.evol_from_pres(p[0], p[ND + 1]);
    double const un = p[1 + dir];
    f[0] = p[0] * un;
#pragma unroll
    for (int d = 0; d < ND; ++d) f[1 + d] = p[0] * un * p[1 + d];
    f[1 + dir] += p[ND + 1];
    f[ND + 1] = un * (E + p[ND + 1]);
    return f;
}

// euler_equations.hpp:93-109: physical flux along `dir` from a conservative state
template <int ND>
__host__ __device__ __forceinline__ State<ND> flux_of_cons(const State<ND>& c, int dir, const PerfectGas& eos)
{
    State<ND> f;
    double const evol = internal_energy(c);
    double const P = eos.pres_from_evol(c[0], evol);
    double const u = c[1 + dir] / c[0];
    f[0] = u * c[0];
#pragma unroll
    for (int d = 0; d < ND; ++d) f[1 + d] = u * c[1 + d];
    f[1 + dir] += P;
    f[ND + 1] = u * (c[ND + 1] + P);
    return f;
}

// ---------------------------------------------------------------- limiters
// slope_limiters.hpp:18-87.  diffR = (v[i+1]-v[i])/dx, diffL = (v[i]-v[i-1])/dx
__host__ __device__ __forceinline__ double limited_slope(int lim, double dR, double dL)
{
    if (lim == kConstant) return 0;
    if (dL * dR > 0) {
        double const r = dR / dL;
        if (lim == kVanLeer) return (1. / 2) * (dR + dL) * (4 * r) / ((r + 1) * (r + 1));
        if (lim == kMinmod) {
            double const mm = 2 * fmin(1., r) / (1 + r);
            return mm * (dL + dR) / 2;
        }
        return (1. / 2) * (dR + dL) * (2 * r) / ((r * r) + 1);
    }
    return 0;
}

template <int LIM>
__host__ __device__ __forceinline__ double limited_slope_t(double dR, double dL)
{
    if constexpr (LIM == kConstant) {
        return 0;
    } else {
        if (dL * dR > 0) {
            double const r = dR / dL;
            if constexpr (LIM == kVanLeer) {
                return (1. / 2) * (dR + dL) * (4 * r) / ((r + 1) * (r + 1));
            } else if constexpr (LIM == kMinmod) {
                double const mm = 2 * fmin(1., r) / (1 + r);
                return mm * (dL + dR) / 2;
            } else {
                return (1. / 2) * (dR + dL) * (2 * r) / ((r * r) + 1);
            }
        }
        return 0;
    }
}

// ---------------------------------------------------------------- Riemann solvers
// riemann_solver.hpp:60-64
__host__ __device__ __forceinline__ double hll_component(double UL, double UR, double FL, double FR, double wsL, double wsR)
{
    return ((wsR * FL) - (wsL * FR) + (wsL * wsR * (UR - UL))) / (wsR - wsL);
}

// riemann_solver.hpp:30-58
template <int ND>
__host__ __device__ __forceinline__ State<ND> riemann_hll(const State<ND>& cL, const State<ND>& cR, int dir, const PerfectGas& eos)
{
    State<ND> const pL = cons_to_prim(cL, eos);
    State<ND> const pR = cons_to_prim(cR, eos);
    double const sL = eos.sound(pL[0], pL[ND + 1]);
    double const sR = eos.sound(pR[0], pR[ND + 1]);
    double const wsL = fmin(pL[1 + dir] - sL, pR[1 + dir] - sR);
    double const wsR = fmax(pL[1 + dir] + sL, pR[1 + dir] + sR);
    double const nL = fmin(wsL, 0.);
    double const pR_ = fmax(wsR, 0.);
    State<ND> const fL = flux_of_prim(pL, dir, eos);
    State<ND> const fR = flux_of_prim(pR, dir, eos);
    State<ND> f;
#pragma unroll
    for (int v = 0; v < ND + 2; ++v) f[v] = hll_component(cL[v], cR[v], fL[v], fR[v], nL, pR_);
    return f;
}

// riemann_solver.hpp:136-140
__host__ __device__ __forceinline__ double hllc_component(double F, double U, double Us, double ws) { return F + (ws * (Us - U)); }

// riemann_solver.hpp:70-134
template <int ND>
__host__ __device__ __forceinline__ State<ND> riemann_hllc(const State<ND>& cL, const State<ND>& cR, int dir, const PerfectGas& eos)
{
    State<ND> const pL = cons_to_prim(cL, eos);
    State<ND> const pR = cons_to_prim(cR, eos);
    double const sL = eos.sound(pL[0], pL[ND + 1]);
    double const sR = eos.sound(pR[0], pR[ND + 1]);
    double const uL = pL[1 + dir];
    double const uR = pR[1 + dir];
    double const wsL = fmin(uL - sL, uR - sR);
    double const wsR = fmax(uL + sL, uR + sR);
    State<ND> const fL = flux_of_prim(pL, dir, eos);
    State<ND> const fR = flux_of_prim(pR, dir, eos);
    double const rcL = pL[0] * (wsL - uL);
    double const rcR = pR[0] * (wsR - uR);
    double const ustar = (pR[ND + 1] - pL[ND + 1] + (rcL * uL) - (rcR * uR)) / (rcL - rcR);
    double const pstar = 1. / 2 * (pL[ND + 1] + pR[ND + 1] + (rcL * (ustar - uL)) + (rcR * (ustar - uR)));
    bool const left = ustar > 0;
    double const S = left ? wsL : wsR;
    State<ND> const& cs = left ? cL : cR;
    State<ND> const& ps = left ? pL : pR;
    State<ND> const& fs = left ? fL : fR;
    double const un = ps[1 + dir];
    if (wsL * wsR > 0) return fs;
    State<ND> f;
    double const rho_star = cs[0] * (S - un) / (S - ustar);
    f[0] = hllc_component(fs[0], cs[0], rho_star, S);
#pragma unroll
    for (int d = 0; d < ND; ++d) f[1 + d] = rho_star * ustar * ps[1 + d];
    f[1 + dir] = (rho_star * ustar * ustar) + pstar;
    double const E_star = ((S - un) / (S - ustar) * cs[ND + 1]) + ((ustar - un) / (S - ustar) * ((ustar * cs[0] * (S - un)) + ps[ND + 1]));
    f[ND + 1] = hllc_component(fs[ND + 1], cs[ND + 1], E_star, S);
    return f;
}

// riemann_solver.hpp:146-180 ("Low Mach" splitting solver)
template <int ND>
__host__ __device__ __forceinline__ State<ND> riemann_lowmach(const State<ND>& cL, const State<ND>& cR, int dir, const PerfectGas& eos)
{
    State<ND> const pL = cons_to_prim(cL, eos);
    State<ND> const pR = cons_to_prim(cR, eos);
    double const sL = eos.sound(pL[0], pL[ND + 1]);
    double const sR = eos.sound(pR[0], pR[ND + 1]);
    double const a = 1.1 * fmax(pL[0] * sL, pR[0] * sR);
    double const ustar = ((pL[1 + dir] + pR[1 + dir]) / 2) - (1 / (2 * a) * (pR[ND + 1] - pL[ND + 1]));
    double const Ma = fabs(ustar) / fmin(sL, sR);
    double const theta = fmin(1., Ma);
    double const Pstar = ((pL[ND + 1] + pR[ND + 1]) / 2) - ((theta * a) / 2 * (pR[1 + dir] - pL[1 + dir]));
    State<ND> const& cs = (ustar > 0) ? cL : cR;
    State<ND> f;
    f[0] = ustar * cs[0];
#pragma unroll
    for (int d = 0; d < ND; ++d) f[1 + d] = ustar * cs[1 + d];
    f[1 + dir] += Pstar;
    f[ND + 1] = (ustar * cs[ND + 1]) + (Pstar * ustar);
    return f;
}

template <int RS, int ND>
__host__ __device__ __forceinline__ State<ND> riemann(const State<ND>& cL, const State<ND>& cR, int dir, const PerfectGas& eos)
{
    if constexpr (RS == kHLL) {
        return riemann_hll(cL, cR, dir, eos);
    } else if constexpr (RS == kHLLC) {
        return riemann_hllc(cL, cR, dir, eos);
    } else {
        return riemann_lowmach(cL, cR, dir, eos);
    }
}

// ---------------------------------------------------------------- spherical sources
// hydro/source_terms.hpp:18-77.  Trigonometry comes from host tables:
// cot_mid = cos(m)/sin(m) with m = (x1 + x1p)/2; sn = sin(x1), snp = sin(x1p).
__host__ __device__ __forceinline__ double src_pressure(double dtodv, double PL, double PR, double dS, double dSp)
{
    return dtodv * (PL + PR) / 2 * (dSp - dS);
}
__host__ __device__ __forceinline__ double src_centrifugal(double dtodv, double rL, double rR, double uL, double uR, double dS, double dSp)
{
    return dtodv * ((rL * uL * uL) + (rR * uR * uR)) / 2 * (dSp - dS) / 2;
}
__host__ __device__ __forceinline__ double src_coriolis_r(double dtodv, double x0, double x0p, double rL, double rR, double uLr, double uRr,
                                                          double uLo, double uRo, double dS, double dSp)
{
    return dtodv * (x0p - x0) / (x0p + x0) * ((rR * uRo * uRr * dSp) + (rL * uLo * uLr * dS));
}
__host__ __device__ __forceinline__ double src_centrifugal_th(double dtodv, double cot_mid, double rL, double rR, double uLp, double uRp, double dS,
                                                              double dSp)
{
    return dtodv * ((rL * uLp * uLp) + (rR * uRp * uRp)) / 2 * cot_mid * (dSp - dS) / 2;
}
__host__ __device__ __forceinline__ double src_coriolis_th(double dtodv, double sn, double snp, double rL, double rR, double uLt, double uRt,
                                                           double uLp, double uRp, double dS, double dSp)
{
    return dtodv * (snp - sn) / (snp + sn) * ((rR * uRp * uRt * dSp) + (rL * uLp * uLt * dS));
}

}  // namespace hclpp
