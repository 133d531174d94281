// physics.hpp -- device-side hydro algebra for the HERACLES++ hydro step on gfx950.
//
// A state is a small register array q[NV] with NV = NDIM + 2: q[0] density,
// q[1..NDIM] momentum (or velocity), q[NDIM+1] total energy (or pressure).
//
// Every routine is generic over an equation-of-state policy that fixes the
// arithmetic:
//  * PerfectGas (FP64): the reference arithmetic -- every expression keeps
//    the reference operation order and rounds like it (built with
//    -ffp-contract=off), so results are bit-identical to the reference;
//  * PerfectGasMixed (FP32): state and fluxes in FP32 with hardware reciprocal
//    divisions and closed-form slope limiters, while the pressure recovery
//    P = (gamma-1)(E - |m|^2 / 2 rho), which cancels at high Mach number, is
//    evaluated in FP64;
//  * PerfectGasFast (FP64, opt-in): the same FP64 state and algorithm without
//    the reference's rounding -- P = (gamma-1) e directly instead of the
//    temperature round trip through kb and mp, divisions as a product with a
//    shared reciprocal, and the slope limiters in closed form.  Not a parity
//    path: its distance to the reference grows with the step count.
// Division is always written through Eos::div (or fdiv below) so the FP64
// policy keeps the correctly-rounded quotient (on normal-range operands, see
// fdiv) and the FP32 policy uses v_rcp_f32.  Halving is a multiplication by
// 0.5, which is exact in IEEE arithmetic.
// Comments cite the reference file:line each function follows.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

namespace hclpp {

// units/units.hpp:60-63 (SI code units, kg = m = s = K = 1)
constexpr double kKb = 1.380649E-23;
constexpr double kMp = 1.672649E-27;
constexpr double kGrav = 6.6743015E-11;
constexpr double kPi = 3.141592653589793238462643383279502884;

// FP64 division a / b.  On the device this is the instruction sequence hipcc emits for an IEEE
// division on gfx950 (v_rcp_f64, two Newton steps on the reciprocal, the quotient, one
// remainder correction, v_div_fixup_f64) without its two v_div_scale_f64 and with
// v_div_fmas_f64 replaced by the plain fma it reduces to.  By the ISA's definition of
// v_div_scale_f64, it returns its operand with VCC clear (so v_div_fmas_f64 is that fma) when
//   * a and b are finite and normal, |a| >= 2^-969, |b| <= 2^1022, a / b is normal and
//     exponent(a) - exponent(b) < 768,
// so there the result is bit-identical to a / b.  v_div_fixup_f64 stays: it sets the sign of
// every quotient and returns what IEEE division gives for zero and infinite operands and NaNs.
// Outside that range:
//   * finite non-zero operands with a numerator below 2^-969, a subnormal quotient or an
//     exponent gap of 768 or more: the result may differ from a / b by an ulp;
//   * a subnormal divisor: v_rcp_f64 overflows to infinity and the result is NaN, also where
//     a / b is finite;
//   * a divisor above 2^1022: the reciprocal is subnormal and the result may be 0 where a / b
//     is not.
// The step divides with fdiv only by densities, density x kb and the EoS constants, cell
// sizes and volumes, wave-speed differences and the HLLC star-state denominators (which are
// of the order of the sound speed times the density).  The slope limiters divide by
// differences of neighbouring values, which do reach the subnormal range where a blast
// wave's far field has decayed, and by squares of their quotient, which overflow: they use
// the IEEE quotient (Eos::ratio).  Where a numerator has decayed below about 1e-290 the
// one-ulp differences stay there: they never reach a value of ordinary size.  Two instructions fewer
// per division, and the reciprocal part depends on b alone, so the compiler computes it once
// for every division sharing a divisor.  The host build divides, and so does a device build
// with HCLPP_IEEE_DIVISION defined (heraclespp_amd/build.py: HCLPP_IEEE_DIVISION=1), for
// results bit-identical to IEEE division everywhere, with FP64 kernels about a fifth slower.
__host__ __device__ __forceinline__ double fdiv(double a, double b)
{
#if defined(__HIP_DEVICE_COMPILE__) && !defined(HCLPP_IEEE_DIVISION)
    double y = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(-b, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-b, y, 1.0);
    y = __builtin_fma(y, e, y);
    double const q = a * y;
    double const r = __builtin_fma(-b, q, a);
    return __builtin_amdgcn_div_fixup(__builtin_fma(r, y, q), b, a);
#else
    return a / b;
#endif
}

enum Limiter { kConstant = 0, kVanLeer = 1, kMinmod = 2, kVanAlbada = 3 };
enum Solver { kHLL = 0, kHLLC = 1, kLowMach = 2 };
enum Geom { kCartesian = 0, kSpherical = 1 };

template <int ND, class T = double>
struct State {
    static constexpr int NV = ND + 2;
    T q[NV];
    __host__ __device__ __forceinline__ T& operator[](int i) { return q[i]; }
    __host__ __device__ __forceinline__ T operator[](int i) const { return q[i]; }
};

// euler_equations.hpp:41-49: kinetic energy of a conservative state (policy-free helper)
template <int ND, class T, class Div>
__host__ __device__ __forceinline__ T kinetic_cons_with(const State<ND, T>& c, Div div)
{
    T n = 0;
#pragma unroll
    for (int d = 0; d < ND; ++d) n += c[1 + d] * c[1 + d];
    return div(n, c[0]) * T(0.5);
}

// thermodynamics/perfect_gas.hpp:17-90 (gamma, gamma-1, mean molecular weight): FP64 reference arithmetic
struct PerfectGas {
    using real = double;
    double gamma;
    double gamma_m1;
    double mmw;

    __host__ __device__ static __forceinline__ double div(double a, double b) { return fdiv(a, b); }
    // IEEE division for the slope limiters, whose operands range over every magnitude (fdiv above)
    __host__ __device__ static __forceinline__ double ratio(double a, double b) { return a / b; }
    // perfect_gas.hpp:69-72
    __host__ __device__ __forceinline__ double temp_from_pres(double rho, double P) const { return fdiv(P * mmw * kMp, rho * kKb); }
    // perfect_gas.hpp:52-56
    __host__ __device__ __forceinline__ double evol_from_temp(double rho, double T) const { return fdiv(rho * kKb * T, mmw * kMp * gamma_m1); }
    // perfect_gas.hpp:46-50
    __host__ __device__ __forceinline__ double evol_from_pres(double rho, double P) const { return evol_from_temp(rho, temp_from_pres(rho, P)); }
    // perfect_gas.hpp:74-78
    __host__ __device__ __forceinline__ double temp_from_evol(double rho, double evol) const
    {
        return fdiv(evol * mmw * kMp * gamma_m1, rho * kKb);
    }
    // perfect_gas.hpp:64-67
    __host__ __device__ __forceinline__ double pres_from_temp(double rho, double T) const { return fdiv(rho * kKb * T, mmw * kMp); }
    // perfect_gas.hpp:58-62
    __host__ __device__ __forceinline__ double pres_from_evol(double rho, double evol) const
    {
        return pres_from_temp(rho, temp_from_evol(rho, evol));
    }
    // perfect_gas.hpp:80-83
    __host__ __device__ __forceinline__ double sound(double rho, double P) const { return sqrt(fdiv(gamma * P, rho)); }
    // euler_equations.hpp:61-65, :111-122: pressure of a conservative state
    template <int ND>
    __host__ __device__ __forceinline__ double pressure_of_cons(const State<ND, double>& c) const
    {
        double const evol = c[ND + 1] - kinetic_cons_with(c, div);
        return pres_from_evol(c[0], evol);
    }
    // the same, also returning the internal energy E - |m|^2 / 2 rho it recovers the pressure from
    template <int ND>
    __host__ __device__ __forceinline__ double pressure_of_cons(const State<ND, double>& c, double& evol) const
    {
        evol = c[ND + 1] - kinetic_cons_with(c, div);
        return pres_from_evol(c[0], evol);
    }
    template <int ND>
    __host__ __device__ __forceinline__ double internal_energy_fp64(const State<ND, double>& c) const
    {
        return c[ND + 1] - kinetic_cons_with(c, div);
    }
};

// FP32 state and arithmetic; FP64 pressure recovery.  Not a parity path: the
// accuracy bar is an L1 distance to the FP64 result.
struct PerfectGasMixed {
    using real = float;
    float gamma;
    float inv_gamma_m1;
    double gamma_m1;
    double gamma_d;

    __host__ __device__ static __forceinline__ float div(float a, float b)
    {
#if defined(__HIP_DEVICE_COMPILE__)
        return a * __builtin_amdgcn_rcpf(b);
#else
        return a / b;
#endif
    }
    __host__ __device__ static __forceinline__ float ratio(float a, float b) { return div(a, b); }
    __host__ __device__ __forceinline__ float evol_from_pres(float, float P) const { return P * inv_gamma_m1; }
    __host__ __device__ __forceinline__ float sound(float rho, float P) const { return sqrtf(gamma * div(P, rho)); }
    // E - |m|^2 / (2 rho) and (gamma - 1) times it, in FP64
    template <int ND>
    __host__ __device__ __forceinline__ double internal_energy_fp64(const State<ND, float>& c) const
    {
        double n = 0;
#pragma unroll
        for (int d = 0; d < ND; ++d) n += (double)c[1 + d] * (double)c[1 + d];
        return (double)c[ND + 1] - fdiv(n, (double)c[0]) * 0.5;
    }
    template <int ND>
    __host__ __device__ __forceinline__ float pressure_of_cons(const State<ND, float>& c) const
    {
        return (float)(gamma_m1 * internal_energy_fp64(c));
    }
    template <int ND>
    __host__ __device__ __forceinline__ float pressure_of_cons(const State<ND, float>& c, double& evol) const
    {
        evol = internal_energy_fp64(c);
        return (float)(gamma_m1 * evol);
    }
};

// FP64 state and arithmetic, without the reference's rounding (opt-in).  Each formula is the
// reference's up to rounding:
//   * pressure and internal energy: P = (gamma-1) e and e = P / (gamma-1) (perfect_gas.hpp:46-78
//     go through the temperature, T = e mmw mp (gamma-1) / (rho kb) and P = rho kb T / (mmw mp),
//     whose constants cancel);
//   * a / b = a * (1 / b), the reciprocal from v_rcp_f64 and two Newton steps (within about two
//     ulp of the quotient for normal operands); every division by one divisor (1 / rho in the
//     velocities and the kinetic energy, 1 / (wsR - wsL) in the HLL flux) shares it;
//   * the slope limiters in closed form (slope_limiter below).
struct PerfectGasFast {
    using real = double;
    double gamma;
    double gamma_m1;
    double inv_gamma_m1;

    __host__ __device__ static __forceinline__ double div(double a, double b)
    {
#if defined(__HIP_DEVICE_COMPILE__)
        double y = __builtin_amdgcn_rcp(b);
        double e = __builtin_fma(-b, y, 1.0);
        y = __builtin_fma(y, e, y);
        e = __builtin_fma(-b, y, 1.0);
        y = __builtin_fma(y, e, y);
        return a * y;
#else
        return a / b;
#endif
    }
    __host__ __device__ static __forceinline__ double ratio(double a, double b) { return div(a, b); }
    __host__ __device__ __forceinline__ double evol_from_pres(double, double P) const { return P * inv_gamma_m1; }
    __host__ __device__ __forceinline__ double sound(double rho, double P) const { return sqrt(gamma * div(P, rho)); }
    template <int ND>
    __host__ __device__ __forceinline__ double internal_energy_fp64(const State<ND, double>& c) const
    {
        return c[ND + 1] - kinetic_cons_with(c, div);
    }
    template <int ND>
    __host__ __device__ __forceinline__ double pressure_of_cons(const State<ND, double>& c) const
    {
        return gamma_m1 * internal_energy_fp64(c);
    }
    template <int ND>
    __host__ __device__ __forceinline__ double pressure_of_cons(const State<ND, double>& c, double& evol) const
    {
        evol = internal_energy_fp64(c);
        return gamma_m1 * evol;
    }
};

// Policies whose state is FP64 (register budgets and kernel shapes follow the storage type).
template <class Eos>
constexpr bool kFp64State = std::is_same_v<typename Eos::real, double>;

// euler_equations.hpp:41-49
template <int ND, class T>
__host__ __device__ __forceinline__ T kinetic_cons(const State<ND, T>& c)
{
    if constexpr (sizeof(T) == 8) {
        return kinetic_cons_with(c, PerfectGas::div);
    } else {
        return kinetic_cons_with(c, PerfectGasMixed::div);
    }
}

// euler_equations.hpp:51-59: kinetic energy of a primitive state
template <int ND, class T>
__host__ __device__ __forceinline__ T kinetic_prim(const State<ND, T>& p)
{
    T n = 0;
#pragma unroll
    for (int d = 0; d < ND; ++d) n += p[1 + d] * p[1 + d];
    return p[0] * n * T(0.5);
}

// euler_equations.hpp:61-65 (FP64 reference arithmetic)
template <int ND>
__host__ __device__ __forceinline__ double internal_energy(const State<ND, double>& c)
{
    return c[ND + 1] - kinetic_cons(c);
}

// euler_equations.hpp:111-122
template <int ND, class Eos>
__host__ __device__ __forceinline__ State<ND, typename Eos::real> cons_to_prim(const State<ND, typename Eos::real>& c, const Eos& eos)
{
    State<ND, typename Eos::real> p;
    p[0] = c[0];
#pragma unroll
    for (int d = 0; d < ND; ++d) p[1 + d] = Eos::div(c[1 + d], c[0]);
    p[ND + 1] = eos.pressure_of_cons(c);
    return p;
}

// cons_to_prim that also returns the internal energy of the state (in FP64 for both policies),
// which the pressure recovery forms anyway (internal_energy.cpp:25-40 computes the same value)
template <int ND, class Eos>
__host__ __device__ __forceinline__ State<ND, typename Eos::real> cons_to_prim_e(const State<ND, typename Eos::real>& c, const Eos& eos,
                                                                                double& evol)
{
    State<ND, typename Eos::real> p;
    p[0] = c[0];
#pragma unroll
    for (int d = 0; d < ND; ++d) p[1 + d] = Eos::div(c[1 + d], c[0]);
    p[ND + 1] = eos.pressure_of_cons(c, evol);
    return p;
}

// euler_equations.hpp:124-134
template <int ND, class Eos>
__host__ __device__ __forceinline__ State<ND, typename Eos::real> prim_to_cons(const State<ND, typename Eos::real>& p, const Eos& eos)
{
    State<ND, typename Eos::real> c;
    c[0] = p[0];
#pragma unroll
    for (int d = 0; d < ND; ++d) c[1 + d] = p[0] * p[1 + d];
    c[ND + 1] = eos.evol_from_pres(p[0], p[ND + 1]) + kinetic_prim(p);
    return c;
}

// euler_equations.hpp:72-86: physical flux along `dir` of a primitive state whose total energy
// E = kinetic + internal is already known (prim_to_cons forms the same sum for the same state,
// evol + kinetic, and IEEE addition commutes, so a caller holding both reuses it)
template <int ND, class T>
__host__ __device__ __forceinline__ State<ND, T> flux_of_prim_e(const State<ND, T>& p, int dir, T E)
{
    State<ND, T> f;
    T const un = p[1 + dir];
    f[0] = p[0] * un;
#pragma unroll
    for (int d = 0; d < ND; ++d) f[1 + d] = p[0] * un * p[1 + d];
    f[1 + dir] += p[ND + 1];
    f[ND + 1] = un * (E + p[ND + 1]);
    return f;
}

// euler_equations.hpp:72-86: physical flux along `dir` from a primitive state
template <int ND, class Eos>
__host__ __device__ __forceinline__ State<ND, typename Eos::real> flux_of_prim(const State<ND, typename Eos::real>& p, int dir, const Eos& eos)
{
    using T = typename Eos::real;
    return flux_of_prim_e<ND, T>(p, dir, kinetic_prim(p) + eos.evol_from_pres(p[0], p[ND + 1]));
}

// euler_equations.hpp:93-109: physical flux along `dir` from a conservative state
template <int ND, class Eos>
__host__ __device__ __forceinline__ State<ND, typename Eos::real> flux_of_cons(const State<ND, typename Eos::real>& c, int dir, const Eos& eos)
{
    using T = typename Eos::real;
    State<ND, T> f;
    T const P = eos.pressure_of_cons(c);
    T const u = Eos::div(c[1 + dir], c[0]);
    f[0] = u * c[0];
#pragma unroll
    for (int d = 0; d < ND; ++d) f[1 + d] = u * c[1 + d];
    f[1 + dir] += P;
    f[ND + 1] = u * (c[ND + 1] + P);
    return f;
}

// ---------------------------------------------------------------- limiters
// slope_limiters.hpp:18-87.  diffR = (v[i+1]-v[i])/dx, diffL = (v[i]-v[i-1])/dx.
// Every division is Eos::ratio: dL can be subnormal and (r + 1)^2 can overflow.
template <int LIM, class Eos, class T>
__host__ __device__ __forceinline__ T limited_slope_t(T dR, T dL)
{
    if constexpr (LIM == kConstant) {
        return 0;
    } else {
        if (dL * dR > 0) {
            T const r = Eos::ratio(dR, dL);
            if constexpr (LIM == kVanLeer) {
                return Eos::ratio(T(0.5) * (dR + dL) * (4 * r), (r + 1) * (r + 1));
            } else if constexpr (LIM == kMinmod) {
                T const mm = Eos::ratio(2 * fmin(T(1), r), 1 + r);
                return mm * (dL + dR) * T(0.5);
            } else {
                return Eos::ratio(T(0.5) * (dR + dL) * (2 * r), (r * r) + 1);
            }
        }
        return 0;
    }
}

template <int LIM>
__host__ __device__ __forceinline__ double limited_slope_t(double dR, double dL)
{
    return limited_slope_t<LIM, PerfectGas, double>(dR, dL);
}

// The limited slope a policy uses.  PerfectGasFast and PerfectGasMixed (FP32: one reciprocal
// instead of two or three; measured +5 % on RT 512^3) evaluate the limiters of
// slope_limiters.hpp in closed form, with r = dR / dL and dL dR > 0:
//   Minmod     2 min(1, r) / (1 + r) (dL + dR) / 2  = dR if |dR| < |dL|, else dL;
//   VanLeer    (dR + dL) / 2 * 4 r / (r + 1)^2       = 2 dR dL / (dR + dL);
//   VanAlbada  (dR + dL) / 2 * 2 r / (r^2 + 1)       = (dR + dL) dR dL / (dR^2 + dL^2);
// PerfectGas follows the reference's expressions.
template <int LIM, class Eos, class T>
__host__ __device__ __forceinline__ T slope_limiter(T dR, T dL)
{
    if constexpr ((std::is_same_v<Eos, PerfectGasFast> || std::is_same_v<Eos, PerfectGasMixed>) && LIM != kConstant) {
        if (!(dL * dR > 0)) return 0;
        if constexpr (LIM == kMinmod) return fabs(dR) < fabs(dL) ? dR : dL;
        else if constexpr (LIM == kVanLeer) return Eos::div(2 * dR * dL, dR + dL);
        else return Eos::div((dR + dL) * dR * dL, (dR * dR) + (dL * dL));
    } else {
        return limited_slope_t<LIM, Eos, T>(dR, dL);
    }
}

// ---------------------------------------------------------------- Riemann solvers
// riemann_solver.hpp:60-64
template <class Eos, class T>
__host__ __device__ __forceinline__ T hll_component(T UL, T UR, T FL, T FR, T wsL, T wsR)
{
    return Eos::div((wsR * FL) - (wsL * FR) + (wsL * wsR * (UR - UL)), wsR - wsL);
}

// riemann_solver.hpp:30-58
template <int ND, class Eos>
__host__ __device__ __forceinline__ State<ND, typename Eos::real> riemann_hll(const State<ND, typename Eos::real>& cL,
                                                                             const State<ND, typename Eos::real>& cR, int dir, const Eos& eos)
{
    using T = typename Eos::real;
    State<ND, T> const pL = cons_to_prim(cL, eos);
    State<ND, T> const pR = cons_to_prim(cR, eos);
    T const sL = eos.sound(pL[0], pL[ND + 1]);
    T const sR = eos.sound(pR[0], pR[ND + 1]);
    T const wsL = fmin(pL[1 + dir] - sL, pR[1 + dir] - sR);
    T const wsR = fmax(pL[1 + dir] + sL, pR[1 + dir] + sR);
    T const nL = fmin(wsL, T(0));
    T const pR_ = fmax(wsR, T(0));
    State<ND, T> const fL = flux_of_prim(pL, dir, eos);
    State<ND, T> const fR = flux_of_prim(pR, dir, eos);
    State<ND, T> f;
#pragma unroll
    for (int v = 0; v < ND + 2; ++v) f[v] = hll_component<Eos>(cL[v], cR[v], fL[v], fR[v], nL, pR_);
    return f;
}

// riemann_solver.hpp:136-140
template <class T>
__host__ __device__ __forceinline__ T hllc_component(T F, T U, T Us, T ws)
{
    return F + (ws * (Us - U));
}

// riemann_solver.hpp:70-134
template <int ND, class Eos>
__host__ __device__ __forceinline__ State<ND, typename Eos::real> riemann_hllc(const State<ND, typename Eos::real>& cL,
                                                                              const State<ND, typename Eos::real>& cR, int dir, const Eos& eos)
{
    using T = typename Eos::real;
    State<ND, T> const pL = cons_to_prim(cL, eos);
    State<ND, T> const pR = cons_to_prim(cR, eos);
    T const sL = eos.sound(pL[0], pL[ND + 1]);
    T const sR = eos.sound(pR[0], pR[ND + 1]);
    T const uL = pL[1 + dir];
    T const uR = pR[1 + dir];
    T const wsL = fmin(uL - sL, uR - sR);
    T const wsR = fmax(uL + sL, uR + sR);
    T const rcL = pL[0] * (wsL - uL);
    T const rcR = pR[0] * (wsR - uR);
    T const ustar = Eos::div(pR[ND + 1] - pL[ND + 1] + (rcL * uL) - (rcR * uR), rcL - rcR);
    T const pstar = T(0.5) * (pL[ND + 1] + pR[ND + 1] + (rcL * (ustar - uL)) + (rcR * (ustar - uR)));
    bool const left = ustar > 0;
    T const S = left ? wsL : wsR;
    State<ND, T> const& cs = left ? cL : cR;
    State<ND, T> const& ps = left ? pL : pR;
    // the reference evaluates both physical fluxes and keeps the upwind one; only that one is
    // evaluated here (same expressions on the same state, so the same value)
    State<ND, T> const fs = flux_of_prim(ps, dir, eos);
    T const un = ps[1 + dir];
    if (wsL * wsR > 0) return fs;
    State<ND, T> f;
    T const rho_star = Eos::div(cs[0] * (S - un), S - ustar);
    f[0] = hllc_component(fs[0], cs[0], rho_star, S);
#pragma unroll
    for (int d = 0; d < ND; ++d) f[1 + d] = rho_star * ustar * ps[1 + d];
    f[1 + dir] = (rho_star * ustar * ustar) + pstar;
    T const E_star = (Eos::div(S - un, S - ustar) * cs[ND + 1]) + (Eos::div(ustar - un, S - ustar) * ((ustar * cs[0] * (S - un)) + ps[ND + 1]));
    f[ND + 1] = hllc_component(fs[ND + 1], cs[ND + 1], E_star, S);
    return f;
}

// riemann_solver.hpp:146-180 ("Low Mach" splitting solver)
template <int ND, class Eos>
__host__ __device__ __forceinline__ State<ND, typename Eos::real> riemann_lowmach(const State<ND, typename Eos::real>& cL,
                                                                                 const State<ND, typename Eos::real>& cR, int dir, const Eos& eos)
{
    using T = typename Eos::real;
    State<ND, T> const pL = cons_to_prim(cL, eos);
    State<ND, T> const pR = cons_to_prim(cR, eos);
    T const sL = eos.sound(pL[0], pL[ND + 1]);
    T const sR = eos.sound(pR[0], pR[ND + 1]);
    T const a = T(1.1) * fmax(pL[0] * sL, pR[0] * sR);
    T const ustar = ((pL[1 + dir] + pR[1 + dir]) * T(0.5)) - (Eos::div(T(1), 2 * a) * (pR[ND + 1] - pL[ND + 1]));
    T const Ma = Eos::div(fabs(ustar), fmin(sL, sR));
    T const theta = fmin(T(1), Ma);
    T const Pstar = ((pL[ND + 1] + pR[ND + 1]) * T(0.5)) - ((theta * a) * T(0.5) * (pR[1 + dir] - pL[1 + dir]));
    State<ND, T> const& cs = (ustar > 0) ? cL : cR;
    State<ND, T> f;
    f[0] = ustar * cs[0];
#pragma unroll
    for (int d = 0; d < ND; ++d) f[1 + d] = ustar * cs[1 + d];
    f[1 + dir] += Pstar;
    f[ND + 1] = (ustar * cs[ND + 1]) + (Pstar * ustar);
    return f;
}

template <int RS, int ND, class Eos>
__host__ __device__ __forceinline__ State<ND, typename Eos::real> riemann(const State<ND, typename Eos::real>& cL,
                                                                         const State<ND, typename Eos::real>& cR, int dir, const Eos& eos)
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
template <class Eos = PerfectGas, class T>
__host__ __device__ __forceinline__ T src_pressure(T dtodv, T PL, T PR, T dS, T dSp)
{
    return dtodv * (PL + PR) * T(0.5) * (dSp - dS);
}
template <class Eos = PerfectGas, class T>
__host__ __device__ __forceinline__ T src_centrifugal(T dtodv, T rL, T rR, T uL, T uR, T dS, T dSp)
{
    return dtodv * ((rL * uL * uL) + (rR * uR * uR)) * T(0.5) * (dSp - dS) * T(0.5);
}
template <class Eos = PerfectGas, class T>
__host__ __device__ __forceinline__ T src_coriolis_r(T dtodv, T x0, T x0p, T rL, T rR, T uLr, T uRr, T uLo, T uRo, T dS, T dSp)
{
    return Eos::div(dtodv * (x0p - x0), x0p + x0) * ((rR * uRo * uRr * dSp) + (rL * uLo * uLr * dS));
}
template <class Eos = PerfectGas, class T>
__host__ __device__ __forceinline__ T src_centrifugal_th(T dtodv, T cot_mid, T rL, T rR, T uLp, T uRp, T dS, T dSp)
{
    return dtodv * ((rL * uLp * uLp) + (rR * uRp * uRp)) * T(0.5) * cot_mid * (dSp - dS) * T(0.5);
}
template <class Eos = PerfectGas, class T>
__host__ __device__ __forceinline__ T src_coriolis_th(T dtodv, T sn, T snp, T rL, T rR, T uLt, T uRt, T uLp, T uRp, T dS, T dSp)
{
    return Eos::div(dtodv * (snp - sn), snp + sn) * ((rR * uRp * uRt * dSp) + (rL * uLp * uLt * dS));
}

}  // namespace hclpp
