// march.hpp -- the fused MUSCL-Hancock + Riemann + Godunov step kernel (k_step_march) and
// the per-cell helpers it shares with the ghost-fill and CFL kernels.
//
// Design (gfx950):
//  * The grid is cut into in-plane tiles (x in 2D; x-y in 3D) and segments
//    along the last active dim (the "march" dim).  A workgroup owns one tile
//    plus a one-cell halo ring ("Hancock region", one thread per column) and
//    marches through its segment one plane at a time.
//  * Per plane h: each column loads the conservative state at h+1, converts
//    to primitives, publishes plane h's primitives (plus a two-cell ring) in
//    LDS, reconstructs all faces (limited slopes along every dim: in-plane
//    neighbours from LDS, march neighbours from registers), converts the face
//    states to conservative form and applies the Hancock half step -- the
//    reference's face_reconstruction / conv_prim_to_cons / Hancock kernels for
//    one cell, in registers.
//  * In-plane faces: each column publishes its right face state in LDS, the
//    right neighbour solves the Riemann problem once per face, fluxes are
//    exchanged back through LDS.  March-dim faces: the right face state of
//    plane h-1 is kept in registers and paired with plane h's left state.
//  * The Godunov update of a cell accumulates dims in the reference order
//    (flux, geometric sources, gravity per dim, march dim last), so with the
//    FP64 policy every conserved value is bit-identical to the reference's
//    unfused sequence (down to the range of physics.hpp fdiv: values below
//    about 1e-290 may differ in their last bits).
//  * Face states never touch HBM: per cell-update the kernel reads the
//    conservative state once (plus halo re-reads served mostly by L2) and
//    writes it once.  The epilogue reduces (in FP64) the CFL rate of the new
//    state, the minimum internal energy of the consumed state and the
//    conserved totals; per-workgroup partials are folded in a fixed order.
//  * The kernels are generic over the precision policy (physics.hpp):
//    PerfectGas streams FP64; PerfectGasMixed streams FP32 (half the bytes)
//    with FP64 pressure recovery and FP64 reductions; PerfectGasFast streams
//    FP64 with the reference's formulas evaluated without its rounding.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "fused.hpp"
#include "launch.hpp"
#include "physics.hpp"
#include "pressure_fix.hpp"
#include "reduce.hpp"

namespace hclpp {

template <int ND>
__device__ __forceinline__ int64_t cell_at(const GridDev& G, int a, int b, int m)
{
    if constexpr (ND == 1) return (int64_t)m;
    else if constexpr (ND == 2) return (int64_t)a + (int64_t)G.nwg[0] * (int64_t)m;
    else return flat(G, a, b, m);
}

template <int ND>
__device__ __forceinline__ void ijk_of(int a, int b, int m, int& i, int& j, int& k)
{
    if constexpr (ND == 1) { i = m; j = 0; k = 0; }
    else if constexpr (ND == 2) { i = a; j = m; k = 0; }
    else { i = a; j = b; k = m; }
}

// Register exchange with the neighbouring column along x.  A tile row of ROW columns occupies
// ROW consecutive lanes (3D: 16, a DPP row, row_shr / row_shl; 2D: 64, the whole wave, wave_shr /
// wave_shl), so the column at x - 1 (FROM_LEFT) or x + 1 is the lane beside this one.  Every lane
// of the wave must execute the shift; the first (FROM_LEFT) or last lane of a row receives 0.
template <int ROW, bool FROM_LEFT>
__device__ __forceinline__ int lane_shift_u32(int v)
{
    static_assert(ROW == 16 || ROW == 64, "tile rows of 16 or 64 columns");
    constexpr int kCtrl = ROW == 16 ? (FROM_LEFT ? 0x111 : 0x101) : (FROM_LEFT ? 0x138 : 0x130);
    return __builtin_amdgcn_update_dpp(0, v, kCtrl, 0xF, 0xF, true);
}

template <int ROW, bool FROM_LEFT>
__device__ __forceinline__ double lane_shift(double v)
{
    int2 p = __builtin_bit_cast(int2, v);
    p.x = lane_shift_u32<ROW, FROM_LEFT>(p.x);
    p.y = lane_shift_u32<ROW, FROM_LEFT>(p.y);
    return __builtin_bit_cast(double, p);
}

template <int ROW, bool FROM_LEFT>
__device__ __forceinline__ float lane_shift(float v)
{
    return __builtin_bit_cast(float, lane_shift_u32<ROW, FROM_LEFT>(__builtin_bit_cast(int, v)));
}

template <int ND, class S>
__device__ __forceinline__ State<ND, S> load_cons(const StepArraysT<S>& A, const GridDev& G, int64_t id)
{
    State<ND, S> c;
    c[0] = A.rho[id];
#pragma unroll
    for (int d = 0; d < ND; ++d) c[1 + d] = A.rhou[id + G.plane * d];
    c[ND + 1] = A.E[id];
    return c;
}

template <int ND, class S>
__device__ __forceinline__ void store_cons(const StepArraysT<S>& A, const GridDev& G, int64_t id, const State<ND, S>& c)
{
    A.rho_new[id] = c[0];
#pragma unroll
    for (int d = 0; d < ND; ++d) A.rhou_new[id + G.plane * d] = c[1 + d];
    A.E_new[id] = c[ND + 1];
}

// CFL rate of a conservative state at (i,j,k) (time_step.hpp:61-76), in FP64.  With the
// FP64 policies this is the policy's conv_cons_to_prim, then the rate (the reference's
// arithmetic for PerfectGas).
template <int ND, int GEOM, class Eos>
__device__ __forceinline__ double state_rate(const GridDev& G, const Eos& eos, int i, int j, int k, const State<ND, typename Eos::real>& c)
{
    if constexpr (kFp64State<Eos>) {
        State<ND> const p = cons_to_prim(c, eos);
        return cfl_rate<ND, GEOM>(G, i, j, k, &p.q[1], eos.sound(p[0], p[ND + 1]));
    } else {
        double const rho = c[0];
        double u[ND];
#pragma unroll
        for (int d = 0; d < ND; ++d) u[d] = fdiv((double)c[1 + d], rho);
        double const P = eos.gamma_m1 * eos.internal_energy_fp64(c);
        return cfl_rate<ND, GEOM>(G, i, j, k, u, sqrt(fdiv(eos.gamma_d * P, rho)));
    }
}

// Internal energy E - |m|^2 / 2 rho of a conservative state in FP64 (internal_energy.cpp:25-40).
template <int ND, class Eos>
__device__ __forceinline__ double state_eint(const Eos& eos, const State<ND, typename Eos::real>& c)
{
    if constexpr (std::is_same_v<Eos, PerfectGas>) return internal_energy(c);
    else return eos.internal_energy_fp64(c);
}

// limited_linear_reconstruction.hpp:53-62 for one variable along one dim
template <int LIM, class Eos, class T>
__device__ __forceinline__ void reconstruct(T vm, T v, T vp, T dlm, T dl, T dlp, T& lo, T& hi)
{
    T const slope = slope_limiter<LIM, Eos, T>(Eos::div(vp - v, (dl + dlp) * T(0.5)), Eos::div(v - vm, (dlm + dl) * T(0.5)));
    lo = v - ((dl * T(0.5)) * slope);
    hi = v + ((dl * T(0.5)) * slope);
}

// Spherical sources of one dim for the cell's own face states (godunov_scheme.hpp:165-256),
// applied to the accumulator in the reference order.
template <class Eos, int ND, int GEOM>
__device__ __forceinline__ void godunov_sources(const GridDev& G, const Eos& eos, int dim, int i, int j, typename Eos::real dtodv,
                                                typename Eos::real dS, typename Eos::real dSp, const State<ND, typename Eos::real>& vL,
                                                const State<ND, typename Eos::real>& vR, State<ND, typename Eos::real>& un)
{
    using T = typename Eos::real;
    if constexpr (GEOM == kSpherical) {
        if (dim > 1) return;
        State<ND, T> const pL = cons_to_prim(vL, eos);
        State<ND, T> const pR = cons_to_prim(vR, eos);
        T const x0 = T(G.x[0][i]), x0p = T(G.x[0][i + 1]);
        if (dim == 0) {
            un[1] += src_pressure<Eos>(dtodv, pL[ND + 1], pR[ND + 1], dS, dSp);
            if constexpr (ND >= 2) {
                un[1] += src_centrifugal<Eos>(dtodv, pL[0], pR[0], pL[2], pR[2], dS, dSp);
                un[2] -= src_coriolis_r<Eos>(dtodv, x0, x0p, pL[0], pR[0], pL[1], pR[1], pL[2], pR[2], dS, dSp);
            }
            if constexpr (ND == 3) {
                un[1] += src_centrifugal<Eos>(dtodv, pL[0], pR[0], pL[3], pR[3], dS, dSp);
                un[3] -= src_coriolis_r<Eos>(dtodv, x0, x0p, pL[0], pR[0], pL[1], pR[1], pL[3], pR[3], dS, dSp);
            }
        } else {
            un[2] += src_pressure<Eos>(dtodv, pL[ND + 1], pR[ND + 1], dS, dSp);
            if constexpr (ND == 3) {
                un[2] += src_centrifugal_th<Eos>(dtodv, T(G.cot_mid[j]), pL[0], pR[0], pL[3], pR[3], dS, dSp);
                un[3] -= src_coriolis_th<Eos>(dtodv, T(G.sin_node[j]), T(G.sin_node[j + 1]), pL[0], pR[0], pL[2], pR[2], pL[3], pR[3], dS,
                                              dSp);
            }
        }
    }
}

// Passive scalars of one cell or face (mass fractions, or partial densities during an update);
// a one-slot placeholder when the step has none.
template <int NFX, class T>
struct Frac {
    T f[NFX > 0 ? NFX : 1];
};

template <int NFX, class T>
__device__ __forceinline__ Frac<NFX, T> zero_frac()
{
    Frac<NFX, T> x;
#pragma unroll
    for (int f = 0; f < (NFX > 0 ? NFX : 1); ++f) x.f[f] = 0;
    return x;
}

// MUSCL-Hancock predictor of one cell (hydro_reconstruction.hpp:99-128 and
// extrapolation_reconstruction.hpp:90-337): nb[dim][side] are the primitive
// neighbours (side 0 = minus, 1 = plus) along each dim; x and xnb the passive
// scalars of the cell and of the same neighbours, xf their face values after the
// half step (mass fractions, as fx_rec leaves the predictor).
template <class Eos, int ND, int GEOM, int LIM, int NFX>
__device__ __forceinline__ void predict_faces(const GridDev& G, const Eos& eos, typename Eos::real dth, typename Eos::real dtodv, int i, int j, int k,
                                              const State<ND, typename Eos::real>& p, const State<ND, typename Eos::real> (&nb)[ND][2],
                                              State<ND, typename Eos::real> (&fs)[2][ND], const Frac<NFX, typename Eos::real>& x,
                                              const Frac<NFX, typename Eos::real> (&xnb)[ND][2], Frac<NFX, typename Eos::real> (&xf)[2][ND])
{
    using T = typename Eos::real;
    constexpr int NV = ND + 2;
    State<ND, T> fp[2][ND];
    Frac<NFX, T> xo[2][ND];  // reconstructed mass fractions, kept for the scalar fluxes (extrapolation_reconstruction.hpp:87-88)
    int const idx[3] = {i, j, k};
#pragma unroll
    for (int dim = 0; dim < ND; ++dim) {
        T const dl = T(G.dx[dim][idx[dim]]);
        T const dlm = T(G.dx[dim][idx[dim] - 1]);
        T const dlp = T(G.dx[dim][idx[dim] + 1]);
#pragma unroll
        for (int v = 0; v < NV; ++v)
            reconstruct<LIM, Eos>(nb[dim][0][v], p[v], nb[dim][1][v], dlm, dl, dlp, fp[0][dim][v], fp[1][dim][v]);
        if constexpr (NFX > 0) {
#pragma unroll
            for (int f = 0; f < NFX; ++f)
                reconstruct<LIM, Eos>(xnb[dim][0].f[f], x.f[f], xnb[dim][1].f[f], dlm, dl, dlp, xo[0][dim].f[f], xo[1][dim].f[f]);
        }
    }
#pragma unroll
    for (int dim = 0; dim < ND; ++dim)
#pragma unroll
        for (int s = 0; s < 2; ++s) fs[s][dim] = prim_to_cons(fp[s][dim], eos);
    T rho_old[2][ND], mom_old[2][ND][ND], E_face[2][ND];
#pragma unroll
    for (int dim = 0; dim < ND; ++dim)
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            rho_old[s][dim] = fs[s][dim][0];
            E_face[s][dim] = fs[s][dim][ND + 1];  // the face's total energy, reused by its flux
#pragma unroll
            for (int v = 0; v < ND; ++v) mom_old[s][dim][v] = fs[s][dim][1 + v];
        }
    if constexpr (NFX > 0) {  // extrapolation_reconstruction.hpp:105-108: partial densities of the faces
#pragma unroll
        for (int dim = 0; dim < ND; ++dim)
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int f = 0; f < NFX; ++f) xf[s][dim].f[f] = xo[s][dim].f[f] * rho_old[s][dim];
    }
    T const x0 = T(G.x[0][i]), x0p = T(G.x[0][i + 1]);
#pragma unroll
    for (int dim = 0; dim < ND; ++dim) {
        State<ND, T> const& pL = fp[0][dim];
        State<ND, T> const& pR = fp[1][dim];
        State<ND, T> const fL = flux_of_prim_e<ND, T>(pL, dim, E_face[0][dim]);
        State<ND, T> const fR = flux_of_prim_e<ND, T>(pR, dim, E_face[1][dim]);
        T const dS = T(face_area<ND, GEOM>(G, dim, i, j, k));
        T const dSp = T(face_area<ND, GEOM>(G, dim, i + (dim == 0), j + (dim == 1), k + (dim == 2)));
#pragma unroll
        for (int pos = 0; pos < ND; ++pos)
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int v = 0; v < NV; ++v) fs[s][pos][v] += dtodv * ((fL[v] * dS) - (fR[v] * dSp));
        if constexpr (GEOM == kSpherical) {
#pragma unroll
            for (int pos = 0; pos < ND; ++pos)
#pragma unroll
                for (int s = 0; s < 2; ++s) {
                    if (dim == 0) {
                        fs[s][pos][1] += src_pressure<Eos>(dtodv, pL[ND + 1], pR[ND + 1], dS, dSp);
                        if constexpr (ND >= 2) {
                            fs[s][pos][1] += src_centrifugal<Eos>(dtodv, pL[0], pR[0], pL[2], pR[2], dS, dSp);
                            fs[s][pos][2] -= src_coriolis_r<Eos>(dtodv, x0, x0p, pL[0], pR[0], pL[1], pR[1], pL[2], pR[2], dS, dSp);
                        }
                        if constexpr (ND == 3) {
                            fs[s][pos][1] += src_centrifugal<Eos>(dtodv, pL[0], pR[0], pL[3], pR[3], dS, dSp);
                            fs[s][pos][3] -= src_coriolis_r<Eos>(dtodv, x0, x0p, pL[0], pR[0], pL[1], pR[1], pL[3], pR[3], dS, dSp);
                        }
                    }
                    if (dim == 1) {
                        fs[s][pos][2] += src_pressure<Eos>(dtodv, pL[ND + 1], pR[ND + 1], dS, dSp);
                        if constexpr (ND == 3) {
                            fs[s][pos][2] += src_centrifugal_th<Eos>(dtodv, T(G.cot_mid[j]), pL[0], pR[0], pL[3], pR[3], dS, dSp);
                            fs[s][pos][3] -= src_coriolis_th<Eos>(dtodv, T(G.sin_node[j]), T(G.sin_node[j + 1]), pL[0], pR[0], pL[2], pR[2],
                                                                  pL[3], pR[3], dS, dSp);
                        }
                    }
                }
        }
        T const g = T(gravity(G, i, dim));
#pragma unroll
        for (int pos = 0; pos < ND; ++pos)
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                fs[s][pos][1 + dim] += dth * g * rho_old[s][pos];
                fs[s][pos][ND + 1] += dth * g * mom_old[s][pos][dim];
            }
        if constexpr (NFX > 0) {  // extrapolation_reconstruction.hpp:315-323
#pragma unroll
            for (int f = 0; f < NFX; ++f) {
                T const ffL = xo[0][dim].f[f] * fL[0];
                T const ffR = xo[1][dim].f[f] * fR[0];
#pragma unroll
                for (int pos = 0; pos < ND; ++pos)
#pragma unroll
                    for (int s = 0; s < 2; ++s) xf[s][pos].f[f] += dtodv * ((ffL * dS) - (ffR * dSp));
            }
        }
    }
    if constexpr (NFX > 0) {  // passive_scalar_extrapolation (extrapolation_reconstruction.hpp:327-337)
#pragma unroll
        for (int dim = 0; dim < ND; ++dim)
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int f = 0; f < NFX; ++f) xf[s][dim].f[f] = Eos::div(xf[s][dim].f[f], fs[s][dim][0]);
    }
}

// Safe stand-in for cells outside the array (never feeds a valid output).
template <int ND, class T>
__device__ __forceinline__ State<ND, T> dummy_state()
{
    State<ND, T> p;
    p[0] = 1;
#pragma unroll
    for (int d = 0; d < ND; ++d) p[1 + d] = 0;
    p[ND + 1] = 1;
    return p;
}

// The columns of a march workgroup.  Workgroup b of nb owns a tile of columns (3D: HX x HY, 2D: a
// row of HX, 1D: NT segments of one column) and a segment of planes of the march dim; thread tid
// owns column (a, bb) (ghost-inclusive coordinates of the in-plane dims; a tile's first and last
// columns are halo) and planes [zs, ze).  Blocks that share an XCD (b % 8) get consecutive tiles,
// so halo re-reads hit that XCD's L2.
struct MarchColumn {
    int hx, hy, ox, oy, a, bb, zs, ze;
    bool in_block;      // the thread owns a column of the tile
    bool col_in_array;  // the column exists in memory
    bool hancock_ok;    // the column needs a predictor
    bool out_col;       // the column holds output cells
    bool seg_ok;        // the segment holds planes
};

template <int ND, int HX, int HY, int NT>
__device__ __forceinline__ MarchColumn march_column(const GridDev& G, const FusedLayout& L, int b, int nb, int tid)
{
    constexpr int MD = ND - 1;
    constexpr int TX = (ND >= 2) ? HX - 2 : 1;
    constexpr int TY = (ND == 3) ? HY - 2 : 1;
    MarchColumn c{};
    int const q = nb / 8, rem = nb % 8, x8 = b % 8;
    int const lb = x8 * q + (x8 < rem ? x8 : rem) + b / 8;
    int seg;
    if constexpr (ND == 1) {
        seg = lb * NT + tid;
    } else {
        c.hx = tid % HX;
        c.hy = tid / HX;
        int const tx = lb % L.ntile[0];
        int const rest = lb / L.ntile[0];
        int const ty = rest % L.ntile[1];
        seg = rest / L.ntile[1];
        c.ox = G.ngh[0] + tx * TX;
        c.oy = (ND == 3) ? G.ngh[1] + ty * TY : 0;
    }
    c.in_block = (ND == 1) || tid < HX * HY;
    c.a = c.ox - 1 + c.hx;
    c.bb = (ND == 3) ? c.oy - 1 + c.hy : 0;
    int const ngm = G.ngh[MD], nm = G.n[MD];
    c.zs = ngm + seg * L.seglen;
    c.ze = min(c.zs + L.seglen, ngm + nm);
    c.col_in_array = c.in_block;
    c.hancock_ok = c.in_block;
    c.out_col = c.in_block;
    if constexpr (ND >= 2) {
        c.col_in_array = c.col_in_array && c.a < G.nwg[0];
        c.hancock_ok = c.hancock_ok && c.a <= G.ngh[0] + G.n[0];
        c.out_col = c.out_col && c.hx >= 1 && c.hx <= TX && c.a < G.ngh[0] + G.n[0];
    }
    if constexpr (ND == 3) {
        c.col_in_array = c.col_in_array && c.bb < G.nwg[1];
        c.hancock_ok = c.hancock_ok && c.bb <= G.ngh[1] + G.n[1];
        c.out_col = c.out_col && c.hy >= 1 && c.hy <= TY && c.bb < G.ngh[1] + G.n[1];
    }
    c.seg_ok = c.zs < ngm + nm;
    if constexpr (ND == 1) {
        c.col_in_array = c.seg_ok;
        c.hancock_ok = c.seg_ok;
        c.out_col = c.seg_ok;
    }
    return c;
}

// Minimum waves per SIMD the register allocator must allow.  Measured on MI355X: the 3D
// FP32-storage kernel gains 23 % at 3 waves (168 VGPRs); the 3D Cartesian FP64 kernel gains 30 %
// at 2 waves (256 VGPRs, 44 B of spills per lane, with the LDS parking below; two workgroups
// fit a CU's LDS); the spherical 3D FP64 kernel loses 14 % at 2 waves (224 B of spills) and the
// 2D kernels lose when forced above their natural 2 or 4.
template <class Eos, int ND, int GEOM>
constexpr int march_min_waves()
{
    if constexpr (ND != 3) return 1;
    else if constexpr (!kFp64State<Eos>) return 3;
    else return GEOM == kCartesian ? 2 : 1;
}

// Load the NFX passive scalars (mass fractions, layout [f][cell]) of one cell.
template <int NFX, class S>
__device__ __forceinline__ Frac<NFX, S> load_frac(const StepArraysT<S>& A, const GridDev& G, int64_t id)
{
    Frac<NFX, S> x;
#pragma unroll
    for (int f = 0; f < NFX; ++f) x.f[f] = A.fx[id + G.plane * f];
    return x;
}

// NFX: passive scalars carried through the step (0, or 1 for the kernels of step_fused_fx.hip).
template <class Eos, int ND, int GEOM, int RS, int LIM, int NFX, int HX, int HY, int NT>
__global__ __launch_bounds__(NT, (march_min_waves<Eos, ND, GEOM>())) void k_step_march(GridDev G, Eos eos, StepArraysT<typename Eos::real> A, FusedLayout L)
{
    using T = typename Eos::real;
    using X = Frac<NFX, T>;
    constexpr int NV = ND + 2;
    constexpr int NL = NV + NFX;             // values per column in the LDS planes
    constexpr int NP = ND - 1;               // in-plane dims
    constexpr int MD = ND - 1;               // march dim
    constexpr int PX = (ND >= 2) ? HX + 2 : 1;
    constexpr int PY = (ND == 3) ? HY + 2 : 1;
    constexpr int TX = (ND >= 2) ? HX - 2 : 1;
    constexpr int TY = (ND == 3) ? HY - 2 : 1;
    constexpr int NXCH = NP > 0 ? NP : 1;
    // Barriers per plane: one after the primitive plane (double-buffered by plane parity) and, in
    // 3D, one after the y face states and one after the y fluxes, which have buffers of their own,
    // so no write waits for the readers of the previous plane's buffer beyond the barrier that
    // already separates them.  x faces are exchanged between lanes (lane_shift), not through LDS.
    __shared__ T s_prim[2][NL][PY][PX];
    constexpr int NYB = ND == 3 ? 1 : 0;  // LDS exchange buffers (the y dim of a 3D tile)
    __shared__ T s_st[NYB > 0 ? NYB : 1][NYB > 0 ? NL : 1][NYB > 0 ? HY : 1][NYB > 0 ? HX : 1];
    __shared__ T s_fl[NYB > 0 ? NYB : 1][NYB > 0 ? NL : 1][NYB > 0 ? HY : 1][NYB > 0 ? HX : 1];
    __shared__ double s_red[NT / 64];
    // Per-thread parking of the values plane h-1 hands to plane h (its right march face state,
    // its partial update and its left march flux, and the scalar's): written at the end of a
    // plane, read back when plane h completes plane h-1.  Only the owning thread touches its
    // slots, so no barrier is involved, and the registers are free across the plane's
    // arithmetic.  Used where it pays (measured): the 3D FP32-storage kernel, whose register
    // budget is set for 3 waves per SIMD (spills go from 40-152 to 0-92 B per lane; RT 512^3
    // +6 %, spherical Sedov +26 %), and the 3D Cartesian FP64 kernel run at 2 waves per SIMD.
    // In spherical FP64 and in 2D it gained nothing or lost.
    constexpr bool kPark = ND == 3 && (!kFp64State<Eos> || GEOM == kCartesian);
    constexpr int NPARK = kPark ? 3 * NV + 3 * NFX : 1;
    __shared__ T s_park[NPARK][kPark ? NT : 1];

    DevClock const* clk = A.clock;
    if (clk->pub.done || clk->pub.status) return;
    double const dt_d = clk->pub.dt;
    double const dth_d = dt_d / 2;
    T const dt = T(dt_d);
    T const dth = T(dth_d);

    int const tid = threadIdx.x;
    MarchColumn const mc = march_column<ND, HX, HY, NT>(G, L, blockIdx.x, gridDim.x, tid);
    int const hx = mc.hx, hy = mc.hy, ox = mc.ox, oy = mc.oy, a = mc.a, bb = mc.bb, zs = mc.zs, ze = mc.ze;
    bool const in_block = mc.in_block, col_in_array = mc.col_in_array, hancock_ok = mc.hancock_ok, out_col = mc.out_col;
    bool const seg_ok = mc.seg_ok;

    // ring cells of the primitive plane (in-plane distance 2 from the output tile)
    constexpr int NRING = (ND == 3) ? 2 * (PX - 2) + 2 * (PY - 2) : ((ND == 2) ? 2 : 0);

    double r_inv = reduce_identity<kMax>();
    double r_emin = reduce_identity<kMin>();
    double r_sum[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) r_sum[v] = 0;

    auto prim_at = [&](int ca, int cb, int m) -> State<ND, T> {
        bool ok = m >= 0 && m < G.nwg[MD];
        if constexpr (ND >= 2) ok = ok && ca >= 0 && ca < G.nwg[0];
        if constexpr (ND == 3) ok = ok && cb >= 0 && cb < G.nwg[1];
        if (!ok) return dummy_state<ND, T>();
        return cons_to_prim(load_cons<ND>(A, G, cell_at<ND>(G, ca, cb, m)), eos);
    };
    auto frac_at = [&](int ca, int cb, int m) -> X {
        bool ok = m >= 0 && m < G.nwg[MD];
        if constexpr (ND >= 2) ok = ok && ca >= 0 && ca < G.nwg[0];
        if constexpr (ND == 3) ok = ok && cb >= 0 && cb < G.nwg[1];
        if (!ok) return zero_frac<NFX, T>();
        return load_frac<NFX>(A, G, cell_at<ND>(G, ca, cb, m));
    };

    // ring cell of this thread (at most one: NRING <= NT), fixed for the whole march
    static_assert(NRING <= NT, "one ring cell per thread");
    bool const has_ring = tid < NRING;
    int ring_px = 0, ring_py = 0;
    if constexpr (ND == 2) {
        ring_px = tid == 0 ? 0 : PX - 1;
    } else if constexpr (ND == 3) {
        int const r = tid;
        if (r < PX - 2) { ring_px = 1 + r; ring_py = 0; }
        else if (r < 2 * (PX - 2)) { ring_px = 1 + r - (PX - 2); ring_py = PY - 1; }
        else if (r < 2 * (PX - 2) + (PY - 2)) { ring_px = 0; ring_py = 1 + r - 2 * (PX - 2); }
        else { ring_px = PX - 1; ring_py = 1 + r - 2 * (PX - 2) - (PY - 2); }
    }
    int const ring_a = ox - 2 + ring_px, ring_b = (ND == 3) ? oy - 2 + ring_py : 0;
    bool ring_in = has_ring;
    if constexpr (ND >= 2) ring_in = ring_in && ring_a >= 0 && ring_a < G.nwg[0];
    if constexpr (ND == 3) ring_in = ring_in && ring_b >= 0 && ring_b < G.nwg[1];

    // pipeline registers.  Global loads run one plane ahead of their use (own column: plane h+2
    // while plane h is processed; ring cell: plane h+1), so their latency hides behind a plane
    // of arithmetic instead of stalling the workgroup at the LDS barrier.
    State<ND, T> pm, p0, pp, c0, cp, cnext, rnext;
    // Carried from plane to plane where registers allow (measured): the internal energies of c0
    // and cp that their pressure recovery forms, and dt / dV of plane h-1 for its completion.
    // The 3D Cartesian FP64 kernel (2 waves per SIMD) and the 3D spherical FP32 one recompute
    // them instead: the extra registers spill there.
    constexpr bool kCarry = !(ND == 3 && (kFp64State<Eos> == (GEOM == kCartesian)));
    double e0 = 1, ep = 1;
    T dtodv_prev = 0;
    State<ND, T> fsR_prev, Fm_prev, part_prev;  // in registers, or in s_park when kPark
    T old_rho_prev = 0, old_mom_prev = 0, src_prev = 0;
    // pressure fix: the cell of plane h-1 is rewritten by k_pressure_fix_fused, which also supplies its rate
    bool pf_prev = false, pf_cur = false;
    // set when the fix rewrites some consumed cell of this workgroup (in LDS, not in a register
    // live across the march)
    __shared__ int s_pf_any;
    if (tid == 0) s_pf_any = 0;
    __syncthreads();
    // passive scalars: mass fractions of planes h-1, h, h+1 and the prefetch; the ring cell's
    // prefetch; the right march face of plane h-1, its partial-density update and the
    // upwinded scalar flux of its left march face (godunov_scheme.hpp:263-290)
    X xm = zero_frac<NFX, T>(), x0 = xm, xp = xm, xnext = xm, rxnext = xm;
    X xfR_prev = xm, xpart_prev = xm, ffm_prev = xm;
    int const zfirst = seg_ok ? zs - 1 : 0;
    int const zlast = seg_ok ? ze : -1;
    bool rnext_ok = false;
    if (col_in_array) {
        pm = prim_at(a, bb, zfirst - 1);
        c0 = load_cons<ND>(A, G, cell_at<ND>(G, a, bb, zfirst));
        if constexpr (kCarry) p0 = cons_to_prim_e(c0, eos, e0);
        else p0 = cons_to_prim(c0, eos);
        if (zfirst <= zlast) cnext = load_cons<ND>(A, G, cell_at<ND>(G, a, bb, zfirst + 1));
        if constexpr (NFX > 0) {
            xm = frac_at(a, bb, zfirst - 1);
            x0 = load_frac<NFX>(A, G, cell_at<ND>(G, a, bb, zfirst));
            if (zfirst <= zlast) xnext = load_frac<NFX>(A, G, cell_at<ND>(G, a, bb, zfirst + 1));
        }
    } else {
        pm = dummy_state<ND, T>();
        p0 = dummy_state<ND, T>();
        c0 = dummy_state<ND, T>();
    }
    if constexpr (ND >= 2) {
        rnext_ok = ring_in && zfirst >= 0 && zfirst < G.nwg[MD] && zfirst <= zlast;
        if (rnext_ok) {
            rnext = load_cons<ND>(A, G, cell_at<ND>(G, ring_a, ring_b, zfirst));
            if constexpr (NFX > 0) rxnext = load_frac<NFX>(A, G, cell_at<ND>(G, ring_a, ring_b, zfirst));
        }
    }

    if constexpr (kPark) {  // plane zfirst-1's primitives of the own column, where the loop reads them
        if (in_block) {
#pragma unroll
            for (int v = 0; v < NV; ++v) s_prim[(zfirst - 1) & 1][v][(ND == 3) ? hy + 1 : 0][hx + 1] = pm[v];
        }
    }
    auto park = [&](int slot, T v) {
        if constexpr (kPark) s_park[slot][tid] = v;
    };
    auto unpark = [&](int slot) -> T {
        if constexpr (kPark) return s_park[slot][tid];
        else return T(0);
    };

    for (int h = zfirst; h <= zlast; ++h) {
        int const pb = h & 1;
        // ---- plane h+1 of this column (prefetched), and the prefetch of plane h+2
        if (col_in_array) {
            cp = cnext;
            xp = xnext;
            if (h + 1 <= zlast) {
                cnext = load_cons<ND>(A, G, cell_at<ND>(G, a, bb, h + 2));
                if constexpr (NFX > 0) xnext = load_frac<NFX>(A, G, cell_at<ND>(G, a, bb, h + 2));
            }
            if constexpr (kCarry) pp = cons_to_prim_e(cp, eos, ep);
            else pp = cons_to_prim(cp, eos);
        } else {
            cp = dummy_state<ND, T>();
            pp = dummy_state<ND, T>();
        }
        // ---- plane h primitives in LDS (own column + ring)
        if constexpr (ND >= 2) {
            State<ND, T> ring_c;
            X const rx = rnext_ok ? rxnext : zero_frac<NFX, T>();
            bool const ring_ok = rnext_ok;
            if (ring_ok) ring_c = rnext;
            rnext_ok = ring_in && h + 1 < G.nwg[MD] && h + 1 <= zlast;
            if (rnext_ok) {
                rnext = load_cons<ND>(A, G, cell_at<ND>(G, ring_a, ring_b, h + 1));
                if constexpr (NFX > 0) rxnext = load_frac<NFX>(A, G, cell_at<ND>(G, ring_a, ring_b, h + 1));
            }
            State<ND, T> const pr = ring_ok ? cons_to_prim(ring_c, eos) : dummy_state<ND, T>();
            // buffer h & 1 was last read at plane h - 2, before the barrier of plane h - 1
            if (in_block) {
#pragma unroll
                for (int v = 0; v < NV; ++v) s_prim[pb][v][(ND == 3) ? hy + 1 : 0][hx + 1] = p0[v];
#pragma unroll
                for (int f = 0; f < NFX; ++f) s_prim[pb][NV + f][(ND == 3) ? hy + 1 : 0][hx + 1] = x0.f[f];
            }
            if (has_ring) {
#pragma unroll
                for (int v = 0; v < NV; ++v) s_prim[pb][v][ring_py][ring_px] = pr[v];
#pragma unroll
                for (int f = 0; f < NFX; ++f) s_prim[pb][NV + f][ring_py][ring_px] = rx.f[f];
            }
            __syncthreads();
        }
        int ci, cj, ck;
        ijk_of<ND>(a, bb, h, ci, cj, ck);
        bool const out_plane = h >= zs && h < ze;

        // ---- predictor of plane h
        State<ND, T> fs[2][ND];
        X xf[2][ND];
        double qv = 0;
        if (hancock_ok) {
            State<ND, T> nb3[ND][2];
            X xnb[ND][2];
            if constexpr (ND >= 2) {
                int const py0 = (ND == 3) ? hy + 1 : 0;
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    nb3[0][0][v] = s_prim[pb][v][py0][hx];
                    nb3[0][1][v] = s_prim[pb][v][py0][hx + 2];
                }
#pragma unroll
                for (int f = 0; f < NFX; ++f) {
                    xnb[0][0].f[f] = s_prim[pb][NV + f][py0][hx];
                    xnb[0][1].f[f] = s_prim[pb][NV + f][py0][hx + 2];
                }
                if constexpr (ND == 3) {
#pragma unroll
                    for (int v = 0; v < NV; ++v) {
                        nb3[1][0][v] = s_prim[pb][v][hy][hx + 1];
                        nb3[1][1][v] = s_prim[pb][v][hy + 2][hx + 1];
                    }
#pragma unroll
                    for (int f = 0; f < NFX; ++f) {
                        xnb[1][0].f[f] = s_prim[pb][NV + f][hy][hx + 1];
                        xnb[1][1].f[f] = s_prim[pb][NV + f][hy + 2][hx + 1];
                    }
                }
            }
            if constexpr (kPark) {  // plane h-1 of the own column, still in the other buffer
#pragma unroll
                for (int v = 0; v < NV; ++v) nb3[MD][0][v] = s_prim[pb ^ 1][v][(ND == 3) ? hy + 1 : 0][hx + 1];
            } else {
                nb3[MD][0] = pm;
            }
            nb3[MD][1] = pp;
            xnb[MD][0] = xm;
            xnb[MD][1] = xp;
            // dt / dV once per cell where registers allow (kCarry): the half step's (dt / 2) / dV is
            // exactly half of it (a power-of-two scaling commutes with rounding in the normal range)
            T dtodv_h;
            if constexpr (kCarry) {
                qv = fdiv(dt_d, cell_volume<ND, GEOM>(G, ci, cj, ck));
                dtodv_h = T(qv * 0.5);
            } else {
                dtodv_h = T(fdiv(dth_d, cell_volume<ND, GEOM>(G, ci, cj, ck)));
            }
            predict_faces<Eos, ND, GEOM, LIM, NFX>(G, eos, dth, dtodv_h, ci, cj, ck, p0, nb3, fs, x0, xnb, xf);
        } else {
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int d = 0; d < ND; ++d) {
                    fs[s][d] = c0;
                    xf[s][d] = x0;
                }
        }

        // ---- in-plane fluxes and the in-plane part of the update of plane h
        State<ND, T> part = c0;
        T old_rho = c0[0], old_mom = c0[1 + MD], src_m = 0;
        X xpart;  // partial densities fx * rho (godunov_scheme.hpp:122) plus the in-plane fluxes
#pragma unroll
        for (int f = 0; f < NFX; ++f) xpart.f[f] = x0.f[f] * c0[0];
        T dtodv = 0;
        if (out_plane) {
            if (out_col) {
                dtodv = kCarry ? T(qv) : T(fdiv(dt_d, cell_volume<ND, GEOM>(G, ci, cj, ck)));
                // min internal energy of the consumed state (internal_energy.cpp:25-40), as its
                // pressure recovery formed it
                r_emin = reduce_combine<kMin>(r_emin, kCarry ? e0 : state_eint<ND>(eos, c0));
                if constexpr (std::is_same_v<Eos, PerfectGas>) {
                    if (A.pf) {
                        pf_cur = pressure_fix_applies(c0, A.eps_pf);  // pressure_fix.hpp:61-68
                        if (pf_cur) s_pf_any = 1;
                    }
                }
            }
            if constexpr (ND >= 2) {
                // Riemann solve of the left face along in-plane dim p, given the right face state
                // (and scalar face values) of the column on the left; a column that needs no
                // such face yields its own state as a placeholder flux
                auto solve = [&](int p, bool need, const State<ND, T>& left, const X& xleft, State<ND, T>& F, X& XF) {
                    XF = zero_frac<NFX, T>();
                    if (need) {
                        F = riemann<RS>(left, fs[0][p], p, eos);
                        // upwind face of the scalar (godunov_scheme.hpp:265-287): this cell's
                        // left face when the mass flux is negative, else the left cell's right face
#pragma unroll
                        for (int f = 0; f < NFX; ++f) XF.f[f] = (F[0] < 0 ? xf[0][p].f[f] : xleft.f[f]) * F[0];
                    } else {
                        F = c0;
                    }
                };
                // the update of dim p (flux, sources, gravity) in the reference order: dims in turn
                auto accumulate = [&](int p, const State<ND, T>& FL, const State<ND, T>& FR, const X& XL, const X& XR) {
                    T const dS = T(face_area<ND, GEOM>(G, p, ci, cj, ck));
                    T const dSp = T(face_area<ND, GEOM>(G, p, ci + (p == 0), cj + (p == 1), ck));
#pragma unroll
                    for (int v = 0; v < NV; ++v) part[v] += dtodv * ((FL[v] * dS) - (FR[v] * dSp));
                    godunov_sources<Eos, ND, GEOM>(G, eos, p, ci, cj, dtodv, dS, dSp, fs[0][p], fs[1][p], part);
                    T const g = T(gravity(G, ci, p));
                    part[1 + p] += dt * g * c0[0];
                    part[ND + 1] += dt * g * c0[1 + p];
#pragma unroll
                    for (int f = 0; f < NFX; ++f) xpart.f[f] += dtodv * ((XL.f[f] * dS) - (XR.f[f] * dSp));
                };
                // y (3D): right face states through LDS; s_st was last read before the flux
                // barrier of the previous plane
                if constexpr (ND == 3) {
                    if (in_block) {
#pragma unroll
                        for (int v = 0; v < NV; ++v) s_st[0][v][hy][hx] = fs[1][1][v];
#pragma unroll
                        for (int f = 0; f < NFX; ++f) s_st[0][NV + f][hy][hx] = xf[1][1].f[f];
                    }
                }
                {  // x: the column at x - 1 is the lane beside (every lane executes the shifts)
                    State<ND, T> left;
                    X xleft;
#pragma unroll
                    for (int v = 0; v < NV; ++v) left[v] = lane_shift<HX, true>(fs[1][0][v]);
#pragma unroll
                    for (int f = 0; f < NFX; ++f) xleft.f[f] = lane_shift<HX, true>(xf[1][0].f[f]);
                    bool const need = hancock_ok && hx >= 1 && (ND == 2 || (hy >= 1 && hy <= TY));
                    State<ND, T> FL, FR;
                    X XL, XR;
                    solve(0, need, left, xleft, FL, XL);
                    // the right face's flux is the left flux of the column at x + 1
#pragma unroll
                    for (int v = 0; v < NV; ++v) FR[v] = lane_shift<HX, false>(FL[v]);
#pragma unroll
                    for (int f = 0; f < NFX; ++f) XR.f[f] = lane_shift<HX, false>(XL.f[f]);
                    if (out_col) accumulate(0, FL, FR, XL, XR);
                }
                if constexpr (ND == 3) {  // y: left states and right fluxes through LDS
                    __syncthreads();
                    bool const need = hancock_ok && hy >= 1 && hx >= 1 && hx <= TX;
                    State<ND, T> left, FL, FR;
                    X xleft = zero_frac<NFX, T>(), XL, XR;
                    if (need) {
#pragma unroll
                        for (int v = 0; v < NV; ++v) left[v] = s_st[0][v][hy - 1][hx];
#pragma unroll
                        for (int f = 0; f < NFX; ++f) xleft.f[f] = s_st[0][NV + f][hy - 1][hx];
                    }
                    solve(1, need, left, xleft, FL, XL);
                    // s_fl was last read before the primitive barrier of this plane
                    if (in_block) {
#pragma unroll
                        for (int v = 0; v < NV; ++v) s_fl[0][v][hy][hx] = FL[v];
#pragma unroll
                        for (int f = 0; f < NFX; ++f) s_fl[0][NV + f][hy][hx] = XL.f[f];
                    }
                    __syncthreads();
                    if (out_col) {
#pragma unroll
                        for (int v = 0; v < NV; ++v) FR[v] = s_fl[0][v][hy + 1][hx];
#pragma unroll
                        for (int f = 0; f < NFX; ++f) XR.f[f] = s_fl[0][NV + f][hy + 1][hx];
                        accumulate(1, FL, FR, XL, XR);
                    }
                }
            }
            // geometric pressure source of the march dim (1D spherical: r; 2D spherical: theta)
            if constexpr (GEOM == kSpherical && MD <= 1) {
                if (out_col) {
                    State<ND, T> const pL = cons_to_prim(fs[0][MD], eos);
                    State<ND, T> const pR = cons_to_prim(fs[1][MD], eos);
                    T const dS = T(face_area<ND, GEOM>(G, MD, ci, cj, ck));
                    T const dSp = T(face_area<ND, GEOM>(G, MD, ci + (MD == 0), cj + (MD == 1), ck));
                    src_m = src_pressure<Eos>(dtodv, pL[ND + 1], pR[ND + 1], dS, dSp);
                }
            }
        }

        // ---- march-dim face h-1/2 and completion of plane h-1
        if (h >= zs && out_col) {
            if constexpr (kPark) {
#pragma unroll
                for (int v = 0; v < NV; ++v) fsR_prev[v] = unpark(v);
            }
            State<ND, T> const Fm = riemann<RS>(fsR_prev, fs[0][MD], MD, eos);
            X XFm;  // upwinded scalar flux of the face h-1/2
#pragma unroll
            for (int f = 0; f < NFX; ++f) {
                T const xr = kPark ? unpark(3 * NV + f) : xfR_prev.f[f];
                XFm.f[f] = (Fm[0] < 0 ? xf[0][MD].f[f] : xr) * Fm[0];
            }
            int const hm = h - 1;
            if (hm >= zs) {
                int pi, pj, pk;
                ijk_of<ND>(a, bb, hm, pi, pj, pk);
                double const vol = cell_volume<ND, GEOM>(G, pi, pj, pk);
                T const dtodv_m = kCarry ? dtodv_prev : T(fdiv(dt_d, vol));
                T const dS = T(face_area<ND, GEOM>(G, MD, pi, pj, pk));
                T const dSp = T(face_area<ND, GEOM>(G, MD, ci, cj, ck));
                if constexpr (kPark) {
#pragma unroll
                    for (int v = 0; v < NV; ++v) {
                        part_prev[v] = unpark(NV + v);
                        Fm_prev[v] = unpark(2 * NV + v);
                    }
#pragma unroll
                    for (int f = 0; f < NFX; ++f) {
                        xpart_prev.f[f] = unpark(3 * NV + NFX + f);
                        ffm_prev.f[f] = unpark(3 * NV + 2 * NFX + f);
                    }
                }
                State<ND, T> un = part_prev;
#pragma unroll
                for (int v = 0; v < NV; ++v) un[v] += dtodv_m * ((Fm_prev[v] * dS) - (Fm[v] * dSp));
                if constexpr (GEOM == kSpherical && MD <= 1) un[1 + MD] += src_prev;
                T const g = T(gravity(G, pi, MD));
                un[1 + MD] += dt * g * old_rho_prev;
                un[ND + 1] += dt * g * old_mom_prev;
                if (!A.scalars_only) store_cons<ND>(A, G, cell_at<ND>(G, a, bb, hm), un);
                if constexpr (NFX > 0) {  // godunov_scheme.hpp:288, then passive_scalar_Godunov (:294-301)
                    int64_t const id = cell_at<ND>(G, a, bb, hm);
#pragma unroll
                    for (int f = 0; f < NFX; ++f) {
                        T const xn = xpart_prev.f[f] + dtodv_m * ((ffm_prev.f[f] * dS) - (XFm.f[f] * dSp));
                        T const y = Eos::div(xn, un[0]);
                        A.fx_new[id + G.plane * f] = y < T(0) ? T(0) : (y > T(1) ? T(1) : y);
                    }
                }
                // epilogue: CFL rate of the new state and the conserved totals (FP64); a cell the
                // pressure fix rewrites gets its rate from k_pressure_fix_fused
                if (!pf_prev) r_inv = reduce_combine<kMax>(r_inv, state_rate<ND, GEOM>(G, eos, pi, pj, pk, un));
#pragma unroll
                for (int v = 0; v < NV; ++v) r_sum[v] += (double)un[v] * vol;
            }
            if constexpr (kPark) {
#pragma unroll
                for (int v = 0; v < NV; ++v) park(2 * NV + v, Fm[v]);
#pragma unroll
                for (int f = 0; f < NFX; ++f) park(3 * NV + 2 * NFX + f, XFm.f[f]);
            } else {
                Fm_prev = Fm;
                ffm_prev = XFm;
            }
        }

        // ---- rotate the pipeline
        if constexpr (kPark) {
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                park(v, fs[1][MD][v]);
                park(NV + v, part[v]);
            }
#pragma unroll
            for (int f = 0; f < NFX; ++f) {
                park(3 * NV + f, xf[1][MD].f[f]);
                park(3 * NV + NFX + f, xpart.f[f]);
            }
        } else {
            fsR_prev = fs[1][MD];
            part_prev = part;
            xfR_prev = xf[1][MD];
            xpart_prev = xpart;
        }
        xm = x0;
        x0 = xp;
        old_rho_prev = old_rho;
        old_mom_prev = old_mom;
        src_prev = src_m;
        pf_prev = pf_cur;
        pf_cur = false;
        pm = p0;
        p0 = pp;
        c0 = cp;
        if constexpr (kCarry) {
            e0 = ep;
            dtodv_prev = dtodv;
        }
    }

    // ---- workgroup partials
    r_inv = block_reduce<kMax>(r_inv, s_red, tid, NT);
    r_emin = block_reduce<kMin>(r_emin, s_red, tid, NT);
#pragma unroll
    for (int v = 0; v < NV; ++v) r_sum[v] = block_reduce<kSum>(r_sum[v], s_red, tid, NT);
    static_assert(2 + NV < kPartialStride, "the pressure-fix flag has a slot of its own");
    if (tid == 0 && !A.scalars_only) {
        double* out = A.partials + (int64_t)blockIdx.x * kPartialStride;
        out[0] = r_inv;
        out[1] = r_emin;
#pragma unroll
        for (int v = 0; v < NV; ++v) out[2 + v] = r_sum[v];
        out[kPartialStride - 1] = s_pf_any ? 1.0 : 0.0;  // the reductions' barriers ordered every write
    }
}

// Runtime dispatch over dimension, geometry, Riemann solver and limiter: calls
// v.template run<ND, GEOM, RS, LIM>() for the configuration given.
template <class V>
static void march_visit(int nd, int geom, int rs, int lim, V&& v)
{
#define HCLPP_F4(NDV, GV, RSV, LV) v.template run<NDV, GV, RSV, LV>()
#define HCLPP_FL(NDV, GV, RSV)                                     \
    switch (lim) {                                                 \
    case kConstant: HCLPP_F4(NDV, GV, RSV, kConstant); break;      \
    case kVanLeer: HCLPP_F4(NDV, GV, RSV, kVanLeer); break;        \
    case kMinmod: HCLPP_F4(NDV, GV, RSV, kMinmod); break;          \
    default: HCLPP_F4(NDV, GV, RSV, kVanAlbada); break;            \
    }
#define HCLPP_FR(NDV, GV)                              \
    switch (rs) {                                      \
    case kHLL: HCLPP_FL(NDV, GV, kHLL); break;         \
    case kHLLC: HCLPP_FL(NDV, GV, kHLLC); break;       \
    default: HCLPP_FL(NDV, GV, kLowMach); break;       \
    }
#define HCLPP_FG(NDV)                                                  \
    if (geom == kSpherical) { HCLPP_FR(NDV, kSpherical); }             \
    else { HCLPP_FR(NDV, kCartesian); }
    switch (nd) {
    case 1: HCLPP_FG(1); break;
    case 2: HCLPP_FG(2); break;
    default: HCLPP_FG(3); break;
    }
#undef HCLPP_F4
#undef HCLPP_FL
#undef HCLPP_FR
#undef HCLPP_FG
}

// Launch of k_step_march with the context's plan (tile counts, segment length, workgroups).
template <class Eos, int NFX>
struct MarchLaunch {
    const GridDev& G;
    const MarchPlan& P;
    const Eos& eos;
    const StepArraysT<typename Eos::real>& A;
    hipStream_t s;
    template <int ND, int GEOM, int RS, int LIM>
    void run() const
    {
        using S = MarchShape<ND>;
        if (P.nb == 0) return;
        hipLaunchKernelGGL((k_step_march<Eos, ND, GEOM, RS, LIM, NFX, S::HX, S::HY, S::NT>), dim3((unsigned)P.nb), dim3(S::NT), 0, s, G, eos,
                           A, P.L);
    }
};

// Resident workgroups per CU of one k_step_march instance (its registers and LDS decide).
template <class Eos, int NFX>
struct MarchOccupancy {
    int* out;
    template <int ND, int GEOM, int RS, int LIM>
    void run() const
    {
        using S = MarchShape<ND>;
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_step_march<Eos, ND, GEOM, RS, LIM, NFX, S::HX, S::HY, S::NT>, S::NT, 0) !=
            hipSuccess)
            n = 0;
        *out = n;
    }
};

}  // namespace hclpp
