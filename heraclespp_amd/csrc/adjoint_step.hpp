// adjoint_step.hpp -- per-cell and per-face bodies of the reverse-mode hydro step.
//
// The forward step is
//   Ug = ghost_fill(U);  p = prim(Ug)                         (all cells)
//   pf[s][d] = p -/+ dl/2 * slope_d(p)                        (cells within 1 of the interior)
//   Uf[s][d] = cons(pf[s][d]) + sum_e [dth/dV (F_e(pf[0][e]) dS - F_e(pf[1][e]) dS+) + Q_e(pf[0][e], pf[1][e])]
//              + gravity(dth)
//   U'(c) = U(c) + sum_d [dt/dV (F*(face c-1/2) dS - F*(face c+1/2) dS+) + Q_d(prim(Uf[0][d]), prim(Uf[1][d]))]
//           + gravity(dt)
// with dth = dt / 2, F* the Riemann flux of (Uf[1][d] of the left cell, Uf[0][d] of the right one)
// and Q_d the spherical geometric sources of dims 0 and 1 (zero in Cartesian geometry), formed
// from the face states of the cell's own faces (hydro/source_terms.hpp:18-77).
// Given the adjoint lambda' of U' on the interior, the reverse step runs five passes; each
// writes every output element exactly once (no atomics, results are deterministic):
//   prim   (cells within 2 of the interior)   p = prim(Ug)
//   faces  (ring-1 cells)       Uf, stored in the face-slot buffer
//   flux   (faces touching the interior, every dim in one pass)
//                               lambda_F = dt/dV (lambda'(right) dS - lambda'(left) dS+), Riemann VJP,
//                               the adjoints of the two face states overwrite their slots
//   hancock (ring-1 cells)      in spherical geometry, first the Godunov sources' adjoint (interior
//                               cells: the face states Uf recomputed, their adjoint from lambda'(c));
//                               then the adjoint of the Hancock predictor (face-state adjoints ->
//                               face-prim adjoints, the predictor's sources included), then of the
//                               limited slopes: the cell's own prim adjoint and the shares owed to
//                               its +-1 neighbours, overwriting the slots
//   gather (interior cells)     cell-prim adjoint = own part + the neighbours' shares, prim -> cons
//                               VJP, then the transposed ghost fill (every ghost cell whose value
//                               was copied from this cell adds its adjoint back, sign-flipped where
//                               the fill negated the momentum), the identity and gravity terms.
// Slot (side s, dim d) of a cell holds NV values at planes v + NV * (s + 2 d).
// Every pass also returns the cell's or face's contribution to dL/dgamma.
#pragma once

#include "adjoint.hpp"
#include "grid.hpp"
#include "launch.hpp"

namespace hclpp::adj {

constexpr int kMaxGhosts = 4;
constexpr int kMaxPre = 1 + 2 * kMaxGhosts;

struct AdjArrays {
    const double* rho;  // step input, ghosts filled
    const double* rhou;
    const double* E;
    double* prim;       // NV planes
    double* face;       // face_planes(ND) planes: face slots, then the slope-adjoint shares
    const double* lrho_new;  // adjoint of the step output (interior cells read)
    const double* lrhou_new;
    const double* lE_new;
    double* lrho;       // adjoint of the step input (interior cells written)
    double* lrhou;
    double* lE;
    double dt;
};

__host__ __device__ __forceinline__ bool inside(const GridDev& G, int d, int a) { return a >= G.ngh[d] && a < G.ngh[d] + G.n[d]; }

template <int ND>
__host__ __device__ __forceinline__ bool ring1(const GridDev& G, const int* c)
{
#pragma unroll
    for (int d = 0; d < ND; ++d)
        if (c[d] < G.ngh[d] - 1 || c[d] > G.ngh[d] + G.n[d]) return false;
    return true;
}

template <int ND>
__host__ __device__ __forceinline__ S<ND> load_planes(const double* a, int64_t id, int64_t pl, int base)
{
    S<ND> q;
#pragma unroll
    for (int v = 0; v < ND + 2; ++v) q[v] = a[id + pl * (base + v)];
    return q;
}

template <int ND>
__host__ __device__ __forceinline__ void store_planes(double* a, int64_t id, int64_t pl, int base, const S<ND>& q)
{
#pragma unroll
    for (int v = 0; v < ND + 2; ++v) a[id + pl * (base + v)] = q[v];
}

template <int ND>
__host__ __device__ __forceinline__ S<ND> load_cons(const double* rho, const double* rhou, const double* E, int64_t id, int64_t pl)
{
    S<ND> q;
    q[0] = rho[id];
#pragma unroll
    for (int d = 0; d < ND; ++d) q[1 + d] = rhou[id + pl * d];
    q[ND + 1] = E[id];
    return q;
}

__host__ __device__ __forceinline__ constexpr int slot(int nv, int s, int d) { return nv * (s + 2 * d); }
// After the Hancock adjoint a ring-1 cell's planes hold: [0, nv) its own prim adjoint, then per dim d
// the share of cell c + e_d (share_up) and of cell c - e_d (share_down): nv (2 nd + 1) planes in all.
__host__ __device__ __forceinline__ constexpr int share_up(int nv, int d) { return nv * (1 + d); }
__host__ __device__ __forceinline__ constexpr int share_down(int nv, int nd, int d) { return nv * (1 + nd + d); }
constexpr int face_planes(int nd) { return (nd + 2) * (2 * nd + 1); }

__host__ __device__ __forceinline__ int64_t shifted(const GridDev& G, const int* c, int d, int by)
{
    int q[3] = {c[0], c[1], c[2]};
    q[d] += by;
    return flat(G, q[0], q[1], q[2]);
}

// Face primitives of cell c (limited_linear_reconstruction.hpp:46-64), same expressions as the forward.
template <int ND, int LIM>
__host__ __device__ __forceinline__ void face_prims(const GridDev& G, const double* prim, const int* c, const S<ND>& p, S<ND> (&pf)[2][ND])
{
    int64_t const pl = G.plane;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        S<ND> const pm = load_planes<ND>(prim, shifted(G, c, d, -1), pl, 0);
        S<ND> const pp = load_planes<ND>(prim, shifted(G, c, d, +1), pl, 0);
        double const dl = G.dx[d][c[d]];
        double const dl_m = G.dx[d][c[d] - 1];
        double const dl_p = G.dx[d][c[d] + 1];
#pragma unroll
        for (int v = 0; v < ND + 2; ++v) {
            double const slope = limited_slope_t<LIM>(fdiv(pp[v] - p[v], (dl + dl_p) / 2), fdiv(p[v] - pm[v], (dl_m + dl) / 2));
            pf[0][d][v] = p[v] - ((dl / 2) * slope);
            pf[1][d][v] = p[v] + ((dl / 2) * slope);
        }
    }
}

// Adjoint of the spherical source terms of dim `dim` (hydro/source_terms.hpp:18-77, added as in
// extrapolation_reconstruction.hpp:145-303 and godunov_scheme.hpp:165-256): lc is the adjoint of
// the conservative components the terms are added to (subtracted for the Coriolis terms), pL and
// pR the primitive face states they are formed from; accumulates into lpL and lpR.
template <int ND>
__host__ __device__ __forceinline__ void spherical_sources_vjp(const GridDev& G, int dim, int i, int j, double dtodv, double dS, double dSp,
                                                               const S<ND>& pL, const S<ND>& pR, const S<ND>& lc, S<ND>& lpL, S<ND>& lpR)
{
    double const w = dtodv * 0.5 * (dSp - dS);  // the pressure term is w (PL + PR)
    // centrifugal term of velocity component o, added to component m: w/2 (rL uL^2 + rR uR^2)
    auto centrifugal = [&](int m, int o, double scale) {
        double const k = 0.5 * w * scale * lc[m];
        lpL[0] += k * pL[o] * pL[o];
        lpR[0] += k * pR[o] * pR[o];
        lpL[o] += k * 2 * pL[0] * pL[o];
        lpR[o] += k * 2 * pR[0] * pR[o];
    };
    // Coriolis term subtracted from component m: q (rR uRo uRr dS+ + rL uLo uLr dS), with r the
    // component along `r` and o the one along `o`
    auto coriolis = [&](int m, int r, int o, double q) {
        double const k = -q * lc[m];
        lpR[0] += k * pR[o] * pR[r] * dSp;
        lpR[o] += k * pR[0] * pR[r] * dSp;
        lpR[r] += k * pR[0] * pR[o] * dSp;
        lpL[0] += k * pL[o] * pL[r] * dS;
        lpL[o] += k * pL[0] * pL[r] * dS;
        lpL[r] += k * pL[0] * pL[o] * dS;
    };
    if (dim == 0) {
        lpL[ND + 1] += w * lc[1];  // pressure: dtodv (PL + PR) / 2 (dS+ - dS) = w (PL + PR) on the radial momentum
        lpR[ND + 1] += w * lc[1];
        if constexpr (ND >= 2) {
            double const x0 = G.x[0][i], x0p = G.x[0][i + 1];
            double const q = fdiv(dtodv * (x0p - x0), x0p + x0);
            centrifugal(1, 2, 1.0);
            coriolis(2, 1, 2, q);
            if constexpr (ND == 3) {
                centrifugal(1, 3, 1.0);
                coriolis(3, 1, 3, q);
            }
        }
    } else if (dim == 1) {
        lpL[ND + 1] += w * lc[2];
        lpR[ND + 1] += w * lc[2];
        if constexpr (ND == 3) {
            double const sn = G.sin_node[j], snp = G.sin_node[j + 1];
            centrifugal(2, 3, G.cot_mid[j]);
            coriolis(3, 2, 3, fdiv(dtodv * (snp - sn), snp + sn));
        }
    }
}

// Hancock face states of cell (i, j, k) from its face primitives pf, dth = dt / 2:
// extrapolation_reconstruction.hpp:90-312 in the forward's accumulation order (per dim: the flux
// difference on every face, the spherical sources of that dim, gravity).
template <int ND, int GEOM>
__host__ __device__ __forceinline__ void hancock_states(const GridDev& G, const PerfectGas& eos, double dth, int i, int j, int k,
                                                        const S<ND> (&pf)[2][ND], S<ND> (&fs)[2][ND])
{
#pragma unroll
    for (int d = 0; d < ND; ++d)
#pragma unroll
        for (int s = 0; s < 2; ++s) fs[s][d] = prim_to_cons(pf[s][d], eos);
    double rho_old[2][ND], rhou_old[2][ND][ND];
#pragma unroll
    for (int d = 0; d < ND; ++d)
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            rho_old[s][d] = fs[s][d][0];
#pragma unroll
            for (int v = 0; v < ND; ++v) rhou_old[s][d][v] = fs[s][d][1 + v];
        }
    double const dtodv = fdiv(dth, cell_volume<ND, GEOM>(G, i, j, k));
    double const x0 = G.x[0][i], x0p = G.x[0][i + 1];
#pragma unroll
    for (int dim = 0; dim < ND; ++dim) {
        S<ND> const& pL = pf[0][dim];
        S<ND> const& pR = pf[1][dim];
        S<ND> const fL = flux_of_prim(pL, dim, eos);
        S<ND> const fR = flux_of_prim(pR, dim, eos);
        double const dS = face_area<ND, GEOM>(G, dim, i, j, k);
        double const dSp = face_area<ND, GEOM>(G, dim, i + (dim == 0), j + (dim == 1), k + (dim == 2));
#pragma unroll
        for (int pos = 0; pos < ND; ++pos)
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int v = 0; v < ND + 2; ++v) fs[s][pos][v] += dtodv * ((fL[v] * dS) - (fR[v] * dSp));
        if constexpr (GEOM == kSpherical) {
#pragma unroll
            for (int pos = 0; pos < ND; ++pos)
#pragma unroll
                for (int s = 0; s < 2; ++s) {
                    if (dim == 0) {
                        fs[s][pos][1] += src_pressure(dtodv, pL[ND + 1], pR[ND + 1], dS, dSp);
                        if constexpr (ND >= 2) {
                            fs[s][pos][1] += src_centrifugal(dtodv, pL[0], pR[0], pL[2], pR[2], dS, dSp);
                            fs[s][pos][2] -= src_coriolis_r(dtodv, x0, x0p, pL[0], pR[0], pL[1], pR[1], pL[2], pR[2], dS, dSp);
                        }
                        if constexpr (ND == 3) {
                            fs[s][pos][1] += src_centrifugal(dtodv, pL[0], pR[0], pL[3], pR[3], dS, dSp);
                            fs[s][pos][3] -= src_coriolis_r(dtodv, x0, x0p, pL[0], pR[0], pL[1], pR[1], pL[3], pR[3], dS, dSp);
                        }
                    }
                    if (dim == 1) {
                        fs[s][pos][2] += src_pressure(dtodv, pL[ND + 1], pR[ND + 1], dS, dSp);
                        if constexpr (ND == 3) {
                            fs[s][pos][2] += src_centrifugal_th(dtodv, G.cot_mid[j], pL[0], pR[0], pL[3], pR[3], dS, dSp);
                            fs[s][pos][3] -= src_coriolis_th(dtodv, G.sin_node[j], G.sin_node[j + 1], pL[0], pR[0], pL[2], pR[2], pL[3], pR[3],
                                                             dS, dSp);
                        }
                    }
                }
        }
        double const g = gravity(G, i, dim);
#pragma unroll
        for (int pos = 0; pos < ND; ++pos)
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                fs[s][pos][1 + dim] += dth * g * rho_old[s][pos];
                fs[s][pos][ND + 1] += dth * g * rhou_old[s][pos][dim];
            }
    }
}

// ---------------------------------------------------------------- pass 1: primitives of every cell
template <int ND>
__host__ __device__ __forceinline__ void cell_prim(const GridDev& G, const PerfectGas& eos, const AdjArrays& A, int i, int j, int k)
{
    int64_t const id = flat(G, i, j, k);
    store_planes<ND>(A.prim, id, G.plane, 0, cons_to_prim(load_cons<ND>(A.rho, A.rhou, A.E, id, G.plane), eos));
}

// ---------------------------------------------------------------- pass 2: Hancock face states (ring-1 cells)
// extrapolation_reconstruction.hpp:90-325 with dt/2, in the forward's accumulation order.
template <int ND, int LIM, int GEOM>
__host__ __device__ __forceinline__ void cell_faces(const GridDev& G, const PerfectGas& eos, const AdjArrays& A, int i, int j, int k)
{
    int const c[3] = {i, j, k};
    int64_t const id = flat(G, i, j, k);
    int64_t const pl = G.plane;
    S<ND> const p = load_planes<ND>(A.prim, id, pl, 0);
    S<ND> pf[2][ND];
    face_prims<ND, LIM>(G, A.prim, c, p, pf);
    S<ND> fs[2][ND];
    hancock_states<ND, GEOM>(G, eos, A.dt / 2, i, j, k, pf, fs);
#pragma unroll
    for (int d = 0; d < ND; ++d)
#pragma unroll
        for (int s = 0; s < 2; ++s) store_planes<ND>(A.face, id, pl, slot(ND + 2, s, d), fs[s][d]);
}

// ---------------------------------------------------------------- pass 3: Riemann flux adjoint (one face)
// Face between cell c - e_d and cell c, with c_d in [ng, ng + n] and c interior along the other dims.
// The face direction is a template parameter: the Riemann VJP indexes state components by it,
// and a run-time index would push the register arrays to scratch memory.
template <int ND, int RS, int d, int GEOM>
__host__ __device__ __forceinline__ double face_flux(const GridDev& G, const PerfectGas& eos, const AdjArrays& A, int i, int j, int k)
{
    constexpr int NV = ND + 2;
    int const c[3] = {i, j, k};
    int64_t const id = flat(G, i, j, k);
    int64_t const idm = shifted(G, c, d, -1);
    int64_t const pl = G.plane;
    S<ND> const cL = load_planes<ND>(A.face, idm, pl, slot(NV, 1, d));
    S<ND> const cR = load_planes<ND>(A.face, id, pl, slot(NV, 0, d));
    S<ND> lF = zero<ND>();
    // godunov_scheme.hpp:110-292: U'(c) += dt/dV F dS;  U'(c - e_d) -= dt/dV F dS+
    if (inside(G, d, c[d])) {
        double const w = fdiv(A.dt, cell_volume<ND, GEOM>(G, i, j, k)) * face_area<ND, GEOM>(G, d, i, j, k);
        S<ND> const l = load_cons<ND>(A.lrho_new, A.lrhou_new, A.lE_new, id, pl);
#pragma unroll
        for (int v = 0; v < NV; ++v) lF[v] += w * l[v];
    }
    if (inside(G, d, c[d] - 1)) {
        int const im = i - (d == 0), jm = j - (d == 1), km = k - (d == 2);
        double const w = fdiv(A.dt, cell_volume<ND, GEOM>(G, im, jm, km)) * face_area<ND, GEOM>(G, d, i, j, k);
        S<ND> const l = load_cons<ND>(A.lrho_new, A.lrhou_new, A.lE_new, idm, pl);
#pragma unroll
        for (int v = 0; v < NV; ++v) lF[v] -= w * l[v];
    }
    S<ND> lcL = zero<ND>(), lcR = zero<ND>();
    double lg = 0;
    riemann_vjp<RS>(cL, cR, d, eos, lF, lcL, lcR, lg);
    store_planes<ND>(A.face, idm, pl, slot(NV, 1, d), lcL);
    store_planes<ND>(A.face, id, pl, slot(NV, 0, d), lcR);
    return lg;
}

// Slot (s, d) of ring-1 cell c carries a flux adjoint iff its face touches an interior cell.
template <int ND>
__host__ __device__ __forceinline__ bool slot_live(const GridDev& G, const int* c, int s, int d)
{
#pragma unroll
    for (int e = 0; e < ND; ++e)
        if (e != d && !inside(G, e, c[e])) return false;
    return inside(G, d, c[d]) || inside(G, d, c[d] + (s ? 1 : -1));
}

// Slope adjoint of ring-1 cell c along d (limited_linear_reconstruction.hpp:46-64 reversed): with
// a = (p(c+1) - p(c)) / hp and b = (p(c) - p(c-1)) / hm, ra = lambda_slope * dslope/da / hp goes to
// p(c+1) (+) and p(c) (-), rb = lambda_slope * dslope/db / hm to p(c) (+) and p(c-1) (-); `own`
// receives the face-prim adjoints l0, l1 that flow to p(c) directly.
template <int ND, int LIM>
__host__ __device__ __forceinline__ void slope_adjoint(const GridDev& G, const double* prim, const int* c, int d, const S<ND>& l0,
                                                       const S<ND>& l1, S<ND>& own, S<ND>& ra, S<ND>& rb)
{
    constexpr int NV = ND + 2;
    int64_t const pl = G.plane;
    int64_t const id = flat(G, c[0], c[1], c[2]);
    S<ND> const p = load_planes<ND>(prim, id, pl, 0);
    S<ND> const pm = load_planes<ND>(prim, shifted(G, c, d, -1), pl, 0);
    S<ND> const pp = load_planes<ND>(prim, shifted(G, c, d, +1), pl, 0);
    double const dl = G.dx[d][c[d]];
    double const hp = (dl + G.dx[d][c[d] + 1]) / 2;
    double const hm = (G.dx[d][c[d] - 1] + dl) / 2;
    double const ihp = fdiv(1, hp), ihm = fdiv(1, hm);
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        double gR, gL;
        // the forward's slope arguments (they decide the limiter branch): its own divisions
        limiter_grad<LIM>(fdiv(pp[v] - p[v], hp), fdiv(p[v] - pm[v], hm), gR, gL);
        double const ls = (dl / 2) * (l1[v] - l0[v]);
        own[v] = l0[v] + l1[v];
        ra[v] = ls * gR * ihp;
        rb[v] = ls * gL * ihm;
    }
}

// ---------------------------------------------------------------- pass 4: Hancock adjoint (ring-1 cells)
template <int ND, int LIM, int GEOM>
__host__ __device__ __forceinline__ double cell_hancock(const GridDev& G, const PerfectGas& eos, const AdjArrays& A, int i, int j, int k)
{
    constexpr int NV = ND + 2;
    int const c[3] = {i, j, k};
    int64_t const id = flat(G, i, j, k);
    int64_t const pl = G.plane;
    S<ND> lam[2][ND];
    bool any = false;
#pragma unroll
    for (int d = 0; d < ND; ++d)
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            bool const live = slot_live<ND>(G, c, s, d);
            any = any || live;
            lam[s][d] = live ? load_planes<ND>(A.face, id, pl, slot(NV, s, d)) : zero<ND>();
        }
    S<ND> lpf[2][ND];
#pragma unroll
    for (int d = 0; d < ND; ++d)
#pragma unroll
        for (int s = 0; s < 2; ++s) lpf[s][d] = zero<ND>();
    double lg = 0;
    if (any) {
        S<ND> const p = load_planes<ND>(A.prim, id, pl, 0);
        S<ND> pf[2][ND];
        face_prims<ND, LIM>(G, A.prim, c, p, pf);
        double const dth = A.dt / 2;
        if constexpr (GEOM == kSpherical) {
            // godunov_scheme.hpp:165-256: the update of an interior cell adds the sources of dims 0
            // and 1 formed from the primitives of its own face states Uf: their adjoint joins the
            // face-state adjoints before the predictor is reversed
            bool interior = true;
#pragma unroll
            for (int d = 0; d < ND; ++d) interior = interior && inside(G, d, c[d]);
            if (interior) {
                S<ND> fs[2][ND];
                hancock_states<ND, GEOM>(G, eos, dth, i, j, k, pf, fs);
                S<ND> const ln = load_cons<ND>(A.lrho_new, A.lrhou_new, A.lE_new, id, pl);
                double const dtodv = fdiv(A.dt, cell_volume<ND, GEOM>(G, i, j, k));
#pragma unroll
                for (int dim = 0; dim < (ND < 2 ? ND : 2); ++dim) {
                    S<ND> const pL = cons_to_prim(fs[0][dim], eos);
                    S<ND> const pR = cons_to_prim(fs[1][dim], eos);
                    double const dS = face_area<ND, GEOM>(G, dim, i, j, k);
                    double const dSp = face_area<ND, GEOM>(G, dim, i + (dim == 0), j + (dim == 1), k + (dim == 2));
                    S<ND> lpL = zero<ND>(), lpR = zero<ND>();
                    spherical_sources_vjp<ND>(G, dim, i, j, dtodv, dS, dSp, pL, pR, ln, lpL, lpR);
                    prim_vjp(fs[0][dim], eos.gamma, lpL, lam[0][dim], lg);
                    prim_vjp(fs[1][dim], eos.gamma, lpR, lam[1][dim], lg);
                }
            }
        }
        S<ND> lD = zero<ND>();
#pragma unroll
        for (int pos = 0; pos < ND; ++pos)
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                S<ND> lc = lam[s][pos];
#pragma unroll
                for (int v = 0; v < NV; ++v) lD[v] += lam[s][pos][v];
                // gravity on the face's own conservative state
#pragma unroll
                for (int dim = 0; dim < ND; ++dim) {
                    double const g = gravity(G, i, dim);
                    lc[0] += dth * g * lam[s][pos][1 + dim];
                    lc[1 + dim] += dth * g * lam[s][pos][ND + 1];
                }
                cons_vjp(pf[s][pos], eos.gamma, lc, lpf[s][pos], lg);
            }
        double const dtodv = fdiv(dth, cell_volume<ND, GEOM>(G, i, j, k));
#pragma unroll
        for (int dim = 0; dim < ND; ++dim) {
            double const dS = face_area<ND, GEOM>(G, dim, i, j, k);
            double const dSp = face_area<ND, GEOM>(G, dim, i + (dim == 0), j + (dim == 1), k + (dim == 2));
            // the predictor's sources of this dim were added to every face state: their adjoint is lD
            if constexpr (GEOM == kSpherical) spherical_sources_vjp<ND>(G, dim, i, j, dtodv, dS, dSp, pf[0][dim], pf[1][dim], lD, lpf[0][dim],
                                                                      lpf[1][dim]);
            S<ND> a, b;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                a[v] = dtodv * dS * lD[v];
                b[v] = -(dtodv * dSp * lD[v]);
            }
            flux_vjp(pf[0][dim], dim, eos.gamma, a, lpf[0][dim], lg);
            flux_vjp(pf[1][dim], dim, eos.gamma, b, lpf[1][dim], lg);
        }
    }
    // slope adjoint: the face-prim adjoints become the cell's own prim adjoint and the shares of
    // its +-1 neighbours along each dim, stored in place of the face slots
    S<ND> self = zero<ND>();
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        S<ND> own = zero<ND>(), ra = zero<ND>(), rb = zero<ND>();
        if (any) slope_adjoint<ND, LIM>(G, A.prim, c, d, lpf[0][d], lpf[1][d], own, ra, rb);
#pragma unroll
        for (int v = 0; v < NV; ++v) self[v] += (own[v] - ra[v]) + rb[v];
        store_planes<ND>(A.face, id, pl, share_up(NV, d), ra);
        store_planes<ND>(A.face, id, pl, share_down(NV, ND, d), rb);
    }
    store_planes<ND>(A.face, id, pl, 0, self);
    return lg;
}

// Adjoint of the primitive state of cell c (any cell of the array): its own part plus the
// shares of its +-1 neighbours along each dim (zero outside the ring-1 cells).
// INTERIOR: c is an interior cell, so c and its +-1 neighbours are ring-1 cells and every test
// below holds (the same sums in the same order, without the tests).
template <int ND, bool INTERIOR = false>
__host__ __device__ __forceinline__ S<ND> cell_prim_adjoint(const GridDev& G, const AdjArrays& A, const int* c)
{
    constexpr int NV = ND + 2;
    int64_t const pl = G.plane;
    S<ND> lp = (INTERIOR || ring1<ND>(G, c)) ? load_planes<ND>(A.face, flat(G, c[0], c[1], c[2]), pl, 0) : zero<ND>();
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        int q[3] = {c[0], c[1], c[2]};
        q[d] = c[d] - 1;
        if (INTERIOR || ring1<ND>(G, q)) {
            S<ND> const ra = load_planes<ND>(A.face, flat(G, q[0], q[1], q[2]), pl, share_up(NV, d));
#pragma unroll
            for (int v = 0; v < NV; ++v) lp[v] += ra[v];
        }
        q[d] = c[d] + 1;
        if (INTERIOR || ring1<ND>(G, q)) {
            S<ND> const rb = load_planes<ND>(A.face, flat(G, q[0], q[1], q[2]), pl, share_down(NV, ND, d));
#pragma unroll
            for (int v = 0; v < NV; ++v) lp[v] -= rb[v];
        }
    }
    return lp;
}

// Cells along dim d whose ghost-filled value is copied from interior index a (a itself first).
__host__ __device__ __forceinline__ int preimages(const GridDev& G, const BcSpec& bc, int d, int a, int* pre, bool* flip)
{
    int m = 0;
    pre[m] = a;
    flip[m] = false;
    ++m;
    int const ng = G.ngh[d], n = G.n[d], ext = G.nwg[d];
    for (int g = 0; g < ng; ++g) {
        bool f = false;
        if (face_map(bc.kind[2 * d], g, ng, n, ext, f) == a) {
            pre[m] = g;
            flip[m] = f;
            ++m;
        }
    }
    for (int g = ng + n; g < ext; ++g) {
        bool f = false;
        if (face_map(bc.kind[2 * d + 1], g, ng, n, ext, f) == a) {
            pre[m] = g;
            flip[m] = f;
            ++m;
        }
    }
    return m;
}

// Interior cells from which no ghost cell is filled, along dim d: [2 ng, n) in ghost-inclusive
// indices (a periodic wrap fills the ghosts from [ng, 2 ng) and [n, n + ng), a Reflexive mirror
// from the same cells, a NullGradient clamp from ng and ng + n - 1).
__host__ __device__ __forceinline__ bool no_ghost_copies(const GridDev& G, int d, int lo, int hi)
{
    return lo >= 2 * G.ngh[d] && hi <= G.n[d];
}

// ---------------------------------------------------------------- pass 5: gather (interior cells)
// DEEP: the cell is interior and no ghost is filled from it along any dim (no_ghost_copies), so
// its only preimage is itself: the same sums in the same order as the general case, without the
// preimage search.
template <int ND, int LIM, bool DEEP = false>
__host__ __device__ __forceinline__ double cell_gather(const GridDev& G, const PerfectGas& eos, const BcSpec& bc, const AdjArrays& A, int i,
                                                       int j, int k)
{
    constexpr int NV = ND + 2;
    int64_t const pl = G.plane;
    int const c[3] = {i, j, k};
    int64_t const id = flat(G, i, j, k);
    S<ND> const ln = load_cons<ND>(A.lrho_new, A.lrhou_new, A.lE_new, id, pl);
    S<ND> out = ln;
    // godunov_scheme.hpp gravity: U'[1+d] += dt g_d rho, U'[E] += dt g_d m_d
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        double const g = gravity(G, i, d);
        out[0] += A.dt * g * ln[1 + d];
        out[1 + d] += A.dt * g * ln[ND + 1];
    }
    if constexpr (DEEP) {
        double lg = 0;
        S<ND> const lp = cell_prim_adjoint<ND, true>(G, A, c);
        S<ND> lu = zero<ND>();
        prim_vjp(load_cons<ND>(A.rho, A.rhou, A.E, id, pl), eos.gamma, lp, lu, lg);
        out[0] += lu[0];
#pragma unroll
        for (int d = 0; d < ND; ++d) out[1 + d] += lu[1 + d];
        out[ND + 1] += lu[ND + 1];
        A.lrho[id] = out[0];
#pragma unroll
        for (int d = 0; d < ND; ++d) A.lrhou[id + pl * d] = out[1 + d];
        A.lE[id] = out[ND + 1];
        return lg;
    }
    int pre[3][kMaxPre];
    bool flip[3][kMaxPre];
    int npre[3] = {1, 1, 1};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        pre[d][0] = c[d];
        flip[d][0] = false;
    }
#pragma unroll
    for (int d = 0; d < ND; ++d) npre[d] = preimages(G, bc, d, c[d], pre[d], flip[d]);
    double lg = 0;
    for (int a = 0; a < npre[0]; ++a)
        for (int b = 0; b < npre[1]; ++b)
            for (int e = 0; e < npre[2]; ++e) {
                int const q[3] = {pre[0][a], pre[1][b], pre[2][e]};
                S<ND> const lp = cell_prim_adjoint<ND>(G, A, q);
                int64_t const qid = flat(G, q[0], q[1], q[2]);
                S<ND> lu = zero<ND>();
                prim_vjp(load_cons<ND>(A.rho, A.rhou, A.E, qid, pl), eos.gamma, lp, lu, lg);
                bool const fl[3] = {flip[0][a], flip[1][b], flip[2][e]};
                out[0] += lu[0];
#pragma unroll
                for (int d = 0; d < ND; ++d) out[1 + d] += fl[d] ? -lu[1 + d] : lu[1 + d];
                out[ND + 1] += lu[ND + 1];
            }
    A.lrho[id] = out[0];
#pragma unroll
    for (int d = 0; d < ND; ++d) A.lrhou[id + pl * d] = out[1 + d];
    A.lE[id] = out[ND + 1];
    return lg;
}

}  // namespace hclpp::adj

namespace hclpp {

// adjoint.hip: number of workgroup partials one reverse step writes at most, and the launcher
// (passes in the order above, then dgamma[0] += the fixed-order sum of the partials).  With
// prims_ready the primitive pass is skipped (a previous step's gather formed them); with
// next_rho the gather pass forms the primitives of (next_rho, next_rhou, next_E) for the next
// reverse step.  pending: the number of partials the previous reverse step left unfolded in
// `partials` (this step's faces pass adds them to dgamma first); fold_now: add this step's
// partials before returning (returns 0), else return their number for the next step's `pending`.
// launch_adjoint_fold adds n partials to dgamma on its own.
int64_t adjoint_partials(const GridDev& G, int nd);
int64_t launch_step_adjoint(int nd, int geom, int rs, int lim, const GridDev& G, const PerfectGas& eos, const BcSpec& bc, const adj::AdjArrays& A,
                            bool prims_ready, const double* next_rho, const double* next_rhou, const double* next_E, int64_t pending,
                            bool fold_now, double* partials, double* dgamma, hipStream_t s);
void launch_adjoint_fold(const double* partials, int64_t n, double* dgamma, hipStream_t s);

}  // namespace hclpp
