// step_fused.hip -- one MUSCL-Hancock + Riemann + Godunov step in a single pass over HBM.
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
//    unfused sequence.
//  * Face states never touch HBM: per cell-update the kernel reads the
//    conservative state once (plus halo re-reads served mostly by L2) and
//    writes it once.  The epilogue reduces (in FP64) the CFL rate of the new
//    state, the minimum internal energy of the consumed state and the
//    conserved totals; per-workgroup partials are folded in a fixed order.
//  * The kernels are generic over the precision policy (physics.hpp):
//    PerfectGas streams FP64; PerfectGasMixed streams FP32 (half the bytes)
//    with FP64 pressure recovery and FP64 reductions.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "fused.hpp"
#include "launch.hpp"
#include "physics.hpp"
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
// FP64 policy this is the reference's arithmetic (conv_cons_to_prim, then the rate).
template <int ND, int GEOM, class Eos>
__device__ __forceinline__ double state_rate(const GridDev& G, const Eos& eos, int i, int j, int k, const State<ND, typename Eos::real>& c)
{
    if constexpr (std::is_same_v<Eos, PerfectGas>) {
        State<ND> const p = cons_to_prim(c, eos);
        return cfl_rate<ND, GEOM>(G, i, j, k, &p.q[1], eos.sound(p[0], p[ND + 1]));
    } else {
        double const rho = c[0];
        double u[ND];
#pragma unroll
        for (int d = 0; d < ND; ++d) u[d] = (double)c[1 + d] / rho;
        double const P = eos.gamma_m1 * eos.internal_energy_fp64(c);
        return cfl_rate<ND, GEOM>(G, i, j, k, u, sqrt(eos.gamma_d * P / rho));
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
    T const slope = limited_slope_t<LIM, Eos, T>(Eos::div(vp - v, (dl + dlp) * T(0.5)), Eos::div(v - vm, (dlm + dl) * T(0.5)));
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

// MUSCL-Hancock predictor of one cell (hydro_reconstruction.hpp:99-128 and
// extrapolation_reconstruction.hpp:90-312 with nfx = 0): nb[dim][side] are
// the primitive neighbours (side 0 = minus, 1 = plus) along each dim.
template <class Eos, int ND, int GEOM, int LIM>
__device__ __forceinline__ void predict_faces(const GridDev& G, const Eos& eos, typename Eos::real dth, double dth_d, int i, int j, int k,
                                              const State<ND, typename Eos::real>& p, const State<ND, typename Eos::real> (&nb)[ND][2],
                                              State<ND, typename Eos::real> (&fs)[2][ND])
{
    using T = typename Eos::real;
    constexpr int NV = ND + 2;
    State<ND, T> fp[2][ND];
    int const idx[3] = {i, j, k};
#pragma unroll
    for (int dim = 0; dim < ND; ++dim) {
        T const dl = T(G.dx[dim][idx[dim]]);
        T const dlm = T(G.dx[dim][idx[dim] - 1]);
        T const dlp = T(G.dx[dim][idx[dim] + 1]);
#pragma unroll
        for (int v = 0; v < NV; ++v)
            reconstruct<LIM, Eos>(nb[dim][0][v], p[v], nb[dim][1][v], dlm, dl, dlp, fp[0][dim][v], fp[1][dim][v]);
    }
#pragma unroll
    for (int dim = 0; dim < ND; ++dim)
#pragma unroll
        for (int s = 0; s < 2; ++s) fs[s][dim] = prim_to_cons(fp[s][dim], eos);
    T rho_old[2][ND], mom_old[2][ND][ND];
#pragma unroll
    for (int dim = 0; dim < ND; ++dim)
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            rho_old[s][dim] = fs[s][dim][0];
#pragma unroll
            for (int v = 0; v < ND; ++v) mom_old[s][dim][v] = fs[s][dim][1 + v];
        }
    T const dtodv = T(dth_d / cell_volume<ND, GEOM>(G, i, j, k));
    T const x0 = T(G.x[0][i]), x0p = T(G.x[0][i + 1]);
#pragma unroll
    for (int dim = 0; dim < ND; ++dim) {
        State<ND, T> const& pL = fp[0][dim];
        State<ND, T> const& pR = fp[1][dim];
        State<ND, T> const fL = flux_of_prim(pL, dim, eos);
        State<ND, T> const fR = flux_of_prim(pR, dim, eos);
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

// Minimum waves per SIMD the register allocator must allow.  Measured on MI355X: the 3D
// FP32-storage kernel gains 23 % at 3 waves (168 VGPRs, 20 B of spills per lane); the FP64
// kernels lose when forced above their natural 1 (3D) or 2 (2D) waves, because they spill.
template <class Eos, int ND>
constexpr int march_min_waves() { return (!std::is_same_v<Eos, PerfectGas> && ND == 3) ? 3 : 1; }

template <class Eos, int ND, int GEOM, int RS, int LIM, int HX, int HY, int NT>
__global__ __launch_bounds__(NT, (march_min_waves<Eos, ND>())) void k_step_march(GridDev G, Eos eos, StepArraysT<typename Eos::real> A, FusedLayout L)
{
    using T = typename Eos::real;
    constexpr int NV = ND + 2;
    constexpr int NP = ND - 1;               // in-plane dims
    constexpr int MD = ND - 1;               // march dim
    constexpr int PX = (ND >= 2) ? HX + 2 : 1;
    constexpr int PY = (ND == 3) ? HY + 2 : 1;
    constexpr int TX = (ND >= 2) ? HX - 2 : 1;
    constexpr int TY = (ND == 3) ? HY - 2 : 1;
    constexpr int NXCH = NP > 0 ? NP : 1;
    __shared__ T s_prim[NV][PY][PX];
    __shared__ T s_xch[NXCH][NV][HY][HX];
    __shared__ double s_red[NT / 64];

    DevClock const* clk = A.clock;
    if (clk->pub.done || clk->pub.status) return;
    double const dt_d = clk->pub.dt;
    double const dth_d = dt_d / 2;
    T const dt = T(dt_d);
    T const dth = T(dth_d);

    int const tid = threadIdx.x;
    // XCD-aware block order: blocks that share an XCD (b % 8) get consecutive tiles.
    int const nb = gridDim.x;
    int const b = blockIdx.x;
    int const q = nb / 8, rem = nb % 8, x8 = b % 8;
    int const lb = x8 * q + (x8 < rem ? x8 : rem) + b / 8;

    int hx = 0, hy = 0, seg;
    int ox = 0, oy = 0;
    if constexpr (ND == 1) {
        seg = lb * NT + tid;
    } else {
        hx = tid % HX;
        hy = tid / HX;
        int const tx = lb % L.ntile[0];
        int const rest = lb / L.ntile[0];
        int const ty = rest % L.ntile[1];
        seg = rest / L.ntile[1];
        ox = G.ngh[0] + tx * TX;
        oy = (ND == 3) ? G.ngh[1] + ty * TY : 0;
    }
    bool const in_block = (ND == 1) || tid < HX * HY;
    int const a = ox - 1 + hx;                  // column coords (ghost-inclusive)
    int const bb = (ND == 3) ? oy - 1 + hy : 0;
    int const ngm = G.ngh[MD], nm = G.n[MD];
    int const zs = ngm + seg * L.seglen;
    int const ze = min(zs + L.seglen, ngm + nm);

    bool col_in_array = in_block;               // column exists in memory
    bool hancock_ok = in_block;                 // column needs a predictor
    bool out_col = in_block;                    // column holds output cells
    if constexpr (ND >= 2) {
        col_in_array = col_in_array && a < G.nwg[0];
        hancock_ok = hancock_ok && a <= G.ngh[0] + G.n[0];
        out_col = out_col && hx >= 1 && hx <= TX && a < G.ngh[0] + G.n[0];
    }
    if constexpr (ND == 3) {
        col_in_array = col_in_array && bb < G.nwg[1];
        hancock_ok = hancock_ok && bb <= G.ngh[1] + G.n[1];
        out_col = out_col && hy >= 1 && hy <= TY && bb < G.ngh[1] + G.n[1];
    }
    bool const seg_ok = zs < ngm + nm;
    if constexpr (ND == 1) {
        col_in_array = seg_ok;
        hancock_ok = seg_ok;
        out_col = seg_ok;
    }

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

    // pipeline registers
    State<ND, T> pm, p0, pp, c0, cp;
    State<ND, T> fsR_prev, Fm_prev, part_prev;
    T old_rho_prev = 0, old_mom_prev = 0, src_prev = 0;
    int const zfirst = seg_ok ? zs - 1 : 0;
    int const zlast = seg_ok ? ze : -1;
    if (col_in_array) {
        pm = prim_at(a, bb, zfirst - 1);
        c0 = load_cons<ND>(A, G, cell_at<ND>(G, a, bb, zfirst));
        p0 = cons_to_prim(c0, eos);
    } else {
        pm = dummy_state<ND, T>();
        p0 = dummy_state<ND, T>();
        c0 = dummy_state<ND, T>();
    }

    for (int h = zfirst; h <= zlast; ++h) {
        // ---- plane h+1 of this column
        if (col_in_array) {
            cp = load_cons<ND>(A, G, cell_at<ND>(G, a, bb, h + 1));
            pp = cons_to_prim(cp, eos);
        } else {
            cp = dummy_state<ND, T>();
            pp = dummy_state<ND, T>();
        }
        // ---- plane h primitives in LDS (own column + ring)
        if constexpr (ND >= 2) {
            __syncthreads();
            if (in_block) {
#pragma unroll
                for (int v = 0; v < NV; ++v) s_prim[v][(ND == 3) ? hy + 1 : 0][hx + 1] = p0[v];
            }
            for (int r = tid; r < NRING; r += NT) {
                int px, py;
                if constexpr (ND == 2) {
                    px = r == 0 ? 0 : PX - 1;
                    py = 0;
                } else {
                    if (r < PX - 2) { px = 1 + r; py = 0; }
                    else if (r < 2 * (PX - 2)) { px = 1 + r - (PX - 2); py = PY - 1; }
                    else if (r < 2 * (PX - 2) + (PY - 2)) { px = 0; py = 1 + r - 2 * (PX - 2); }
                    else { px = PX - 1; py = 1 + r - 2 * (PX - 2) - (PY - 2); }
                }
                State<ND, T> const pr = prim_at(ox - 2 + px, (ND == 3) ? oy - 2 + py : 0, h);
#pragma unroll
                for (int v = 0; v < NV; ++v) s_prim[v][py][px] = pr[v];
            }
            __syncthreads();
        }
        int ci, cj, ck;
        ijk_of<ND>(a, bb, h, ci, cj, ck);
        bool const out_plane = h >= zs && h < ze;

        // ---- predictor of plane h
        State<ND, T> fs[2][ND];
        if (hancock_ok) {
            State<ND, T> nb3[ND][2];
            if constexpr (ND >= 2) {
                int const py0 = (ND == 3) ? hy + 1 : 0;
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    nb3[0][0][v] = s_prim[v][py0][hx];
                    nb3[0][1][v] = s_prim[v][py0][hx + 2];
                }
                if constexpr (ND == 3) {
#pragma unroll
                    for (int v = 0; v < NV; ++v) {
                        nb3[1][0][v] = s_prim[v][hy][hx + 1];
                        nb3[1][1][v] = s_prim[v][hy + 2][hx + 1];
                    }
                }
            }
            nb3[MD][0] = pm;
            nb3[MD][1] = pp;
            predict_faces<Eos, ND, GEOM, LIM>(G, eos, dth, dth_d, ci, cj, ck, p0, nb3, fs);
        } else {
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int d = 0; d < ND; ++d) fs[s][d] = c0;
        }

        // ---- in-plane fluxes and the in-plane part of the update of plane h
        State<ND, T> part = c0;
        T old_rho = c0[0], old_mom = c0[1 + MD], src_m = 0;
        T dtodv = 0;
        if (out_plane) {
            if (out_col) {
                dtodv = T(dt_d / cell_volume<ND, GEOM>(G, ci, cj, ck));
                // min internal energy of the consumed state (internal_energy.cpp:25-40)
                r_emin = reduce_combine<kMin>(r_emin, state_eint<ND>(eos, c0));
            }
            if constexpr (ND >= 2) {
                State<ND, T> FLp[NXCH], FRp[NXCH];
#pragma unroll
                for (int p = 0; p < NP; ++p) {
                    __syncthreads();
                    if (in_block) {
#pragma unroll
                        for (int v = 0; v < NV; ++v) s_xch[p][v][hy][hx] = fs[1][p][v];
                    }
                    __syncthreads();
                    bool need;
                    if (p == 0) need = hancock_ok && hx >= 1 && (ND == 2 || (hy >= 1 && hy <= TY));
                    else need = hancock_ok && hy >= 1 && hx >= 1 && hx <= TX;
                    State<ND, T> F;
                    if (need) {
                        State<ND, T> left;
#pragma unroll
                        for (int v = 0; v < NV; ++v) left[v] = (p == 0) ? s_xch[p][v][hy][hx - 1] : s_xch[p][v][hy - 1][hx];
                        F = riemann<RS>(left, fs[0][p], p, eos);
                    } else {
                        F = c0;
                    }
                    FLp[p] = F;
                    __syncthreads();
                    if (in_block) {
#pragma unroll
                        for (int v = 0; v < NV; ++v) s_xch[p][v][hy][hx] = F[v];
                    }
                    __syncthreads();
                    if (out_col) {
#pragma unroll
                        for (int v = 0; v < NV; ++v) FRp[p][v] = (p == 0) ? s_xch[p][v][hy][hx + 1] : s_xch[p][v][hy + 1][hx];
                    }
                }
                if (out_col) {
#pragma unroll
                    for (int p = 0; p < NP; ++p) {
                        T const dS = T(face_area<ND, GEOM>(G, p, ci, cj, ck));
                        T const dSp = T(face_area<ND, GEOM>(G, p, ci + (p == 0), cj + (p == 1), ck));
#pragma unroll
                        for (int v = 0; v < NV; ++v) part[v] += dtodv * ((FLp[p][v] * dS) - (FRp[p][v] * dSp));
                        godunov_sources<Eos, ND, GEOM>(G, eos, p, ci, cj, dtodv, dS, dSp, fs[0][p], fs[1][p], part);
                        T const g = T(gravity(G, ci, p));
                        part[1 + p] += dt * g * c0[0];
                        part[ND + 1] += dt * g * c0[1 + p];
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
            State<ND, T> const Fm = riemann<RS>(fsR_prev, fs[0][MD], MD, eos);
            int const hm = h - 1;
            if (hm >= zs) {
                int pi, pj, pk;
                ijk_of<ND>(a, bb, hm, pi, pj, pk);
                double const vol = cell_volume<ND, GEOM>(G, pi, pj, pk);
                T const dtodv_m = T(dt_d / vol);
                T const dS = T(face_area<ND, GEOM>(G, MD, pi, pj, pk));
                T const dSp = T(face_area<ND, GEOM>(G, MD, ci, cj, ck));
                State<ND, T> un = part_prev;
#pragma unroll
                for (int v = 0; v < NV; ++v) un[v] += dtodv_m * ((Fm_prev[v] * dS) - (Fm[v] * dSp));
                if constexpr (GEOM == kSpherical && MD <= 1) un[1 + MD] += src_prev;
                T const g = T(gravity(G, pi, MD));
                un[1 + MD] += dt * g * old_rho_prev;
                un[ND + 1] += dt * g * old_mom_prev;
                store_cons<ND>(A, G, cell_at<ND>(G, a, bb, hm), un);
                // epilogue: CFL rate of the new state and the conserved totals (FP64)
                r_inv = reduce_combine<kMax>(r_inv, state_rate<ND, GEOM>(G, eos, pi, pj, pk, un));
#pragma unroll
                for (int v = 0; v < NV; ++v) r_sum[v] += (double)un[v] * vol;
            }
            Fm_prev = Fm;
        }

        // ---- rotate the pipeline
        fsR_prev = fs[1][MD];
        part_prev = part;
        old_rho_prev = old_rho;
        old_mom_prev = old_mom;
        src_prev = src_m;
        pm = p0;
        p0 = pp;
        c0 = cp;
    }

    // ---- workgroup partials
    r_inv = block_reduce<kMax>(r_inv, s_red, tid, NT);
    r_emin = block_reduce<kMin>(r_emin, s_red, tid, NT);
#pragma unroll
    for (int v = 0; v < NV; ++v) r_sum[v] = block_reduce<kSum>(r_sum[v], s_red, tid, NT);
    if (tid == 0) {
        double* out = A.partials + (int64_t)blockIdx.x * kPartialStride;
        out[0] = r_inv;
        out[1] = r_emin;
#pragma unroll
        for (int v = 0; v < NV; ++v) out[2 + v] = r_sum[v];
    }
}

// ------------------------------------------------------------ ghost fill + ghost CFL rates
__device__ __forceinline__ int ghost_map(int kind, int idx, int ng, int n, int ext, bool& flip)
{
    if (idx < ng) {
        if (kind == kPeriodic) return idx + n;
        if (kind == kNullGradient) return ng;
        flip = !flip;
        return (2 * ng) - 1 - idx;
    }
    if (idx >= ng + n) {
        if (kind == kPeriodic) return idx - n;
        if (kind == kNullGradient) return ext - ng - 1;
        flip = !flip;
        return (2 * (ext - ng)) - 1 - idx;
    }
    return idx;
}

template <class Eos, int ND, int GEOM>
__global__ __launch_bounds__(256) void k_ghosts_rate(GridDev G, Eos eos, BcSpec bc, GhostShell B, StepArraysT<typename Eos::real> A)
{
    using T = typename Eos::real;
    __shared__ double s_red[4];
    DevClock const* clk = A.clock;
    if (clk->pub.done || clk->pub.status) return;
    int64_t const t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double rate = reduce_identity<kMax>();
    if (t < B.start[B.nbox]) {
        int b = 0;
        while (t >= B.start[b + 1]) ++b;
        int64_t loc = t - B.start[b];
        int const nx = B.hi[b][0] - B.lo[b][0], ny = B.hi[b][1] - B.lo[b][1];
        int const i = B.lo[b][0] + (int)(loc % nx);
        loc /= nx;
        int const j = B.lo[b][1] + (int)(loc % ny);
        int const k = B.lo[b][2] + (int)(loc / ny);
        int const idx[3] = {i, j, k};
        int src[3] = {i, j, k};
        bool flip[3] = {false, false, false};
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            int const side = idx[d] < G.ngh[d] ? 0 : 1;
            src[d] = ghost_map(bc.kind[2 * d + side], idx[d], G.ngh[d], G.n[d], G.nwg[d], flip[d]);
        }
        int64_t const dst = flat(G, i, j, k);
        int64_t const s = flat(G, src[0], src[1], src[2]);
        State<ND, T> c;
        c[0] = A.rho_new[s];
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            T const m = A.rhou_new[s + G.plane * d];
            c[1 + d] = flip[d] ? -m : m;
        }
        c[ND + 1] = A.E_new[s];
        A.rho_new[dst] = c[0];
#pragma unroll
        for (int d = 0; d < ND; ++d) A.rhou_new[dst + G.plane * d] = c[1 + d];
        A.E_new[dst] = c[ND + 1];
        rate = state_rate<ND, GEOM>(G, eos, i, j, k, c);
    }
    rate = block_reduce<kMax>(rate, s_red, threadIdx.x, 256);
    if (threadIdx.x == 0) A.ghost_partials[blockIdx.x] = rate;
}

// Max CFL rate over every cell (ghosts included) of an FP32-storage state.
template <int ND, int GEOM>
__global__ __launch_bounds__(256) void k_state_rate_f32(GridDev G, PerfectGasMixed eos, Range3 r, const float* __restrict__ rho,
                                                        const float* __restrict__ rhou, const float* __restrict__ E, double* __restrict__ partials)
{
    __shared__ double s_red[4];
    int i, j, k;
    double v = reduce_identity<kMax>();
    if (range_cell(r, i, j, k)) {
        int64_t const id = flat(G, i, j, k);
        State<ND, float> c;
        c[0] = rho[id];
#pragma unroll
        for (int d = 0; d < ND; ++d) c[1 + d] = rhou[id + G.plane * d];
        c[ND + 1] = E[id];
        v = state_rate<ND, GEOM>(G, eos, i, j, k, c);
    }
    int const tid = threadIdx.x + kBX * threadIdx.y;
    v = block_reduce<kMax>(v, s_red, tid, kBX * kBY);
    if (tid == 0) partials[blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)] = v;
}

template <class In, class Out>
__global__ __launch_bounds__(256) void k_cast(const In* __restrict__ in, Out* __restrict__ out, int64_t n)
{
    int64_t const t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t < n) out[t] = (Out)in[t];
}

// ------------------------------------------------------------ clock
// main.cpp:355-385 (dt, t_end clip, max_iter) and :397-425 (internal energy check)
__device__ __forceinline__ void schedule_next(DevClock* c, double inv)
{
    double dt = c->pub.cfl * (1 / inv);
    int pending = 0;
    if ((c->pub.t + dt) >= c->pub.t_end) {
        dt = c->pub.t_end - c->pub.t;
        pending = 1;
    }
    if ((c->pub.iter + 1) >= c->pub.max_iter) pending = 1;
    c->pub.dt = dt;
    c->pending_exit = pending;
}

__global__ __launch_bounds__(64) void k_clock_start(DevClock* c, const double* inv)
{
    if (threadIdx.x == 0) schedule_next(c, inv[0]);
}
__global__ __launch_bounds__(64) void k_clock_set_dt(DevClock* c, double dt)
{
    if (threadIdx.x != 0 || c->pub.done || c->pub.status) return;
    int pending = 0;
    if ((c->pub.t + dt) >= c->pub.t_end) pending = 1;
    if ((c->pub.iter + 1) >= c->pub.max_iter) pending = 1;
    c->pub.dt = dt;
    c->pending_exit = pending;
}
__global__ __launch_bounds__(64) void k_clock_set_log(DevClock* c, double* log, int cap)
{
    if (threadIdx.x != 0) return;
    c->dt_log = log;
    c->dt_log_cap = log ? cap : 0;
}

template <int NV>
__global__ __launch_bounds__(1024) void k_clock_advance(DevClock* c, const double* __restrict__ partials, int64_t nblocks,
                                                        const double* __restrict__ ghost_partials, int64_t nghost)
{
    __shared__ double lds[16];
    if (c->pub.done || c->pub.status) return;
    double inv = reduce_identity<kMax>(), emin = reduce_identity<kMin>();
    double s[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) s[v] = 0;
    for (int64_t t = threadIdx.x; t < nblocks; t += 1024) {
        double const* p = partials + t * kPartialStride;
        inv = reduce_combine<kMax>(inv, p[0]);
        emin = reduce_combine<kMin>(emin, p[1]);
#pragma unroll
        for (int v = 0; v < NV; ++v) s[v] += p[2 + v];
    }
    for (int64_t t = threadIdx.x; t < nghost; t += 1024) inv = reduce_combine<kMax>(inv, ghost_partials[t]);
    inv = block_reduce<kMax>(inv, lds, threadIdx.x, 1024);
    emin = block_reduce<kMin>(emin, lds, threadIdx.x, 1024);
#pragma unroll
    for (int v = 0; v < NV; ++v) s[v] = block_reduce<kSum>(s[v], lds, threadIdx.x, 1024);
    if (threadIdx.x != 0) return;
    c->pub.min_eint = emin;
    if (isnan(emin) || emin < 0) {
        c->pub.status = 3;  // HCLPP_ERR_STATE: the step is discarded
        return;
    }
    if (c->dt_log && c->pub.iter < c->dt_log_cap) c->dt_log[c->pub.iter] = c->pub.dt;
    c->pub.t += c->pub.dt;
    c->pub.iter += 1;
    c->pub.mass = s[0];
    for (int d = 0; d < 3; ++d) c->pub.momentum[d] = d < NV - 2 ? s[1 + d] : 0.;
    c->pub.energy = s[NV - 1];
    if (c->pending_exit) {
        c->pub.done = 1;
        return;
    }
    schedule_next(c, inv);
}

// ------------------------------------------------------------ launchers
int64_t fused_blocks(const GridDev& G, int nd, FusedLayout* layout)
{
    FusedLayout L{};
    int64_t nb = 0;
    auto shape = [&](int HX, int HY, int NT, int SEG) {
        L.seglen = SEG;
        int const nm = G.n[nd - 1];
        int const nseg = (nm + SEG - 1) / SEG;
        if (nd == 1) {
            L.ntile[0] = 1;
            L.ntile[1] = 1;
            nb = (nseg + NT - 1) / NT;
        } else {
            int const TX = HX - 2;
            int const TY = (nd == 3) ? HY - 2 : 1;
            L.ntile[0] = (G.n[0] + TX - 1) / TX;
            L.ntile[1] = (nd == 3) ? (G.n[1] + TY - 1) / TY : 1;
            nb = (int64_t)L.ntile[0] * L.ntile[1] * nseg;
        }
    };
    // Shorter segments for small grids: halve the march length until the launch holds at least
    // kFillWaves wavefronts (two per SIMD of the 1024 SIMDs), down to kMinSeg planes.  Every
    // cell is computed by the same expressions whatever the split, so results do not change.
    constexpr int64_t kFillWaves = 2048;
    constexpr int kMinSeg = 4;
    auto fit = [&](int HX, int HY, int NT, int SEG) {
        int seg = SEG;
        shape(HX, HY, NT, seg);
        while (nd > 1 && nb * (NT / 64) < kFillWaves && seg > kMinSeg) {
            seg /= 2;
            shape(HX, HY, NT, seg);
        }
    };
    if (nd == 1) shape(MarchShape<1>::HX, MarchShape<1>::HY, MarchShape<1>::NT, MarchShape<1>::SEG);
    else if (nd == 2) fit(MarchShape<2>::HX, MarchShape<2>::HY, MarchShape<2>::NT, MarchShape<2>::SEG);
    else fit(MarchShape<3>::HX, MarchShape<3>::HY, MarchShape<3>::NT, MarchShape<3>::SEG);
    if (layout) *layout = L;
    return nb;
}

template <class Eos, int ND, int GEOM, int RS, int LIM>
static void launch_march_t(const GridDev& G, const Eos& eos, const StepArraysT<typename Eos::real>& A, hipStream_t s)
{
    using S = MarchShape<ND>;
    FusedLayout L;
    int64_t const nb = fused_blocks(G, ND, &L);
    if (nb == 0) return;
    hipLaunchKernelGGL((k_step_march<Eos, ND, GEOM, RS, LIM, S::HX, S::HY, S::NT>), dim3((unsigned)nb), dim3(S::NT), 0, s, G, eos, A, L);
}

template <class Eos>
static void launch_fused_any(int nd, int geom, int rs, int lim, const GridDev& G, const Eos& eos, const StepArraysT<typename Eos::real>& A,
                             hipStream_t s)
{
#define HCLPP_F4(NDV, GV, RSV, LV) launch_march_t<Eos, NDV, GV, RSV, LV>(G, eos, A, s)
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

void launch_fused_step(int nd, int geom, int rs, int lim, const GridDev& G, const PerfectGas& eos, const StepArrays& A, hipStream_t s)
{
    launch_fused_any(nd, geom, rs, lim, G, eos, A, s);
}

void launch_fused_step(int nd, int geom, int rs, int lim, const GridDev& G, const PerfectGasMixed& eos, const StepArraysF& A, hipStream_t s)
{
    launch_fused_any(nd, geom, rs, lim, G, eos, A, s);
}

GhostShell make_ghost_shell(const GridDev& G, int nd)
{
    GhostShell B{};
    int64_t total = 0;
    for (int d = 0; d < nd; ++d)
        for (int side = 0; side < 2; ++side) {
            int const b = B.nbox++;
            for (int e = 0; e < 3; ++e) {
                int lo, hi;
                if (e < d) { lo = 0; hi = G.nwg[e]; }
                else if (e > d) { lo = G.ngh[e]; hi = G.ngh[e] + G.n[e]; }
                else if (side == 0) { lo = 0; hi = G.ngh[e]; }
                else { lo = G.ngh[e] + G.n[e]; hi = G.nwg[e]; }
                B.lo[b][e] = lo;
                B.hi[b][e] = hi;
            }
            B.start[b] = total;
            total += (int64_t)(B.hi[b][0] - B.lo[b][0]) * (B.hi[b][1] - B.lo[b][1]) * (B.hi[b][2] - B.lo[b][2]);
        }
    B.start[B.nbox] = total;
    return B;
}

int64_t ghost_blocks(const GhostShell& B) { return (B.start[B.nbox] + 255) / 256; }

template <class Eos>
static void launch_ghosts_any(int nd, int geom, const GridDev& G, const Eos& eos, const BcSpec& bc, const GhostShell& B,
                              const StepArraysT<typename Eos::real>& A, hipStream_t s)
{
    int64_t const nb = ghost_blocks(B);
    if (nb == 0) return;
#define HCLPP_GR(NDV)                                                                                                                    \
    if (geom == kSpherical) hipLaunchKernelGGL((k_ghosts_rate<Eos, NDV, kSpherical>), dim3((unsigned)nb), dim3(256), 0, s, G, eos, bc, B, A); \
    else hipLaunchKernelGGL((k_ghosts_rate<Eos, NDV, kCartesian>), dim3((unsigned)nb), dim3(256), 0, s, G, eos, bc, B, A);
    switch (nd) {
    case 1: HCLPP_GR(1); break;
    case 2: HCLPP_GR(2); break;
    default: HCLPP_GR(3); break;
    }
#undef HCLPP_GR
}

void launch_ghosts_rate(int nd, int geom, const GridDev& G, const PerfectGas& eos, const BcSpec& bc, const GhostShell& B, const StepArrays& A,
                        hipStream_t s)
{
    launch_ghosts_any(nd, geom, G, eos, bc, B, A, s);
}

void launch_ghosts_rate(int nd, int geom, const GridDev& G, const PerfectGasMixed& eos, const BcSpec& bc, const GhostShell& B,
                        const StepArraysF& A, hipStream_t s)
{
    launch_ghosts_any(nd, geom, G, eos, bc, B, A, s);
}

void launch_state_rate_f32(int nd, int geom, const GridDev& G, const PerfectGasMixed& eos, const float* rho, const float* rhou, const float* E,
                           double* partials, double* out, hipStream_t s)
{
    Range3 all;
    for (int d = 0; d < 3; ++d) {
        all.b[d] = 0;
        all.e[d] = G.nwg[d];
    }
    dim3 const grid = range_grid(all);
#define HCLPP_SR(NDV)                                                                                                                   \
    if (geom == kSpherical) hipLaunchKernelGGL((k_state_rate_f32<NDV, kSpherical>), grid, dim3(kBX, kBY), 0, s, G, eos, all, rho, rhou, E, partials); \
    else hipLaunchKernelGGL((k_state_rate_f32<NDV, kCartesian>), grid, dim3(kBX, kBY), 0, s, G, eos, all, rho, rhou, E, partials);
    switch (nd) {
    case 1: HCLPP_SR(1); break;
    case 2: HCLPP_SR(2); break;
    default: HCLPP_SR(3); break;
    }
#undef HCLPP_SR
    launch_fold_max(partials, (int64_t)grid.x * grid.y * grid.z, out, s);
}

void launch_cast(const double* in, float* out, int64_t n, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL((k_cast<double, float>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in, out, n);
}

void launch_cast(const float* in, double* out, int64_t n, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL((k_cast<float, double>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in, out, n);
}

void launch_clock_start(DevClock* c, const double* inv, hipStream_t s) { hipLaunchKernelGGL(k_clock_start, dim3(1), dim3(64), 0, s, c, inv); }
void launch_clock_set_dt(DevClock* c, double dt, hipStream_t s) { hipLaunchKernelGGL(k_clock_set_dt, dim3(1), dim3(64), 0, s, c, dt); }
void launch_clock_set_log(DevClock* c, double* log, int cap, hipStream_t s)
{
    hipLaunchKernelGGL(k_clock_set_log, dim3(1), dim3(64), 0, s, c, log, cap);
}

void launch_clock_advance(int nd, DevClock* c, const double* partials, int64_t nblocks, const double* ghost_partials, int64_t nghost,
                          hipStream_t s)
{
    switch (nd) {
    case 1: hipLaunchKernelGGL(k_clock_advance<3>, dim3(1), dim3(1024), 0, s, c, partials, nblocks, ghost_partials, nghost); break;
    case 2: hipLaunchKernelGGL(k_clock_advance<4>, dim3(1), dim3(1024), 0, s, c, partials, nblocks, ghost_partials, nghost); break;
    default: hipLaunchKernelGGL(k_clock_advance<5>, dim3(1), dim3(1024), 0, s, c, partials, nblocks, ghost_partials, nghost); break;
    }
}

}  // namespace hclpp
