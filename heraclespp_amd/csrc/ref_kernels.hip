// ref_kernels.hip -- HIP kernels behind the reference operator interfaces.
//
// One kernel per reference Kokkos kernel label, with the same cell ranges and
// per-cell arithmetic, so each C-ABI entry point is a drop-in for the
// reference operator it replaces and is bit-identical to it in FP64.  These
// kernels also serve as the parity-bisection path for the fused step kernel.
#include <hip/hip_runtime.h>

#include "launch.hpp"
#include "physics.hpp"
#include "pressure_fix.hpp"
#include "reduce.hpp"

namespace hclpp {

// ------------------------------------------------------------ conversions
// hydro/array_conversion.hpp:46-62 ("conv_prim_to_cons_array")
template <int ND>
__global__ __launch_bounds__(256) void k_prim_to_cons(GridDev G, Range3 r, PerfectGas eos, const double* __restrict__ rho,
                                                      const double* __restrict__ u, const double* __restrict__ P,
                                                      double* __restrict__ rhou, double* __restrict__ E)
{
    int i, j, k;
    if (!range_cell(r, i, j, k)) return;
    int64_t const id = flat(G, i, j, k);
    State<ND> p;
    p[0] = rho[id];
#pragma unroll
    for (int d = 0; d < ND; ++d) p[1 + d] = u[id + G.plane * d];
    p[ND + 1] = P[id];
    State<ND> const c = prim_to_cons(p, eos);
#pragma unroll
    for (int d = 0; d < ND; ++d) rhou[id + G.plane * d] = c[1 + d];
    E[id] = c[ND + 1];
}

// hydro/array_conversion.hpp:121-137 ("conv_cons_to_prim_array")
template <int ND>
__global__ __launch_bounds__(256) void k_cons_to_prim(GridDev G, Range3 r, PerfectGas eos, const double* __restrict__ rho,
                                                      const double* __restrict__ rhou, const double* __restrict__ E,
                                                      double* __restrict__ u, double* __restrict__ P)
{
    int i, j, k;
    if (!range_cell(r, i, j, k)) return;
    int64_t const id = flat(G, i, j, k);
    State<ND> c;
    c[0] = rho[id];
#pragma unroll
    for (int d = 0; d < ND; ++d) c[1 + d] = rhou[id + G.plane * d];
    c[ND + 1] = E[id];
    State<ND> const p = cons_to_prim(c, eos);
#pragma unroll
    for (int d = 0; d < ND; ++d) u[id + G.plane * d] = p[1 + d];
    P[id] = p[ND + 1];
}

// ------------------------------------------------------------ reconstruction
// hydro/limited_linear_reconstruction.hpp:46-64 ("face_reconstruction");
// var_rec layout [wg, side(2), idim(ND)]
template <int ND, int LIM>
__global__ __launch_bounds__(256) void k_face_reconstruction(GridDev G, Range3 r, const double* __restrict__ var, double* __restrict__ rec)
{
    int i, j, k;
    if (!range_cell(r, i, j, k)) return;
    int64_t const id = flat(G, i, j, k);
    double const v = var[id];
#pragma unroll
    for (int dim = 0; dim < ND; ++dim) {
        int const di = dim == 0, dj = dim == 1, dk = dim == 2;
        int const a = dim == 0 ? i : (dim == 1 ? j : k);
        double const dl = G.dx[dim][a];
        double const dl_m = G.dx[dim][a - 1];
        double const dl_p = G.dx[dim][a + 1];
        double const vp = var[flat(G, i + di, j + dj, k + dk)];
        double const vm = var[flat(G, i - di, j - dj, k - dk)];
        double const slope = limited_slope_t<LIM>(fdiv(vp - v, (dl + dl_p) / 2), fdiv(v - vm, (dl_m + dl) / 2));
        rec[id + G.plane * (0 + 2 * dim)] = v - ((dl / 2) * slope);
        rec[id + G.plane * (1 + 2 * dim)] = v + ((dl / 2) * slope);
    }
}

// hydro/hydro_reconstruction.hpp:111-126: face primitive -> conservative for every (dim, side)
template <int ND>
__global__ __launch_bounds__(256) void k_face_prim_to_cons(GridDev G, Range3 r, PerfectGas eos, const double* __restrict__ rho_rec,
                                                           const double* __restrict__ u_rec, const double* __restrict__ P_rec,
                                                           double* __restrict__ rhou_rec, double* __restrict__ E_rec)
{
    int i, j, k;
    if (!range_cell(r, i, j, k)) return;
    int64_t const id = flat(G, i, j, k);
    int64_t const pl = G.plane;
#pragma unroll
    for (int fd = 0; fd < 2 * ND; ++fd) {  // fd = side + 2 * dim
        int64_t const off = id + pl * fd;
        State<ND> p;
        p[0] = rho_rec[off];
#pragma unroll
        for (int v = 0; v < ND; ++v) p[1 + v] = u_rec[off + pl * 2 * ND * v];
        p[ND + 1] = P_rec[off];
        State<ND> const c = prim_to_cons(p, eos);
#pragma unroll
        for (int v = 0; v < ND; ++v) rhou_rec[off + pl * 2 * ND * v] = c[1 + v];
        E_rec[off] = c[ND + 1];
    }
}

// hydro/extrapolation_reconstruction.hpp:90-325 ("Hancock_extrapolation")
template <int ND, int GEOM>
__global__ __launch_bounds__(256) void k_hancock(GridDev G, Range3 r, PerfectGas eos, double dt, const double* __restrict__ u_rec,
                                                 const double* __restrict__ P_rec, double* __restrict__ rho_rec,
                                                 double* __restrict__ rhou_rec, double* __restrict__ E_rec, double* __restrict__ fx_rec,
                                                 const double* __restrict__ fx_old, int nfx)
{
    int i, j, k;
    if (!range_cell(r, i, j, k)) return;
    int64_t const id = flat(G, i, j, k);
    int64_t const pl = G.plane;
    // face states held in registers: fs[side][face dim] (cons)
    State<ND> fs[2][ND];
    double rho_old[2][ND], rhou_old[2][ND][ND];
#pragma unroll
    for (int dim = 0; dim < ND; ++dim)
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            int64_t const off = id + pl * (s + 2 * dim);
            fs[s][dim][0] = rho_rec[off];
#pragma unroll
            for (int v = 0; v < ND; ++v) fs[s][dim][1 + v] = rhou_rec[off + pl * 2 * ND * v];
            fs[s][dim][ND + 1] = E_rec[off];
            rho_old[s][dim] = fs[s][dim][0];
#pragma unroll
            for (int v = 0; v < ND; ++v) rhou_old[s][dim][v] = fs[s][dim][1 + v];
        }
    for (int f = 0; f < nfx; ++f)
#pragma unroll
        for (int dim = 0; dim < ND; ++dim)
#pragma unroll
            for (int s = 0; s < 2; ++s) fx_rec[id + pl * (s + 2 * (dim + ND * f))] *= rho_old[s][dim];

#pragma unroll
    for (int dim = 0; dim < ND; ++dim) {
        int const ip = i + (dim == 0), jp = j + (dim == 1), kp = k + (dim == 2);
        State<ND> pL, pR;
        pL[0] = rho_old[0][dim];
        pR[0] = rho_old[1][dim];
#pragma unroll
        for (int v = 0; v < ND; ++v) {
            pL[1 + v] = u_rec[id + pl * (0 + 2 * (dim + ND * v))];
            pR[1 + v] = u_rec[id + pl * (1 + 2 * (dim + ND * v))];
        }
        pL[ND + 1] = P_rec[id + pl * (0 + 2 * dim)];
        pR[ND + 1] = P_rec[id + pl * (1 + 2 * dim)];
        State<ND> const fL = flux_of_prim(pL, dim, eos);
        State<ND> const fR = flux_of_prim(pR, dim, eos);
        double const dtodv = fdiv(dt, cell_volume<ND, GEOM>(G, i, j, k));
        double const dS = face_area<ND, GEOM>(G, dim, i, j, k);
        double const dSp = face_area<ND, GEOM>(G, dim, ip, jp, kp);
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
                            fs[s][pos][2] -= src_coriolis_r(dtodv, G.x[0][i], G.x[0][i + 1], pL[0], pR[0], pL[1], pR[1], pL[2], pR[2], dS, dSp);
                        }
                        if constexpr (ND == 3) {
                            fs[s][pos][1] += src_centrifugal(dtodv, pL[0], pR[0], pL[3], pR[3], dS, dSp);
                            fs[s][pos][3] -= src_coriolis_r(dtodv, G.x[0][i], G.x[0][i + 1], pL[0], pR[0], pL[1], pR[1], pL[3], pR[3], dS, dSp);
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
                fs[s][pos][1 + dim] += dt * g * rho_old[s][pos];
                fs[s][pos][ND + 1] += dt * g * rhou_old[s][pos][dim];
            }
        for (int f = 0; f < nfx; ++f) {
            double const ffL = fx_old[id + pl * (0 + 2 * (dim + ND * f))] * fL[0];
            double const ffR = fx_old[id + pl * (1 + 2 * (dim + ND * f))] * fR[0];
#pragma unroll
            for (int pos = 0; pos < ND; ++pos)
#pragma unroll
                for (int s = 0; s < 2; ++s) fx_rec[id + pl * (s + 2 * (pos + ND * f))] += dtodv * ((ffL * dS) - (ffR * dSp));
        }
    }
#pragma unroll
    for (int dim = 0; dim < ND; ++dim)
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            int64_t const off = id + pl * (s + 2 * dim);
            rho_rec[off] = fs[s][dim][0];
#pragma unroll
            for (int v = 0; v < ND; ++v) rhou_rec[off + pl * 2 * ND * v] = fs[s][dim][1 + v];
            E_rec[off] = fs[s][dim][ND + 1];
            // passive_scalar_extrapolation (:327-337): mass fraction from the updated density
            for (int f = 0; f < nfx; ++f) fx_rec[id + pl * (s + 2 * (dim + ND * f))] /= fs[s][dim][0];
        }
}

// hydro/godunov_scheme.hpp:110-292 ("Riemann_based_Godunov_scheme") + :294-301
template <int ND, int GEOM, int RS>
__global__ __launch_bounds__(256) void k_godunov(GridDev G, Range3 r, PerfectGas eos, double dt, const double* __restrict__ rho,
                                                 const double* __restrict__ rhou, const double* __restrict__ E, const double* __restrict__ fx,
                                                 const double* __restrict__ rho_rec, const double* __restrict__ rhou_rec,
                                                 const double* __restrict__ E_rec, const double* __restrict__ fx_rec,
                                                 double* __restrict__ rho_new, double* __restrict__ rhou_new, double* __restrict__ E_new,
                                                 double* __restrict__ fx_new, int nfx)
{
    int i, j, k;
    if (!range_cell(r, i, j, k)) return;
    int64_t const id = flat(G, i, j, k);
    int64_t const pl = G.plane;
    auto face = [&](int64_t cell, int s, int dim) {
        State<ND> q;
        int64_t const off = cell + pl * (s + 2 * dim);
        q[0] = rho_rec[off];
#pragma unroll
        for (int v = 0; v < ND; ++v) q[1 + v] = rhou_rec[off + pl * 2 * ND * v];
        q[ND + 1] = E_rec[off];
        return q;
    };
    State<ND> un;
    un[0] = rho[id];
#pragma unroll
    for (int d = 0; d < ND; ++d) un[1 + d] = rhou[id + pl * d];
    un[ND + 1] = E[id];
    State<ND> const old = un;
    for (int f = 0; f < nfx; ++f) fx_new[id + pl * f] = fx[id + pl * f] * old[0];
#pragma unroll
    for (int dim = 0; dim < ND; ++dim) {
        int const di = dim == 0, dj = dim == 1, dk = dim == 2;
        int64_t const idm = flat(G, i - di, j - dj, k - dk);
        int64_t const idp = flat(G, i + di, j + dj, k + dk);
        State<ND> const vL = face(id, 0, dim);
        State<ND> const vR = face(id, 1, dim);
        State<ND> const FL = riemann<RS>(face(idm, 1, dim), vL, dim, eos);
        State<ND> const FR = riemann<RS>(vR, face(idp, 0, dim), dim, eos);
        double const dtodv = fdiv(dt, cell_volume<ND, GEOM>(G, i, j, k));
        double const dS = face_area<ND, GEOM>(G, dim, i, j, k);
        double const dSp = face_area<ND, GEOM>(G, dim, i + di, j + dj, k + dk);
#pragma unroll
        for (int v = 0; v < ND + 2; ++v) un[v] += dtodv * ((FL[v] * dS) - (FR[v] * dSp));
        if constexpr (GEOM == kSpherical) {
            if (dim <= 1) {
                State<ND> const pL = cons_to_prim(vL, eos);
                State<ND> const pR = cons_to_prim(vR, eos);
                if (dim == 0) {
                    un[1] += src_pressure(dtodv, pL[ND + 1], pR[ND + 1], dS, dSp);
                    if constexpr (ND >= 2) {
                        un[1] += src_centrifugal(dtodv, pL[0], pR[0], pL[2], pR[2], dS, dSp);
                        un[2] -= src_coriolis_r(dtodv, G.x[0][i], G.x[0][i + 1], pL[0], pR[0], pL[1], pR[1], pL[2], pR[2], dS, dSp);
                    }
                    if constexpr (ND == 3) {
                        un[1] += src_centrifugal(dtodv, pL[0], pR[0], pL[3], pR[3], dS, dSp);
                        un[3] -= src_coriolis_r(dtodv, G.x[0][i], G.x[0][i + 1], pL[0], pR[0], pL[1], pR[1], pL[3], pR[3], dS, dSp);
                    }
                } else {
                    un[2] += src_pressure(dtodv, pL[ND + 1], pR[ND + 1], dS, dSp);
                    if constexpr (ND == 3) {
                        un[2] += src_centrifugal_th(dtodv, G.cot_mid[j], pL[0], pR[0], pL[3], pR[3], dS, dSp);
                        un[3] -= src_coriolis_th(dtodv, G.sin_node[j], G.sin_node[j + 1], pL[0], pR[0], pL[2], pR[2], pL[3], pR[3], dS, dSp);
                    }
                }
            }
        }
        double const g = gravity(G, i, dim);
        un[1 + dim] += dt * g * old[0];
        un[ND + 1] += dt * g * old[1 + dim];
        for (int f = 0; f < nfx; ++f) {
            int64_t const cl = FL[0] < 0 ? id : idm;
            int const sl = FL[0] < 0 ? 0 : 1;
            int64_t const cr = FR[0] < 0 ? idp : id;
            int const sr = FR[0] < 0 ? 0 : 1;
            double const ffL = fx_rec[cl + pl * (sl + 2 * (dim + ND * f))] * FL[0];
            double const ffR = fx_rec[cr + pl * (sr + 2 * (dim + ND * f))] * FR[0];
            fx_new[id + pl * f] += dtodv * ((ffL * dS) - (ffR * dSp));
        }
    }
    rho_new[id] = un[0];
#pragma unroll
    for (int d = 0; d < ND; ++d) rhou_new[id + pl * d] = un[1 + d];
    E_new[id] = un[ND + 1];
    for (int f = 0; f < nfx; ++f) {
        double const y = fdiv(fx_new[id + pl * f], un[0]);
        fx_new[id + pl * f] = y < 0. ? 0. : (y > 1. ? 1. : y);  // Kokkos::clamp(y, 0, 1)
    }
}

// ------------------------------------------------------------ reductions
constexpr int kRedThreads = kBX * kBY;

__device__ __forceinline__ int block_linear() { return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z); }

// hydro/time_step.hpp:55-77 ("time_step"): per-block max of the CFL rate
template <int ND, int GEOM>
__global__ __launch_bounds__(256) void k_time_step(GridDev G, Range3 r, PerfectGas eos, const double* __restrict__ rho,
                                                   const double* __restrict__ u, const double* __restrict__ P, double* __restrict__ partials)
{
    __shared__ double lds[kRedThreads / 64];
    int i, j, k;
    double v = reduce_identity<kMax>();
    if (range_cell(r, i, j, k)) {
        int64_t const id = flat(G, i, j, k);
        double const c = eos.sound(rho[id], P[id]);
        double uu[ND];
#pragma unroll
        for (int d = 0; d < ND; ++d) uu[d] = u[id + G.plane * d];
        v = cfl_rate<ND, GEOM>(G, i, j, k, uu, c);
    }
    int const tid = threadIdx.x + kBX * threadIdx.y;
    v = block_reduce<kMax>(v, lds, tid, kRedThreads);
    if (tid == 0) partials[block_linear()] = v;
}

// The same CFL rate computed from conservative arrays (the fused path keeps no
// primitive arrays); cons -> prim follows array_conversion.hpp:121-137 exactly.
template <int ND, int GEOM>
__global__ __launch_bounds__(256) void k_time_step_cons(GridDev G, Range3 r, PerfectGas eos, const double* __restrict__ rho,
                                                        const double* __restrict__ rhou, const double* __restrict__ E,
                                                        double* __restrict__ partials)
{
    __shared__ double lds[kRedThreads / 64];
    int i, j, k;
    double v = reduce_identity<kMax>();
    if (range_cell(r, i, j, k)) {
        int64_t const id = flat(G, i, j, k);
        State<ND> c;
        c[0] = rho[id];
#pragma unroll
        for (int d = 0; d < ND; ++d) c[1 + d] = rhou[id + G.plane * d];
        c[ND + 1] = E[id];
        State<ND> const p = cons_to_prim(c, eos);
        v = cfl_rate<ND, GEOM>(G, i, j, k, &p.q[1], eos.sound(p[0], p[ND + 1]));
    }
    int const tid = threadIdx.x + kBX * threadIdx.y;
    v = block_reduce<kMax>(v, lds, tid, kRedThreads);
    if (tid == 0) partials[block_linear()] = v;
}

// analysis/internal_energy.cpp:25-40 ("min_internal_energy")
template <int ND>
__global__ __launch_bounds__(256) void k_min_eint(GridDev G, Range3 r, const double* __restrict__ rho, const double* __restrict__ rhou,
                                                  const double* __restrict__ E, double* __restrict__ partials)
{
    __shared__ double lds[kRedThreads / 64];
    int i, j, k;
    double v = reduce_identity<kMin>();
    if (range_cell(r, i, j, k)) {
        int64_t const id = flat(G, i, j, k);
        State<ND> c;
        c[0] = rho[id];
#pragma unroll
        for (int d = 0; d < ND; ++d) c[1 + d] = rhou[id + G.plane * d];
        c[ND + 1] = E[id];
        v = internal_energy(c);
    }
    int const tid = threadIdx.x + kBX * threadIdx.y;
    v = block_reduce<kMin>(v, lds, tid, kRedThreads);
    if (tid == 0) partials[block_linear()] = v;
}

// analysis/integration.cpp:28-32 ("integration")
template <int ND, int GEOM>
__global__ __launch_bounds__(256) void k_integrate(GridDev G, Range3 r, const double* __restrict__ var, double* __restrict__ partials)
{
    __shared__ double lds[kRedThreads / 64];
    int i, j, k;
    double v = 0;
    if (range_cell(r, i, j, k)) v = var[flat(G, i, j, k)] * cell_volume<ND, GEOM>(G, i, j, k);
    int const tid = threadIdx.x + kBX * threadIdx.y;
    v = block_reduce<kSum>(v, lds, tid, kRedThreads);
    if (tid == 0) partials[block_linear()] = v;
}

// Fold `n` partials in a fixed order with one workgroup of 1024 threads.
template <int OP>
__global__ __launch_bounds__(1024) void k_fold(const double* __restrict__ partials, int64_t n, double* __restrict__ out)
{
    __shared__ double lds[16];
    double v = reduce_identity<OP>();
    for (int64_t t = threadIdx.x; t < n; t += 1024) v = reduce_combine<OP>(v, partials[t]);
    v = block_reduce<OP>(v, lds, threadIdx.x, 1024);
    if (threadIdx.x == 0) out[0] = v;
}

// ------------------------------------------------------------ ghost fill
// Single-rank semantics of boundary_distribute.hpp:42-76: a periodic
// self-exchange of every face in dim order, then NullGradient / Reflexive
// (boundary.cpp:33-108) on the border faces in priority order.  Every
// operation copies whole slabs over the full transverse extent and reads
// only interior indices along its own dim, so the composite result is: a
// ghost cell takes the value of the interior cell found by applying each
// dim's face map to its index, with the normal momentum negated for every
// Reflexive map applied.  One kernel writes every ghost cell that way.
struct GhostBoxes {
    int nbox;
    int lo[6][3];
    int hi[6][3];
    int64_t start[7];
};

template <int ND>
__global__ __launch_bounds__(256) void k_fill_ghosts(GridDev G, BcSpec bc, GhostBoxes B, double* __restrict__ rho, double* __restrict__ rhou,
                                                     double* __restrict__ E, double* __restrict__ fx, int nfx)
{
    int64_t const t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= B.start[B.nbox]) return;
    int b = 0;
    while (t >= B.start[b + 1]) ++b;
    int64_t loc = t - B.start[b];
    int const nx = B.hi[b][0] - B.lo[b][0], ny = B.hi[b][1] - B.lo[b][1];
    int const i = B.lo[b][0] + (int)(loc % nx);
    loc /= nx;
    int const j = B.lo[b][1] + (int)(loc % ny);
    int const k = B.lo[b][2] + (int)(loc / ny);
    int idx[3] = {i, j, k};
    int src[3] = {i, j, k};
    bool flip[3] = {false, false, false};
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        int const side = idx[d] < G.ngh[d] ? 0 : 1;
        src[d] = face_map(bc.kind[2 * d + side], idx[d], G.ngh[d], G.n[d], G.nwg[d], flip[d]);
    }
    int64_t const dst = flat(G, i, j, k);
    int64_t const s = flat(G, src[0], src[1], src[2]);
    rho[dst] = rho[s];
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        double const m = rhou[s + G.plane * d];
        rhou[dst + G.plane * d] = flip[d] ? -m : m;
    }
    E[dst] = E[s];
    for (int f = 0; f < nfx; ++f) fx[dst + G.plane * f] = fx[s + G.plane * f];
}

// ------------------------------------------------------------ launchers
#define HCLPP_ND_SWITCH(nd, BODY)                \
    switch (nd) {                                \
    case 1: { constexpr int ND = 1; BODY; } break; \
    case 2: { constexpr int ND = 2; BODY; } break; \
    default: { constexpr int ND = 3; BODY; } break; \
    }

#define HCLPP_GEOM_SWITCH(geom, BODY)                        \
    if ((geom) == kSpherical) { constexpr int GEOM = kSpherical; BODY; } \
    else { constexpr int GEOM = kCartesian; BODY; }

static inline bool empty(const Range3& r) { return r.count() <= 0; }

void launch_prim_to_cons(int nd, const GridDev& G, const Range3& r, const PerfectGas& eos, const double* rho, const double* u, const double* P,
                         double* rhou, double* E, hipStream_t s)
{
    if (empty(r)) return;
    HCLPP_ND_SWITCH(nd, hipLaunchKernelGGL(k_prim_to_cons<ND>, range_grid(r), dim3(kBX, kBY), 0, s, G, r, eos, rho, u, P, rhou, E));
}

void launch_cons_to_prim(int nd, const GridDev& G, const Range3& r, const PerfectGas& eos, const double* rho, const double* rhou,
                         const double* E, double* u, double* P, hipStream_t s)
{
    if (empty(r)) return;
    HCLPP_ND_SWITCH(nd, hipLaunchKernelGGL(k_cons_to_prim<ND>, range_grid(r), dim3(kBX, kBY), 0, s, G, r, eos, rho, rhou, E, u, P));
}

void launch_face_reconstruction(int nd, int lim, const GridDev& G, const Range3& r, const double* var, double* rec, hipStream_t s)
{
    if (empty(r)) return;
#define HCLPP_LIM_CASE(L) \
    case L: HCLPP_ND_SWITCH(nd, hipLaunchKernelGGL((k_face_reconstruction<ND, L>), range_grid(r), dim3(kBX, kBY), 0, s, G, r, var, rec)); break;
    switch (lim) {
        HCLPP_LIM_CASE(kConstant)
        HCLPP_LIM_CASE(kVanLeer)
        HCLPP_LIM_CASE(kMinmod)
        HCLPP_LIM_CASE(kVanAlbada)
    }
#undef HCLPP_LIM_CASE
}

void launch_face_prim_to_cons(int nd, const GridDev& G, const Range3& r, const PerfectGas& eos, const double* rho_rec, const double* u_rec,
                              const double* P_rec, double* rhou_rec, double* E_rec, hipStream_t s)
{
    if (empty(r)) return;
    HCLPP_ND_SWITCH(nd, hipLaunchKernelGGL(k_face_prim_to_cons<ND>, range_grid(r), dim3(kBX, kBY), 0, s, G, r, eos, rho_rec, u_rec, P_rec,
                                           rhou_rec, E_rec));
}

void launch_hancock(int nd, int geom, const GridDev& G, const Range3& r, const PerfectGas& eos, double dt, const double* u_rec,
                    const double* P_rec, double* rho_rec, double* rhou_rec, double* E_rec, double* fx_rec, const double* fx_old, int nfx,
                    hipStream_t s)
{
    if (empty(r)) return;
    HCLPP_GEOM_SWITCH(geom, HCLPP_ND_SWITCH(nd, hipLaunchKernelGGL((k_hancock<ND, GEOM>), range_grid(r), dim3(kBX, kBY), 0, s, G, r, eos, dt,
                                                                   u_rec, P_rec, rho_rec, rhou_rec, E_rec, fx_rec, fx_old, nfx)));
}

void launch_godunov(int nd, int geom, int rs, const GridDev& G, const Range3& r, const PerfectGas& eos, double dt, const double* rho,
                    const double* rhou, const double* E, const double* fx, const double* rho_rec, const double* rhou_rec, const double* E_rec,
                    const double* fx_rec, double* rho_new, double* rhou_new, double* E_new, double* fx_new, int nfx, hipStream_t s)
{
    if (empty(r)) return;
#define HCLPP_RS_CASE(RSV)                                                                                                         \
    case RSV:                                                                                                                      \
        HCLPP_GEOM_SWITCH(geom, HCLPP_ND_SWITCH(nd, hipLaunchKernelGGL((k_godunov<ND, GEOM, RSV>), range_grid(r), dim3(kBX, kBY), 0, s, G, r, \
                                                                       eos, dt, rho, rhou, E, fx, rho_rec, rhou_rec, E_rec, fx_rec,      \
                                                                       rho_new, rhou_new, E_new, fx_new, nfx)));                         \
        break;
    switch (rs) {
        HCLPP_RS_CASE(kHLL)
        HCLPP_RS_CASE(kHLLC)
        HCLPP_RS_CASE(kLowMach)
    }
#undef HCLPP_RS_CASE
}

int64_t reduce_blocks(const Range3& r)
{
    dim3 const g = range_grid(r);
    return (int64_t)g.x * g.y * g.z;
}

void launch_time_step_rate(int nd, int geom, const GridDev& G, const Range3& r, const PerfectGas& eos, const double* rho, const double* u,
                           const double* P, double* partials, double* out, hipStream_t s)
{
    HCLPP_GEOM_SWITCH(geom, HCLPP_ND_SWITCH(nd, hipLaunchKernelGGL((k_time_step<ND, GEOM>), range_grid(r), dim3(kBX, kBY), 0, s, G, r, eos, rho,
                                                                   u, P, partials)));
    hipLaunchKernelGGL(k_fold<kMax>, dim3(1), dim3(1024), 0, s, partials, reduce_blocks(r), out);
}

void launch_time_step_cons(int nd, int geom, const GridDev& G, const Range3& r, const PerfectGas& eos, const double* rho, const double* rhou,
                           const double* E, double* partials, double* out, hipStream_t s)
{
    HCLPP_GEOM_SWITCH(geom, HCLPP_ND_SWITCH(nd, hipLaunchKernelGGL((k_time_step_cons<ND, GEOM>), range_grid(r), dim3(kBX, kBY), 0, s, G, r, eos,
                                                                   rho, rhou, E, partials)));
    hipLaunchKernelGGL(k_fold<kMax>, dim3(1), dim3(1024), 0, s, partials, reduce_blocks(r), out);
}

// analysis/temperature.hpp:18-26 with perfect_gas.hpp:69-72 (IEEE division, as the reference)
__global__ __launch_bounds__(256) void k_temperature(GridDev G, Range3 r, PerfectGas eos, const double* __restrict__ rho,
                                                     const double* __restrict__ P, double* __restrict__ T)
{
    int i, j, k;
    if (!range_cell(r, i, j, k)) return;
    int64_t const id = flat(G, i, j, k);
    T[id] = (P[id] * eos.mmw * kMp) / (rho[id] * kKb);
}

void launch_temperature(const GridDev& G, const Range3& r, const PerfectGas& eos, const double* rho, const double* P, double* T, hipStream_t s)
{
    if (empty(r)) return;
    hipLaunchKernelGGL(k_temperature, range_grid(r), dim3(kBX, kBY), 0, s, G, r, eos, rho, P, T);
}

// k_fold<kMax> writing the reciprocal of the maximum: time_step.hpp:81 returns 1 / max
__global__ __launch_bounds__(1024) void k_fold_max_inverse(const double* __restrict__ partials, int64_t n, double* __restrict__ out)
{
    __shared__ double lds[16];
    double v = reduce_identity<kMax>();
    for (int64_t t = threadIdx.x; t < n; t += 1024) v = reduce_combine<kMax>(v, partials[t]);
    v = block_reduce<kMax>(v, lds, threadIdx.x, 1024);
    if (threadIdx.x == 0) out[0] = 1 / v;
}

void launch_time_step_inverse(int nd, int geom, const GridDev& G, const Range3& r, const PerfectGas& eos, const double* rho, const double* u,
                              const double* P, double* partials, double* dt_out, hipStream_t s)
{
    HCLPP_GEOM_SWITCH(geom, HCLPP_ND_SWITCH(nd, hipLaunchKernelGGL((k_time_step<ND, GEOM>), range_grid(r), dim3(kBX, kBY), 0, s, G, r, eos, rho,
                                                                   u, P, partials)));
    hipLaunchKernelGGL(k_fold_max_inverse, dim3(1), dim3(1024), 0, s, partials, reduce_blocks(r), dt_out);
}

void launch_fold_max(const double* partials, int64_t n, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_fold<kMax>, dim3(1), dim3(1024), 0, s, partials, n, out);
}

void launch_min_internal_energy(int nd, const GridDev& G, const Range3& r, const double* rho, const double* rhou, const double* E,
                                double* partials, double* out, hipStream_t s)
{
    HCLPP_ND_SWITCH(nd, hipLaunchKernelGGL(k_min_eint<ND>, range_grid(r), dim3(kBX, kBY), 0, s, G, r, rho, rhou, E, partials));
    hipLaunchKernelGGL(k_fold<kMin>, dim3(1), dim3(1024), 0, s, partials, reduce_blocks(r), out);
}

void launch_integrate(const GridDev& G, int nd, int geom, const Range3& r, const double* var, double* partials, double* out, hipStream_t s)
{
    HCLPP_GEOM_SWITCH(geom, HCLPP_ND_SWITCH(nd, hipLaunchKernelGGL((k_integrate<ND, GEOM>), range_grid(r), dim3(kBX, kBY), 0, s, G, r, var,
                                                                   partials)));
    hipLaunchKernelGGL(k_fold<kSum>, dim3(1), dim3(1024), 0, s, partials, reduce_blocks(r), out);
}

void launch_fill_ghosts(int nd, const GridDev& G, const BcSpec& bc, double* rho, double* rhou, double* E, double* fx, int nfx, hipStream_t s)
{
    // Partition the ghost shell: slab d holds cells that are ghosts along d,
    // interior along every later dim and anything along earlier dims.
    GhostBoxes B{};
    int64_t total = 0;
    for (int d = 0; d < nd; ++d) {
        for (int side = 0; side < 2; ++side) {
            int lo[3], hi[3];
            for (int e = 0; e < 3; ++e) {
                if (e < d) { lo[e] = 0; hi[e] = G.nwg[e]; }
                else if (e > d) { lo[e] = G.ngh[e]; hi[e] = G.ngh[e] + G.n[e]; }
                else if (side == 0) { lo[e] = 0; hi[e] = G.ngh[e]; }
                else { lo[e] = G.ngh[e] + G.n[e]; hi[e] = G.nwg[e]; }
            }
            int const b = B.nbox++;
            for (int e = 0; e < 3; ++e) { B.lo[b][e] = lo[e]; B.hi[b][e] = hi[e]; }
            B.start[b] = total;
            total += (int64_t)(hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]);
        }
    }
    B.start[B.nbox] = total;
    if (total == 0) return;
    unsigned const blocks = (unsigned)((total + 255) / 256);
    HCLPP_ND_SWITCH(nd, hipLaunchKernelGGL(k_fill_ghosts<ND>, dim3(blocks), dim3(256), 0, s, G, bc, B, rho, rhou, E, fx, nfx));
}

// ------------------------------------------------------------ pressure fix
// hydro/pressure_fix.hpp:19-172.  For a cell whose internal energy is below eps times its
// kinetic energy, E_new is rebuilt from the internal-energy equation de = dt (-P div u -
// div(u e)) (face velocities from the predicted face states, upwinded cell internal
// energies), blended with the Godunov change by alpha = (e / (eps ek))^2 (0 below eps / 10).
// The double loop over (idim, ipos) and the face areas of direction ipos are the reference's.
template <int ND, int GEOM>
__global__ __launch_bounds__(256) void k_pressure_fix(GridDev G, Range3 r, PerfectGas eos, double dt, double eps,
                                                      const double* __restrict__ rho, const double* __restrict__ rhou,
                                                      const double* __restrict__ E, const double* __restrict__ rho_rec,
                                                      const double* __restrict__ rhou_rec, const double* __restrict__ E_rec,
                                                      const double* __restrict__ rho_new, const double* __restrict__ rhou_new,
                                                      double* __restrict__ E_new)
{
    int i, j, k;
    if (!range_cell(r, i, j, k)) return;
    int64_t const pl = G.plane;
    int64_t const id = flat(G, i, j, k);
    auto cell = [&](int64_t c) {
        State<ND> q;
        q[0] = rho[c];
#pragma unroll
        for (int d = 0; d < ND; ++d) q[1 + d] = rhou[c + pl * d];
        q[ND + 1] = E[c];
        return q;
    };
    auto face = [&](int64_t c, int s, int idim) {
        State<ND> q;
        int64_t const off = c + pl * (s + 2 * idim);
        q[0] = rho_rec[off];
#pragma unroll
        for (int v = 0; v < ND; ++v) q[1 + v] = rhou_rec[off + pl * 2 * ND * v];
        q[ND + 1] = E_rec[off];
        return q;
    };
    State<ND> const cons = cell(id);
    State<ND> cons_new;
    cons_new[0] = rho_new[id];
#pragma unroll
    for (int d = 0; d < ND; ++d) cons_new[1 + d] = rhou_new[id + pl * d];
    cons_new[ND + 1] = E_new[id];
    if (!pressure_fix_applies(cons, eps)) return;
    double const out = pressure_fix_energy<ND, GEOM>(G, eos, dt, eps, i, j, k, cons, cons_new, cell, face);
    E_new[id] = out;
}

void launch_pressure_fix(int nd, int geom, const GridDev& G, const Range3& r, const PerfectGas& eos, double dt, double eps, const double* rho,
                         const double* rhou, const double* E, const double* rho_rec, const double* rhou_rec, const double* E_rec,
                         const double* rho_new, const double* rhou_new, double* E_new, hipStream_t s)
{
    if (empty(r)) return;
    HCLPP_GEOM_SWITCH(geom, HCLPP_ND_SWITCH(nd, hipLaunchKernelGGL((k_pressure_fix<ND, GEOM>), range_grid(r), dim3(kBX, kBY), 0, s, G, r, eos,
                                                                   dt, eps, rho, rhou, E, rho_rec, rhou_rec, E_rec, rho_new, rhou_new,
                                                                   E_new)));
}

// ------------------------------------------------------------ internal-mass gravity
// gravity/gravity.cpp:59-128 on one rank along dim 0.  Pass 1: one workgroup per radial
// layer s (cells i = s + ng, s < n + ng, i.e. the interior and the outer ghosts) sums
// rho dV and dV over the interior (j, k) in a fixed order.  Pass 2 (one thread): the
// enclosed mass at each cell centre by the reference's sequential scan, then g = -G M / r^2.
template <int ND, int GEOM, class S>
__global__ __launch_bounds__(256) void k_shell_sums(GridDev G, const S* __restrict__ rho, double* __restrict__ shells, const DevFlag* run)
{
    __shared__ double lds[4];
    if (run && !run->go()) return;
    int const s = blockIdx.x;
    int const i = s + G.ngh[0];
    int const ny = G.n[1], nz = G.n[2];
    int64_t const cnt = (int64_t)ny * nz;
    double sm = 0, sv = 0;
    for (int64_t t = threadIdx.x; t < cnt; t += 256) {
        int const j = G.ngh[1] + (int)(t % ny);
        int const k = G.ngh[2] + (int)(t / ny);
        double const dv = cell_volume<ND, GEOM>(G, i, j, k);
        sm += (double)rho[flat(G, i, j, k)] * dv;
        sv += dv;
    }
    sm = block_reduce<kSum>(sm, lds, threadIdx.x, 256);
    sv = block_reduce<kSum>(sv, lds, threadIdx.x, 256);
    if (threadIdx.x == 0) {
        shells[2 * s] = sm;
        shells[2 * s + 1] = sv;
    }
}

__global__ __launch_bounds__(64) void k_mass_table(GridDev G, double central_mass, const double* __restrict__ shells, double* __restrict__ gtab,
                                                   const DevFlag* run)
{
    if (threadIdx.x != 0) return;
    if (run && !run->go()) return;
    int const ng0 = G.ngh[0], n0 = G.n[0];
    double const* x0 = G.x[0];
    double const* xc = G.xc[0];
    double const vol_s3 = 4 * kPi / 3;
    double partial = 0;
    for (int i = 0; i < G.nwg[0]; ++i) {
        double M;
        if (i < ng0) {
            M = central_mass;
        } else {
            int const off = i - ng0;
            double const x3 = x0[i] * x0[i] * x0[i];
            double const rm = shells[2 * off] / shells[2 * off + 1];
            if (i == ng0) {
                partial += central_mass + (vol_s3 * ((xc[i] * xc[i] * xc[i]) - x3) * rm);
            } else {
                double const rmm = shells[2 * (off - 1)] / shells[2 * (off - 1) + 1];
                partial += (vol_s3 * (x3 - (xc[i - 1] * xc[i - 1] * xc[i - 1])) * rmm) + (vol_s3 * ((xc[i] * xc[i] * xc[i]) - x3) * rm);
            }
            M = partial;
        }
        gtab[i] = -kGrav * M / (xc[i] * xc[i]);
    }
    (void)n0;
}

template <class S>
void launch_internal_mass_gravity(int nd, int geom, const GridDev& G, double central_mass, const S* rho, double* shells, double* gtab,
                                  const DevFlag* run, hipStream_t s)
{
    unsigned const nsh = (unsigned)(G.n[0] + G.ngh[0]);
    HCLPP_GEOM_SWITCH(geom, HCLPP_ND_SWITCH(nd, hipLaunchKernelGGL((k_shell_sums<ND, GEOM, S>), dim3(nsh), dim3(256), 0, s, G, rho, shells, run)));
    hipLaunchKernelGGL(k_mass_table, dim3(1), dim3(64), 0, s, G, central_mass, shells, gtab, run);
}
template void launch_internal_mass_gravity<double>(int, int, const GridDev&, double, const double*, double*, double*, const DevFlag*, hipStream_t);
template void launch_internal_mass_gravity<float>(int, int, const GridDev&, double, const float*, double*, double*, const DevFlag*, hipStream_t);

}  // namespace hclpp
