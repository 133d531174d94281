// pressure_fix_fused.hip -- the pressure fix (hydro/pressure_fix.hpp:19-172, main.cpp:459-475)
// inside the fused step.
//
// The fused march kernel keeps face states in registers and never writes them to memory, but
// the fix needs the Hancock face states of the cell it rewrites and of its neighbours.  Whether
// a cell is rewritten depends on its state at time n alone (E - e_k < eps e_k), so this kernel,
// launched after the march kernel on the same input state, tests every interior cell and, for
// the cells the fix applies to, recomputes those face states with the march kernel's own
// predictor (same function, same inputs: the same bits), then applies the reference's formula
// to E_new.  The march kernel leaves these cells out of its CFL maximum; this kernel supplies
// their rate from the fixed state and the change of their energy times dV, per workgroup, for
// the device clock's fold.  Cells the fix does not touch cost one read of their state, and a
// workgroup with no such cell writes identity partials without reducing.
#include <hip/hip_runtime.h>

#include "march.hpp"
#include "pressure_fix.hpp"

namespace hclpp {

namespace {

// The Hancock face states fs[side][dim] of cell (i, j, k), recomputed from the state at time n
// (the march kernel's predictor: predict_faces over the cell's primitive and its face neighbours').
template <int ND, int GEOM, int LIM>
__device__ __forceinline__ void predictor_at(const GridDev& G, const PerfectGas& eos, double dth, const StepArrays& A, int i, int j, int k,
                                             State<ND> (&fs)[2][ND])
{
    auto prim = [&](int a, int b, int c) { return cons_to_prim(load_cons<ND>(A, G, flat(G, a, b, c)), eos); };
    State<ND> const p = prim(i, j, k);
    State<ND> nb[ND][2];
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        int const di = d == 0, dj = d == 1, dk = d == 2;
        nb[d][0] = prim(i - di, j - dj, k - dk);
        nb[d][1] = prim(i + di, j + dj, k + dk);
    }
    double const dtodv_h = fdiv(dth, cell_volume<ND, GEOM>(G, i, j, k));
    Frac<0, double> const x = zero_frac<0, double>();
    Frac<0, double> xnb[ND][2];
    Frac<0, double> xf[2][ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) xnb[d][0] = xnb[d][1] = x;
    predict_faces<PerfectGas, ND, GEOM, LIM, 0>(G, eos, dth, dtodv_h, i, j, k, p, nb, fs, x, xnb, xf);
}

template <int ND, int GEOM, int LIM, int HX, int HY, int NT>
__global__ __launch_bounds__(NT) void k_pressure_fix_fused(GridDev G, PerfectGas eos, StepArrays A, FusedLayout L,
                                                           double* __restrict__ fix_partials)
{
    __shared__ double s_red[NT / 64];
    DevClock const* clk = A.clock;
    if (clk->pub.done || clk->pub.status) return;
    int const b = blockIdx.x, tid = threadIdx.x;
    double rate = reduce_identity<kMax>();
    double dE = 0;
    // the march workgroup b flagged whether any of its cells is rewritten
    if (A.partials[(int64_t)b * kPartialStride + kPartialStride - 1] == 0.0) {
        if (tid == 0) {
            fix_partials[2 * b] = rate;
            fix_partials[2 * b + 1] = dE;
        }
        return;
    }
    double const dt = clk->pub.dt;
    double const dth = dt / 2;
    double const eps = A.eps_pf;
    MarchColumn const mc = march_column<ND, HX, HY, NT>(G, L, b, gridDim.x, tid);
    if (mc.out_col) {
        for (int z = mc.zs; z < mc.ze; ++z) {
            int i, j, k;
            ijk_of<ND>(mc.a, mc.bb, z, i, j, k);
            int64_t const id = flat(G, i, j, k);
            State<ND> const cons = load_cons<ND>(A, G, id);
            if (!pressure_fix_applies(cons, eps)) continue;
            State<ND> un;
            un[0] = A.rho_new[id];
#pragma unroll
            for (int d = 0; d < ND; ++d) un[1 + d] = A.rhou_new[id + G.plane * d];
            un[ND + 1] = A.E_new[id];
            // face states of this cell, and of its neighbours along each dim as the fix reads them
            State<ND> own[2][ND];
            predictor_at<ND, GEOM, LIM>(G, eos, dth, A, i, j, k, own);
            State<ND> nbf[ND][2];  // [idim][0]: right face of cell - 1; [idim][1]: left face of cell + 1
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                int const di = d == 0, dj = d == 1, dk = d == 2;
                State<ND> f[2][ND];
                predictor_at<ND, GEOM, LIM>(G, eos, dth, A, i - di, j - dj, k - dk, f);
                nbf[d][0] = f[1][d];
                predictor_at<ND, GEOM, LIM>(G, eos, dth, A, i + di, j + dj, k + dk, f);
                nbf[d][1] = f[0][d];
            }
            auto cell = [&](int64_t c) { return load_cons<ND>(A, G, c); };
            auto face = [&](int64_t c, int s, int d) -> State<ND> {
                if (c == id) return own[s][d];
                return s == 1 ? nbf[d][0] : nbf[d][1];  // (cell - 1, right face) or (cell + 1, left face)
            };
            double const E_fixed = pressure_fix_energy<ND, GEOM>(G, eos, dt, eps, i, j, k, cons, un, cell, face);
            double const vol = cell_volume<ND, GEOM>(G, i, j, k);
            dE += (E_fixed - un[ND + 1]) * vol;
            A.E_new[id] = E_fixed;
            un[ND + 1] = E_fixed;
            rate = reduce_combine<kMax>(rate, state_rate<ND, GEOM>(G, eos, i, j, k, un));
        }
    }
    rate = block_reduce<kMax>(rate, s_red, tid, NT);
    dE = block_reduce<kSum>(dE, s_red, tid, NT);
    if (tid == 0) {
        fix_partials[2 * b] = rate;
        fix_partials[2 * b + 1] = dE;
    }
}

struct FixLaunch {
    const GridDev& G;
    const PerfectGas& eos;
    const StepArrays& A;
    const MarchPlan& P;
    double* partials;
    hipStream_t s;
    template <int ND, int GEOM, int RS, int LIM>
    void run() const
    {
        using S = MarchShape<ND>;
        if constexpr (RS == kHLL) {  // the fix does not depend on the Riemann solver: one instance per limiter
            if (P.nb == 0) return;
            hipLaunchKernelGGL((k_pressure_fix_fused<ND, GEOM, LIM, S::HX, S::HY, S::NT>), dim3((unsigned)P.nb), dim3(S::NT), 0, s, G, eos,
                               A, P.L, partials);
        }
    }
};

}  // namespace

// The fused step's pressure fix over the march kernel's workgroups (plan P), after k_step_march;
// fix_partials receives two doubles per workgroup.
void launch_pressure_fix_fused(int nd, int geom, int lim, const GridDev& G, const PerfectGas& eos, const StepArrays& A, const MarchPlan& P,
                               double* fix_partials, hipStream_t s)
{
    march_visit(nd, geom, kHLL, lim, FixLaunch{G, eos, A, P, fix_partials, s});
}

}  // namespace hclpp
