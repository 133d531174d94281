// step_fused.hip -- launchers of the fused step (k_step_march, march.hpp, without passive
// scalars), the ghost fill of the new state with its CFL rates, and the device clock.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "march.hpp"

namespace hclpp {

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
        for (int f = 0; f < A.nfx; ++f) A.fx_new[dst + G.plane * f] = A.fx_new[s + G.plane * f];  // no sign flip
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
// main.cpp:355-385 (dt, output clipping, t_end clip, max_iter) and :397-425 (internal energy check)
__device__ __forceinline__ void schedule_next(DevClock* c, double inv)
{
    double dt = c->pub.cfl * (1 / inv);
    int pending = 0, output = 0;
    if (c->iter_output_frequency > 0) {  // :362-368 (iter_ini = 0)
        int const next_output = (c->pub.iter_output_id + 1) * c->iter_output_frequency;
        if ((c->pub.iter + 1) >= next_output) {
            output = 1;
            c->pub.iter_output_id += 1;
        }
    }
    if (c->time_output_frequency > 0) {  // :369-375
        double const next_output = c->time_first_output + ((c->pub.time_output_id + 1) * c->time_output_frequency);
        if ((c->pub.t + dt) >= next_output) {
            dt = next_output - c->pub.t;
            if (dt <= 0) {
                c->pub.status = HCLPP_ERR_NEGATIVE_DT;
                return;
            }
            output = 1;
            c->pub.time_output_id += 1;
        }
    }
    if ((c->pub.t + dt) >= c->pub.t_end) {
        dt = c->pub.t_end - c->pub.t;
        output = 1;
        pending = 1;
    }
    if ((c->pub.iter + 1) >= c->pub.max_iter) {
        output = 1;
        pending = 1;
    }
    c->pub.dt = dt;
    c->pending_exit = pending;
    c->pending_output = output;
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
    c->pending_output = 0;
}
__global__ __launch_bounds__(64) void k_clock_set_log(DevClock* c, double* log, int cap)
{
    if (threadIdx.x != 0) return;
    c->dt_log = log;
    c->dt_log_cap = log ? cap : 0;
}

template <int NV>
__global__ __launch_bounds__(1024) void k_clock_advance(DevClock* c, const double* __restrict__ partials, int64_t nblocks,
                                                        const double* __restrict__ ghost_partials, int64_t nghost,
                                                        const double* __restrict__ fix_partials, int64_t nfix)
{
    __shared__ double lds[16][NV + 2];  // per-wave results of every reduced value: one LDS round
    if (c->pub.done || c->pub.status) return;
    double inv = reduce_identity<kMax>(), emin = reduce_identity<kMin>();
    double s[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) s[v] = 0;
    // rows t, t + 1024, ... of this thread, folded in that order; the loads of kR rows are issued
    // before their folds, so a launch pays about one memory latency per kR rows, not per row
    constexpr int kR = 4;
    for (int64_t t0 = threadIdx.x; t0 < nblocks; t0 += kR * 1024) {
        double row[kR][NV + 2];
#pragma unroll
        for (int r = 0; r < kR; ++r) {
            int64_t const t = t0 + r * 1024;
#pragma unroll
            for (int v = 0; v < NV + 2; ++v) row[r][v] = t < nblocks ? partials[t * kPartialStride + v] : 0.0;
        }
#pragma unroll
        for (int r = 0; r < kR; ++r) {
            if (t0 + r * 1024 >= nblocks) break;
            inv = reduce_combine<kMax>(inv, row[r][0]);
            emin = reduce_combine<kMin>(emin, row[r][1]);
#pragma unroll
            for (int v = 0; v < NV; ++v) s[v] += row[r][2 + v];
        }
    }
    for (int64_t t = threadIdx.x; t < nghost; t += 1024) inv = reduce_combine<kMax>(inv, ghost_partials[t]);
    // pressure fix: rates of the rewritten cells and the change of their energy times dV
    for (int64_t t = threadIdx.x; t < nfix; t += 1024) {
        inv = reduce_combine<kMax>(inv, fix_partials[2 * t]);
        s[NV - 1] += fix_partials[2 * t + 1];
    }
    // all NV + 2 values reduced together: wave shuffles, one barrier, then thread 0 folds the
    // 16 wave results in wave order (a fixed order: the totals are deterministic)
    inv = wave_reduce<kMax>(inv);
    emin = wave_reduce<kMin>(emin);
#pragma unroll
    for (int v = 0; v < NV; ++v) s[v] = wave_reduce<kSum>(s[v]);
    int const lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        lds[wave][0] = inv;
        lds[wave][1] = emin;
#pragma unroll
        for (int v = 0; v < NV; ++v) lds[wave][2 + v] = s[v];
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    inv = reduce_identity<kMax>();
    emin = reduce_identity<kMin>();
#pragma unroll
    for (int v = 0; v < NV; ++v) s[v] = 0;
    for (int w = 0; w < 16; ++w) {
        inv = reduce_combine<kMax>(inv, lds[w][0]);
        emin = reduce_combine<kMin>(emin, lds[w][1]);
#pragma unroll
        for (int v = 0; v < NV; ++v) s[v] += lds[w][2 + v];
    }
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
    c->pub.output_due = c->pending_output;
    if (c->pending_exit) {
        c->pub.done = 1;
        c->pub.exiting = 1;
        return;
    }
    if (c->pending_output) {  // pause: the caller writes the output, then starts the clock again
        c->pub.done = 1;
        return;
    }
    schedule_next(c, inv);
}

// ------------------------------------------------------------ launchers
// Launch plan.  Every workgroup of the march kernel carries the same work: one tile through a
// segment of `seg` planes, plus the two halo planes zs-1 and ze its predictor also visits.  The
// chip holds `resident` = blocks_per_cu x ncu workgroups at once, so a launch of nb workgroups
// runs in ceil(nb / resident) rounds of (seg + 2) planes each.  The plan takes the segment
// length (kMinSeg planes at least) with the fewest rounds x (seg + 2): e.g. 2D 2048^2 in FP64
// (34 tiles, two resident single-wave workgroups per SIMD) gets 35-plane segments in one round
// of 2006 workgroups, where 32-plane segments need 2176 workgroups and a second round for 6 %
// of the work.  The split changes which workgroup computes a cell, not how: results do not
// depend on it (the conserved totals, summed per workgroup, only change their summation order).
MarchPlan plan_march(const GridDev& G, int nd, int blocks_per_cu, int ncu)
{
    MarchPlan P{};
    int const nm = G.n[nd - 1];
    if (nd == 1) {  // one segment per thread: the same cost model over NT-segment workgroups
        using S = MarchShape<1>;
        P.L.ntile[0] = 1;
        P.L.ntile[1] = 1;
        int64_t const resident = (int64_t)std::max(blocks_per_cu, 1) * std::max(ncu, 1);
        int64_t best_cost = -1;
        for (int seg = std::min(4, nm); seg <= std::min(std::max(nm, 1), 1 << 16); ++seg) {
            int64_t const nb = ((nm + seg - 1) / seg + S::NT - 1) / S::NT;
            int64_t const cost = ((nb + resident - 1) / resident) * (seg + 2);
            if (best_cost < 0 || cost < best_cost) {
                best_cost = cost;
                P.L.seglen = seg;
                P.nb = nb;
            }
        }
        return P;
    }
    int const HX = nd == 2 ? MarchShape<2>::HX : MarchShape<3>::HX;
    int const HY = nd == 2 ? MarchShape<2>::HY : MarchShape<3>::HY;
    int const TX = HX - 2;
    int const TY = (nd == 3) ? HY - 2 : 1;
    P.L.ntile[0] = (G.n[0] + TX - 1) / TX;
    P.L.ntile[1] = (nd == 3) ? (G.n[1] + TY - 1) / TY : 1;
    int64_t const tiles = (int64_t)P.L.ntile[0] * P.L.ntile[1];
    int64_t const resident = (int64_t)std::max(blocks_per_cu, 1) * std::max(ncu, 1);
    constexpr int kMinSeg = 4;
    int64_t best_cost = -1;
    for (int seg = std::min(kMinSeg, nm); seg <= nm; ++seg) {
        int64_t const nb = tiles * ((nm + seg - 1) / seg);
        int64_t const cost = ((nb + resident - 1) / resident) * (seg + 2);
        if (best_cost < 0 || cost < best_cost) {
            best_cost = cost;
            P.L.seglen = seg;
            P.nb = nb;
        }
    }
    return P;
}

int march_occupancy(int nd, int geom, int rs, int lim, bool fp32, bool fx)
{
    if (fx) return march_occupancy_fx(nd, geom, rs, lim, fp32);
    int n = 0;
    if (fp32) march_visit(nd, geom, rs, lim, MarchOccupancy<PerfectGasMixed, 0>{&n});
    else march_visit(nd, geom, rs, lim, MarchOccupancy<PerfectGas, 0>{&n});
    return n;
}

void launch_fused_step(int nd, int geom, int rs, int lim, const GridDev& G, const MarchPlan& P, const PerfectGas& eos, const StepArrays& A,
                       hipStream_t s)
{
    march_visit(nd, geom, rs, lim, MarchLaunch<PerfectGas, 0>{G, P, eos, A, s});
}

void launch_fused_step(int nd, int geom, int rs, int lim, const GridDev& G, const MarchPlan& P, const PerfectGasMixed& eos,
                       const StepArraysF& A, hipStream_t s)
{
    march_visit(nd, geom, rs, lim, MarchLaunch<PerfectGasMixed, 0>{G, P, eos, A, s});
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
                          const double* fix_partials, int64_t nfix, hipStream_t s)
{
    switch (nd) {
    case 1:
        hipLaunchKernelGGL(k_clock_advance<3>, dim3(1), dim3(1024), 0, s, c, partials, nblocks, ghost_partials, nghost, fix_partials, nfix);
        break;
    case 2:
        hipLaunchKernelGGL(k_clock_advance<4>, dim3(1), dim3(1024), 0, s, c, partials, nblocks, ghost_partials, nghost, fix_partials, nfix);
        break;
    default:
        hipLaunchKernelGGL(k_clock_advance<5>, dim3(1), dim3(1024), 0, s, c, partials, nblocks, ghost_partials, nghost, fix_partials, nfix);
        break;
    }
}

}  // namespace hclpp
