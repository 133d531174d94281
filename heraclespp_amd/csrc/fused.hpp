// fused.hpp -- structures shared by the fused step kernels and the C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "grid.hpp"
#include "hclpp_hip.h"
#include "launch.hpp"

namespace hclpp {

// Device-resident clock: the public hclpp_clock plus the exit flag of the step in flight.
struct DevClock {
    hclpp_clock pub;
    int pending_exit;
    int pending_output;  // the scheduled step ends at an output
    int dt_log_cap;   // entries of dt_log
    double* dt_log;   // when set, step `iter` records its dt at dt_log[iter] (iter < dt_log_cap)
    // output schedule (main.cpp:362-375); frequencies of 0 disable it
    int iter_output_frequency;
    double time_first_output;
    double time_output_frequency;
};

// Per-workgroup partials of the fused step: [0] max CFL rate, [1] min internal
// energy, [2..2+NV) sums of the conserved variables times dV, [kPartialStride - 1] 1 when the
// pressure fix rewrites some cell of the workgroup (else 0).
constexpr int kPartialStride = 8;

// Arrays of one fused step; S is the storage type (double, or float for FP32 storage).
template <class S>
struct StepArraysT {
    const S* __restrict__ rho;
    const S* __restrict__ rhou;
    const S* __restrict__ E;
    S* __restrict__ rho_new;
    S* __restrict__ rhou_new;
    S* __restrict__ E_new;
    const S* __restrict__ fx;  // passive scalars (mass fractions, [f][cell]); unused when nfx = 0
    S* __restrict__ fx_new;
    int nfx;
    // launches after the first of a step with several passive scalars (one scalar per launch, fx
    // and fx_new pointing at it): the kernel writes that scalar's update only, not the hydro
    // update or the workgroup partials the first launch wrote
    int scalars_only;
    // pressure fix (FP64 only): cells whose state at time n has E - e_k < eps_pf e_k are left out
    // of the march kernel's CFL maximum; k_pressure_fix_fused rewrites their energy and rate
    int pf;
    double eps_pf;
    double* __restrict__ partials;
    double* __restrict__ ghost_partials;
    DevClock* clock;
};
using StepArrays = StepArraysT<double>;
using StepArraysF = StepArraysT<float>;

struct FusedLayout {
    int ntile[2];  // in-plane tile counts
    int seglen;    // planes per workgroup along the march dim
};

// Ghost shell split into 2*ndim disjoint boxes (slab d: ghost along d,
// interior along later dims, any index along earlier dims).
struct GhostShell {
    int nbox;
    int lo[6][3];
    int hi[6][3];
    int64_t start[7];
};

// Workgroup shape of the fused kernel per dimensionality.
template <int ND>
struct MarchShape;
template <>
struct MarchShape<1> {  // 64 independent segments per workgroup (plan_march sets their length)
    static constexpr int HX = 1, HY = 1, NT = 64;
};
template <>
struct MarchShape<2> {  // a 64-column row (62 outputs + halo) marching along y
    static constexpr int HX = 64, HY = 1, NT = 64;
};
template <>
struct MarchShape<3> {  // 16 x 16 columns (14 x 14 outputs + halo) marching along z
    // The halo columns repeat the predictor of their neighbours' faces: 16 x 16 computes 256
    // predictors for 196 outputs where 32 x 8 computed 256 for 180 (measured +11 % on RT 512^3).
    static constexpr int HX = 16, HY = 16, NT = 256;
};

// Launch plan of the fused march kernel for one context.
struct MarchPlan {
    FusedLayout L;
    int64_t nb;  // workgroups (one partials row each)
};
// Tile counts and segment length for a kernel of which blocks_per_cu workgroups are resident
// per CU, on ncu CUs (see step_fused.hip for the cost model).
MarchPlan plan_march(const GridDev& G, int nd, int blocks_per_cu, int ncu);
// Resident workgroups per CU of the k_step_march instance a configuration selects
// (FP32 storage or not, with the one-scalar kernel or without scalars).
int march_occupancy(int nd, int geom, int rs, int lim, bool fp32, bool fx);
void launch_fused_step(int nd, int geom, int rs, int lim, const GridDev& G, const MarchPlan& P, const PerfectGas& eos, const StepArrays& A,
                       hipStream_t s);
void launch_fused_step(int nd, int geom, int rs, int lim, const GridDev& G, const MarchPlan& P, const PerfectGasMixed& eos,
                       const StepArraysF& A, hipStream_t s);
// the fused step with the opt-in fast FP64 arithmetic (step_fused_fast.hip; no passive scalars)
void launch_fused_step(int nd, int geom, int rs, int lim, const GridDev& G, const MarchPlan& P, const PerfectGasFast& eos, const StepArrays& A,
                       hipStream_t s);
int march_occupancy_fast(int nd, int geom, int rs, int lim);
// the fused step carrying one passive scalar (step_fused_fx.hip); A.fx and A.fx_new point at it
void launch_fused_step_fx(int nd, int geom, int rs, int lim, const GridDev& G, const MarchPlan& P, const PerfectGas& eos, const StepArrays& A,
                          hipStream_t s);
void launch_fused_step_fx(int nd, int geom, int rs, int lim, const GridDev& G, const MarchPlan& P, const PerfectGasMixed& eos,
                          const StepArraysF& A, hipStream_t s);
int march_occupancy_fx(int nd, int geom, int rs, int lim, bool fp32);
// the fused step's pressure fix (pressure_fix_fused.hip): after k_step_march, over the march
// kernel's workgroups (plan P); fix_partials receives two doubles per workgroup
void launch_pressure_fix_fused(int nd, int geom, int lim, const GridDev& G, const PerfectGas& eos, const StepArrays& A, const MarchPlan& P,
                               double* fix_partials, hipStream_t s);
GhostShell make_ghost_shell(const GridDev& G, int nd);
int64_t ghost_blocks(const GhostShell& B);
void launch_ghosts_rate(int nd, int geom, const GridDev& G, const PerfectGas& eos, const BcSpec& bc, const GhostShell& B, const StepArrays& A,
                        hipStream_t s);
void launch_ghosts_rate(int nd, int geom, const GridDev& G, const PerfectGasMixed& eos, const BcSpec& bc, const GhostShell& B,
                        const StepArraysF& A, hipStream_t s);
// max CFL rate of a whole FP32-storage conservative state (FP64 arithmetic) into out[0]
void launch_state_rate_f32(int nd, int geom, const GridDev& G, const PerfectGasMixed& eos, const float* rho, const float* rhou, const float* E,
                           double* partials, double* out, hipStream_t s);
void launch_cast(const double* in, float* out, int64_t n, hipStream_t s);
void launch_cast(const float* in, double* out, int64_t n, hipStream_t s);
void launch_clock_start(DevClock* c, const double* inv, hipStream_t s);
// replace the scheduled dt of the next step (the t_end / max_iter exit rules are re-evaluated)
void launch_clock_set_dt(DevClock* c, double dt, hipStream_t s);
void launch_clock_set_log(DevClock* c, double* log, int cap, hipStream_t s);
void launch_clock_advance(int nd, DevClock* c, const double* partials, int64_t nblocks, const double* ghost_partials, int64_t nghost,
                          const double* fix_partials, int64_t nfix,
                          hipStream_t s);

}  // namespace hclpp
