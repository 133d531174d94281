// launch.hpp -- host-side declarations shared by the kernel translation units and the C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "grid.hpp"

namespace hclpp {

// Cell range [b, e) per dim, as produced by mesh/range.cpp:120-135.
struct Range3 {
    int b[3];
    int e[3];
    __host__ __device__ int64_t count() const
    {
        return (int64_t)(e[0] - b[0]) * (int64_t)(e[1] - b[1]) * (int64_t)(e[2] - b[2]);
    }
};

// Block of 64 x 4 threads over (x, y); grid z walks k.  x is the fastest
// (contiguous) index, so a wavefront reads 64 consecutive doubles.
constexpr int kBX = 64;
constexpr int kBY = 4;

inline dim3 range_grid(const Range3& r)
{
    int const nx = r.e[0] - r.b[0], ny = r.e[1] - r.b[1], nz = r.e[2] - r.b[2];
    return dim3((unsigned)((nx + kBX - 1) / kBX), (unsigned)((ny + kBY - 1) / kBY), (unsigned)nz);
}

__device__ __forceinline__ bool range_cell(const Range3& r, int& i, int& j, int& k)
{
    i = r.b[0] + (int)(blockIdx.x * kBX + threadIdx.x);
    j = r.b[1] + (int)(blockIdx.y * kBY + threadIdx.y);
    k = r.b[2] + (int)blockIdx.z;
    return i < r.e[0] && j < r.e[1];
}

// Boundary kinds per face (boundary_factory.hpp:31-47)
enum BcKind { kPeriodic = 0, kNullGradient = 1, kReflexive = 2 };

struct BcSpec {
    int kind[6];  // per face 2*dim + side
};

// Source index along one dim of a ghost-filled array: the index of the
// interior cell whose value a ghost at `idx` receives (periodic wrap,
// NullGradient clamp, Reflexive mirror with `flip` toggled for the normal
// momentum); interior indices map to themselves.  ng ghosts, n interior
// cells, ext = n + 2 ng.
__host__ __device__ __forceinline__ int face_map(int kind, int idx, int ng, int n, int ext, bool& flip)
{
    if (idx < ng) {  // left ghosts
        if (kind == kPeriodic) return idx + n;
        if (kind == kNullGradient) return ng;
        flip = !flip;
        return (2 * ng) - 1 - idx;
    }
    if (idx >= ng + n) {  // right ghosts
        if (kind == kPeriodic) return idx - n;
        if (kind == kNullGradient) return ext - ng - 1;
        flip = !flip;
        return (2 * (ext - ng)) - 1 - idx;
    }
    return idx;
}

// Optional gate read by a kernel before it does any work: a kernel enqueued behind the
// device clock skips itself once the clock has finished or failed (the step it belongs to
// is then a no-op too).  A null pointer means "always run".
struct DevFlag {
    int done;
    int status;
    __device__ __forceinline__ bool go() const { return !done && !status; }
};

// ---- launchers (ref_kernels.hip): reference-shaped kernels, one per reference kernel label
void launch_prim_to_cons(int nd, const GridDev& G, const Range3& r, const PerfectGas& eos, const double* rho, const double* u,
                         const double* P, double* rhou, double* E, hipStream_t s);
void launch_cons_to_prim(int nd, const GridDev& G, const Range3& r, const PerfectGas& eos, const double* rho, const double* rhou,
                         const double* E, double* u, double* P, hipStream_t s);
void launch_face_reconstruction(int nd, int lim, const GridDev& G, const Range3& r, const double* var, double* var_rec, hipStream_t s);
void launch_face_prim_to_cons(int nd, const GridDev& G, const Range3& r, const PerfectGas& eos, const double* rho_rec,
                              const double* u_rec, const double* P_rec, double* rhou_rec, double* E_rec, hipStream_t s);
void launch_hancock(int nd, int geom, const GridDev& G, const Range3& r, const PerfectGas& eos, double dt, const double* u_rec,
                    const double* P_rec, double* rho_rec, double* rhou_rec, double* E_rec, double* fx_rec, const double* fx_rec_old,
                    int nfx, hipStream_t s);
void launch_godunov(int nd, int geom, int rs, const GridDev& G, const Range3& r, const PerfectGas& eos, double dt, const double* rho,
                    const double* rhou, const double* E, const double* fx, const double* rho_rec, const double* rhou_rec,
                    const double* E_rec, const double* fx_rec, double* rho_new, double* rhou_new, double* E_new, double* fx_new, int nfx,
                    hipStream_t s);
// Reductions write one partial per block into `partials` (capacity from reduce_blocks) then fold
// them in a fixed order into out[0] (device).  op: 0 max, 1 min, 2 sum.
int64_t reduce_blocks(const Range3& r);
void launch_time_step_rate(int nd, int geom, const GridDev& G, const Range3& r, const PerfectGas& eos, const double* rho, const double* u,
                           const double* P, double* partials, double* out, hipStream_t s);
void launch_min_internal_energy(int nd, const GridDev& G, const Range3& r, const double* rho, const double* rhou, const double* E,
                                double* partials, double* out, hipStream_t s);
void launch_integrate(const GridDev& G, int nd, int geom, const Range3& r, const double* var, double* partials, double* out, hipStream_t s);
// analysis/temperature.hpp:18-26
void launch_temperature(const GridDev& G, const Range3& r, const PerfectGas& eos, const double* rho, const double* P, double* T, hipStream_t s);
// Fold n partials (max) into out[0] in a fixed order.
void launch_fold_max(const double* partials, int64_t n, double* out, hipStream_t s);
// launch_time_step_rate folding to the reciprocal of the maximum (time_step.hpp:81) in device memory
void launch_time_step_inverse(int nd, int geom, const GridDev& G, const Range3& r, const PerfectGas& eos, const double* rho, const double* u,
                              const double* P, double* partials, double* dt_out, hipStream_t s);
// hydro/pressure_fix.hpp:19-172 (E_new updated in place)
void launch_pressure_fix(int nd, int geom, const GridDev& G, const Range3& r, const PerfectGas& eos, double dt, double eps, const double* rho,
                         const double* rhou, const double* E, const double* rho_rec, const double* rhou_rec, const double* E_rec,
                         const double* rho_new, const double* rhou_new, double* E_new, hipStream_t s);
// gravity/gravity.cpp:59-128: radial gravity table gtab (nwg[0]) from rho; shells holds 2 (n0 + ng) doubles
template <class S>
void launch_internal_mass_gravity(int nd, int geom, const GridDev& G, double central_mass, const S* rho, double* shells, double* gtab,
                                  const DevFlag* run, hipStream_t s);
// Ghost fill on conservative arrays (boundary_distribute.hpp:42-76 semantics, single rank)
void launch_fill_ghosts(int nd, const GridDev& G, const BcSpec& bc, double* rho, double* rhou, double* E, double* fx, int nfx, hipStream_t s);

}  // namespace hclpp
