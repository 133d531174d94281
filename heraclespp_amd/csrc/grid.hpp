// grid.hpp -- grid description shared by all kernels.
//
// The reference stores ds (4D) and dv (3D) arrays; here they are evaluated
// on the fly from 1D tables (node coordinates, cell sizes, and host-computed
// trigonometric tables), with the same multiplication order as
// mesh/geometry.cpp:32-126, so the values are bit-identical and cost no HBM
// traffic.  Arrays are Kokkos::LayoutLeft over the ghost-inclusive extents:
// flat index = i + nx * (j + ny * k), component planes slowest.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "physics.hpp"

namespace hclpp {

enum GravityKind { kUniformGravity = 0, kPointMassGravity = 1 };

struct GridDev {
    int nwg[3];        // extents including ghosts (1 on inactive dims)
    int ngh[3];        // ghost layers per dim (0 on inactive dims)
    int n[3];          // interior cells per dim
    int64_t plane;     // nwg0 * nwg1 * nwg2
    const double* x[3];       // node coordinates, nwg[d] + 1 entries
    const double* dx[3];      // cell sizes, nwg[d]
    const double* xc[3];      // cell centers, nwg[d]
    const double* dcos;       // cos(x1[j]) - cos(x1[j+1]), nwg[1]
    const double* sin_node;   // sin(x1[j]), nwg[1] + 1
    const double* cot_mid;    // cos(m)/sin(m), m = (x1[j] + x1[j+1]) / 2, nwg[1]
    const double* sin_center; // sin(xc1[j]), nwg[1]
    const double* gpm;        // point-mass gravity per radial cell, nwg[0]
    double g[3];              // uniform gravity
    int gravity;              // GravityKind
};

__host__ __device__ __forceinline__ int64_t flat(const GridDev& G, int i, int j, int k)
{
    return (int64_t)i + (int64_t)G.nwg[0] * ((int64_t)j + (int64_t)G.nwg[1] * (int64_t)k);
}

// gravity/gravity.hpp:36-40 (uniform) and :63-70 (point mass)
__host__ __device__ __forceinline__ double gravity(const GridDev& G, int i, int dir)
{
    if (G.gravity == kUniformGravity) return G.g[dir];
    return dir == 0 ? G.gpm[i] : 0.;
}

// mesh/geometry.cpp:32-126: cell volume
template <int ND, int GEOM>
__host__ __device__ __forceinline__ double cell_volume(const GridDev& G, int i, int j, int k)
{
    if constexpr (GEOM == kCartesian) {
        return G.dx[0][i] * G.dx[1][j] * G.dx[2][k];
    } else {
        double const r = G.x[0][i];
        double const rp = G.x[0][i + 1];
        if constexpr (ND == 1) {
            return (4 * kPi * ((rp * rp * rp) - (r * r * r))) / 3;
        } else if constexpr (ND == 2) {
            return (2 * kPi / 3) * ((rp * rp * rp) - (r * r * r)) * G.dcos[j];
        } else {
            return (1. / 3) * ((rp * rp * rp) - (r * r * r)) * G.dcos[j] * G.dx[2][k];
        }
    }
}

// mesh/geometry.cpp:32-126: area of the face at the lower side of cell (i,j,k) normal to `dir`
template <int ND, int GEOM>
__host__ __device__ __forceinline__ double face_area(const GridDev& G, int dir, int i, int j, int k)
{
    if constexpr (GEOM == kCartesian) {
        if (dir == 0) return G.dx[1][j] * G.dx[2][k];
        if (dir == 1) return G.dx[0][i] * G.dx[2][k];
        return G.dx[0][i] * G.dx[1][j];
    } else {
        double const r = G.x[0][i];
        if constexpr (ND == 1) {
            return 4 * kPi * r * r;
        } else if constexpr (ND == 2) {
            if (dir == 0) return 2 * kPi * r * r * G.dcos[j];
            double const rp = G.x[0][i + 1];
            return (2 * kPi / 2) * ((rp * rp) - (r * r)) * G.sin_node[j];
        } else {
            if (dir == 0) return r * r * G.dcos[j] * G.dx[2][k];
            if (dir == 1) {
                double const rp = G.x[0][i + 1];
                return (1. / 2) * ((rp * rp) - (r * r)) * G.sin_node[j] * G.dx[2][k];
            }
            return r * G.dx[0][i] * G.dx[1][j];
        }
    }
}

// hydro/time_step.hpp:61-76: the cell's CFL rate sum_d (|u_d| + c) / dx_geom_d
template <int ND, int GEOM>
__device__ __forceinline__ double cfl_rate(const GridDev& G, int i, int j, int k, const double* u, double c)
{
    double dxg[3] = {G.dx[0][i], G.dx[1][j], G.dx[2][k]};
    if constexpr (GEOM == kSpherical) {
        if constexpr (ND == 2) dxg[1] *= G.xc[0][i];
        if constexpr (ND == 3) {
            dxg[1] *= G.xc[0][i];
            dxg[2] *= G.xc[0][i] * G.sin_center[j];
        }
    }
    double rate = 0;
#pragma unroll
    for (int d = 0; d < ND; ++d) rate += fdiv(fabs(u[d]) + c, dxg[d]);
    return rate;
}

}  // namespace hclpp
