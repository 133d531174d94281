// reduce.hpp -- deterministic wavefront + workgroup reductions (fixed combination order).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace hclpp {

enum ReduceOp { kMax = 0, kMin = 1, kSum = 2 };

template <int OP>
__device__ __forceinline__ double reduce_identity()
{
    if constexpr (OP == kMax) return -INFINITY;
    if constexpr (OP == kMin) return INFINITY;
    return 0.0;
}

// kMax follows Kokkos::max (a < b ? b : a); kMin propagates NaN so an invalid
// cell is never hidden from the internal-energy check.
template <int OP>
__device__ __forceinline__ double reduce_combine(double a, double b)
{
    if constexpr (OP == kMax) return a < b ? b : a;
    if constexpr (OP == kMin) return (isnan(a) || isnan(b)) ? NAN : (b < a ? b : a);
    return a + b;
}

template <int OP>
__device__ __forceinline__ double wave_reduce(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = reduce_combine<OP>(v, __shfl_xor(v, o, 64));
    return v;
}

// Reduce over a workgroup of `nthreads` (multiple of 64, <= 1024).  Every thread
// returns the workgroup result.  `lds` must hold nthreads/64 doubles.
template <int OP>
__device__ __forceinline__ double block_reduce(double v, double* lds, int tid, int nthreads)
{
    v = wave_reduce<OP>(v);
    int const lane = tid & 63, wave = tid >> 6, nw = nthreads >> 6;
    __syncthreads();
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    double r = reduce_identity<OP>();
    for (int w = 0; w < nw; ++w) r = reduce_combine<OP>(r, lds[w]);
    return r;
}

}  // namespace hclpp
