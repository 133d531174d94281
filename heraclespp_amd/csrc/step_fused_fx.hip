// step_fused_fx.hip -- launchers of the fused step (k_step_march, march.hpp) carrying one
// passive scalar: reconstructed and extrapolated with the state, upwinded with the mass flux
// and clamped to [0, 1] as the reference's predictor and Godunov scheme do
// (extrapolation_reconstruction.hpp:87-88, 105-108, 315-337; godunov_scheme.hpp:122, 263-301).
// A translation unit of its own so that its instantiations compile in parallel with the
// scalar-free ones.
#include <hip/hip_runtime.h>

#include "march.hpp"

namespace hclpp {

void launch_fused_step_fx(int nd, int geom, int rs, int lim, const GridDev& G, const MarchPlan& P, const PerfectGas& eos, const StepArrays& A,
                          hipStream_t s)
{
    march_visit(nd, geom, rs, lim, MarchLaunch<PerfectGas, 1>{G, P, eos, A, s});
}

void launch_fused_step_fx(int nd, int geom, int rs, int lim, const GridDev& G, const MarchPlan& P, const PerfectGasMixed& eos,
                          const StepArraysF& A, hipStream_t s)
{
    march_visit(nd, geom, rs, lim, MarchLaunch<PerfectGasMixed, 1>{G, P, eos, A, s});
}

int march_occupancy_fx(int nd, int geom, int rs, int lim, bool fp32)
{
    int n = 0;
    if (fp32) march_visit(nd, geom, rs, lim, MarchOccupancy<PerfectGasMixed, 1>{&n});
    else march_visit(nd, geom, rs, lim, MarchOccupancy<PerfectGas, 1>{&n});
    return n;
}

}  // namespace hclpp
