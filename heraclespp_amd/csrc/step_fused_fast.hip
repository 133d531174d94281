// step_fused_fast.hip -- launchers of the fused step (k_step_march, march.hpp) with the opt-in
// fast FP64 arithmetic (PerfectGasFast, physics.hpp): FP64 state and algorithm, the reference's
// formulas without its rounding.  A translation unit of its own so that its instantiations
// compile in parallel with the others.
#include <hip/hip_runtime.h>

#include "march.hpp"

namespace hclpp {

void launch_fused_step(int nd, int geom, int rs, int lim, const GridDev& G, const MarchPlan& P, const PerfectGasFast& eos, const StepArrays& A,
                       hipStream_t s)
{
    march_visit(nd, geom, rs, lim, MarchLaunch<PerfectGasFast, 0>{G, P, eos, A, s});
}

int march_occupancy_fast(int nd, int geom, int rs, int lim)
{
    int n = 0;
    march_visit(nd, geom, rs, lim, MarchOccupancy<PerfectGasFast, 0>{&n});
    return n;
}

}  // namespace hclpp
