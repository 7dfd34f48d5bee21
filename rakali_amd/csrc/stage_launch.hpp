// stage_launch.hpp -- launchers of the stage kernels, one translation unit
// per set of source terms (stage_feat*.hip), so that the instantiations
// compile in parallel.  Not part of the ABI.
#pragma once

#include <hip/hip_runtime.h>

#include "swe_kernels.hpp"

namespace rk {

// Stage `stage` (1 or 2) of the fused SSP-RK2 step on one grid, with the
// FEAT source terms (FEAT_FORCING | FEAT_SPONGE).
template <int FEAT>
void launch_fused_stage(int stage, hipStream_t stream, const GridDesc& d, int ntiles, int parity, int mode,
                        double tlim, bool hllc, bool nf);

// One stage of a generic SSP scheme on one grid.
template <int FEAT>
void launch_generic_stage(hipStream_t stream, const GridDesc& d, const GenStage& gs, int ntiles, int parity,
                          int mode, double tlim, bool hllc, bool nf);

}  // namespace rk
