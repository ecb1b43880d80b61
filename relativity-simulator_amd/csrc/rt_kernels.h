// rt_kernels.h — host-visible launchers of the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "rt_params.h"

namespace rt {

// Camera frame: one lane per pixel, lanes of a wave on an 8x8 tile.
hipError_t launch_trace_camera(const TraceParams& P, const OutPtrs& O, uint32_t block, hipStream_t stream);
// Explicit initial states (O.in_x / O.in_p).
hipError_t launch_trace_rays(const TraceParams& P, const OutPtrs& O, uint32_t block, hipStream_t stream);

}  // namespace rt
