// rt_kernels.h — host-visible launchers of the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "rt_params.h"

namespace rt {

// Launch shape chosen by the autotuner: threads per block (multiple of 64,
// at most 256) and the occupancy the kernel is compiled for (1 or 2 waves per
// SIMD; 2 caps registers at 256 and spills in the once-per-ray event code).
struct LaunchShape {
    uint32_t block = 256;
    uint32_t waves = 1;
};

// Camera frame: one lane per pixel, lanes of a wave on an 8x8 tile.
hipError_t launch_trace_camera(const TraceParams& P, const OutPtrs& O, const LaunchShape& L, hipStream_t stream);
// Explicit initial states (O.in_x / O.in_p).
hipError_t launch_trace_rays(const TraceParams& P, const OutPtrs& O, const LaunchShape& L, hipStream_t stream);

}  // namespace rt
