// rt_kernels.h — host-visible launchers of the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "rt_params.h"

namespace rt {

// Launch shape chosen by the autotuner: threads per block (multiple of 64,
// at most 256), the occupancy the kernel is compiled for (launch bounds of 1
// or 2 waves per SIMD; both compile to 2 waves per SIMD for the plain FP64
// kernels) and, for the work-queue kernels, the queue granularity and how
// many lanes of a wave must be idle before the lane-refill kernel refills
// them.
struct LaunchShape {
    uint32_t block = 64;
    uint32_t waves = 2;
    uint32_t refill_min = 16;
    uint32_t queue_kind = 0;   // batches: 0 tile queue, 1 lane-refill queue, 2 per-request kernels
};

// Defaults, the fastest shapes measured on MI355X:
//  - per-frame kernels: one-wave workgroups (the hardware refills a wave slot
//    as soon as its 8x8 tile is done);
//  - batches of a few large frames (4 x 2048^2 Kerr): one per-frame kernel per
//    request on concurrent streams (queue_kind 2), 58 ms against 69.5 ms for
//    the shared tile queue;
//  - the serving loop's batches of many uneven requests: the shared tile queue
//    at two waves per SIMD (512 uneven requests: 656 ms against 734 ms for
//    per-request kernels);
//  - the superposition kernel (2..16 sources, 512^2): two waves per SIMD, a
//    refill once the whole wave is idle.
constexpr LaunchShape kPlainShape{64, 1, 16, 0};
constexpr LaunchShape kQueueShape{64, 1, 16, 2};
constexpr LaunchShape kServerShape{64, 2, 16, 0};
constexpr LaunchShape kMultiShape{64, 2, 64, 0};

// Camera frame (single Kerr): one lane per pixel, lanes of a wave on an 8x8 tile.
hipError_t launch_trace_camera(const TraceParams& P, const OutPtrs& O, const LaunchShape& L, hipStream_t stream);
// The same frame as a persistent kernel of `grid` one-wave workgroups over its
// own tile queue (*next zero at launch; grid is cut to the number of tiles).
// limit/active as for launch_trace_multi.
hipError_t launch_trace_camera_queue(const TraceParams& P, const OutPtrs& O, unsigned long long* next, uint32_t grid,
                                     hipStream_t stream, const uint32_t* limit = nullptr,
                                     unsigned long long* active = nullptr);
// Block sizes the kernels accept: one to four waves.
inline bool shape_ok(const LaunchShape& L) { return L.block != 0 && L.block <= 256 && (L.block & 63) == 0; }

// Workgroups of the superposition kernel resident on one compute unit (rt_multi.hip).
hipError_t multi_blocks_per_cu(int* blocks);

// Wave slots of the device for the serving loop's kernels: one-wave
// workgroups of the persistent camera and superposition kernels that can be
// resident at once (compute units x workgroups per compute unit).
hipError_t wave_slots(uint32_t* slots);
// RT_METHOD_RENDER_RK4 camera frame, single Kerr (rt_render.hip).
hipError_t launch_trace_render(const TraceParams& P, const OutPtrs& O, const LaunchShape& L, hipStream_t stream);
// Explicit initial states (O.in_x / O.in_p), single Kerr.
hipError_t launch_trace_rays(const TraceParams& P, const OutPtrs& O, const LaunchShape& L, hipStream_t stream);
// Superposition (camera or explicit rays per P.mode), table T in device memory.
// Lanes per ray of the superposition kernel (2, 4 or 8; one or two sources per lane).
uint32_t multi_ray_lanes(uint32_t n_src);
// max_grid > 0: at most that many workgroups (the serving loop's grant of
// wave slots); 0: fill the device once. limit/active (serving loop): waves
// leave when more than *limit of them run (*active counts them, zero at launch).
hipError_t launch_trace_multi(const TraceParams& P, const OutPtrs& O, const MultiDev* T, unsigned long long* next,
                              const LaunchShape& L, hipStream_t stream, uint32_t max_grid = 0,
                              const uint32_t* limit = nullptr, unsigned long long* active = nullptr);
// Persistent work-queue kernel over many single-Kerr requests (tiles or rays
// per L.queue_kind 0 or 1; Q.req_offset counts tiles or rays accordingly).
// grid 0: computed from the occupancy (not allowed inside a stream capture).
hipError_t launch_trace_queue(const QueueDev& Q, const LaunchShape& L, bool mixed, bool carter, uint32_t grid,
                              hipStream_t stream);
// Grid that fills the device at the kernel's occupancy.
hipError_t queue_grid(const LaunchShape& L, bool mixed, bool carter, uint32_t* grid);
// Zero n 64-bit counters with a kernel (no memset node in captured graphs).
hipError_t launch_zero_u64(unsigned long long* p, uint32_t n, hipStream_t stream);
// Byte ranges copied by one k_copy_segments launch; sources and destinations
// are device memory or mapped pinned host memory, 4-byte aligned, sizes a
// multiple of 4. The serving loop uploads a superposition's source table with
// it, in the request's stream, instead of an asynchronous copy: the copy
// engine's queue is shared by every stream, so a copy issued behind a long
// kernel would hold back the copies of requests issued after it.
constexpr int kMaxCopySegs = 8;
struct CopySeg {
    const void* src;
    void* dst;
    uint64_t bytes;
};
struct CopySegs {
    CopySeg seg[kMaxCopySegs];
    uint32_t n;
};
// At most max_blocks workgroups of 256 threads.
hipError_t launch_copy_segments(const CopySegs& S, hipStream_t stream, uint32_t max_blocks);
// Pack n results into 32-byte pixel records (rt_pixel layout) at out.
hipError_t launch_pack_pixels(const OutPtrs& O, uint64_t n, void* out, hipStream_t stream);

}  // namespace rt
