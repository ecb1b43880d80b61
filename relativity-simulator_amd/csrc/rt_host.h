// rt_host.h — host-side helpers shared by the context, batch and server code.
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/rt_tracer.h"
#include "rt_kernels.h"
#include "rt_params.h"

struct rt_context;

namespace rt {

rt_status fail(rt_status st, const std::string& msg);
rt_status hip_fail(hipError_t e, const char* what);

#define RT_HIP(call)                                          \
    do {                                                      \
        hipError_t e_ = (call);                               \
        if (e_ != hipSuccess) return ::rt::hip_fail(e_, #call); \
    } while (0)

bool valid_request(const rt_metric* mt, const rt_integrator* ig, const rt_termination* tm, std::string& why);
bool valid_camera(const rt_camera* c, std::string& why);
void fill_params(const rt_metric* mt, const rt_integrator* ig, const rt_termination* tm, TraceParams& P);
// Camera constants; for a superposition also the source table and the
// observer frame (T must be non-null then).
void fill_camera(const rt_camera* c, TraceParams& P, MultiDev* T, const rt_tile* win = nullptr);
// Source table of a superposition (sources only, no observer frame).
void fill_sources(const rt_metric* mt, const rt_termination* tm, MultiDev& T);
double kerr_isco_host(double m, double a_spin);

// Bytes of device memory one traced ray needs (results + optional inputs).
constexpr size_t kResultBytesPerRay = 4 + 32 + 32 + 24 + 8 + 8 + 4 + 4;
constexpr size_t kInputBytesPerRay = 64;

// Grow-only device allocation.
struct DeviceArena {
    void* base = nullptr;
    size_t bytes = 0;
    ~DeviceArena();
    rt_status ensure(size_t need);
    void release();
};

// SoA result (and optional input) arrays of n rays carved from a cursor
// (256-byte aligned sub-arrays); advances the cursor.
OutPtrs carve_results(char*& cursor, size_t n, bool with_inputs);
size_t results_bytes(size_t n, bool with_inputs);

rt_status copy_results_to_host(const OutPtrs& d, size_t n, const rt_results* out, hipStream_t s);

// Device address of this device's allocation-free scratch array, taken once
// per device with the device current (rt_noalloc.hip).
rt_status scratch_setup(int device);

struct GraphCache;
// Drop the captured batch graph of a context (on destroy).
void forget_graphs(const rt_context* ctx);

}  // namespace rt

// Device context: stream, events, result arena, statistics, launch shape.
struct rt_context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    rt::DeviceArena arena;       // results of the last trace
    rt::DeviceArena aux;         // superposition table, queue tables, counters
    rt::DeviceArena pix;         // packed pixel records (rt_fetch_pixels)
    rt::OutPtrs last{};          // device results of the last resident trace
    size_t n_last = 0;
    rt_stats stats{};
    rt::LaunchShape shape;        // plain kernels
    rt::LaunchShape qshape;       // batches (work-queue kernels, or per-request kernels)
    // Side streams and fork/join events: a batch with queue_kind 2 spreads its
    // per-request kernels over the context stream and these (graph branches).
    static constexpr int kSide = 3;
    hipStream_t side[kSide] = {nullptr, nullptr, nullptr};
    hipEvent_t fork = nullptr;
    hipEvent_t join[kSide] = {nullptr, nullptr, nullptr};
    bool tuned = false;
    std::mutex mu;
};
