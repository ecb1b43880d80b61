// rt_noalloc.hip — rt_trace_tiles_noalloc: a batch of decoded tile tasks
// traced without any device allocation.
//
// The work-queue kernel of a tile batch reads per-task tables (TraceParams,
// output views, queue segment offsets) and updates counters and a queue
// head. rt_trace_tiles keeps those in the context's arena, which grows with
// hipMalloc. Here they live in `scratch`, a file-scope device array of fixed
// size; the caller brings the result arrays (device memory). Callers: code
// under HIP graph capture, serving loops that must not allocate, harnesses
// that forbid hipMalloc.
//
// Every GPU holds its own copy of `scratch` (at its own address) until the
// process ends. Host code never passes the array's name to a kernel: the
// device address of each GPU's copy is taken once per device, at context
// creation with the device current (scratch_setup), and kept in g_scratch.
// Calls on one device are serialised by a per-device mutex, because they
// share that device's copy; kernels launched by other means that overlap a
// call on the same device would race on it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>

#include "../../include/rt_tracer.h"
#include "rt_host.h"
#include "rt_kernels.h"
#include "rt_params.h"

namespace rt {
namespace {

constexpr size_t kScratchBytes = 64 * 1024;
constexpr int kMaxDevices = 64;
constexpr size_t kCounterWords = 16;  // [0..4] statistics, [8] queue head
constexpr size_t kCounterBytes = kCounterWords * 8;

constexpr size_t al256(size_t v) { return (v + 255) & ~size_t(255); }

// Scratch layout for n tasks: counters | TraceParams[n] | OutPtrs[n] | offsets[2n + 1].
constexpr size_t off_params() { return al256(kCounterBytes); }
constexpr size_t off_outs(size_t n) { return off_params() + al256(n * sizeof(TraceParams)); }
constexpr size_t off_offsets(size_t n) { return off_outs(n) + al256(n * sizeof(OutPtrs)); }
constexpr size_t scratch_need(size_t n) { return off_offsets(n) + al256((2 * n + 1) * sizeof(uint64_t)); }
static_assert(scratch_need(RT_NOALLOC_MAX_TASKS) <= kScratchBytes, "scratch holds RT_NOALLOC_MAX_TASKS tasks");

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_setup_mu;
void* g_scratch[kMaxDevices] = {};
std::mutex g_call_mu[kMaxDevices];

bool aligned(const void* p, size_t a) { return (reinterpret_cast<uintptr_t>(p) % a) == 0; }

}  // namespace

// Device address of this device's copy of scratch; once per device, with
// the device current (called by rt_context_create).
rt_status scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(RT_ERR_INVALID_ARG, "device index above 63");
    std::lock_guard<std::mutex> lk(g_setup_mu);
    if (g_scratch[device]) return RT_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_scratch[device] = p;
    return RT_OK;
}

}  // namespace rt

using namespace rt;

extern "C" rt_status rt_trace_tiles_noalloc(rt_context* ctx, const rt_tile_task* tasks, size_t n_tasks,
                                            const rt_results* dev_out, size_t capacity, size_t* n_rays) {
    const auto t0 = std::chrono::steady_clock::now();
    if (!ctx || !tasks || !dev_out) return fail(RT_ERR_INVALID_ARG, "null context, tasks or results");
    if (n_tasks == 0) return fail(RT_ERR_INVALID_ARG, "no tasks");
    if (n_tasks > RT_NOALLOC_MAX_TASKS)
        return fail(RT_ERR_INVALID_ARG, "rt_trace_tiles_noalloc: " + std::to_string(n_tasks) +
                                            " tasks exceed the scratch capacity of " +
                                            std::to_string(RT_NOALLOC_MAX_TASKS) + " tasks per call");
    const rt_results& o = *dev_out;
    if (!o.cls || !o.x || !o.p || !o.bl || !o.lambda || !o.g || !o.steps || !o.rejected)
        return fail(RT_ERR_INVALID_ARG, "rt_trace_tiles_noalloc: every result array must be given (device memory)");
    if (!aligned(o.cls, 4) || !aligned(o.steps, 4) || !aligned(o.rejected, 4) || !aligned(o.x, 8) ||
        !aligned(o.p, 8) || !aligned(o.bl, 8) || !aligned(o.lambda, 8) || !aligned(o.g, 8))
        return fail(RT_ERR_INVALID_ARG, "rt_trace_tiles_noalloc: result arrays must be naturally aligned");

    // host-side tables on the stack (no heap), copied into scratch in one piece
    alignas(256) unsigned char stage[kScratchBytes];
    TraceParams* P = reinterpret_cast<TraceParams*>(stage + off_params());
    OutPtrs* outs = reinterpret_cast<OutPtrs*>(stage + off_outs(n_tasks));
    uint64_t* offs = reinterpret_cast<uint64_t*>(stage + off_offsets(n_tasks));
    uint64_t hot[RT_NOALLOC_MAX_TASKS], cold[RT_NOALLOC_MAX_TASKS];
    size_t first_ray[RT_NOALLOC_MAX_TASKS];
    bool mixed = false, carter = false;
    size_t total = 0;
    for (size_t i = 0; i < n_tasks; ++i) {
        const rt_tile_task& t = tasks[i];
        std::string why;
        if (!valid_request(&t.metric, &t.integrator, &t.termination, why) || !valid_camera(&t.camera, why))
            return fail(RT_ERR_INVALID_ARG, "task " + std::to_string(i) + ": " + why);
        const rt_tile& w = t.tile;
        if (w.width == 0 || w.height == 0 || w.origin_x >= t.camera.width || w.origin_y >= t.camera.height ||
            w.width > t.camera.width - w.origin_x || w.height > t.camera.height - w.origin_y)
            return fail(RT_ERR_INVALID_ARG, "task " + std::to_string(i) + ": tile empty or not inside the frame");
        if (t.metric.n_sources >= 2 || t.integrator.method != RT_METHOD_DP45)
            return fail(RT_ERR_INVALID_ARG, "task " + std::to_string(i) +
                                                ": superpositions and RT_METHOD_RENDER_RK4 are not traced by this entry point");
        fill_params(&t.metric, &t.integrator, &t.termination, P[i]);
        fill_camera(&t.camera, P[i], nullptr, &w);
        const uint64_t n_tiles = static_cast<uint64_t>(P[i].tiles_x) * ((P[i].n_rays / P[i].cols + 7) / 8);
        hot[i] = static_cast<uint64_t>(P[i].hot_w) * P[i].hot_h;
        cold[i] = n_tiles - hot[i];
        first_ray[i] = total;
        total += P[i].n_rays;
        mixed = mixed || P[i].mixed;
        carter = carter || P[i].carter;
    }
    if (total > capacity)
        return fail(RT_ERR_INVALID_ARG, "rt_trace_tiles_noalloc: the tasks hold " + std::to_string(total) +
                                            " rays, the result arrays " + std::to_string(capacity));
    // queue segments: every task's hot 8x8 blocks first, then the other blocks
    offs[0] = 0;
    for (size_t i = 0; i < n_tasks; ++i) offs[1 + i] = offs[i] + hot[i];
    for (size_t i = 0; i < n_tasks; ++i) offs[1 + n_tasks + i] = offs[n_tasks + i] + cold[i];

    std::lock_guard<std::mutex> lk(ctx->mu);
    if (ctx->device < 0 || ctx->device >= kMaxDevices || !g_scratch[ctx->device])
        return fail(RT_ERR_DEVICE, "scratch of this device not set up");
    std::lock_guard<std::mutex> dev_lk(g_call_mu[ctx->device]);
    RT_HIP(hipSetDevice(ctx->device));
    char* base = static_cast<char*>(g_scratch[ctx->device]);
    unsigned long long* counters = reinterpret_cast<unsigned long long*>(base);
    // results task after task in the caller's arrays; leading dimension = capacity
    for (size_t i = 0; i < n_tasks; ++i) {
        OutPtrs v{};
        v.cls = o.cls + first_ray[i];
        v.x = o.x + first_ray[i];
        v.p = o.p + first_ray[i];
        v.bl = o.bl + first_ray[i];
        v.lambda = o.lambda + first_ray[i];
        v.g = o.g + first_ray[i];
        v.steps = o.steps + first_ray[i];
        v.rejected = o.rejected + first_ray[i];
        v.counters = counters;
        v.n = capacity;
        outs[i] = v;
    }
    QueueDev Q{};
    Q.reqs = reinterpret_cast<const TraceParams*>(base + off_params());
    Q.outs = reinterpret_cast<const OutPtrs*>(base + off_outs(n_tasks));
    Q.req_offset = reinterpret_cast<const uint64_t*>(base + off_offsets(n_tasks));
    Q.next = counters + 8;
    Q.counters = counters;
    Q.total = offs[2 * n_tasks];
    Q.n_req = static_cast<uint32_t>(n_tasks);
    Q.refill_min = kServerShape.refill_min;
    const size_t table_bytes = scratch_need(n_tasks) - off_params();
    hipStream_t s = ctx->stream;
    // the stage buffer stays untouched until the stream has been synchronised below
    RT_HIP(hipMemcpyAsync(base + off_params(), stage + off_params(), table_bytes, hipMemcpyHostToDevice, s));
    RT_HIP(launch_zero_u64(counters, kCounterWords, s));
    RT_HIP(hipEventRecord(ctx->ev0, s));
    RT_HIP(launch_trace_queue(Q, kServerShape, mixed, carter, 0, s));
    RT_HIP(hipEventRecord(ctx->ev1, s));
    RT_HIP(hipEventSynchronize(ctx->ev1));
    float ms = 0.0f;
    RT_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    unsigned long long cnt[5] = {0, 0, 0, 0, 0};
    RT_HIP(hipMemcpy(cnt, counters, sizeof(cnt), hipMemcpyDeviceToHost));
    rt_stats& st = ctx->stats;
    st.kernel_ms = ms;
    st.rays = total;
    st.accepted_steps = cnt[0];
    st.rejected_steps = cnt[1];
    st.rhs_evals = cnt[2];
    st.fp32_steps = cnt[3];
    st.block_size = kServerShape.block;
    st.grid_size = 0;
    st.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (n_rays) *n_rays = total;
    return RT_OK;
}
