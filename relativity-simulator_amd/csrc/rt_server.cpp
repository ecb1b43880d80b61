// rt_server.cpp — the serving loop (rt_server_*).
//
// A worker thread drives the serve core (rt_serve_core.hpp): it expires
// requests past their deadline, admits pending ones as device memory, pinned
// staging, wave slots and a lane (stream) free up, and launches each admitted
// request at once, next to whatever is running, as a persistent kernel whose
// workgroup count is the slots granted to it: single-Kerr requests as the
// camera kernel over the frame's own tile queue, superpositions as the
// superposition kernel. Every workgroup of a granted kernel is resident from
// the start. (A kernel with more workgroups than free slots would keep the
// hardware dispatching its workgroups as slots free up, and kernels launched
// on other streams would wait behind it: measured, a 3 ms request waited the
// whole 136 ms of a large superposition.)
// The tracing kernel writes the outputs the caller asked for straight into
// the request's block of device-mapped pinned staging, and one event is
// recorded after it; the worker polls the events and never blocks. (An
// asynchronous copy afterwards would go through the copy engine's one queue,
// where a copy waiting for a long kernel holds back every later request's
// copies; a copy kernel afterwards would need its own dispatch, and waits
// when the device is full.) rt_server_take copies staging into the caller's
// arrays on the caller's thread. Four lanes, one per
// hardware queue, each running one request at a time; the last takes
// single-Kerr requests only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/rt_tracer.h"
#include "rt_host.h"
#include "rt_kernels.h"
#include "rt_serve_core.hpp"

using namespace rt;

namespace {

constexpr size_t kCounterBytes = 256;      // 16 statistics counters; the queue head at [8]
constexpr uint32_t kLanes = 4;             // one stream per hardware queue
constexpr uint32_t kKerrOnlyLanes = 1;     // the last lane takes single-Kerr requests only
constexpr size_t kCtlBytes = 256;          // staging head: the request's wave limit (u32), read by its kernel
constexpr size_t kDefaultStagingMiB = 1024;

size_t al256(size_t v) { return (v + 255) & ~size_t(255); }

// The outputs, in the order of rt_results and of a staging block.
struct Field {
    uint32_t bit;
    size_t bytes_per_ray;
};
constexpr Field kFields[8] = {{RT_OUT_CLS, 4},     {RT_OUT_X, 32}, {RT_OUT_P, 32},     {RT_OUT_BL, 24},
                              {RT_OUT_LAMBDA, 8},  {RT_OUT_G, 8},  {RT_OUT_STEPS, 4}, {RT_OUT_REJECTED, 4}};

void set_field(OutPtrs& O, int f, char* p) {
    switch (f) {
        case 0: O.cls = reinterpret_cast<uint32_t*>(p); break;
        case 1: O.x = reinterpret_cast<double*>(p); break;
        case 2: O.p = reinterpret_cast<double*>(p); break;
        case 3: O.bl = reinterpret_cast<double*>(p); break;
        case 4: O.lambda = reinterpret_cast<double*>(p); break;
        case 5: O.g = reinterpret_cast<double*>(p); break;
        case 6: O.steps = reinterpret_cast<uint32_t*>(p); break;
        default: O.rejected = reinterpret_cast<uint32_t*>(p); break;
    }
}

void* out_field(const rt_results& o, int f) {
    switch (f) {
        case 0: return o.cls;
        case 1: return o.x;
        case 2: return o.p;
        case 3: return o.bl;
        case 4: return o.lambda;
        case 5: return o.g;
        case 6: return o.steps;
        default: return o.rejected;
    }
}

struct ReqData {
    TraceParams P;
    std::unique_ptr<MultiDev> T;  // superpositions
    size_t n = 0;
    uint32_t outputs = RT_OUT_ALL;
    size_t table_bytes = 0;  // upload table at the head of the staging block
};

struct Live {
    uint64_t id;
    hipEvent_t done;
    uint32_t* limit;  // host address of the request's wave limit
};

void store_limit(uint32_t* limit, uint32_t v) { __atomic_store_n(limit, v, __ATOMIC_RELEASE); }

}  // namespace

struct rt_server {
    int device = 0;
    void* pool = nullptr;
    size_t pool_bytes = 0;
    void* staging = nullptr;      // host address of the pinned staging pool
    void* staging_dev = nullptr;  // its device address
    size_t staging_bytes = 0;
    hipStream_t lanes[kLanes] = {};
    uint32_t slots = 0;  // the device's wave slots
    std::unique_ptr<serve::Core> core;
    std::mutex mu;  // guards data
    std::map<uint64_t, ReqData> data;
    std::thread worker;
    std::atomic<bool> stop{false};

    void run();
    rt_status launch(const serve::Admission& a, Live* out);
};

rt_status rt_server::launch(const serve::Admission& a, Live* out) {
    char* dbase = static_cast<char*>(pool) + a.dev_off;
    char* hbase = static_cast<char*>(staging) + a.host_off;
    TraceParams P;
    size_t n = 0, table = 0;
    uint32_t outputs = 0;
    bool multi = false;
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = data.find(a.id);
        if (it == data.end()) return fail(RT_ERR_NOT_FOUND, "server: request data missing");
        const ReqData& d = it->second;
        P = d.P;
        n = d.n;
        outputs = d.outputs;
        table = d.table_bytes;
        multi = d.T != nullptr;
        if (multi) std::memcpy(hbase + kCtlBytes, d.T.get(), sizeof(MultiDev));  // pinned upload source
    }
    auto* counters = reinterpret_cast<unsigned long long*>(dbase);  // [8] tile/ray queue, [9] running waves
    auto* Td = reinterpret_cast<MultiDev*>(dbase + kCounterBytes);
    char* cur = dbase + kCounterBytes + table;
    OutPtrs O = carve_results(cur, n, false);
    O.counters = counters;
    hipStream_t s = lanes[a.lane];
    char* hdev = static_cast<char*>(staging_dev) + a.host_off;  // the staging block as the device sees it
    // the outputs asked for are written by the tracing kernel straight into
    // staging (device-mapped pinned memory); the others go to the device block
    char* h = hdev + kCtlBytes + table;
    for (int f = 0; f < 8; ++f) {
        if (!(outputs & kFields[f].bit)) continue;
        set_field(O, f, h);
        h += al256(n * kFields[f].bytes_per_ray);
    }
    auto* limit_host = reinterpret_cast<uint32_t*>(hbase);
    const auto* limit_dev = reinterpret_cast<const uint32_t*>(hdev);
    store_limit(limit_host, std::max<uint32_t>(a.slots, 1));
    if (multi) {
        CopySegs up{};
        up.seg[0] = CopySeg{hdev + kCtlBytes, Td, sizeof(MultiDev)};
        up.n = 1;
        RT_HIP(launch_copy_segments(up, s, 1));
    }
    RT_HIP(launch_zero_u64(counters, 16, s));
    if (multi)
        RT_HIP(launch_trace_multi(P, O, Td, counters + 8, kMultiShape, s, a.slots, limit_dev, counters + 9));
    else if (P.method == RT_METHOD_RENDER_RK4)
        RT_HIP(launch_trace_render(P, O, kPlainShape, s));
    else
        RT_HIP(launch_trace_camera_queue(P, O, counters + 8, a.slots, s, limit_dev, counters + 9));
    hipEvent_t ev = nullptr;
    RT_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    const hipError_t e = hipEventRecord(ev, s);
    if (e != hipSuccess) {
        (void)hipEventDestroy(ev);
        return hip_fail(e, "hipEventRecord(server request)");
    }
    out->id = a.id;
    out->done = ev;
    out->limit = limit_host;
    return RT_OK;
}

void rt_server::run() {
    (void)hipSetDevice(device);
    std::vector<Live> live;
    std::vector<serve::Admission> adms;
    std::vector<serve::Resize> resizes;
    while (!stop.load(std::memory_order_acquire)) {
        core->expire();
        adms.clear();
        resizes.clear();
        core->admit(&adms, &resizes);
        // lower the limits of running requests first: their waves leave
        // between tiles and make room for the new kernels
        for (const serve::Resize& r : resizes)
            for (const Live& L : live)
                if (L.id == r.id) store_limit(L.limit, r.slots);
        bool progressed = !adms.empty() || !resizes.empty();
        for (const serve::Admission& a : adms) {
            Live L{};
            const rt_status st = launch(a, &L);
            if (st == RT_OK)
                live.push_back(L);
            else
                core->launch_failed(a.id, st);
        }
        for (size_t i = 0; i < live.size();) {
            const hipError_t q = hipEventQuery(live[i].done);
            if (q == hipErrorNotReady) {
                ++i;
                continue;
            }
            core->finished(live[i].id, q == hipSuccess ? RT_OK : RT_ERR_DEVICE);
            (void)hipEventDestroy(live[i].done);
            live[i] = live.back();
            live.pop_back();
            progressed = true;
        }
        if (!progressed) core->wait_work(live.empty() ? 20000 : 100);
    }
    // drain: every request still on the device finishes before the pools go
    for (const Live& L : live) {
        const hipError_t e = hipEventSynchronize(L.done);
        core->finished(L.id, e == hipSuccess ? RT_OK : RT_ERR_DEVICE);
        (void)hipEventDestroy(L.done);
    }
}

extern "C" {

rt_status rt_server_create(int device, const rt_server_config* cfg, rt_server** out) {
    if (!out || !cfg) return fail(RT_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (cfg->hbm_budget_bytes < (1u << 20) || cfg->max_inflight == 0)
        return fail(RT_ERR_INVALID_ARG, "server: hbm_budget_bytes >= 1 MiB and max_inflight > 0");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(RT_ERR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= n) return fail(RT_ERR_INVALID_ARG, "device index out of range");
    std::unique_ptr<rt_server> s(new (std::nothrow) rt_server());
    if (!s) return fail(RT_ERR_OOM, "server allocation");
    s->device = device;
    RT_HIP(hipSetDevice(device));
    s->pool_bytes = al256(cfg->hbm_budget_bytes);
    const size_t staging_mib = cfg->staging_mib ? cfg->staging_mib : kDefaultStagingMiB;
    s->staging_bytes = std::min<size_t>(staging_mib << 20, s->pool_bytes);
    hipError_t e = hipMalloc(&s->pool, s->pool_bytes);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(server pool)");
    e = hipHostMalloc(&s->staging, s->staging_bytes, hipHostMallocMapped);
    if (e == hipSuccess) e = hipHostGetDevicePointer(&s->staging_dev, s->staging, 0);
    if (e != hipSuccess) {
        if (s->staging) (void)hipHostFree(s->staging);
        (void)hipFree(s->pool);
        return hip_fail(e, "hipHostMalloc(server staging)");
    }
    for (auto& st : s->lanes) {
        e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (e != hipSuccess) break;
    }
    if (e != hipSuccess) {
        for (auto& st : s->lanes)
            if (st) (void)hipStreamDestroy(st);
        (void)hipHostFree(s->staging);
        (void)hipFree(s->pool);
        return hip_fail(e, "server streams");
    }
    e = wave_slots(&s->slots);
    if (e != hipSuccess) {
        for (auto& st : s->lanes) (void)hipStreamDestroy(st);
        (void)hipHostFree(s->staging);
        (void)hipFree(s->pool);
        return hip_fail(e, "server: wave slots");
    }
    serve::Config cc;
    cc.dev_bytes = s->pool_bytes;
    cc.host_bytes = s->staging_bytes;
    cc.max_inflight = cfg->max_inflight;
    cc.lanes = kLanes;
    cc.kerr_only_lanes = kKerrOnlyLanes;
    cc.slots = s->slots;
    s->core.reset(new serve::Core(cc, serve::steady_now_us));
    rt_server* raw = s.release();
    raw->worker = std::thread([raw] { raw->run(); });
    *out = raw;
    return RT_OK;
}

rt_status rt_server_destroy(rt_server* s) {
    if (!s) return RT_OK;
    s->core->shutdown();
    s->stop.store(true, std::memory_order_release);
    s->core->notify_worker();
    if (s->worker.joinable()) s->worker.join();
    (void)hipSetDevice(s->device);
    for (hipStream_t st : s->lanes)
        if (st) (void)hipStreamDestroy(st);
    if (s->staging) (void)hipHostFree(s->staging);
    if (s->pool) (void)hipFree(s->pool);
    delete s;
    return RT_OK;
}

rt_status rt_server_submit_ex(rt_server* s, const rt_request* req, const rt_submit_options* opt, uint64_t* id) {
    if (!s || !req || !id) return fail(RT_ERR_INVALID_ARG, "null argument");
    const uint32_t outputs = opt ? opt->outputs : RT_OUT_ALL;
    if (outputs == 0 || (outputs & ~RT_OUT_ALL) != 0) return fail(RT_ERR_INVALID_ARG, "outputs: a non-empty subset of RT_OUT_ALL");
    if (opt && opt->deadline_ms < 0) return fail(RT_ERR_INVALID_ARG, "deadline_ms must be >= 0");
    std::string why;
    if (!valid_request(&req->metric, &req->integrator, &req->termination, why) || !valid_camera(&req->camera, why))
        return fail(RT_ERR_INVALID_ARG, why);
    ReqData d;
    fill_params(&req->metric, &req->integrator, &req->termination, d.P);
    if (d.P.n_src >= 2) {
        d.T.reset(new (std::nothrow) MultiDev());
        if (!d.T) return fail(RT_ERR_OOM, "superposition table");
        fill_sources(&req->metric, &req->termination, *d.T);
        fill_camera(&req->camera, d.P, d.T.get());
        d.table_bytes = al256(sizeof(MultiDev));
    } else {
        fill_camera(&req->camera, d.P, nullptr);
    }
    d.n = d.P.n_rays;
    if (d.n == 0) return fail(RT_ERR_INVALID_ARG, "no rays to trace");
    d.outputs = outputs;
    serve::Job job;
    job.multi = d.T != nullptr;
    job.dev_bytes = kCounterBytes + d.table_bytes + results_bytes(d.n, false);
    job.host_bytes = kCtlBytes + d.table_bytes;
    for (const Field& f : kFields)
        if (outputs & f.bit) job.host_bytes += al256(d.n * f.bytes_per_ray);
    // wave slots it can use: one per 8x8 tile, or one per 64 lanes of G-lane groups
    const uint64_t tiles = static_cast<uint64_t>(d.P.tiles_x) * ((d.n / d.P.cols + 7) / 8);
    const uint64_t need = job.multi ? (tiles * 64 * multi_ray_lanes(d.P.n_src) + 63) / 64 : tiles;
    job.slots_need = static_cast<uint32_t>(std::min<uint64_t>(need, 0xffffffffu));
    if (opt && opt->deadline_ms > 0) job.deadline_us = serve::steady_now_us() + opt->deadline_ms * 1000;
    std::lock_guard<std::mutex> lk(s->mu);
    uint64_t rid = 0;
    const rt_status st = s->core->submit(job, &rid);
    if (st == RT_ERR_BUSY)
        return fail(RT_ERR_BUSY, "request footprint (" + std::to_string(job.dev_bytes) + " B on the device, " +
                                     std::to_string(job.host_bytes) + " B of staging) exceeds the server's budget");
    if (st != RT_OK) return fail(st, "server is shutting down");
    s->data.emplace(rid, std::move(d));
    *id = rid;
    return RT_OK;
}

rt_status rt_server_submit(rt_server* s, const rt_request* req, uint64_t* id) {
    return rt_server_submit_ex(s, req, nullptr, id);
}

rt_status rt_server_poll(rt_server* s, uint64_t id, int* done) {
    if (!s || !done) return fail(RT_ERR_INVALID_ARG, "null argument");
    const rt_status st = s->core->poll(id, done);
    if (st != RT_OK) return fail(st, "unknown request id");
    return RT_OK;
}

static rt_status status_message(rt_status st) {
    switch (st) {
        case RT_OK: return RT_OK;
        case RT_ERR_TIMEOUT: return fail(st, "request not finished before the wait's time limit");
        case RT_ERR_CANCELLED: return fail(st, "request cancelled");
        case RT_ERR_EXPIRED: return fail(st, "request deadline passed before it finished");
        case RT_ERR_NOT_FOUND: return fail(st, "unknown request id");
        case RT_ERR_BUSY: return fail(st, "request not finished");
        default: return fail(st, "request failed on the device");
    }
}

rt_status rt_server_wait(rt_server* s, uint64_t id, int64_t timeout_ms) {
    if (!s) return fail(RT_ERR_INVALID_ARG, "null server");
    return status_message(s->core->wait(id, timeout_ms));
}

rt_status rt_server_take(rt_server* s, uint64_t id, rt_results* out) {
    if (!s || !out) return fail(RT_ERR_INVALID_ARG, "null argument");
    size_t off = 0;
    const rt_status st = s->core->take_begin(id, &off);
    if (st != RT_OK) {
        if (st != RT_ERR_BUSY) {
            std::lock_guard<std::mutex> lk(s->mu);
            s->data.erase(id);
        }
        return status_message(st);
    }
    size_t n = 0, table = 0;
    uint32_t outputs = 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->data.find(id);
        if (it != s->data.end()) {
            n = it->second.n;
            table = it->second.table_bytes;
            outputs = it->second.outputs;
        }
    }
    for (int f = 0; f < 8; ++f)
        if (out_field(*out, f) && !(outputs & kFields[f].bit)) {
            s->core->take_end(id, /*give_back=*/true);
            return fail(RT_ERR_INVALID_ARG, "take asks for an output that was not requested at submit");
        }
    const char* h = static_cast<const char*>(s->staging) + off + kCtlBytes + table;
    for (int f = 0; f < 8; ++f) {
        if (!(outputs & kFields[f].bit)) continue;
        const size_t bytes = n * kFields[f].bytes_per_ray;
        if (void* dst = out_field(*out, f)) std::memcpy(dst, h, bytes);
        h += al256(bytes);
    }
    s->core->take_end(id);
    std::lock_guard<std::mutex> lk(s->mu);
    s->data.erase(id);
    return RT_OK;
}

rt_status rt_server_cancel(rt_server* s, uint64_t id) {
    if (!s) return fail(RT_ERR_INVALID_ARG, "null server");
    const rt_status st = s->core->cancel(id);
    if (st != RT_OK) return fail(st, "unknown request id");
    return RT_OK;
}

}  // extern "C"
