// rt_batch.cpp — batched frames (rt_trace_batch, rt_trace_tiles), launch
// shapes and the autotuner (rt_*_launch_shape, rt_autotune).
//
// A batch is laid out as one work queue: every single-Kerr request becomes a
// contiguous range of queue indices (8x8 tiles, or the rays of those tiles in
// tile order) served by one persistent kernel, k_trace_tiles (a wave takes a
// whole tile) or k_trace_queue (idle lanes refill ray by ray); superposition
// requests run as their own
// k_trace_multi kernels after it. The launch sequence of a batch is captured
// into a hipGraph and replayed while the batch layout (pointers, sizes,
// by-value kernel arguments, launch shape) stays the same.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/rt_tracer.h"
#include "rt_host.h"
#include "rt_kernels.h"

namespace rt {

namespace {

constexpr size_t kCounterBytes = 256;  // 16 statistics counters + queue head at [8]
constexpr uint32_t kQueueHead = 8;

size_t al256(size_t v) { return (v + 255) & ~size_t(255); }

struct BatchPlan {
    std::vector<uint32_t> q_req;         // request index of each queue entry
    std::vector<TraceParams> q_params;
    std::vector<uint64_t> q_offset;      // 2 n_q + 1 segment starts, in rays (lane refill) or tiles
    std::vector<uint32_t> m_req;         // request index of each superposition entry
    std::vector<TraceParams> m_params;
    std::vector<MultiDev> m_tables;
    std::vector<uint32_t> r_req;         // request index of each RT_METHOD_RENDER_RK4 entry
    std::vector<TraceParams> r_params;
    std::vector<size_t> n_rays;          // per request
    bool mixed = false;   // some request uses the mixed precision policy
    bool carter = false;  // some request asks for RT_TERM_CARTER_CAPTURE
    bool concat = false;  // results of all requests in one array, request after request
    size_t result_bytes = 0, table_bytes = 0;
};

// Is rectangle w non-empty and inside the camera's frame?
bool valid_window(const rt_camera& c, const rt_tile& w, std::string& why) {
    if (w.width == 0 || w.height == 0 || w.origin_x >= c.width || w.origin_y >= c.height ||
        w.width > c.width - w.origin_x || w.height > c.height - w.origin_y) {
        why = "region: empty, or not inside the camera's frame";
        return false;
    }
    return true;
}

// Validate and lay out a batch (no device work). tiles: queue entries are
// 8x8 tiles (k_trace_tiles) rather than rays (k_trace_queue). wins (may be
// null): request i traces rectangle wins[i] of its frame; concat: the results
// of all requests share one array, request after request.
rt_status plan_batch(const rt_request* reqs, size_t n, BatchPlan& plan, bool tiles, const rt_tile* wins = nullptr,
                     bool concat = false) {
    plan = BatchPlan{};
    plan.n_rays.resize(n);
    plan.concat = concat;
    std::vector<uint64_t> hot_len, cold_len;  // queue entries per request, hot and other tiles
    size_t total = 0;
    for (size_t i = 0; i < n; ++i) {
        const rt_request& r = reqs[i];
        const rt_tile* win = wins ? &wins[i] : nullptr;
        std::string why;
        if (!valid_request(&r.metric, &r.integrator, &r.termination, why) || !valid_camera(&r.camera, why) ||
            (win && !valid_window(r.camera, *win, why)))
            return fail(RT_ERR_INVALID_ARG, "request " + std::to_string(i) + ": " + why);
        TraceParams P;
        fill_params(&r.metric, &r.integrator, &r.termination, P);
        if (P.n_src >= 2) {
            MultiDev T;
            fill_sources(&r.metric, &r.termination, T);
            fill_camera(&r.camera, P, &T, win);
            plan.m_req.push_back(static_cast<uint32_t>(i));
            plan.m_params.push_back(P);
            plan.m_tables.push_back(T);
        } else if (P.method == RT_METHOD_RENDER_RK4) {
            // the renderer's tracer: its own per-request kernel
            fill_camera(&r.camera, P, nullptr, win);
            plan.r_req.push_back(static_cast<uint32_t>(i));
            plan.r_params.push_back(P);
        } else {
            fill_camera(&r.camera, P, nullptr, win);
            plan.q_req.push_back(static_cast<uint32_t>(i));
            plan.q_params.push_back(P);
            const uint64_t tiles_y = (P.n_rays / P.cols + 7) / 8;
            const uint64_t n_tiles = static_cast<uint64_t>(P.tiles_x) * tiles_y;
            const uint64_t n_hot = static_cast<uint64_t>(P.hot_w) * P.hot_h;
            const uint64_t unit = tiles ? 1 : 64;
            hot_len.push_back(n_hot * unit);
            cold_len.push_back((n_tiles - n_hot) * unit);
            plan.mixed = plan.mixed || P.mixed;
            plan.carter = plan.carter || P.carter;
        }
        plan.n_rays[i] = P.n_rays;
        total += P.n_rays;
        if (!concat) plan.result_bytes += results_bytes(P.n_rays, false);
    }
    if (concat) {
        if (total > (1ull << 31)) return fail(RT_ERR_INVALID_ARG, "more than 2^31 rays in one batch");
        plan.result_bytes = results_bytes(total, false);
    }
    // queue segments: every request's hot tiles first, then the other tiles
    plan.q_offset.assign(1, 0);
    for (uint64_t v : hot_len) plan.q_offset.push_back(plan.q_offset.back() + v);
    for (uint64_t v : cold_len) plan.q_offset.push_back(plan.q_offset.back() + v);
    const size_t nq = plan.q_params.size(), nm = plan.m_params.size();
    plan.table_bytes = kCounterBytes + al256(nm * 8) + al256(nq * sizeof(TraceParams)) + al256(nq * sizeof(OutPtrs)) +
                       al256((2 * nq + 1) * sizeof(uint64_t)) + al256(nm * sizeof(MultiDev));
    return RT_OK;
}

// Device pointers of a placed batch.
struct Placed {
    std::vector<OutPtrs> outs;           // per request (index = request index)
    OutPtrs all{};                       // concatenated batches: the whole array
    QueueDev Q{};
    unsigned long long* counters = nullptr;
    unsigned long long* heads = nullptr; // queue head of each superposition request
    size_t n_zero = 0;                   // u64 words cleared before every launch (counters + heads)
    TraceParams* d_params = nullptr;
    OutPtrs* d_outs = nullptr;
    uint64_t* d_offset = nullptr;
    MultiDev* d_tables = nullptr;
};

void place_batch(const BatchPlan& plan, char* results, char* tables, const LaunchShape& shape, Placed& pl) {
    const size_t n = plan.n_rays.size();
    pl.outs.assign(n, OutPtrs{});
    char* cur = results;
    char* tb = tables;
    const size_t nq = plan.q_params.size(), nm = plan.m_params.size();
    pl.counters = reinterpret_cast<unsigned long long*>(tb);
    tb += kCounterBytes;
    pl.heads = reinterpret_cast<unsigned long long*>(tb);
    tb += al256(nm * 8);
    pl.n_zero = (kCounterBytes + al256(nm * 8)) / 8;
    if (plan.concat) {
        // one array of all rays; request i's view starts at its first ray and
        // keeps the whole array's leading dimension
        size_t total = 0;
        for (size_t v : plan.n_rays) total += v;
        pl.all = carve_results(cur, total, false);
        pl.all.counters = pl.counters;
        size_t off = 0;
        for (size_t i = 0; i < n; ++i) {
            OutPtrs o = pl.all;
            o.cls += off;
            o.x += off;
            o.p += off;
            o.bl += off;
            o.lambda += off;
            o.g += off;
            o.steps += off;
            o.rejected += off;
            pl.outs[i] = o;
            off += plan.n_rays[i];
        }
    } else {
        for (size_t i = 0; i < n; ++i) {
            pl.outs[i] = carve_results(cur, plan.n_rays[i], false);
            pl.outs[i].counters = pl.counters;
        }
    }
    pl.d_params = reinterpret_cast<TraceParams*>(tb);
    tb += al256(nq * sizeof(TraceParams));
    pl.d_outs = reinterpret_cast<OutPtrs*>(tb);
    tb += al256(nq * sizeof(OutPtrs));
    pl.d_offset = reinterpret_cast<uint64_t*>(tb);
    tb += al256((2 * nq + 1) * sizeof(uint64_t));
    pl.d_tables = reinterpret_cast<MultiDev*>(tb);
    pl.Q.reqs = pl.d_params;
    pl.Q.outs = pl.d_outs;
    pl.Q.req_offset = pl.d_offset;
    pl.Q.next = pl.counters + kQueueHead;
    pl.Q.counters = pl.counters;
    pl.Q.total = plan.q_offset.back();
    pl.Q.n_req = static_cast<uint32_t>(nq);
    pl.Q.refill_min = std::max<uint32_t>(1, std::min<uint32_t>(shape.refill_min, 64));
}

// Host->device copies of the batch tables (outside any captured graph).
rt_status upload_batch(const BatchPlan& plan, const Placed& pl, hipStream_t s) {
    const size_t nq = plan.q_params.size(), nm = plan.m_params.size();
    if (nq) {
        std::vector<OutPtrs> qo(nq);
        for (size_t j = 0; j < nq; ++j) qo[j] = pl.outs[plan.q_req[j]];
        RT_HIP(hipMemcpyAsync(pl.d_params, plan.q_params.data(), nq * sizeof(TraceParams), hipMemcpyHostToDevice, s));
        RT_HIP(hipMemcpyAsync(pl.d_outs, qo.data(), nq * sizeof(OutPtrs), hipMemcpyHostToDevice, s));
        RT_HIP(hipMemcpyAsync(pl.d_offset, plan.q_offset.data(), (2 * nq + 1) * sizeof(uint64_t),
                              hipMemcpyHostToDevice, s));
        // the kernels read these tables; wait for the pageable copies to land
        RT_HIP(hipStreamSynchronize(s));
    }
    if (nm) {
        RT_HIP(hipMemcpyAsync(pl.d_tables, plan.m_tables.data(), nm * sizeof(MultiDev), hipMemcpyHostToDevice, s));
        RT_HIP(hipStreamSynchronize(s));
    }
    return RT_OK;
}

// Streams a batch may fork onto (queue_kind 2): side streams and events.
struct Branches {
    const hipStream_t* side = nullptr;
    hipEvent_t fork = nullptr;
    const hipEvent_t* join = nullptr;
    int n = 0;
};

// The launch sequence of a batch (captured into a graph by rt_trace_batch).
// queue_kind 0 and 1: one persistent kernel over the shared queue.
// queue_kind 2: one per-frame kernel (kPlainShape) per request, spread over s
// and the side streams; captured, the streams become parallel graph branches,
// so one request's slow rays overlap the other requests.
hipError_t launch_batch(const BatchPlan& plan, const Placed& pl, const LaunchShape& shape, uint32_t grid,
                        hipStream_t s, const Branches& br) {
    hipError_t e = launch_zero_u64(pl.counters, static_cast<uint32_t>(pl.n_zero), s);
    if (e != hipSuccess) return e;
    const bool per_request = !plan.q_params.empty() && shape.queue_kind == 2;
    const bool fork = br.n > 0 && (per_request || !plan.r_params.empty());
    size_t next_branch = 0;
    auto branch = [&]() {
        const size_t b = (next_branch++) % static_cast<size_t>(br.n + 1);
        return b == 0 ? s : br.side[b - 1];
    };
    if (fork) {
        e = hipEventRecord(br.fork, s);
        for (int i = 0; i < br.n && e == hipSuccess; ++i) e = hipStreamWaitEvent(br.side[i], br.fork, 0);
        if (e != hipSuccess) return e;
    }
    if (per_request) {
        for (size_t j = 0; j < plan.q_params.size() && e == hipSuccess; ++j)
            e = launch_trace_camera(plan.q_params[j], pl.outs[plan.q_req[j]], kPlainShape, fork ? branch() : s);
    } else if (!plan.q_params.empty()) {
        e = launch_trace_queue(pl.Q, shape, plan.mixed, plan.carter, grid, s);
    }
    for (size_t j = 0; j < plan.r_params.size() && e == hipSuccess; ++j)
        e = launch_trace_render(plan.r_params[j], pl.outs[plan.r_req[j]], kPlainShape, fork ? branch() : s);
    if (e != hipSuccess) return e;
    if (fork) {
        for (int i = 0; i < br.n && e == hipSuccess; ++i) {
            e = hipEventRecord(br.join[i], br.side[i]);
            if (e == hipSuccess) e = hipStreamWaitEvent(s, br.join[i], 0);
        }
        if (e != hipSuccess) return e;
    }
    for (size_t j = 0; j < plan.m_params.size(); ++j) {
        e = launch_trace_multi(plan.m_params[j], pl.outs[plan.m_req[j]], pl.d_tables + j, pl.heads + j, kMultiShape, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Signature of everything a captured graph bakes in.
std::vector<unsigned char> graph_key(const BatchPlan& plan, const Placed& pl, const LaunchShape& shape) {
    std::vector<unsigned char> k;
    auto put = [&](const void* p, size_t b) {
        const unsigned char* c = static_cast<const unsigned char*>(p);
        k.insert(k.end(), c, c + b);
    };
    put(&pl.Q, sizeof(pl.Q));
    put(&shape, sizeof(shape));
    if (shape.queue_kind == 2)  // per-request kernels take their parameters by value
        for (size_t j = 0; j < plan.q_params.size(); ++j) {
            put(&plan.q_params[j], sizeof(TraceParams));
            put(&pl.outs[plan.q_req[j]], sizeof(OutPtrs));
        }
    const bool variant[2] = {plan.mixed, plan.carter};
    put(variant, sizeof(variant));
    for (size_t j = 0; j < plan.m_params.size(); ++j) {
        put(&plan.m_params[j], sizeof(TraceParams));
        put(&pl.outs[plan.m_req[j]], sizeof(OutPtrs));
    }
    for (size_t j = 0; j < plan.r_params.size(); ++j) {
        put(&plan.r_params[j], sizeof(TraceParams));
        put(&pl.outs[plan.r_req[j]], sizeof(OutPtrs));
    }
    return k;
}

}  // namespace

// Graph cache of a context (one batch layout).
struct GraphCache {
    std::vector<unsigned char> key;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void reset() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        key.clear();
    }
    ~GraphCache() { reset(); }
};

static std::mutex g_graph_mu;
static std::map<const rt_context*, GraphCache*> g_graphs;

void forget_graphs(const rt_context* ctx) {
    std::lock_guard<std::mutex> lk(g_graph_mu);
    auto it = g_graphs.find(ctx);
    if (it != g_graphs.end()) {
        delete it->second;
        g_graphs.erase(it);
    }
}

static GraphCache& graph_cache(const rt_context* ctx) {
    std::lock_guard<std::mutex> lk(g_graph_mu);
    GraphCache*& g = g_graphs[ctx];
    if (!g) g = new GraphCache();
    return *g;
}

// Run a batch on the context: place, upload, replay or capture the graph, time it.
// wins (may be null): per-request rectangles, results concatenated in request
// order and left resident as the context's last trace; when the context's
// shape asks for per-request kernels they run on the shared tile queue at two
// waves per SIMD instead (the serving loop's shape), a rectangle being too
// small to fill the device alone.
static rt_status run_batch(rt_context* ctx, const rt_request* reqs, size_t n_req, rt_results* outs,
                           const rt_tile* wins = nullptr) {
    const auto t0 = std::chrono::steady_clock::now();
    RT_HIP(hipSetDevice(ctx->device));
    LaunchShape shape = ctx->qshape;
    if (wins && shape.queue_kind == 2) shape = kServerShape;
    BatchPlan plan;
    rt_status st = plan_batch(reqs, n_req, plan, shape.queue_kind != 1, wins, wins != nullptr);
    if (st != RT_OK) return st;
    st = ctx->arena.ensure(std::max<size_t>(plan.result_bytes, 256));
    if (st != RT_OK) return st;
    st = ctx->aux.ensure(plan.table_bytes + sizeof(MultiDev) + 256);
    if (st != RT_OK) return st;
    Placed pl;
    place_batch(plan, static_cast<char*>(ctx->arena.base), static_cast<char*>(ctx->aux.base), shape, pl);
    st = upload_batch(plan, pl, ctx->stream);
    if (st != RT_OK) return st;
    GraphCache& gc = graph_cache(ctx);
    (void)0;
    std::vector<unsigned char> key = graph_key(plan, pl, shape);
    if (!gc.exec || gc.key != key) {
        gc.reset();
        uint32_t grid = 0;
        RT_HIP(queue_grid(shape, plan.mixed, plan.carter, &grid));
        RT_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        Branches br;
        br.side = ctx->side;
        br.fork = ctx->fork;
        br.join = ctx->join;
        br.n = rt_context::kSide;
        const hipError_t le = launch_batch(plan, pl, shape, grid, ctx->stream, br);
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(ctx->stream, &g);
        if (le != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return hip_fail(le, "batch launch (capture)");
        }
        if (ce != hipSuccess) return hip_fail(ce, "hipStreamEndCapture");
        gc.graph = g;
        RT_HIP(hipGraphInstantiate(&gc.exec, gc.graph, nullptr, nullptr, 0));
        gc.key = key;
    }
    RT_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    RT_HIP(hipGraphLaunch(gc.exec, ctx->stream));
    RT_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    RT_HIP(hipEventSynchronize(ctx->ev1));
    float ms = 0.0f;
    RT_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    unsigned long long cnt[5] = {0, 0, 0, 0, 0};
    RT_HIP(hipMemcpy(cnt, pl.counters, sizeof(cnt), hipMemcpyDeviceToHost));
    size_t total = 0;
    for (size_t i = 0; i < n_req; ++i) total += plan.n_rays[i];
    if (outs)
        for (size_t i = 0; i < n_req; ++i) {
            st = copy_results_to_host(pl.outs[i], plan.n_rays[i], &outs[i], ctx->stream);
            if (st != RT_OK) return st;
        }
    if (plan.concat) {
        ctx->last = pl.all;
        ctx->n_last = total;
    } else {
        ctx->last = n_req ? pl.outs[0] : OutPtrs{};
        ctx->n_last = n_req ? plan.n_rays[0] : 0;
    }
    rt_stats& s = ctx->stats;
    s.kernel_ms = ms;
    s.rays = total;
    s.accepted_steps = cnt[0];
    s.rejected_steps = cnt[1];
    s.rhs_evals = cnt[2];
    s.fp32_steps = cnt[3];
    s.block_size = shape.block;
    uint32_t grid = 0;
    (void)queue_grid(shape, plan.mixed, plan.carter, &grid);
    s.grid_size = grid;
    s.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return RT_OK;
}

}  // namespace rt

using namespace rt;

extern "C" {

rt_status rt_get_launch_shape(const rt_context* ctx, rt_launch_shape* plain, rt_launch_shape* queue) {
    if (!ctx) return fail(RT_ERR_INVALID_ARG, "null context");
    if (plain) *plain = rt_launch_shape{ctx->shape.block, ctx->shape.waves, ctx->shape.refill_min, ctx->shape.queue_kind};
    if (queue) *queue = rt_launch_shape{ctx->qshape.block, ctx->qshape.waves, ctx->qshape.refill_min,
                                        ctx->qshape.queue_kind};
    return RT_OK;
}

static bool shape_valid(const rt_launch_shape* s) {
    return s->block >= 64 && s->block <= 256 && (s->block % 64) == 0 && (s->waves == 1 || s->waves == 2) &&
           s->refill_min >= 1 && s->refill_min <= 64 && s->queue_kind <= 2;
}

rt_status rt_set_launch_shape(rt_context* ctx, const rt_launch_shape* plain, const rt_launch_shape* queue) {
    if (!ctx) return fail(RT_ERR_INVALID_ARG, "null context");
    if ((plain && !shape_valid(plain)) || (queue && !shape_valid(queue)))
        return fail(RT_ERR_INVALID_ARG, "launch shape: block in {64,128,192,256}, waves in {1,2}, refill_min 1..64, queue_kind in {0,1,2}");
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (plain) ctx->shape = LaunchShape{plain->block, plain->waves, plain->refill_min, plain->queue_kind};
    if (queue) ctx->qshape = LaunchShape{queue->block, queue->waves, queue->refill_min, queue->queue_kind};
    return RT_OK;
}

rt_status rt_trace_batch(rt_context* ctx, const rt_request* reqs, size_t n_req, rt_results* outs) {
    if (!ctx || !reqs || n_req == 0) return fail(RT_ERR_INVALID_ARG, "null context/requests or n_req == 0");
    std::lock_guard<std::mutex> lk(ctx->mu);
    return run_batch(ctx, reqs, n_req, outs);
}

rt_status rt_trace_tiles(rt_context* ctx, const rt_tile_task* tasks, size_t n_tasks, size_t* n_rays) {
    if (!ctx || !tasks || n_tasks == 0) return fail(RT_ERR_INVALID_ARG, "null context or tasks, or no tasks");
    std::vector<rt_request> reqs(n_tasks);
    std::vector<rt_tile> wins(n_tasks);
    for (size_t i = 0; i < n_tasks; ++i) {
        reqs[i].camera = tasks[i].camera;
        reqs[i].metric = tasks[i].metric;
        reqs[i].integrator = tasks[i].integrator;
        reqs[i].termination = tasks[i].termination;
        wins[i] = tasks[i].tile;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    rt_status st = run_batch(ctx, reqs.data(), n_tasks, nullptr, wins.data());
    if (st != RT_OK) return st;
    if (n_rays) *n_rays = ctx->n_last;
    return RT_OK;
}

rt_status rt_autotune(rt_context* ctx, const rt_request* sample, size_t n_sample, rt_launch_shape* best) {
    if (!ctx || !sample || n_sample == 0) return fail(RT_ERR_INVALID_ARG, "null context/sample");
    for (size_t i = 0; i < n_sample; ++i)
        if (sample[i].metric.n_sources >= 2) return fail(RT_ERR_INVALID_ARG, "autotune takes single-Kerr requests");
    std::lock_guard<std::mutex> lk(ctx->mu);
    // candidates: tile queue at each block size and occupancy, lane-refill
    // queue at each block size with three refill thresholds, per-request
    // kernels on concurrent streams
    LaunchShape keep = ctx->qshape, win = ctx->qshape;
    std::vector<LaunchShape> cands;
    for (uint32_t b : {64u, 128u, 256u})
        for (uint32_t w : {1u, 2u}) {
            cands.push_back(LaunchShape{b, w, 16, 0});
            for (uint32_t r : {8u, 16u, 32u}) cands.push_back(LaunchShape{b, w, r, 1});
        }
    cands.push_back(LaunchShape{kPlainShape.block, kPlainShape.waves, 16, 2});
    double best_ms = 1e300;
    for (const LaunchShape& c : cands) {
        ctx->qshape = c;
        rt_status st = run_batch(ctx, sample, n_sample, nullptr);  // capture + run
        if (st == RT_OK) st = run_batch(ctx, sample, n_sample, nullptr);  // replay, timed
        if (st != RT_OK) {
            ctx->qshape = keep;
            return st;
        }
        if (ctx->stats.kernel_ms < best_ms) {
            best_ms = ctx->stats.kernel_ms;
            win = ctx->qshape;
        }
    }
    ctx->qshape = win;
    ctx->tuned = true;
    if (best) *best = rt_launch_shape{win.block, win.waves, win.refill_min, win.queue_kind};
    return RT_OK;
}

}  // extern "C"
