// rt_host.cpp — C++ host runtime of librtsim.so: parameter set-up from the
// C-ABI structs, per-context device buffers and stream, timing, and the
// extern "C" entry points declared in include/rt_tracer.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>

#include "../../include/rt_tracer.h"
#include "rt_host.h"
#include "rt_kernels.h"
#include "rt_params.h"

namespace rt {

// ------------------------------------------------------------- errors
static thread_local std::string g_last_error;

rt_status fail(rt_status st, const std::string& msg) {
    g_last_error = msg;
    return st;
}

rt_status hip_fail(hipError_t e, const char* what) {
    return fail(e == hipErrorOutOfMemory ? RT_ERR_OOM : RT_ERR_DEVICE,
                std::string(what) + ": " + hipGetErrorString(e));
}

#define RT_HIP(call)                                         \
    do {                                                     \
        hipError_t e_ = (call);                              \
        if (e_ != hipSuccess) return hip_fail(e_, #call);    \
    } while (0)

// ------------------------------------------------------ host geometry
namespace {

double horizon_radius(double m, double a) { return m + std::sqrt(std::max(m * m - a * a, 0.0)); }

struct HField {
    double r, H, k[3];
};

// Kerr–Schild field of one hole (core/kerr_schild_null_integrator.hpp:30-44).
HField ks_field_host(const double x[3], double M, double a) {
    const double a2 = a * a;
    const double z2 = x[2] * x[2];
    const double diff = x[0] * x[0] + x[1] * x[1] + z2 - a2;
    const double r2 = std::max(0.5 * (diff + std::sqrt(diff * diff + 4.0 * a2 * z2)), 1e-24);
    const double r = std::sqrt(r2);
    const double den = r2 * r2 + a2 * z2;
    const double rpa = r2 + a2;
    HField f;
    f.r = r;
    f.H = M * r2 * r / den;
    f.k[0] = (r * x[0] + a * x[1]) / rpa;
    f.k[1] = (r * x[1] - a * x[0]) / rpa;
    f.k[2] = x[2] / r;
    return f;
}

}  // namespace

bool valid_request(const rt_metric* mt, const rt_integrator* ig, const rt_termination* tm, std::string& why) {
    if (!mt || !ig || !tm) {
        why = "null metric/integrator/termination";
        return false;
    }
    if (!(ig->rtol > 0.0) || !(ig->atol >= 0.0) || !(ig->h_min > 0.0) || !(ig->h_max >= ig->h_min) ||
        ig->max_steps == 0 || !(ig->h0 >= 0.0) || !(ig->invariant_tol > 0.0)) {
        why = "integrator: need rtol > 0, atol >= 0, 0 < h_min <= h_max, h0 >= 0, max_steps > 0, invariant_tol > 0";
        return false;
    }
    if (ig->precision > RT_PRECISION_MIXED) {
        why = "integrator: unknown precision policy";
        return false;
    }
    if (!(tm->escape_radius > 0.0) || !(tm->horizon_eps >= 0.0)) {
        why = "termination: need escape_radius > 0 and horizon_eps >= 0";
        return false;
    }
    if (!(mt->mass > 0.0) && mt->n_sources < 2) {
        why = "metric: mass must be > 0";
        return false;
    }
    if (mt->n_sources > RT_MAX_SOURCES) {
        why = "metric: at most 16 sources";
        return false;
    }
    return true;
}

// Metric, integrator and termination parts of the parameter block
// (restating dispatch_fp64_scalar's set-up, render/software_compute_engine.hpp:3801-3809).
void fill_params(const rt_metric* mt, const rt_integrator* ig, const rt_termination* tm, TraceParams& P) {
    std::memset(&P, 0, sizeof(P));
    P.M = std::max(mt->mass, 1e-4);
    P.a = std::clamp(mt->spin, -0.999 * P.M, 0.999 * P.M);
    P.a2 = P.a * P.a;
    P.rh = horizon_radius(P.M, P.a);
    P.r_hz = P.rh + tm->horizon_eps;
    P.r_esc = tm->escape_radius;
    P.disk_spin = std::clamp(P.a, -0.999 * P.M, 0.999 * P.M);
    P.disk_on = tm->disk_enabled ? 1u : 0u;
    P.clamp_g = (tm->flags & RT_TERM_NO_G_CLAMP) ? 0u : 1u;
    const double isco = kerr_isco_host(P.M, P.a);
    P.disk_in = (tm->disk_inner > 0.0) ? tm->disk_inner : isco;
    P.disk_out = std::max((tm->disk_outer > 0.0) ? tm->disk_outer : 20.0 * P.M, P.disk_in * 1.05);
    P.rh102 = P.rh * 1.02;
    P.near_r = 3.0 * (2.0 * P.M);
    P.n_src = (mt->n_sources >= 2) ? mt->n_sources : 0u;
    P.rtol = ig->rtol;
    P.atol = ig->atol;
    P.h0 = ig->h0;
    P.h_min = ig->h_min;
    P.h_max = ig->h_max;
    P.inv_tol = ig->invariant_tol;
    P.max_steps = ig->max_steps;
    P.precision = ig->precision;
    const double sw = std::max(ig->mixed_switch_radius, 3.0 * P.rh);
    P.switch_r = sw;
}

double kerr_isco_host(double m, double a_spin) {
    const double a_star = std::clamp(a_spin / std::max(m, 1e-12), -0.9999999, 0.9999999);
    const double abs_a = std::abs(a_star);
    const double orbit_sign = (a_star >= 0.0) ? -1.0 : 1.0;
    const double cbrt_term = std::cbrt(std::max(1.0 - abs_a * abs_a, 0.0));
    const double z1 = 1.0 + cbrt_term * (std::cbrt(1.0 + abs_a) + std::cbrt(1.0 - abs_a));
    const double z2 = std::sqrt(3.0 * abs_a * abs_a + z1 * z1);
    return (3.0 + z2 + orbit_sign * std::sqrt(std::max((3.0 - z1) * (3.0 + z1 + 2.0 * z2), 0.0))) * m;
}

bool valid_camera(const rt_camera* c, std::string& why) {
    if (!c || c->width == 0 || c->height == 0) {
        why = "camera: width and height must be > 0";
        return false;
    }
    if (!(c->fov_y > 0.0 && c->fov_y < 3.14159)) {
        why = "camera: fov_y must be in (0, pi)";
        return false;
    }
    if (!(c->observer[1] > 0.0)) {
        why = "camera: observer radius must be > 0";
        return false;
    }
    return true;
}

// Camera part: per-frame constants of the pixel loop and the observer's
// Gram–Schmidt frame (core/kerr_schild_null_integrator.hpp:149-177), computed
// once on the host because they are the same for every pixel.
void fill_camera(const rt_camera* c, TraceParams& P) {
    const uint32_t first = std::min(c->row_offset, c->height);
    const uint32_t last = c->row_count ? std::min(first + c->row_count, c->height) : c->height;
    P.width = c->width;
    P.height = c->height;
    P.row_first = first;
    P.n_rays = (last - first) * c->width;
    P.tiles_x = (c->width + 7) / 8;
    P.mode = 0;
    P.aspect = static_cast<double>(c->width) / static_cast<double>(c->height);
    P.tan_half = std::tan(c->fov_y * 0.5);
    for (int i = 0; i < 3; ++i) {
        P.fwd[i] = c->forward[i];
        P.right[i] = c->right[i];
        P.up[i] = c->up[i];
    }
    const double r_obs = std::max(c->observer[1], P.rh * 1.02);
    const double th = std::clamp(c->observer[2], 0.001, 3.14159265358979323846 - 0.001);
    const double ph = c->observer[3];
    P.sth = std::sin(th);
    P.cth = std::cos(th);
    P.sph = std::sin(ph);
    P.cph = std::cos(ph);
    if (P.n_src >= 2) {
        P.X0[0] = r_obs * P.sth * P.cph;
        P.X0[1] = r_obs * P.sth * P.sph;
        P.X0[2] = r_obs * P.cth;
        return;  // superposition frame is filled by fill_sources()
    }
    // kerr_schild_position_from_boyer_lindquist (core/kerr_schild_null_integrator.hpp:46-54)
    const double st = std::sin(th);
    P.X0[0] = (r_obs * std::cos(ph) - P.a * std::sin(ph)) * st;
    P.X0[1] = (r_obs * std::sin(ph) + P.a * std::cos(ph)) * st;
    P.X0[2] = r_obs * std::cos(th);
    const HField f = ks_field_host(P.X0, P.M, P.a);
    const double two_h = 2.0 * f.H;
    const double* k = f.k;
    auto inner = [&](const double* A, const double* B) {
        const double la = A[0] + k[0] * A[1] + k[1] * A[2] + k[2] * A[3];
        const double lb = B[0] + k[0] * B[1] + k[1] * B[2] + k[2] * B[3];
        return -A[0] * B[0] + A[1] * B[1] + A[2] * B[2] + A[3] * B[3] + two_h * la * lb;
    };
    const double u0 = 1.0 / std::sqrt(std::max(1.0 - two_h, 1e-6));
    double e[3][4];
    for (int i = 0; i < 3; ++i) {
        e[i][0] = two_h * k[i] * u0 * u0;
        e[i][1] = e[i][2] = e[i][3] = 0.0;
        e[i][1 + i] = 1.0;
    }
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < i; ++j) {
            const double proj = inner(e[i], e[j]);
            for (int cc = 0; cc < 4; ++cc) e[i][cc] -= proj * e[j][cc];
        }
        const double inv_norm = 1.0 / std::sqrt(std::max(inner(e[i], e[i]), 1e-30));
        for (int cc = 0; cc < 4; ++cc) e[i][cc] *= inv_norm;
    }
    std::memcpy(P.frame, e, sizeof(e));
    P.u0 = u0;
    P.kobs[0] = k[0];
    P.kobs[1] = k[1];
    P.kobs[2] = k[2];
    P.two_h_obs = two_h;
}

// ------------------------------------------------------------ buffers
DeviceResults::~DeviceResults() { release(); }

void DeviceResults::release() {
    if (base) (void)hipFree(base);
    base = nullptr;
    capacity = 0;
    std::memset(&ptrs, 0, sizeof(ptrs));
}

size_t DeviceResults::bytes_per_ray() { return 4 + 8 * 4 + 8 * 4 + 8 * 3 + 8 + 8 + 4 + 4 + 2 * 8 * 4; }

rt_status DeviceResults::ensure(size_t n) {
    if (n <= capacity && base) return RT_OK;
    release();
    const size_t cap = std::max<size_t>(n, 1024);
    // one allocation, 256-byte aligned sub-arrays
    auto align = [](size_t v) { return (v + 255) & ~size_t(255); };
    size_t off = 0;
    const size_t o_cls = off;
    off = align(off + cap * 4);
    const size_t o_x = off;
    off = align(off + cap * 32);
    const size_t o_p = off;
    off = align(off + cap * 32);
    const size_t o_bl = off;
    off = align(off + cap * 24);
    const size_t o_lam = off;
    off = align(off + cap * 8);
    const size_t o_g = off;
    off = align(off + cap * 8);
    const size_t o_st = off;
    off = align(off + cap * 4);
    const size_t o_rj = off;
    off = align(off + cap * 4);
    const size_t o_ix = off;
    off = align(off + cap * 32);
    const size_t o_ip = off;
    off = align(off + cap * 32);
    const size_t o_cnt = off;
    off = align(off + 16 * 8);
    hipError_t e = hipMalloc(&base, off);
    if (e != hipSuccess) {
        base = nullptr;
        return hip_fail(e, "hipMalloc(results)");
    }
    char* b = static_cast<char*>(base);
    ptrs.cls = reinterpret_cast<uint32_t*>(b + o_cls);
    ptrs.x = reinterpret_cast<double*>(b + o_x);
    ptrs.p = reinterpret_cast<double*>(b + o_p);
    ptrs.bl = reinterpret_cast<double*>(b + o_bl);
    ptrs.lambda = reinterpret_cast<double*>(b + o_lam);
    ptrs.g = reinterpret_cast<double*>(b + o_g);
    ptrs.steps = reinterpret_cast<uint32_t*>(b + o_st);
    ptrs.rejected = reinterpret_cast<uint32_t*>(b + o_rj);
    ptrs.in_x = reinterpret_cast<double*>(b + o_ix);
    ptrs.in_p = reinterpret_cast<double*>(b + o_ip);
    ptrs.counters = reinterpret_cast<unsigned long long*>(b + o_cnt);
    capacity = cap;
    return RT_OK;
}

// Copy n results with leading dimension ld (device) into the caller's arrays.
rt_status copy_results_to_host(const OutPtrs& d, size_t n, const rt_results* out, hipStream_t s) {
    if (!out) return fail(RT_ERR_INVALID_ARG, "null results");
    const size_t ld = d.n;
    auto cp = [&](void* dst, const void* src, size_t bytes) -> hipError_t {
        if (!dst || bytes == 0) return hipSuccess;
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s);
    };
    RT_HIP(cp(out->cls, d.cls, n * 4));
    for (int c = 0; c < 4; ++c) {
        if (out->x) RT_HIP(cp(out->x + c * n, d.x + c * ld, n * 8));
        if (out->p) RT_HIP(cp(out->p + c * n, d.p + c * ld, n * 8));
    }
    for (int c = 0; c < 3; ++c)
        if (out->bl) RT_HIP(cp(out->bl + c * n, d.bl + c * ld, n * 8));
    RT_HIP(cp(out->lambda, d.lambda, n * 8));
    RT_HIP(cp(out->g, d.g, n * 8));
    RT_HIP(cp(out->steps, d.steps, n * 4));
    RT_HIP(cp(out->rejected, d.rejected, n * 4));
    RT_HIP(hipStreamSynchronize(s));
    return RT_OK;
}

}  // namespace rt

// ===================================================================== context
struct rt_context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    rt::DeviceResults res;
    size_t n_last = 0;
    rt_stats stats{};
    rt::LaunchShape shape;
    std::mutex mu;
};

using namespace rt;

namespace {

rt_status begin(rt_context* ctx) {
    RT_HIP(hipSetDevice(ctx->device));
    return RT_OK;
}

// Launch one trace on the context's stream, timing it with HIP events.
// count rays are traced; ld is the leading dimension of the SoA arrays.
template <class Launch>
rt_status run_trace(rt_context* ctx, const TraceParams& P, size_t count, size_t ld, Launch&& launch) {
    const auto t0 = std::chrono::steady_clock::now();
    rt_status st = ctx->res.ensure(ld);
    if (st != RT_OK) return st;
    OutPtrs O = ctx->res.ptrs;
    O.n = ld;
    RT_HIP(hipMemsetAsync(O.counters, 0, 16 * 8, ctx->stream));
    RT_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    RT_HIP(launch(P, O));
    RT_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    RT_HIP(hipEventSynchronize(ctx->ev1));
    float ms = 0.0f;
    RT_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    unsigned long long cnt[5] = {0, 0, 0, 0, 0};
    RT_HIP(hipMemcpy(cnt, O.counters, sizeof(cnt), hipMemcpyDeviceToHost));
    ctx->n_last = count;
    ctx->res.ptrs.n = ld;
    rt_stats& s = ctx->stats;
    s.kernel_ms = ms;
    s.rays = count;
    s.accepted_steps = cnt[0];
    s.rejected_steps = cnt[1];
    s.rhs_evals = cnt[2];
    s.fp32_steps = cnt[3];
    s.block_size = ctx->shape.block;
    s.grid_size = 0;
    s.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return RT_OK;
}

}  // namespace

extern "C" {

void rt_default_integrator(rt_integrator* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->rtol = 1e-10;
    o->atol = 1e-12;
    o->h0 = 0.0;
    o->h_min = 1e-12;
    o->h_max = 1e13;
    o->invariant_tol = 1e-12;
    o->mixed_switch_radius = 0.0;
    o->max_steps = 20000;
    o->precision = RT_PRECISION_FP64;
}

void rt_default_termination(rt_termination* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->horizon_eps = 1e-6;
    o->escape_radius = 1000.0;
    o->disk_inner = 0.0;
    o->disk_outer = 20.0;
    o->disk_enabled = 1;
    o->flags = 0;
}

void rt_default_camera(rt_camera* c, double r_obs, double th, double ph, double fov_y, uint32_t w, uint32_t h) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    const double st = std::sin(th), ct = std::cos(th), sp = std::sin(ph), cp = std::cos(ph);
    c->observer[0] = 0.0;
    c->observer[1] = r_obs;
    c->observer[2] = th;
    c->observer[3] = ph;
    c->forward[0] = -st * cp;
    c->forward[1] = -st * sp;
    c->forward[2] = -ct;
    c->right[0] = -sp;
    c->right[1] = cp;
    c->right[2] = 0.0;
    c->up[0] = -ct * cp;
    c->up[1] = -ct * sp;
    c->up[2] = st;
    c->fov_y = fov_y;
    c->width = w;
    c->height = h;
}

void rt_kerr_metric(rt_metric* m, double mass, double spin) {
    if (!m) return;
    std::memset(m, 0, sizeof(*m));
    m->mass = mass;
    m->spin = spin;
}

uint32_t rt_abi_version(void) { return RT_ABI_VERSION; }
const char* rt_version(void) { return "librtsim 0.1 (gfx950, DP45 Kerr-Schild null tracer)"; }
const char* rt_last_error(void) { return rt::g_last_error.c_str(); }

rt_status rt_context_create(int device, rt_context** out) {
    if (!out) return fail(RT_ERR_INVALID_ARG, "null out");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(RT_ERR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= n) return fail(RT_ERR_INVALID_ARG, "device index out of range");
    rt_context* ctx = new (std::nothrow) rt_context();
    if (!ctx) return fail(RT_ERR_OOM, "context allocation");
    ctx->device = device;
    // Launch-shape override for experiments: RTSIM_BLOCK (64..256), RTSIM_WAVES (1|2).
    if (const char* b = std::getenv("RTSIM_BLOCK")) ctx->shape.block = static_cast<uint32_t>(std::atoi(b));
    if (const char* w = std::getenv("RTSIM_WAVES")) ctx->shape.waves = static_cast<uint32_t>(std::atoi(w));
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev0);
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev1);
    if (e != hipSuccess) {
        rt_context_destroy(ctx);
        return hip_fail(e, "context set-up");
    }
    *out = ctx;
    return RT_OK;
}

rt_status rt_context_destroy(rt_context* ctx) {
    if (!ctx) return RT_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->res.release();
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return RT_OK;
}

rt_status rt_get_stats(const rt_context* ctx, rt_stats* out) {
    if (!ctx || !out) return fail(RT_ERR_INVALID_ARG, "null argument");
    *out = ctx->stats;
    return RT_OK;
}

rt_status rt_trace_camera_resident(rt_context* ctx, const rt_camera* cam, const rt_metric* metric,
                                   const rt_integrator* integ, const rt_termination* term, size_t* n_rays) {
    if (!ctx) return fail(RT_ERR_INVALID_ARG, "null context");
    std::string why;
    if (!valid_request(metric, integ, term, why) || !valid_camera(cam, why)) return fail(RT_ERR_INVALID_ARG, why);
    if (metric->n_sources >= 2) return fail(RT_ERR_INVALID_ARG, "superposition metrics are traced by rt_trace_batch");
    std::lock_guard<std::mutex> lk(ctx->mu);
    rt_status st = begin(ctx);
    if (st != RT_OK) return st;
    TraceParams P;
    fill_params(metric, integ, term, P);
    fill_camera(cam, P);
    const size_t n = P.n_rays;
    st = run_trace(ctx, P, n, n, [&](const TraceParams& p, const OutPtrs& o) {
        return launch_trace_camera(p, o, ctx->shape, ctx->stream);
    });
    if (st != RT_OK) return st;
    if (n_rays) *n_rays = n;
    return RT_OK;
}

rt_status rt_fetch_results(rt_context* ctx, rt_results* out) {
    if (!ctx || !out) return fail(RT_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    rt_status st = begin(ctx);
    if (st != RT_OK) return st;
    return copy_results_to_host(ctx->res.ptrs, ctx->n_last, out, ctx->stream);
}

rt_status rt_trace_camera(rt_context* ctx, const rt_camera* cam, const rt_metric* metric, const rt_integrator* integ,
                          const rt_termination* term, rt_results* out) {
    if (!out) return fail(RT_ERR_INVALID_ARG, "null results");
    size_t n = 0;
    rt_status st = rt_trace_camera_resident(ctx, cam, metric, integ, term, &n);
    if (st != RT_OK) return st;
    return rt_fetch_results(ctx, out);
}

rt_status rt_trace_rays(rt_context* ctx, const double* x_soa, const double* p_soa, size_t n, const rt_metric* metric,
                        const rt_integrator* integ, const rt_termination* term, rt_results* out) {
    if (!ctx || !x_soa || !p_soa || !out || n == 0) return fail(RT_ERR_INVALID_ARG, "null argument or n == 0");
    std::string why;
    if (!valid_request(metric, integ, term, why)) return fail(RT_ERR_INVALID_ARG, why);
    if (metric->n_sources >= 2) return fail(RT_ERR_INVALID_ARG, "superposition metrics are traced by rt_trace_batch");
    std::lock_guard<std::mutex> lk(ctx->mu);
    rt_status st = begin(ctx);
    if (st != RT_OK) return st;
    st = ctx->res.ensure(n);
    if (st != RT_OK) return st;
    OutPtrs O = ctx->res.ptrs;
    for (int c = 0; c < 4; ++c) {
        RT_HIP(hipMemcpyAsync(const_cast<double*>(O.in_x) + c * ctx->res.capacity, x_soa + c * n, n * 8,
                              hipMemcpyHostToDevice, ctx->stream));
        RT_HIP(hipMemcpyAsync(const_cast<double*>(O.in_p) + c * ctx->res.capacity, p_soa + c * n, n * 8,
                              hipMemcpyHostToDevice, ctx->stream));
    }
    TraceParams P;
    fill_params(metric, integ, term, P);
    P.mode = 1;
    P.n_rays = static_cast<uint32_t>(n);
    // Inputs are staged with the buffer capacity as leading dimension, so trace
    // with that leading dimension and copy back the first n of each row.
    const size_t ld = ctx->res.capacity;
    st = run_trace(ctx, P, n, ld, [&](const TraceParams& p, const OutPtrs& o) {
        return launch_trace_rays(p, o, ctx->shape, ctx->stream);
    });
    if (st != RT_OK) return st;
    return copy_results_to_host(ctx->res.ptrs, n, out, ctx->stream);
}

}  // extern "C"
