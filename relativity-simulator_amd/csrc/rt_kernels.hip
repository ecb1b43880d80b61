// rt_kernels.hip — HIP kernels of the null-geodesic tracer for gfx950.
//
// One lane traces one ray: adaptive DP45 on the Kerr–Schild Hamiltonian
// (first-same-as-last), events located on the dense output (horizon,
// equatorial disk, escape), results written once as SoA.
//
//   k_trace_camera_kerr  one camera frame, one lane per pixel, the lanes of a
//                        wave on an 8x8 pixel tile (similar step counts)
//   k_trace_rays_kerr    explicit initial states
//   k_trace_tiles        persistent kernel over a device work queue of 8x8
//                        tiles of many requests (one atomic per tile)
//   k_trace_queue        persistent kernel over a device work queue of rays:
//                        a wave refills its idle lanes with one atomic
//                        (ballot + prefix count)
//
// The superposition kernel (k_trace_multi) is in rt_multi.hip.
#include <hip/hip_runtime.h>

#include "rt_camera.hpp"
#include "rt_device.hpp"
#include "rt_kernels.h"

namespace rt {

// ------------------------------------------------------------ lane state
struct Lane {
    double y[8], k1[8];
    double ah, lambda, r0, xi, g;
    uint32_t acc, rej, flags, cls;
    uint32_t rhs, proj, f32;
};


// Render-step heuristic (software_compute_engine.hpp:1677-1683 with the
// GpuCameraPushConstants defaults) as the initial |h|.
__device__ __forceinline__ double initial_step(double r, double rh, double m) {
    const double scale = sqrt(r * fmax(r - rh, 0.02 * m));
    const double blend = fmin(fmax((r - 20.0 * rh) / (80.0 * rh), 0.0), 1.0);
    const double far = 1.0 + (7.8 - 1.0) * blend;
    return fmin(fmax(0.5 * scale * far, 0.004), 3.5 * 7.8);
}

// Stages k2..k7 of one DP45 attempt from (y, k1, h): the 5th-order solution
// y5, k7 = f(y5), the error combination ep = sum_{j<7} e_j k_j and, when
// WithDense, dn = sum_{j<7} d_j k_j for the continuous extension.
// The stage inputs only need x, y, z and the spatial momentum (the RHS reads
// no t, and p_t is constant): they are fused multiply-add chains on y with the
// coefficients scaled by h once per stage. y5 keeps the form y + h (sum),
// whose rounding does not grow with the number of terms.
template <bool WithDense, class R>
__device__ __forceinline__ Field dp45_stages(const double y[8], const double k1[8], double h, const KerrConst& K,
                                             double y5[8], double ep[8], double k7[8], double dn[8]) {
    double k2[8], k3[8], k4[8], k5[8], k6[8], Y[8];
    Y[0] = y[0];
    Y[4] = y[4];
    {
        const double c1 = h * dp::a21;
#pragma unroll
        for (int i = 1; i < 8; ++i)
            if (i != 4) Y[i] = fma(c1, k1[i], y[i]);
    }
    rhs_kerr<R>(Y, K, k2);
    {
        const double c1 = h * dp::a31, c2 = h * dp::a32;
#pragma unroll
        for (int i = 1; i < 8; ++i)
            if (i != 4) Y[i] = fma(c2, k2[i], fma(c1, k1[i], y[i]));
    }
    rhs_kerr<R>(Y, K, k3);
    {
        const double c1 = h * dp::a41, c2 = h * dp::a42, c3 = h * dp::a43;
#pragma unroll
        for (int i = 1; i < 8; ++i)
            if (i != 4) Y[i] = fma(c3, k3[i], fma(c2, k2[i], fma(c1, k1[i], y[i])));
    }
    rhs_kerr<R>(Y, K, k4);
    {
        const double c1 = h * dp::a51, c2 = h * dp::a52, c3 = h * dp::a53, c4 = h * dp::a54;
#pragma unroll
        for (int i = 1; i < 8; ++i)
            if (i != 4) Y[i] = fma(c4, k4[i], fma(c3, k3[i], fma(c2, k2[i], fma(c1, k1[i], y[i]))));
    }
    rhs_kerr<R>(Y, K, k5);
    {
        const double c1 = h * dp::a61, c2 = h * dp::a62, c3 = h * dp::a63, c4 = h * dp::a64, c5 = h * dp::a65;
#pragma unroll
        for (int i = 1; i < 8; ++i)
            if (i != 4) Y[i] = fma(c5, k5[i], fma(c4, k4[i], fma(c3, k3[i], fma(c2, k2[i], fma(c1, k1[i], y[i])))));
    }
    rhs_kerr<R>(Y, K, k6);
    y5[4] = y[4];
    ep[4] = 0.0;
    if (WithDense) dn[4] = 0.0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        if (i == 4) continue;
        y5[i] = y[i] + h * (dp::a71 * k1[i] + dp::a73 * k3[i] + dp::a74 * k4[i] + dp::a75 * k5[i] + dp::a76 * k6[i]);
        ep[i] = dp::e1 * k1[i] + dp::e3 * k3[i] + dp::e4 * k4[i] + dp::e5 * k5[i] + dp::e6 * k6[i];
        if (WithDense) dn[i] = dp::d1 * k1[i] + dp::d3 * k3[i] + dp::d4 * k4[i] + dp::d5 * k5[i] + dp::d6 * k6[i];
    }
    return rhs_kerr<R>(y5, K, k7);
}

// Event location inside an accepted step [y, y5] of size h (single Kerr):
// horizon r = r_hz, equatorial disk z = 0 with r in [disk_in, disk_out],
// escape r = r_esc while outgoing; earliest first. The stages are recomputed
// with the dense-output sums from (y, k1, h). Fills the end state and returns
// the class (0: no event).
__device__ __forceinline__ uint32_t locate_event(const TraceParams& P, const KerrConst& K, const double y[8],
                                                 const double k1[8], double h, double r0, double r1, bool hz, bool dk,
                                                 bool es, double xi, bool f32, double yend[8], double* s_out,
                                                 double* g_out) {
    const double a2 = P.a2;
    double y5[8], ep[8], k7[8], dn[8];
    if (f32)
        dp45_stages<true, float>(y, k1, h, K, y5, ep, k7, dn);
    else
        dp45_stages<true, double>(y, k1, h, K, y5, ep, k7, dn);
    // dense-output coefficients of the position only; other components are
    // evaluated one at a time at the end (keeps the event path's registers low)
    auto coef = [&](int i) {
        Dense q;
        q.D = y5[i] - y[i];
        q.B = h * k1[i] - q.D;
        q.C = q.D - h * k7[i] - q.B;
        q.E = h * (dn[i] + dp::d7 * k7[i]);
        return q;
    };
    const Dense qx = coef(1), qy = coef(2), qz = coef(3);
    auto rad_at = [&](double ss) {
        return ks_radius(dense_at(qx, y[1], ss), dense_at(qy, y[2], ss), dense_at(qz, y[3], ss), a2);
    };
    double best = 2.0, best_g = 0.0;
    uint32_t best_cls = 0;
    if (hz) {
        best = illinois_root([&](double ss) { return rad_at(ss) - P.r_hz; }, r0 - P.r_hz, r1 - P.r_hz);
        best_cls = RT_CLS_CAPTURED;
    }
    if (dk) {
        const double sd = illinois_root([&](double ss) { return dense_at(qz, y[3], ss); }, y[3], y5[3]);
        const double rc = rad_at(sd);
        if (rc >= P.disk_in && rc <= P.disk_out && rc > P.rh102 && sd < best) {
            best = sd;
            best_cls = RT_CLS_DISK;
            best_g = disk_redshift(P.M, P.disk_spin, rc, xi, P.clamp_g != 0);
        }
    }
    if (es) {
        double se = 1.0;
        if (r0 < P.r_esc)
            se = illinois_root([&](double ss) { return rad_at(ss) - P.r_esc; }, r0 - P.r_esc, r1 - P.r_esc);
        if (se < best) {
            best = se;
            best_cls = RT_CLS_ESCAPED;
        }
    }
    if (best_cls == 0) return 0;
    if (best >= 1.0) {
#pragma unroll
        for (int i = 0; i < 8; ++i) yend[i] = y5[i];
    } else {
#pragma unroll
        for (int i = 0; i < 8; ++i) yend[i] = dense_at(coef(i), y[i], best);
    }
    *s_out = best;
    *g_out = best_g;
    return best_cls;
}

// Step-size controller for both outcomes of an attempt (one FP32 log2 and
// exp2 for the power): accepted, 0.9 err^(-1/5) within [0.2, 5] (5 for
// err = 0) and |h| within [h_min, h_max]; rejected, 0.9 err^(-1/4) within
// [0.1, 0.9] (integrators/step_controller.hpp:44-78). FP64: the power is
// refined by one Newton step on y^n e = 1 (about 6e-13 relative); Fast
// (mixed policy): the FP32 power as it is.
template <bool Fast>
__device__ __forceinline__ double next_step_any(double ah, double err, bool acc, double h_min, double h_max) {
    const double e = fmax(err, 1e-12);
    const float le = __builtin_amdgcn_logf(static_cast<float>(e));  // log2(e)
    double yv = static_cast<double>(__builtin_amdgcn_exp2f((acc ? -0.2f : -0.25f) * le));
    if (!Fast) {
        const double y2 = yv * yv;
        const double y4 = y2 * y2;
        const double yn = acc ? y4 * yv : y4;
        yv = yv * fma(1.0 - yn * e, acc ? 0.2 : 0.25, 1.0);
    }
    const double factor = (acc && err == 0.0) ? 5.0 : yv;
    const double s = acc ? fmin(fmax(0.9 * factor, 0.2), 5.0) : fmin(fmax(0.9 * factor, 0.1), 0.9);
    const double nh = ah * s;
    return acc ? fmin(fmax(nh, h_min), h_max) : nh;
}

// Start a ray: first RHS (FSAL seed) and initial step.
template <bool Carter>
__device__ __forceinline__ void lane_begin(const TraceParams& P, const KerrConst& K, const double y0[8], double xi,
                                           Lane& L) {
#pragma unroll
    for (int i = 0; i < 8; ++i) L.y[i] = y0[i];
    const Field f0 = rhs_kerr<double>(L.y, K, L.k1);
    L.r0 = f0.r;
    const double ah = (P.h0 > 0.0) ? P.h0 : initial_step(L.r0, P.rh, P.M);
    L.ah = fmin(fmax(ah, P.h_min), P.h_max);
    L.xi = xi;
    L.lambda = 0.0;
    L.g = 0.0;
    L.acc = L.rej = L.flags = L.cls = 0;
    L.rhs = 1;
    L.proj = L.f32 = 0;
    if constexpr (Carter) {
        if (P.carter && carter_certain_capture(L.y, L.r0, P.M, P.a, P.rh, kCarterMargin)) L.flags = kCarterNoTurn;
    }
}

// One DP45 attempt of a single-Kerr ray. Returns true when the ray ended:
// L.y then holds the end state and L.cls its class. Mixed: an attempt that
// starts beyond switch_r evaluates the RHS in FP32 (state, stage sums, error
// test, projection and event location stay FP64).
//
// An accepted attempt whose end shows an event (r below r_hz, a sign change
// of z with the disk on, r beyond r_esc moving out) locates it from the state
// before the attempt; the attempt's own results are not kept live across the
// location (they would double the registers of the whole loop). When the
// sign change of z is a crossing off the disk, the lane repeats the attempt,
// which gives the same step bit for bit, and takes it with the event test off.
//
// Every outcome (out of budget, non-finite, event, Carter capture, plain
// accept or reject) is decided as flags and the lane state is merged once at
// the end: with a return per outcome the compiler keeps a copy of the whole
// lane state on each exit path, and those copies run on every attempt. A lane
// out of budget still computes the attempt (its wave runs it anyway) and
// discards it.
template <bool Mixed>
__device__ __forceinline__ bool lane_step(const TraceParams& P, const KerrConst& K, Lane& L) {
    const double ah = L.ah;
    const bool over = (L.acc + L.rej >= P.max_steps) || (ah < P.h_min);
    const double h = -ah;
    double y5[8], ep[8], k7[8];
    const bool f32 = Mixed && P.mixed && (L.r0 > P.switch_r);
    Field f7 = f32 ? dp45_stages<false, float>(L.y, L.k1, h, K, y5, ep, k7, nullptr)
                   : dp45_stages<false, double>(L.y, L.k1, h, K, y5, ep, k7, nullptr);
    // err = |h| max_i |sum_j e_j k_j,i| / (rtol max(|y_i|, |y5_i|) + atol), p_t excluded.
    // The seven ratios share one reciprocal: e_j / sc_j = e_j prod_{k != j} sc_k / prod_k sc_k
    // (prefix and suffix products). The plain form takes over when the products leave the
    // normal range.
    double e[7], sc[7];
    bool finite = true;
#pragma unroll
    for (int i = 0, j = 0; i < 8; ++i) {
        finite = finite && isfinite(y5[i]);
        if (i == 4) continue;  // dp_t/dl = 0: no error in p_t
        e[j] = fabs(fma(dp::e7, k7[i], ep[i]));
        sc[j] = fma(P.rtol, fmax(fabs(L.y[i]), fabs(y5[i])), P.atol);
        ++j;
    }
    double pre[7], suf[7];
    pre[1] = sc[0];
#pragma unroll
    for (int j = 2; j < 7; ++j) pre[j] = pre[j - 1] * sc[j - 1];
    suf[5] = sc[6];
#pragma unroll
    for (int j = 4; j >= 0; --j) suf[j] = suf[j + 1] * sc[j + 1];
    double num = e[0] * suf[0];
#pragma unroll
    for (int j = 1; j < 6; ++j) num = fmax(num, e[j] * (pre[j] * suf[j]));
    num = fmax(num, e[6] * pre[6]);
    const double den = pre[6] * sc[6];
    double err;
    if (den >= 1e-290 && den <= 1e290 && num <= 1e290) {
        err = (ah * num) * rcp64(den);
    } else {
        err = 0.0;
#pragma unroll
        for (int j = 0; j < 7; ++j) err = fmax(err, e[j] / sc[j]);
        err *= ah;
    }
    finite = finite && isfinite(err);
    const bool acc = !over && finite && err <= 1.0;
    const double r1 = f7.r;
    bool ended = over || (!over && !finite);
    bool retry = false;
    uint32_t cls = RT_CLS_CAPTURED;
    double dl = ah;
    uint32_t nrhs = 6;
    if (acc && !(L.flags & kNoEventRetry)) {
        const double z0 = L.y[3], z1 = y5[3];
        const bool hz = (r1 <= P.r_hz);
        const bool dk = P.disk_on && ((z0 < 0.0 && z1 >= 0.0) || (z0 > 0.0 && z1 <= 0.0));
        const bool es = (r1 >= P.r_esc) && (-(y5[1] * k7[1] + y5[2] * k7[2] + y5[3] * k7[3]) > 0.0);
        if (hz || dk || es) {
            double s_ev = 1.0, g_ev = 0.0;
            const uint32_t c = locate_event(P, K, L.y, L.k1, h, L.r0, r1, hz, dk, es, L.xi, f32, y5, &s_ev, &g_ev);
            if (c == 0) {
                retry = true;
            } else {
                ended = true;
                cls = c;
                L.g = g_ev;
                dl = s_ev * ah;
                nrhs = 12;  // the attempt and the dense recomputation
                f7 = ks_field(y5[1], y5[2], y5[3], P.M, P.a, P.a2);
            }
        }
    }
    if (acc && !retry) {
        L.flags &= ~kNoEventRetry;
        ++L.acc;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            L.y[i] = y5[i];
            L.k1[i] = k7[i];
        }
        L.lambda += dl;
        if (project_pt(L.y, f7.H, f7.k0 * L.y[5] + f7.k1 * L.y[6] + f7.k2 * L.y[7], P.inv_tol)) {
            L.flags |= RT_FLAG_PROJECTED;
            ++L.proj;
            if (!ended) {
                if (f32)
                    rhs_kerr<float>(L.y, K, L.k1);
                else
                    rhs_kerr<double>(L.y, K, L.k1);
                ++L.rhs;
            }
        }
        // RT_TERM_CARTER_CAPTURE: no turning point, moving inward inside the gate
        if (!ended && (L.flags & kCarterNoTurn) && r1 < L.r0 && r1 < P.carter_gate) {
            ended = true;
            L.flags |= RT_FLAG_CARTER;
        }
        L.r0 = r1;
    }
    if (over) {
        L.flags |= (L.acc + L.rej >= P.max_steps) ? RT_FLAG_BUDGET : RT_FLAG_MIN_STEP;
        cls = (L.r0 < P.near_r) ? RT_CLS_CAPTURED : RT_CLS_ESCAPED;
    } else if (!finite) {
        L.flags |= RT_FLAG_NONFINITE;
    } else if (!acc) {
        ++L.rej;
    }
    if (retry) {
        L.flags |= kNoEventRetry;
    } else if (!over) {
        L.f32 += f32 ? 1u : 0u;
        L.rhs += nrhs;
    }
    if (ended) {
        L.cls = cls;
    } else if (!retry) {
        L.ah = (Mixed && P.mixed) ? next_step_any<true>(ah, err, acc, P.h_min, P.h_max)
                                  : next_step_any<false>(ah, err, acc, P.h_min, P.h_max);
    }
    return ended;
}

// Redshift per class and the final flag word.
__device__ __forceinline__ void lane_finish(Lane& L) {
    const uint32_t c = L.cls & RT_CLS_MASK;
    if (c == RT_CLS_ESCAPED) L.g = 1.0 / (-L.y[4]);
    if (c == RT_CLS_CAPTURED) L.g = 0.0;
    L.cls |= L.flags & ~(kCarterNoTurn | kNoEventRetry);
    if (L.f32) L.cls |= RT_FLAG_FP32_STEPS;
}

__device__ __forceinline__ void store_lane(const OutPtrs& O, uint64_t idx, const Lane& L, double a, double a2) {
    const uint64_t n = O.n;
    O.cls[idx] = L.cls;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        O.x[c * n + idx] = L.y[c];
        O.p[c * n + idx] = L.y[4 + c];
    }
    double bl[3];
    bl_from_ks(L.y[1], L.y[2], L.y[3], a, a2, bl);
#pragma unroll
    for (int c = 0; c < 3; ++c) O.bl[c * n + idx] = bl[c];
    O.lambda[idx] = L.lambda;
    O.g[idx] = L.g;
    O.steps[idx] = L.acc;
    O.rejected[idx] = L.rej;
}

// Find the queue segment holding queue index g (req_offset[s] <= g < req_offset[s + 1]).
__device__ __forceinline__ uint32_t find_segment(const uint64_t* req_offset, uint32_t n_seg, uint64_t g) {
    uint32_t lo = 0, hi = n_seg;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (req_offset[mid] <= g)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// ------------------------------------------------------- plain kernels
// Camera frame, single Kerr. Thread t -> 8x8 tile t/64, lane t%64.
template <bool Mixed, int Waves, bool Carter>
__global__ __launch_bounds__(256, Waves) void k_trace_camera_kerr(TraceParams P, OutPtrs O) {
    const uint64_t t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint32_t lane = static_cast<uint32_t>(t & 63);
    uint32_t tx, ty;
    tile_xy(P, static_cast<uint32_t>(t >> 6), tx, ty);
    const uint32_t px = tx * 8 + (lane & 7);
    const uint32_t py = ty * 8 + (lane >> 3);
    const uint32_t rows = P.n_rays / P.cols;
    Tally tally;
    if (px < P.cols && py < rows) {
        double y0[8], xi;
        camera_state_kerr(P, P.col_first + px, P.row_first + py, y0, &xi);
        const KerrConst K = kerr_const(P.M, P.a, P.a2);
        Lane L;
        lane_begin<Carter>(P, K, y0, xi, L);
        while (!lane_step<Mixed>(P, K, L)) {
        }
        lane_finish(L);
        store_lane(O, static_cast<uint64_t>(py) * P.cols + px, L, P.a, P.a2);
        tally.add(L);
    }
    tally.flush(O.counters);
}

// Explicit initial states, single Kerr.
template <bool Mixed, int Waves, bool Carter>
__global__ __launch_bounds__(256, Waves) void k_trace_rays_kerr(TraceParams P, OutPtrs O) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    Tally tally;
    if (i < P.n_rays) {
        double y0[8];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            y0[c] = O.in_x[c * O.n + i];
            y0[4 + c] = O.in_p[c * O.n + i];
        }
        const double xi = (y0[1] * y0[6] - y0[2] * y0[5]) / (-y0[4]);
        const KerrConst K = kerr_const(P.M, P.a, P.a2);
        Lane L;
        lane_begin<Carter>(P, K, y0, xi, L);
        while (!lane_step<Mixed>(P, K, L)) {
        }
        lane_finish(L);
        store_lane(O, i, L, P.a, P.a2);
        tally.add(L);
    }
    tally.flush(O.counters);
}

// ------------------------------------------------------ work-queue kernel
// Persistent: every wave loops until the queue is drained and its lanes are
// idle. Idle lanes are refilled together: one ballot, one atomicAdd on the
// queue head for the whole wave, each idle lane taking base + (its rank among
// the idle lanes). A refill happens once at least refill_min lanes are idle
// (or none is active), which bounds the atomics and the divergent set-up code.
// Each lane carries its own request; the request's parameters are read from
// the device table (cached), so lanes of different requests share a wave.
template <bool Mixed, int Waves, bool Carter>
__global__ __launch_bounds__(256, Waves) void k_trace_queue(QueueDev Q) {
    const uint32_t lane_id = threadIdx.x & 63;
    const uint64_t lt_mask = (lane_id == 0) ? 0ull : (~0ull >> (64 - lane_id));
    Lane L;
    bool active = false;
    bool drained = false;
    uint32_t req = 0;
    uint64_t ray = 0;
    Tally tally;
    for (;;) {
        const uint64_t idle = __ballot(!active);
        const uint32_t n_idle = __popcll(idle);
        if (!drained && n_idle != 0 && (n_idle >= Q.refill_min || n_idle == 64)) {
            const uint32_t leader = __ffsll(static_cast<unsigned long long>(idle)) - 1;
            unsigned long long base = 0;
            if (lane_id == leader) base = atomicAdd(Q.next, static_cast<unsigned long long>(n_idle));
            base = __shfl(base, leader, 64);
            if (base + n_idle >= Q.total) drained = true;
            if (!active) {
                const uint64_t g = base + __popcll(idle & lt_mask);
                if (g < Q.total) {
                    const uint32_t seg = find_segment(Q.req_offset, 2 * Q.n_req, g);
                    const bool hot = seg < Q.n_req;
                    req = hot ? seg : seg - Q.n_req;
                    ray = g - Q.req_offset[seg];
                    const TraceParams& P = Q.reqs[req];
                    double y0[8], xi = 0.0;
                    bool ok = true;
                    if (P.mode == 0) {
                        const uint32_t l = static_cast<uint32_t>(ray & 63);
                        uint32_t tx, ty;
                        tile_xy(P, static_cast<uint32_t>(ray >> 6) + (hot ? 0u : P.hot_w * P.hot_h), tx, ty);
                        const uint32_t px = tx * 8 + (l & 7);
                        const uint32_t py = ty * 8 + (l >> 3);
                        ok = (px < P.cols) && (py < P.n_rays / P.cols);
                        if (ok) {
                            camera_state_kerr<true>(P, P.col_first + px, P.row_first + py, y0, &xi);
                            ray = static_cast<uint64_t>(py) * P.cols + px;
                        }
                    } else {
                        const OutPtrs& O = Q.outs[req];
#pragma unroll
                        for (int c = 0; c < 4; ++c) {
                            y0[c] = O.in_x[c * O.n + ray];
                            y0[4 + c] = O.in_p[c * O.n + ray];
                        }
                        xi = (y0[1] * y0[6] - y0[2] * y0[5]) / (-y0[4]);
                    }
                    if (ok) {
                        lane_begin<Carter>(P, kerr_const(P.M, P.a, P.a2), y0, xi, L);
                        active = true;
                    }
                }
            }
        }
        if (__ballot(active) == 0) {
            if (drained) break;
            continue;
        }
        if (active) {
            const TraceParams& P = Q.reqs[req];
            if (lane_step<Mixed>(P, kerr_const(P.M, P.a, P.a2), L)) {
                lane_finish(L);
                store_lane(Q.outs[req], ray, L, P.a, P.a2);
                tally.add(L);
                active = false;
            }
        }
    }
    tally.flush(Q.counters);
}

// Tile-granular work queue: a wave takes one 8x8 tile (64 rays of one
// request) per atomicAdd and traces it to the end before taking the next. The
// request index is wave-uniform (readfirstlane), so its parameters load into
// scalar registers as in the single-request kernel; the persistent loop
// refills a wave's slot the moment its tile is done. Q.req_offset counts tiles.
__device__ __forceinline__ uint64_t uniform_u64(uint64_t v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v >> 32));
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

// One camera request as a persistent kernel over its own tile queue: a wave
// takes one 8x8 tile per atomicAdd on *next (zero at launch) and traces it as
// k_trace_camera_kerr would, until the frame's tiles are gone. The grid is
// whatever number of workgroups the caller grants (the serving loop's share
// of the device's wave slots); every workgroup is resident at once, so the
// kernel never holds the dispatcher while other requests' kernels wait. With
// a limit, waves leave between tiles when the host lowers it (wave_yields).
template <bool Mixed, bool Carter>
__global__ __launch_bounds__(256, 2) void k_trace_camera_queue(TraceParams P, OutPtrs O,
                                                               unsigned long long* __restrict__ next,
                                                               uint32_t n_tiles, const uint32_t* limit,
                                                               unsigned long long* active) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t rows = P.n_rays / P.cols;
    Tally tally;
    wave_joins(limit, active);
    for (;;) {
        if (wave_yields(limit, active)) break;
        unsigned long long t = 0;
        if (lane == 0) t = atomicAdd(next, 1ull);
        t = uniform_u64(__shfl(t, 0, 64));
        if (t >= n_tiles) break;
        uint32_t tx, ty;
        tile_xy(P, static_cast<uint32_t>(t), tx, ty);
        const uint32_t px = tx * 8 + (lane & 7);
        const uint32_t py = ty * 8 + (lane >> 3);
        if (px < P.cols && py < rows) {
            double y0[8], xi;
            camera_state_kerr<true>(P, P.col_first + px, P.row_first + py, y0, &xi);
            const KerrConst K = kerr_const(P.M, P.a, P.a2);
            Lane L;
            lane_begin<Carter>(P, K, y0, xi, L);
            while (!lane_step<Mixed>(P, K, L)) {
            }
            lane_finish(L);
            store_lane(O, static_cast<uint64_t>(py) * P.cols + px, L, P.a, P.a2);
            tally.add(L);
        }
    }
    tally.flush(O.counters);
}

// The request tables come in as separate read-only, non-aliased arguments so
// that their wave-uniform loads stay scalar loads the compiler may repeat
// instead of holding the parameters in vector registers.
template <bool Mixed, int Waves, bool Carter>
__global__ __launch_bounds__(256, Waves) void k_trace_tiles(QueueDev Q, const TraceParams* __restrict__ reqs,
                                                            const OutPtrs* __restrict__ outs,
                                                            const uint64_t* __restrict__ req_offset) {
    const uint32_t lane = threadIdx.x & 63;
    // statistics accumulate per tile as wave sums, in scalar registers
    unsigned long long s_acc = 0, s_rej = 0, s_rhs = 0, s_f32 = 0, s_proj = 0;
    for (;;) {
        unsigned long long t = 0;
        if (lane == 0) t = atomicAdd(Q.next, 1ull);
        t = uniform_u64(__shfl(t, 0, 64));
        if (t >= Q.total) break;
        // batches hold camera requests only (plan_batch): segment seg, tile t of it
        const uint32_t seg = find_segment(req_offset, 2 * Q.n_req, t);
        const bool hot = seg < Q.n_req;
        const uint32_t req = hot ? seg : seg - Q.n_req;
        const TraceParams& P = reqs[req];
        uint32_t tx, ty;
        tile_xy(P, static_cast<uint32_t>(t - req_offset[seg]) + (hot ? 0u : P.hot_w * P.hot_h), tx, ty);
        const uint32_t px = tx * 8 + (lane & 7);
        const uint32_t py = ty * 8 + (lane >> 3);
        uint32_t c_acc = 0, c_rej = 0, c_rhs = 0, c_f32 = 0, c_proj = 0;
        if (px < P.cols && py < P.n_rays / P.cols) {
            double y0[8], xi = 0.0;
            camera_state_kerr<true>(P, P.col_first + px, P.row_first + py, y0, &xi);
            const KerrConst K = kerr_const(P.M, P.a, P.a2);
            Lane L;
            lane_begin<Carter>(P, K, y0, xi, L);
            while (!lane_step<Mixed>(P, K, L)) {
            }
            lane_finish(L);
            store_lane(outs[req], static_cast<uint64_t>(py) * P.cols + px, L, P.a, P.a2);
            c_acc = L.acc;
            c_rej = L.rej;
            c_rhs = L.rhs;
            c_f32 = L.f32;
            c_proj = L.proj;
        }
        s_acc += uniform_u64(wave_sum(c_acc));
        s_rej += uniform_u64(wave_sum(c_rej));
        s_rhs += uniform_u64(wave_sum(c_rhs));
        s_f32 += uniform_u64(wave_sum(c_f32));
        s_proj += uniform_u64(wave_sum(c_proj));
    }
    if (lane == 0) {
        atomicAdd(&Q.counters[0], s_acc);
        atomicAdd(&Q.counters[1], s_rej);
        atomicAdd(&Q.counters[2], s_rhs);
        atomicAdd(&Q.counters[3], s_f32);
        atomicAdd(&Q.counters[4], s_proj);
    }
}

// ------------------------------------------------------------ launchers

template <bool Mixed, bool Carter>
static hipError_t launch_camera_t(const TraceParams& P, const OutPtrs& O, const LaunchShape& L, hipStream_t stream) {
    const uint64_t tiles_y = (P.n_rays / P.cols + 7) / 8;
    const uint64_t threads = static_cast<uint64_t>(P.tiles_x) * tiles_y * 64;
    const uint32_t grid = static_cast<uint32_t>((threads + L.block - 1) / L.block);
    if (L.waves >= 2)
        hipLaunchKernelGGL((k_trace_camera_kerr<Mixed, 2, Carter>), dim3(grid), dim3(L.block), 0, stream, P, O);
    else
        hipLaunchKernelGGL((k_trace_camera_kerr<Mixed, 1, Carter>), dim3(grid), dim3(L.block), 0, stream, P, O);
    return hipGetLastError();
}

hipError_t launch_trace_camera(const TraceParams& P, const OutPtrs& O, const LaunchShape& L, hipStream_t stream) {
    if (!shape_ok(L) || P.cols == 0 || P.n_rays % P.cols != 0) return hipErrorInvalidValue;
    if (P.carter)
        return P.mixed ? launch_camera_t<true, true>(P, O, L, stream) : launch_camera_t<false, true>(P, O, L, stream);
    return P.mixed ? launch_camera_t<true, false>(P, O, L, stream) : launch_camera_t<false, false>(P, O, L, stream);
}

hipError_t launch_trace_camera_queue(const TraceParams& P, const OutPtrs& O, unsigned long long* next,
                                     uint32_t grid, hipStream_t stream, const uint32_t* limit,
                                     unsigned long long* active) {
    if (P.cols == 0 || P.n_rays % P.cols != 0 || grid == 0) return hipErrorInvalidValue;
    const uint64_t tiles = static_cast<uint64_t>(P.tiles_x) * ((P.n_rays / P.cols + 7) / 8);
    if (tiles > 0xffffffffull) return hipErrorInvalidValue;
    const uint32_t n_tiles = static_cast<uint32_t>(tiles);
    grid = std::min(grid, n_tiles);
    if (P.carter) {
        if (P.mixed)
            hipLaunchKernelGGL((k_trace_camera_queue<true, true>), dim3(grid), dim3(64), 0, stream, P, O, next, n_tiles,
                               limit, active);
        else
            hipLaunchKernelGGL((k_trace_camera_queue<false, true>), dim3(grid), dim3(64), 0, stream, P, O, next, n_tiles,
                               limit, active);
    } else if (P.mixed) {
        hipLaunchKernelGGL((k_trace_camera_queue<true, false>), dim3(grid), dim3(64), 0, stream, P, O, next, n_tiles,
                               limit, active);
    } else {
        hipLaunchKernelGGL((k_trace_camera_queue<false, false>), dim3(grid), dim3(64), 0, stream, P, O, next, n_tiles,
                               limit, active);
    }
    return hipGetLastError();
}

hipError_t wave_slots(uint32_t* slots) {
    int dev = 0, cus = 0, a = 0, b = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &a, reinterpret_cast<const void*>(&k_trace_camera_queue<false, false>), 64, 0);
    if (e == hipSuccess) e = multi_blocks_per_cu(&b);
    if (e != hipSuccess) return e;
    *slots = static_cast<uint32_t>(cus) * static_cast<uint32_t>(std::max(1, std::min(a, b)));
    return hipSuccess;
}

template <bool Mixed, bool Carter>
static hipError_t launch_rays_t(const TraceParams& P, const OutPtrs& O, const LaunchShape& L, hipStream_t stream) {
    const uint32_t grid = static_cast<uint32_t>((static_cast<uint64_t>(P.n_rays) + L.block - 1) / L.block);
    if (L.waves >= 2)
        hipLaunchKernelGGL((k_trace_rays_kerr<Mixed, 2, Carter>), dim3(grid), dim3(L.block), 0, stream, P, O);
    else
        hipLaunchKernelGGL((k_trace_rays_kerr<Mixed, 1, Carter>), dim3(grid), dim3(L.block), 0, stream, P, O);
    return hipGetLastError();
}

hipError_t launch_trace_rays(const TraceParams& P, const OutPtrs& O, const LaunchShape& L, hipStream_t stream) {
    if (!shape_ok(L)) return hipErrorInvalidValue;
    if (P.carter)
        return P.mixed ? launch_rays_t<true, true>(P, O, L, stream) : launch_rays_t<false, true>(P, O, L, stream);
    return P.mixed ? launch_rays_t<true, false>(P, O, L, stream) : launch_rays_t<false, false>(P, O, L, stream);
}

template <bool Mixed, bool Carter>
static const void* queue_fn(const LaunchShape& L) {
    if (L.queue_kind == 1)
        return (L.waves >= 2) ? reinterpret_cast<const void*>(&k_trace_queue<Mixed, 2, Carter>)
                              : reinterpret_cast<const void*>(&k_trace_queue<Mixed, 1, Carter>);
    return (L.waves >= 2) ? reinterpret_cast<const void*>(&k_trace_tiles<Mixed, 2, Carter>)
                          : reinterpret_cast<const void*>(&k_trace_tiles<Mixed, 1, Carter>);
}

template <bool Mixed, bool Carter>
static hipError_t launch_queue_t(const QueueDev& Q, const LaunchShape& L, uint32_t grid, hipStream_t stream) {
    if (L.queue_kind == 1) {
        if (L.waves >= 2)
            hipLaunchKernelGGL((k_trace_queue<Mixed, 2, Carter>), dim3(grid), dim3(L.block), 0, stream, Q);
        else
            hipLaunchKernelGGL((k_trace_queue<Mixed, 1, Carter>), dim3(grid), dim3(L.block), 0, stream, Q);
    } else {
        if (L.waves >= 2)
            hipLaunchKernelGGL((k_trace_tiles<Mixed, 2, Carter>), dim3(grid), dim3(L.block), 0, stream, Q, Q.reqs,
                               Q.outs, Q.req_offset);
        else
            hipLaunchKernelGGL((k_trace_tiles<Mixed, 1, Carter>), dim3(grid), dim3(L.block), 0, stream, Q, Q.reqs,
                               Q.outs, Q.req_offset);
    }
    return hipGetLastError();
}

hipError_t queue_grid(const LaunchShape& L, bool mixed, bool carter, uint32_t* grid) {
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    const void* fn = carter ? (mixed ? queue_fn<true, true>(L) : queue_fn<false, true>(L))
                            : (mixed ? queue_fn<true, false>(L) : queue_fn<false, false>(L));
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, static_cast<int>(L.block), 0);
    if (e != hipSuccess) return e;
    *grid = static_cast<uint32_t>(cus) * static_cast<uint32_t>(per_cu > 0 ? per_cu : 1);
    return hipSuccess;
}

hipError_t launch_trace_queue(const QueueDev& Q, const LaunchShape& L, bool mixed, bool carter, uint32_t grid,
                              hipStream_t stream) {
    if (!shape_ok(L) || Q.n_req == 0) return hipErrorInvalidValue;
    if (grid == 0) {
        const hipError_t e = queue_grid(L, mixed, carter, &grid);
        if (e != hipSuccess) return e;
    }
    if (carter)
        return mixed ? launch_queue_t<true, true>(Q, L, grid, stream) : launch_queue_t<false, true>(Q, L, grid, stream);
    return mixed ? launch_queue_t<true, false>(Q, L, grid, stream) : launch_queue_t<false, false>(Q, L, grid, stream);
}

// Zero a small counter array with a kernel (safe inside a captured graph).
__global__ void k_zero_u64(unsigned long long* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// Results -> the reference's 32-byte pixel record: (0, 0, 0, 1) colour,
// float redshift and affine parameter, PixelFlags bits, accepted steps. Two
// 16-byte stores per ray.
__global__ void k_pack_pixels(OutPtrs O, uint64_t n, uint4* __restrict__ out) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = static_cast<float>(O.g[i]);
    const float lam = static_cast<float>(O.lambda[i]);
    out[2 * i] = make_uint4(0u, 0u, 0u, __float_as_uint(1.0f));
    out[2 * i + 1] = make_uint4(__float_as_uint(g), __float_as_uint(lam), O.cls[i] & 0xFFu, O.steps[i]);
}

hipError_t launch_pack_pixels(const OutPtrs& O, uint64_t n, void* out, hipStream_t stream) {
    const uint64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(k_pack_pixels, dim3(static_cast<uint32_t>(blocks)), dim3(256), 0, stream, O, n,
                       static_cast<uint4*>(out));
    return hipGetLastError();
}

hipError_t launch_zero_u64(unsigned long long* p, uint32_t n, hipStream_t stream) {
    hipLaunchKernelGGL(k_zero_u64, dim3((n + 63) / 64), dim3(64), 0, stream, p, n);
    return hipGetLastError();
}

// Copy up to kMaxCopySegs byte ranges (device <-> mapped pinned host) in one
// launch: 16-byte moves where source and destination are both 16-byte
// aligned, the rest in 4-byte moves. Grid-stride over each segment in turn.
__global__ void k_copy_segments(CopySegs S) {
    const uint64_t tid = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint32_t k = 0; k < S.n; ++k) {
        const char* src = static_cast<const char*>(S.seg[k].src);
        char* dst = static_cast<char*>(S.seg[k].dst);
        const uint64_t bytes = S.seg[k].bytes;
        uint64_t vec = 0;
        if (((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 15) == 0) {
            vec = bytes / 16;
            const uint4* s4 = reinterpret_cast<const uint4*>(src);
            uint4* d4 = reinterpret_cast<uint4*>(dst);
            for (uint64_t i = tid; i < vec; i += stride) d4[i] = s4[i];
        }
        const uint64_t done = vec * 16;
        const uint64_t words = (bytes - done) / 4;
        const uint32_t* s1 = reinterpret_cast<const uint32_t*>(src + done);
        uint32_t* d1 = reinterpret_cast<uint32_t*>(dst + done);
        for (uint64_t i = tid; i < words; i += stride) d1[i] = s1[i];
    }
}

hipError_t launch_copy_segments(const CopySegs& S, hipStream_t stream, uint32_t max_blocks) {
    if (S.n > static_cast<uint32_t>(kMaxCopySegs)) return hipErrorInvalidValue;
    uint64_t total = 0;
    for (uint32_t k = 0; k < S.n; ++k) {
        if ((S.seg[k].bytes & 3) != 0 || (reinterpret_cast<uintptr_t>(S.seg[k].src) & 3) != 0 ||
            (reinterpret_cast<uintptr_t>(S.seg[k].dst) & 3) != 0)
            return hipErrorInvalidValue;
        total += S.seg[k].bytes;
    }
    if (total == 0) return hipSuccess;
    const uint64_t want = (total / 16 + 255) / 256;
    const uint32_t grid = static_cast<uint32_t>(
        std::max<uint64_t>(1, std::min<uint64_t>(want, std::max<uint32_t>(1, std::min<uint32_t>(max_blocks, 256)))));
    hipLaunchKernelGGL(k_copy_segments, dim3(grid), dim3(256), 0, stream, S);
    return hipGetLastError();
}

}  // namespace rt
