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
//   k_trace_multi        superposition of up to 16 sources: G lanes per ray
//                        (one source and a share of the state components
//                        each, sums by DPP butterflies), persistent, idle
//                        groups refilled ray by ray
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

// Lane-internal flag bit (not exported): the event test found a sign change
// in the last accepted attempt but no event inside it (a crossing of z = 0
// off the disk); the attempt is repeated and taken as a plain step.
constexpr uint32_t kNoEventRetry = 1u << 29;

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

struct Tally {
    unsigned long long acc = 0, rej = 0, rhs = 0, f32 = 0, proj = 0;
    template <class LaneT>
    __device__ void add(const LaneT& L) {
        acc += L.acc;
        rej += L.rej;
        rhs += L.rhs;
        f32 += L.f32;
        proj += L.proj;
    }
    __device__ void flush(unsigned long long* counters) {
        const unsigned long long sa = wave_sum(acc), sr = wave_sum(rej), sh = wave_sum(rhs), sf = wave_sum(f32),
                                 sp = wave_sum(proj);
        if ((threadIdx.x & 63) == 0) {
            atomicAdd(&counters[0], sa);
            atomicAdd(&counters[1], sr);
            atomicAdd(&counters[2], sh);
            atomicAdd(&counters[3], sf);
            atomicAdd(&counters[4], sp);
        }
    }
};

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

// Wave-level yield of the serving loop's persistent kernels. *active counts
// the request's running waves; a wave leaves when more are running than the
// host's current limit for the request (a word in mapped pinned host memory,
// read at system scope, lowered by the host to make room for other requests).
// The host keeps the limit at 1 or more, so the request's last waves always
// drain its queue. limit == nullptr: never yield.
__device__ __forceinline__ bool wave_yields(const uint32_t* limit, unsigned long long* active) {
    if (limit == nullptr) return false;
    int go = 0;
    if ((threadIdx.x & 63) == 0) {
        const uint32_t lim = __hip_atomic_load(limit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        const unsigned long long a = __hip_atomic_load(active, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (a > lim && atomicCAS(active, a, a - 1) == a) go = 1;
    }
    return __shfl(go, 0, 64) != 0;
}

__device__ __forceinline__ void wave_joins(const uint32_t* limit, unsigned long long* active) {
    if (limit != nullptr && (threadIdx.x & 63) == 0) atomicAdd(active, 1ull);
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

// ------------------------------------------------------- superposition
// Component-distributed layout. The G lanes of a group (G = 2, 4 or 8; see
// multi_ray_lanes) trace one ray together. Lane l of the group evaluates
// source l (and source l + 8 in groups of 8), and owns the state components
// c = q + Gc m (m < C) of the DP45 step, with q = l, Gc = G and C = 8 / G: it
// keeps their values, their seven stages and their stage sums, so no lane
// holds the whole step. Per right-hand side the group gathers the position
// and momentum from their owners (every lane needs them for its sources),
// each lane adds its sources' terms, and the sums over the sources are
// reduced onto the owners of the components (reduce-scatter by DPP
// butterflies, no memory). Values every lane needs
// (p_t, the current position and momentum, the step size, counters) are kept
// in every lane. Floating-point addition is commutative bit for bit, so both
// partners of a butterfly level hold identical bits: every group-wide value
// is the same in all lanes of the group, and the lanes take the same
// accept/reject and event decisions in lock step. A ray's latency per step is
// that of its lane's sources and one share of the stage arithmetic, which
// matters because a frame with few sources is bound by its slowest rays (rays
// that graze a horizon take thousands of steps).

// DPP lane moves inside rows of 16: xor 1 and xor 2 (quad permutes), mirror
// inside 8 lanes and inside 16 lanes. Each pairs every lane with a partner in
// the same aligned group of 2, 4, 8 or 16 lanes.
constexpr int kDppXor1 = 0xB1;        // quad_perm [1,0,3,2]
constexpr int kDppXor2 = 0x4E;        // quad_perm [2,3,0,1]
constexpr int kDppHalfMirror = 0x141; // row_half_mirror
constexpr int kDppMirror = 0x140;     // row_mirror

template <int Ctrl>
__device__ __forceinline__ double dpp_f64(double v) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(static_cast<uint32_t>(b)), Ctrl, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(static_cast<uint32_t>(b >> 32)), Ctrl, 0xF, 0xF,
                                               false);
    return __builtin_bit_cast(double, (static_cast<uint64_t>(static_cast<uint32_t>(hi)) << 32) |
                                          static_cast<uint32_t>(lo));
}

template <int Ctrl>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v) {
    return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), Ctrl, 0xF, 0xF, false));
}

// The value of v in the lane whose byte address (4 x lane) is addr.
__device__ __forceinline__ double bperm_f64(int addr, double v) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const int lo = __builtin_amdgcn_ds_bpermute(addr, static_cast<int>(static_cast<uint32_t>(b)));
    const int hi = __builtin_amdgcn_ds_bpermute(addr, static_cast<int>(static_cast<uint32_t>(b >> 32)));
    return __builtin_bit_cast(double, (static_cast<uint64_t>(static_cast<uint32_t>(hi)) << 32) |
                                          static_cast<uint32_t>(lo));
}

// Sum over the group (identical bits in every lane of the group).
template <int G>
__device__ __forceinline__ double group_sum(double v) {
    if constexpr (G == 1) return v;
    v = v + dpp_f64<kDppXor1>(v);
    if constexpr (G >= 4) v = v + dpp_f64<kDppXor2>(v);
    if constexpr (G >= 8) v = v + dpp_f64<kDppHalfMirror>(v);
    if constexpr (G >= 16) v = v + dpp_f64<kDppMirror>(v);
    return v;
}

template <int G>
__device__ __forceinline__ double group_min(double v) {
    if constexpr (G == 1) return v;
    v = fmin(v, dpp_f64<kDppXor1>(v));
    if constexpr (G >= 4) v = fmin(v, dpp_f64<kDppXor2>(v));
    if constexpr (G >= 8) v = fmin(v, dpp_f64<kDppHalfMirror>(v));
    if constexpr (G >= 16) v = fmin(v, dpp_f64<kDppMirror>(v));
    return v;
}

// Earliest event of the group: smallest s, ties to the lower source index
// (the order in which a sequential scan over the sources would keep them).
template <int Ctrl>
__device__ __forceinline__ void argmin_step(double& s, uint32_t& i) {
    const double so = dpp_f64<Ctrl>(s);
    const uint32_t io = dpp_u32<Ctrl>(i);
    if (so < s || (so == s && io < i)) {
        s = so;
        i = io;
    }
}

template <int G>
__device__ __forceinline__ void group_argmin(double& s, uint32_t& i) {
    if constexpr (G == 1) return;
    argmin_step<kDppXor1>(s, i);
    if constexpr (G >= 4) argmin_step<kDppXor2>(s, i);
    if constexpr (G >= 8) argmin_step<kDppHalfMirror>(s, i);
    if constexpr (G >= 16) argmin_step<kDppMirror>(s, i);
}

// Layout constants of a group of G lanes.
template <int G>
struct MG {
    static_assert(G == 2 || G == 4 || G == 8, "groups of 2, 4 or 8 lanes");
    static constexpr int Gc = G;      // lanes owning components
    static constexpr int C = 8 / G;   // components per lane
};

// Per-lane constants of the group layout.
struct MCtx {
    uint32_t q;   // component slot: the lane owns components q + Gc m
    int bp;       // 4 x (first lane of the group): ds_bpermute address of the group's first lane
};

// Component c (compile time) of a group-distributed vector v, in every lane of
// the group: read from its owner (lane q_c = c mod Gc of the group, register
// m = c / Gc). Pairs and quads broadcast by a quad permute; groups of 8 read
// the owner through the LDS crossbar.
template <int G, int c>
__device__ __forceinline__ double from_owner(const double* v, const MCtx& cx) {
    constexpr int Gc = MG<G>::Gc;
    constexpr int qc = c % Gc, m = c / Gc;
    if constexpr (Gc == 2)
        return dpp_f64<(qc == 0) ? 0xA0 : 0xF5>(v[m]);  // quad_perm [0,0,2,2] / [1,1,3,3]
    else if constexpr (Gc == 4)
        return dpp_f64<qc * 0x55>(v[m]);  // quad_perm [qc,qc,qc,qc]
    else
        return bperm_f64(cx.bp + 4 * qc, v[m]);
}

// Sum over the group of the 8-slot vectors v (slot = component), delivered to
// the owners: S[m] = sum over the group's lanes of v[q + Gc m]. Each level
// halves the slots a lane keeps: it sends its partner the half the partner
// keeps and adds the half it receives to its own.
template <int G>
__device__ __forceinline__ void reduce_scatter(double v[8], uint32_t q, double* S) {
    constexpr int Gc = MG<G>::Gc;
    if constexpr (Gc == 8) {
        // lanes 0-3 keep components 0-3, lanes 4-7 components 4-7 (partner 7 - q)
        const bool lo = q < 4;
        double w[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const double own = lo ? v[k] : v[4 + k];
            const double send = lo ? v[4 + k] : v[k];
            w[k] = own + dpp_f64<kDppHalfMirror>(send);
        }
        // partner q ^ 2: keep w[0..1] or w[2..3]
        const bool b1 = (q & 2) == 0;
        double u[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const double own = b1 ? w[k] : w[2 + k];
            const double send = b1 ? w[2 + k] : w[k];
            u[k] = own + dpp_f64<kDppXor2>(send);
        }
        // partner q ^ 1: component q
        const bool b0 = (q & 1) == 0;
        S[0] = (b0 ? u[0] : u[1]) + dpp_f64<kDppXor1>(b0 ? u[1] : u[0]);
    } else if constexpr (Gc == 4) {
        // partner q ^ 2: lanes 0, 1 keep components {0, 1, 4, 5}, lanes 2, 3 {2, 3, 6, 7}
        const bool b1 = (q & 2) == 0;
        double w[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int ca = (k & 1) + 4 * (k >> 1);
            const double own = b1 ? v[ca] : v[ca + 2];
            const double send = b1 ? v[ca + 2] : v[ca];
            w[k] = own + dpp_f64<kDppXor2>(send);
        }
        // partner q ^ 1: components q and q + 4
        const bool b0 = (q & 1) == 0;
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            const double own = b0 ? w[2 * m] : w[2 * m + 1];
            const double send = b0 ? w[2 * m + 1] : w[2 * m];
            S[m] = own + dpp_f64<kDppXor1>(send);
        }
    } else {
        // partner q ^ 1: lane 0 keeps the even components, lane 1 the odd ones
        const bool b0 = (q & 1) == 0;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const double own = b0 ? v[2 * m] : v[2 * m + 1];
            const double send = b0 ? v[2 * m + 1] : v[2 * m];
            S[m] = own + dpp_f64<kDppXor1>(send);
        }
    }
}

// Maximum over the lanes of the group.
template <int G>
__device__ __forceinline__ double owners_max(double v) {
    constexpr int Gc = MG<G>::Gc;
    v = fmax(v, dpp_f64<kDppXor1>(v));
    if constexpr (Gc >= 4) v = fmax(v, dpp_f64<kDppXor2>(v));
    if constexpr (Gc >= 8) v = fmax(v, dpp_f64<kDppHalfMirror>(v));
    return v;
}

// Component q + Gc m of a whole state vector y (select chain: the index
// depends on the lane).
template <int G, int m>
__device__ __forceinline__ double own_component(const double y[8], uint32_t q) {
    constexpr int Gc = MG<G>::Gc;
    double v = y[Gc * m];
#pragma unroll
    for (int j = 1; j < Gc; ++j) v = (q == static_cast<uint32_t>(j)) ? y[Gc * m + j] : v;
    return v;
}

// The components q + Gc m (m < C) of a whole state vector y.
template <int G>
__device__ __forceinline__ void own_components(const double y[8], uint32_t q, double* out) {
    constexpr int C = MG<G>::C;
    out[0] = own_component<G, 0>(y, q);
    if constexpr (C > 1) out[1] = own_component<G, 1>(y, q);
    if constexpr (C > 2) {
        out[2] = own_component<G, 2>(y, q);
        out[3] = own_component<G, 3>(y, q);
    }
}

__device__ __forceinline__ void to_local(const SourceDev& s, double x, double y, double z, double l[3]) {
    const double dx = x - s.c[0], dy = y - s.c[1], dz = z - s.c[2];
    l[0] = dx * s.fx[0] + dy * s.fx[1] + dz * s.fx[2];
    l[1] = dx * s.fy[0] + dy * s.fy[1] + dz * s.fy[2];
    l[2] = dx * s.ax[0] + dy * s.ax[1] + dz * s.ax[2];
}

// The sources of one lane: sources si and, in groups of 8 lanes, si + 8 of
// the table in LDS, those that exist (si: the lane's index in its group of G).
struct GroupSrc {
    const SourceDev* S;
    uint32_t first, n;
};

// Sources per lane in a group of G lanes (multi_ray_lanes picks G).
template <int G>
constexpr int lane_sources() {
    return (G == 8) ? 2 : 1;
}

// f(V, i) for each source of the lane, in order.
template <int G, class F>
__device__ __forceinline__ void for_sources(const GroupSrc& gs, F&& f) {
#pragma unroll
    for (int k = 0; k < lane_sources<G>(); ++k) {
        const uint32_t i = gs.first + static_cast<uint32_t>(k * G);
        if (i < gs.n) f(gs.S + i, i);
    }
}

// One source's terms in the superposed Hamiltonian (ms_rhs,
// software_compute_engine.hpp:2857-2870, with a carried p_t) at world
// position X, p_t and world momentum Pm, added to v: (2 H s, -2 H s k, 0, dp)
// with k and dp rotated to world axes.
__device__ __forceinline__ void add_source_terms(const SourceDev* V, const double X[3], double pt, const double Pm[3],
                                                 double v[8]) {
    double ths, c0, c1, c2, q0, q1, q2;
    {
        const double dx = X[0] - V->c[0], dy = X[1] - V->c[1], dz = X[2] - V->c[2];
        double lx, ly, lz, px, py, pz;
        {
            const double f0 = V->fx[0], f1 = V->fx[1], f2 = V->fx[2];
            lx = dx * f0 + dy * f1 + dz * f2;
            px = Pm[0] * f0 + Pm[1] * f1 + Pm[2] * f2;
        }
        {
            const double f0 = V->fy[0], f1 = V->fy[1], f2 = V->fy[2];
            ly = dx * f0 + dy * f1 + dz * f2;
            py = Pm[0] * f0 + Pm[1] * f1 + Pm[2] * f2;
        }
        {
            const double f0 = V->ax[0], f1 = V->ax[1], f2 = V->ax[2];
            lz = dx * f0 + dy * f1 + dz * f2;
            pz = Pm[0] * f0 + Pm[1] * f1 + Pm[2] * f2;
        }
        double t, c[3], q[3];
        ks_terms<double>(lx, ly, lz, pt, px, py, pz, kerr_const(V->m, V->a, V->a2), t, c, q);
        ths = t;
        {
            const double f0 = V->fx[0], f1 = V->fx[1], f2 = V->fx[2];
            c0 = c[0] * f0;
            c1 = c[0] * f1;
            c2 = c[0] * f2;
            q0 = q[0] * f0;
            q1 = q[0] * f1;
            q2 = q[0] * f2;
        }
        {
            const double f0 = V->fy[0], f1 = V->fy[1], f2 = V->fy[2];
            c0 += c[1] * f0;
            c1 += c[1] * f1;
            c2 += c[1] * f2;
            q0 += q[1] * f0;
            q1 += q[1] * f1;
            q2 += q[1] * f2;
        }
        {
            const double f0 = V->ax[0], f1 = V->ax[1], f2 = V->ax[2];
            c0 += c[2] * f0;
            c1 += c[2] * f1;
            c2 += c[2] * f2;
            q0 += q[2] * f0;
            q1 += q[2] * f1;
            q2 += q[2] * f2;
        }
    }
    v[0] += ths;
    v[1] += c0;
    v[2] += c1;
    v[3] += c2;
    v[5] += q0;
    v[6] += q1;
    v[7] += q2;
}

// Superposed right-hand side of the group-distributed state Y (p_t given):
// gathers the position X and momentum Pm (returned; the same in every lane),
// adds the lane's source and reduces onto the owners, d[m] being the rate of
// component q + Gc m: t' = -p_t + sum 2Hs, x' = p - sum 2Hs k, p_t' = 0,
// p' = sum dp.
template <int G>
__device__ __forceinline__ void mrhs(const GroupSrc& gs, const MCtx& cx, const double* Y, double pt, double* d,
                                     double X[3], double Pm[3]) {
    constexpr int Gc = MG<G>::Gc, C = MG<G>::C;
    X[0] = from_owner<G, 1>(Y, cx);
    X[1] = from_owner<G, 2>(Y, cx);
    X[2] = from_owner<G, 3>(Y, cx);
    Pm[0] = from_owner<G, 5>(Y, cx);
    Pm[1] = from_owner<G, 6>(Y, cx);
    Pm[2] = from_owner<G, 7>(Y, cx);
    double v[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for_sources<G>(gs, [&](const SourceDev* V, uint32_t) { add_source_terms(V, X, pt, Pm, v); });
    double S[C];
    reduce_scatter<G>(v, cx.q, S);
#pragma unroll
    for (int m = 0; m < C; ++m) {
        if (Gc * m < 4) {
            const uint32_t c = cx.q + Gc * m;
            double base = 0.0;
            base = (c == 1) ? Pm[0] : base;
            base = (c == 2) ? Pm[1] : base;
            base = (c == 3) ? Pm[2] : base;
            base = (c == 0) ? -pt : base;
            d[m] = base + S[m];
        } else {
            d[m] = S[m];
        }
    }
}

// The seven stages of one DP45 attempt on the group's components (as
// dp45_stages for one hole; WithDense adds the dense-output combination).
// X5, P5: position and momentum of y5 in every lane.
template <bool WithDense, int G>
__device__ __forceinline__ void mstages(const GroupSrc& gs, const MCtx& cx, const double* y, const double* k1,
                                        double pt, double h, double* y5, double* ep, double* k7, double* dn,
                                        double X5[3], double P5[3]) {
    constexpr int C = MG<G>::C;
    double k2[C], k3[C], k4[C], k5[C], k6[C], Y[C], X[3], Pm[3];
#pragma unroll
    for (int i = 0; i < C; ++i) Y[i] = y[i] + h * (dp::a21 * k1[i]);
    mrhs<G>(gs, cx, Y, pt, k2, X, Pm);
#pragma unroll
    for (int i = 0; i < C; ++i) Y[i] = y[i] + h * (dp::a31 * k1[i] + dp::a32 * k2[i]);
    mrhs<G>(gs, cx, Y, pt, k3, X, Pm);
#pragma unroll
    for (int i = 0; i < C; ++i) Y[i] = y[i] + h * (dp::a41 * k1[i] + dp::a42 * k2[i] + dp::a43 * k3[i]);
    mrhs<G>(gs, cx, Y, pt, k4, X, Pm);
#pragma unroll
    for (int i = 0; i < C; ++i)
        Y[i] = y[i] + h * (dp::a51 * k1[i] + dp::a52 * k2[i] + dp::a53 * k3[i] + dp::a54 * k4[i]);
    mrhs<G>(gs, cx, Y, pt, k5, X, Pm);
#pragma unroll
    for (int i = 0; i < C; ++i)
        Y[i] = y[i] + h * (dp::a61 * k1[i] + dp::a62 * k2[i] + dp::a63 * k3[i] + dp::a64 * k4[i] + dp::a65 * k5[i]);
    mrhs<G>(gs, cx, Y, pt, k6, X, Pm);
#pragma unroll
    for (int i = 0; i < C; ++i) {
        y5[i] = y[i] + h * (dp::a71 * k1[i] + dp::a73 * k3[i] + dp::a74 * k4[i] + dp::a75 * k5[i] + dp::a76 * k6[i]);
        ep[i] = dp::e1 * k1[i] + dp::e3 * k3[i] + dp::e4 * k4[i] + dp::e5 * k5[i] + dp::e6 * k6[i];
        if (WithDense) dn[i] = dp::d1 * k1[i] + dp::d3 * k3[i] + dp::d4 * k4[i] + dp::d5 * k5[i] + dp::d6 * k6[i];
    }
    mrhs<G>(gs, cx, y5, pt, k7, X5, P5);
}

// Local coordinates of a world point in the frame of source V.
__device__ __forceinline__ void to_local_v(const SourceDev* V, double x, double y, double z, double l[3]) {
    const double dx = x - V->c[0], dy = y - V->c[1], dz = z - V->c[2];
    l[0] = dx * V->fx[0] + dy * V->fx[1] + dz * V->fx[2];
    l[1] = dx * V->fy[0] + dy * V->fy[1] + dz * V->fy[2];
    l[2] = dx * V->ax[0] + dy * V->ax[1] + dz * V->ax[2];
}

// A superposition ray in its group: own components of the state and of its
// derivative (first-same-as-last), and what every lane keeps.
template <int C>
struct MLane {
    double y[C], k1[C];
    double xs[3], ps[3], pt;  // position, momentum and p_t of the state
    double ah, lambda, g;
    uint32_t acc, rej, flags, cls;
    uint32_t rhs, proj, f32;
};

// Superposed null constraint as a quadratic in p_t at (X, pt, Pm); projection
// as for one hole. Updates pt; true if it fired.
template <int G>
__device__ __forceinline__ bool mproject(const GroupSrc& gs, const double X[3], const double Pm[3], double& pt,
                                         double tol) {
    double h = 0.0, hq = 0.0, hq2 = 0.0;
    for_sources<G>(gs, [&](const SourceDev* V, uint32_t) {
        double l[3];
        to_local_v(V, X[0], X[1], X[2], l);
        const double px = Pm[0] * V->fx[0] + Pm[1] * V->fx[1] + Pm[2] * V->fx[2];
        const double py = Pm[0] * V->fy[0] + Pm[1] * V->fy[1] + Pm[2] * V->fy[2];
        const double pz = Pm[0] * V->ax[0] + Pm[1] * V->ax[1] + Pm[2] * V->ax[2];
        const Field f = ks_field(l[0], l[1], l[2], V->m, V->a, V->a2);
        const double q = f.k0 * px + f.k1 * py + f.k2 * pz;
        h += f.H;
        hq += f.H * q;
        hq2 += f.H * q * q;
    });
    const double sh = group_sum<G>(h), shq = group_sum<G>(hq), shq2 = group_sum<G>(hq2);
    const double P2 = Pm[0] * Pm[0] + Pm[1] * Pm[1] + Pm[2] * Pm[2];
    const double A = -(1.0 + 2.0 * sh);
    const double B = 4.0 * shq;
    const double C0 = P2 - 2.0 * shq2;
    const double res = (A * pt + B) * pt + C0;
    const double scale = fmax(1.0, pt * pt + P2);
    if (!(fabs(res) >= tol * scale)) return false;
    const double disc = B * B - 4.0 * A * C0;
    if (disc < 0.0) return false;
    const double sd = sqrt(disc);
    const double r1 = (-B + sd) / (2.0 * A);
    const double r2 = (-B - sd) / (2.0 * A);
    pt = (fabs(r1 - pt) <= fabs(r2 - pt)) ? r1 : r2;
    return true;
}

// Keep the owner's copy of component 4 equal to p_t after a projection.
template <int G>
__device__ __forceinline__ void sync_pt(const MCtx& cx, double* y, double pt) {
    constexpr int Gc = MG<G>::Gc;
    constexpr int m = 4 / Gc;
    if (cx.q == static_cast<uint32_t>(4 % Gc)) y[m] = pt;
}

// Camera pixel -> initial state in a superposition (trace_multi_source set-up,
// software_compute_engine.hpp:3054-3066, and ms_initial_momentum :2919-2964):
// flat spherical observer position, superposed-metric frame precomputed.
__device__ __forceinline__ void camera_state_multi(const TraceParams& P, const MultiDev& T, uint32_t px, uint32_t py,
                                                   double y[8]) {
    double o[3];
    camera_outward<true>(P, px, py, o);
    double vv[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) vv[c] = (c == 0 ? T.u0 : 0.0) - o[0] * T.e[0][c] - o[1] * T.e[1][c] - o[2] * T.e[2][c];
    double mom0 = vv[1], mom1 = vv[2], mom2 = vv[3];
    double tc = -vv[0];
    for (uint32_t i = 0; i < T.n; ++i) {
        const double lv = vv[0] + T.dir[i][0] * vv[1] + T.dir[i][1] * vv[2] + T.dir[i][2] * vv[3];
        const double w = 2.0 * T.pot[i] * lv;
        tc += w;
        mom0 += w * T.dir[i][0];
        mom1 += w * T.dir[i][1];
        mom2 += w * T.dir[i][2];
    }
    const double inv_pt = 1.0 / fmax(-tc, 1e-12);
    y[0] = 0.0;
    y[1] = P.X0[0];
    y[2] = P.X0[1];
    y[3] = P.X0[2];
    y[4] = -1.0;
    y[5] = mom0 * inv_pt;
    y[6] = mom1 * inv_pt;
    y[7] = mom2 * inv_pt;
}

// A ray through a superposition (FP64): DP45 as for one hole; events per
// source (horizon of its local KS radius, crossing of its local equatorial
// plane inside its disk), escape on the Euclidean radius while outgoing.
// mbegin sets a group up from the whole initial state y0 (every lane), mstep
// makes one attempt and returns true (in every lane of the group) once the
// ray has terminated.
template <int G>
__device__ __forceinline__ void mbegin(const TraceParams& P, const GroupSrc& gs, const MCtx& cx, const double y0[8],
                                       MLane<MG<G>::C>& L) {
    own_components<G>(y0, cx.q, L.y);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        L.xs[i] = y0[1 + i];
        L.ps[i] = y0[5 + i];
    }
    L.pt = y0[4];
    double X[3], Pm[3];
    mrhs<G>(gs, cx, L.y, L.pt, L.k1, X, Pm);
    L.rhs = 1;
    L.proj = L.f32 = 0;
    L.acc = L.rej = L.flags = L.cls = 0;
    L.lambda = 0.0;
    L.g = 0.0;
    double scale = 1e300, blend = 1.0;
    for_sources<G>(gs, [&](const SourceDev* V, uint32_t) {
        double l[3];
        to_local_v(V, L.xs[0], L.xs[1], L.xs[2], l);
        const double r = ks_radius(l[0], l[1], l[2], V->a2);
        const double rh = V->rh;
        scale = fmin(scale, sqrt(r * fmax(r - rh, 0.02 * V->m)));
        blend = fmin(blend, fmin(fmax((r - 20.0 * rh) / (80.0 * rh), 0.0), 1.0));
    });
    scale = group_min<G>(scale);
    blend = group_min<G>(blend);
    const double ah = (P.h0 > 0.0) ? P.h0 : fmin(fmax(0.5 * scale * (1.0 + 6.8 * blend), 0.004), 3.5 * 7.8);
    L.ah = fmin(fmax(ah, P.h_min), P.h_max);
}

// Dense-output coefficients of component c (compile time), in every lane.
template <int G, int c>
__device__ __forceinline__ Dense dense_from_owner(const double* D, const double* B, const double* Cc, const double* E,
                                                  const MCtx& cx) {
    Dense q;
    q.D = from_owner<G, c>(D, cx);
    q.B = from_owner<G, c>(B, cx);
    q.C = from_owner<G, c>(Cc, cx);
    q.E = from_owner<G, c>(E, cx);
    return q;
}

// Event location inside an accepted step [y, y5] of size h: every lane the
// horizons and disks of its own sources, then the group's earliest (ties to
// the lower source index, as a scan over the sources in order keeps them),
// then escape on the Euclidean radius while outgoing. Recomputes the step
// with the dense-output combination; the dense coefficients of the position
// and momentum are gathered into every lane. Fills the end state (own
// components ye; position Xe and momentum Pe in every lane); returns the
// class (0: no event), the same in every lane of the group.
template <int G>
__device__ __forceinline__ uint32_t mlocate_event(const TraceParams& P, const GroupSrc& gs, const MCtx& cx,
                                                  uint32_t lane0, const MLane<MG<G>::C>& L, double h,
                                                  bool es, double* ye, double Xe[3], double Pe[3], double* s_out,
                                                  double* g_out) {
    constexpr int C = MG<G>::C;
    double y5[C], ep[C], k7[C], dn[C], X5[3], P5[3];
    mstages<true, G>(gs, cx, L.y, L.k1, L.pt, h, y5, ep, k7, dn, X5, P5);
    double qD[C], qB[C], qC[C], qE[C];
#pragma unroll
    for (int i = 0; i < C; ++i) {
        qD[i] = y5[i] - L.y[i];
        qB[i] = h * L.k1[i] - qD[i];
        qC[i] = qD[i] - h * k7[i] - qB[i];
        qE[i] = h * (dn[i] + dp::d7 * k7[i]);
    }
    const Dense qx = dense_from_owner<G, 1>(qD, qB, qC, qE, cx);
    const Dense qy = dense_from_owner<G, 2>(qD, qB, qC, qE, cx);
    const Dense qz = dense_from_owner<G, 3>(qD, qB, qC, qE, cx);
    const Dense q5 = dense_from_owner<G, 5>(qD, qB, qC, qE, cx);
    const Dense q6 = dense_from_owner<G, 6>(qD, qB, qC, qE, cx);
    const Dense q7 = dense_from_owner<G, 7>(qD, qB, qC, qE, cx);
    const double* y = L.xs;
    double best = 2.0, best_g = 0.0;
    uint32_t best_cls = 0, best_src = 0xFFFFFFFFu;
    for_sources<G>(gs, [&](const SourceDev* V, uint32_t i) {
        auto local_at = [&](double ss, double l[3]) {
            to_local_v(V, dense_at(qx, y[0], ss), dense_at(qy, y[1], ss), dense_at(qz, y[2], ss), l);
        };
        double l0[3], l1[3];
        to_local_v(V, y[0], y[1], y[2], l0);
        to_local_v(V, X5[0], X5[1], X5[2], l1);
        const double a2 = V->a2, r_hz = V->r_hz;
        const double r0 = ks_radius(l0[0], l0[1], l0[2], a2);
        const double r1 = ks_radius(l1[0], l1[1], l1[2], a2);
        if (r1 <= r_hz && r0 > r_hz) {
            const double sh = illinois_root(
                [&](double ss) {
                    double l[3];
                    local_at(ss, l);
                    return ks_radius(l[0], l[1], l[2], a2) - r_hz;
                },
                r0 - r_hz, r1 - r_hz);
            if (sh < best) {
                best = sh;
                best_cls = RT_CLS_CAPTURED;
                best_src = i;
            }
        }
        const double z0 = l0[2], z1 = l1[2];
        if (V->disk_on && ((z0 < 0.0 && z1 >= 0.0) || (z0 > 0.0 && z1 <= 0.0))) {
            const double sd = illinois_root(
                [&](double ss) {
                    double l[3];
                    local_at(ss, l);
                    return l[2];
                },
                z0, z1);
            double l[3];
            local_at(sd, l);
            const double rc = ks_radius(l[0], l[1], l[2], a2);
            if (rc >= V->disk_in && rc <= V->disk_out && rc > V->rh102 && sd < best) {
                const double pw0 = dense_at(q5, L.ps[0], sd), pw1 = dense_at(q6, L.ps[1], sd),
                             pw2 = dense_at(q7, L.ps[2], sd);
                const double pl0 = pw0 * V->fx[0] + pw1 * V->fx[1] + pw2 * V->fx[2];
                const double pl1 = pw0 * V->fy[0] + pw1 * V->fy[1] + pw2 * V->fy[2];
                const double pl2 = pw0 * V->ax[0] + pw1 * V->ax[1] + pw2 * V->ax[2];
                const double pm = sqrt(pl0 * pl0 + pl1 * pl1 + pl2 * pl2);
                const double lzv = (l[0] * pl1 - l[1] * pl0) / fmax(pm, 1e-30);
                best = sd;
                best_cls = RT_CLS_DISK;
                best_src = i;
                best_g = disk_redshift(V->m, V->a, rc, lzv, P.clamp_g != 0);
            }
        }
    });
    {
        uint32_t win = best_src;
        group_argmin<G>(best, win);
        if (best < 2.0) {
            const int from = static_cast<int>(lane0 + (win & (G - 1)));
            best_cls = static_cast<uint32_t>(__shfl(static_cast<int>(best_cls), from, 64));
            best_g = __shfl(best_g, from, 64);
        }
    }
    if (!(best < 2.0)) best_cls = 0;
    if (es) {
        const double R0 = sqrt(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
        const double R1 = sqrt(X5[0] * X5[0] + X5[1] * X5[1] + X5[2] * X5[2]);
        double se = 1.0;
        if (R0 < P.r_esc)
            se = illinois_root(
                [&](double ss) {
                    const double x = dense_at(qx, y[0], ss), yy = dense_at(qy, y[1], ss), z = dense_at(qz, y[2], ss);
                    return sqrt(x * x + yy * yy + z * z) - P.r_esc;
                },
                R0 - P.r_esc, R1 - P.r_esc);
        if (se < best) {
            best = se;
            best_cls = RT_CLS_ESCAPED;
        }
    }
    if (best_cls == 0) return 0;
    if (best >= 1.0) {
#pragma unroll
        for (int i = 0; i < C; ++i) ye[i] = y5[i];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            Xe[i] = X5[i];
            Pe[i] = P5[i];
        }
    } else {
#pragma unroll
        for (int i = 0; i < C; ++i) {
            Dense q;
            q.D = qD[i];
            q.B = qB[i];
            q.C = qC[i];
            q.E = qE[i];
            ye[i] = dense_at(q, L.y[i], best);
        }
        Xe[0] = dense_at(qx, y[0], best);
        Xe[1] = dense_at(qy, y[1], best);
        Xe[2] = dense_at(qz, y[2], best);
        Pe[0] = dense_at(q5, L.ps[0], best);
        Pe[1] = dense_at(q6, L.ps[1], best);
        Pe[2] = dense_at(q7, L.ps[2], best);
    }
    *s_out = best;
    *g_out = best_g;
    return best_cls;
}

template <int G>
__device__ __forceinline__ bool mstep(const TraceParams& P, const GroupSrc& gs, const MCtx& cx, uint32_t lane0,
                                      MLane<MG<G>::C>& L) {
    constexpr int Gc = MG<G>::Gc, C = MG<G>::C;
    if (L.acc + L.rej >= P.max_steps || L.ah < P.h_min) {
        L.flags |= (L.acc + L.rej >= P.max_steps) ? RT_FLAG_BUDGET : RT_FLAG_MIN_STEP;
        double near = 0.0;
        for_sources<G>(gs, [&](const SourceDev* V, uint32_t) {
            const double dx = L.xs[0] - V->c[0], dy = L.xs[1] - V->c[1], dz = L.xs[2] - V->c[2];
            if (sqrt(dx * dx + dy * dy + dz * dz) < 6.0 * V->m) near = 1.0;
        });
        L.cls = (group_min<G>(-near) < 0.0) ? RT_CLS_CAPTURED : RT_CLS_ESCAPED;
        return true;
    }
    const double ah = L.ah, h = -ah;
    double y5[C], ep[C], k7[C], X5[3], P5[3];
    mstages<false, G>(gs, cx, L.y, L.k1, L.pt, h, y5, ep, k7, nullptr, X5, P5);
    L.rhs += 6;
    // err = |h| max_i |sum_j e_j k_j,i| / (rtol max(|y_i|, |y5_i|) + atol) over the
    // components other than p_t; infinite when a component is not finite
    double err = 0.0;
    bool ok = true;
#pragma unroll
    for (int m = 0; m < C; ++m) {
        const uint32_t c = cx.q + Gc * m;
        const double e = ah * fabs(ep[m] + dp::e7 * k7[m]);
        const double sc = P.rtol * fmax(fabs(L.y[m]), fabs(y5[m])) + P.atol;
        const double r = e * rcp64(sc);
        ok = ok && isfinite(y5[m]);
        if (c != 4) {
            ok = ok && isfinite(r);
            err = fmax(err, r);
        }
    }
    err = owners_max<G>(ok ? err : __builtin_huge_val());
    if (!(err < __builtin_huge_val())) {
        L.flags |= RT_FLAG_NONFINITE;
        L.cls = RT_CLS_CAPTURED;
        return true;
    }
    if (!(err <= 1.0)) {
        ++L.rej;
        L.ah = next_step(ah, err, false, P.h_min, P.h_max);
        return false;
    }
    ++L.acc;
    // any event in this step? (a horizon reached or a disk plane crossed by one
    // of the group's sources, escape)
    double mine = 0.0;
    for_sources<G>(gs, [&](const SourceDev* V, uint32_t) {
        double l0[3], l1[3];
        to_local_v(V, L.xs[0], L.xs[1], L.xs[2], l0);
        to_local_v(V, X5[0], X5[1], X5[2], l1);
        const bool hz = ks_radius(l1[0], l1[1], l1[2], V->a2) <= V->r_hz;
        const bool dk = V->disk_on && ((l0[2] < 0.0 && l1[2] >= 0.0) || (l0[2] > 0.0 && l1[2] <= 0.0));
        if (hz || dk) mine = 1.0;
    });
    const bool src_ev = group_min<G>(-mine) < 0.0;
    const double R1 = sqrt(X5[0] * X5[0] + X5[1] * X5[1] + X5[2] * X5[2]);
    bool es = false;
    if (R1 >= P.r_esc) {
        const double kx = from_owner<G, 1>(k7, cx), ky = from_owner<G, 2>(k7, cx), kz = from_owner<G, 3>(k7, cx);
        es = -(X5[0] * kx + X5[1] * ky + X5[2] * kz) > 0.0;
    }
    if (src_ev || es) {
        double s_ev = 1.0, g_ev = 0.0, ye[C], Xe[3], Pe[3];
        const uint32_t c = mlocate_event<G>(P, gs, cx, lane0, L, h, es, ye, Xe, Pe, &s_ev, &g_ev);
        if (c != 0) {
            L.rhs += 6;
            if (mproject<G>(gs, Xe, Pe, L.pt, P.inv_tol)) {
                L.flags |= RT_FLAG_PROJECTED;
                ++L.proj;
            }
#pragma unroll
            for (int i = 0; i < C; ++i) L.y[i] = ye[i];
            sync_pt<G>(cx, L.y, L.pt);
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                L.xs[i] = Xe[i];
                L.ps[i] = Pe[i];
            }
            L.lambda += s_ev * ah;
            L.cls = c;
            L.g = g_ev;
            return true;
        }
    }
#pragma unroll
    for (int i = 0; i < C; ++i) L.y[i] = y5[i];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        L.xs[i] = X5[i];
        L.ps[i] = P5[i];
    }
    L.lambda += ah;
    if (mproject<G>(gs, L.xs, L.ps, L.pt, P.inv_tol)) {
        L.flags |= RT_FLAG_PROJECTED;
        ++L.proj;
        sync_pt<G>(cx, L.y, L.pt);
        double X[3], Pm[3];
        mrhs<G>(gs, cx, L.y, L.pt, L.k1, X, Pm);
        ++L.rhs;
    } else {
#pragma unroll
        for (int i = 0; i < C; ++i) L.k1[i] = k7[i];
    }
    L.ah = next_step(ah, err, true, P.h_min, P.h_max);
    return false;
}

// Redshift per class and the final flag word of a superposition ray.
template <int C>
__device__ __forceinline__ void mlane_finish(MLane<C>& L) {
    const uint32_t c = L.cls & RT_CLS_MASK;
    if (c == RT_CLS_ESCAPED) L.g = 1.0 / (-L.pt);
    if (c == RT_CLS_CAPTURED) L.g = 0.0;
    L.cls |= L.flags & ~(kCarterNoTurn | kNoEventRetry);
}

// The per-ray outputs other than the state (not inlined, for the same reason
// as camera_state_multi: atan2's constants).
__device__ __noinline__ void store_multi_rest(const OutPtrs& O, uint64_t idx, uint32_t cls, double x, double y,
                                              double z, double lambda, double g, uint32_t acc, uint32_t rej,
                                              const SourceDev& s0) {
    const uint64_t n = O.n;
    O.cls[idx] = cls;
    double lc[3], bl[3];
    to_local(s0, x, y, z, lc);
    bl_from_ks(lc[0], lc[1], lc[2], s0.a, s0.a2, bl);
#pragma unroll
    for (int c = 0; c < 3; ++c) O.bl[c * n + idx] = bl[c];
    O.lambda[idx] = lambda;
    O.g[idx] = g;
    O.steps[idx] = acc;
    O.rejected[idx] = rej;
}

// Store: every owner its components, lane 0 of the group the rest, with BL coordinates taken in the frame of source 0 (the
// central hole).
template <int G>
__device__ __forceinline__ void store_mlane(const OutPtrs& O, uint64_t idx, const MLane<MG<G>::C>& L,
                                            const MultiDev& T, const MCtx& cx, uint32_t l) {
    constexpr int Gc = MG<G>::Gc, C = MG<G>::C;
    const uint64_t n = O.n;
    if (l < static_cast<uint32_t>(Gc)) {
#pragma unroll
        for (int m = 0; m < C; ++m) {
            const uint32_t c = cx.q + Gc * m;
            double* dst = (c < 4) ? O.x : O.p;
            dst[(c & 3) * n + idx] = L.y[m];
        }
    }
    if (l == 0) store_multi_rest(O, idx, L.cls, L.xs[0], L.xs[1], L.xs[2], L.lambda, L.g, L.acc, L.rej, T.src[0]);
}

// Superposition request: a persistent kernel, G lanes per ray (G = 2, 4 or
// 8; lane s of a group takes its sources and a share of the components).
// Idle groups are refilled from a device counter (ballot over the group
// leaders, one atomicAdd per refill, rank by popcount), so a group whose ray
// terminates takes the next ray while the slow rays of the wave continue.
// Camera rays are handed out in 8x8-tile order (ray g: tile g / 64, pixel
// g % 64), the tiles around the holes' images first (tile_xy): the rays that
// take the most steps start at once instead of forming the frame's tail. The
// source table is staged into LDS. *next starts at zero.
template <int G, int Waves>
__global__ __launch_bounds__(256, Waves) void k_trace_multi(TraceParams P, OutPtrs O,
                                                            const MultiDev* __restrict__ Tg,
                                                            unsigned long long* __restrict__ next,
                                                            uint32_t refill_min, const uint32_t* limit,
                                                            unsigned long long* active_waves) {
    constexpr int C = MG<G>::C;
    __shared__ MultiDev sm;
    {
        const uint32_t words = sizeof(MultiDev) / 8;
        const double* src = reinterpret_cast<const double*>(Tg);
        double* dst = reinterpret_cast<double*>(&sm);
        for (uint32_t w = threadIdx.x; w < words; w += blockDim.x) dst[w] = src[w];
        __syncthreads();
    }
    const MultiDev& T = sm;
    const uint32_t lane_id = threadIdx.x & 63;
    const uint32_t si = lane_id & (G - 1);
    const uint32_t lane0 = lane_id & ~static_cast<uint32_t>(G - 1);
    const uint64_t below = (lane0 == 0) ? 0ull : (~0ull >> (64 - lane0));
    const GroupSrc gs{sm.src, si, T.n};
    const MCtx cx{si & static_cast<uint32_t>(MG<G>::Gc - 1), static_cast<int>(4 * lane0)};
    const uint64_t total = (P.mode == 0) ? static_cast<uint64_t>(P.tiles_x) * ((P.n_rays / P.cols + 7) / 8) * 64
                                         : static_cast<uint64_t>(P.n_rays);
    MLane<C> L;
    bool active = false, drained = false;
    uint64_t idx = 0;
    Tally tally;
    wave_joins(limit, active_waves);
    for (;;) {
        const uint64_t idle = __ballot(!active && si == 0);
        const uint32_t n_idle = __popcll(idle);
        // the host asks for fewer waves: take no more rays, leave once idle
        if (!drained && n_idle != 0 && (n_idle * G >= refill_min || n_idle * G == 64) &&
            wave_yields(limit, active_waves))
            drained = true;
        if (!drained && n_idle != 0 && (n_idle * G >= refill_min || n_idle * G == 64)) {
            const uint32_t first = __ffsll(static_cast<unsigned long long>(idle)) - 1;
            unsigned long long base = 0;
            if (lane_id == first) base = atomicAdd(next, static_cast<unsigned long long>(n_idle));
            base = __shfl(base, first, 64);
            if (base + n_idle >= total) drained = true;
            if (!active) {
                const uint64_t g = base + __popcll(idle & below);
                if (g < total) {
                    double y0[8];
                    bool ok = true;
                    if (P.mode == 0) {
                        const uint32_t l = static_cast<uint32_t>(g & 63);
                        uint32_t tx, ty;
                        tile_xy(P, static_cast<uint32_t>(g >> 6), tx, ty);
                        const uint32_t px = tx * 8 + (l & 7);
                        const uint32_t py = ty * 8 + (l >> 3);
                        ok = (px < P.cols) && (py < P.n_rays / P.cols);
                        if (ok) {
                            camera_state_multi(P, T, P.col_first + px, P.row_first + py, y0);
                            idx = static_cast<uint64_t>(py) * P.cols + px;
                        }
                    } else {
                        idx = g;
#pragma unroll
                        for (int c = 0; c < 4; ++c) {
                            y0[c] = O.in_x[c * O.n + g];
                            y0[4 + c] = O.in_p[c * O.n + g];
                        }
                    }
                    if (ok && P.mode == 0 && T.observer_inside) {
                        // observer inside a horizon: captured where it stands
                        own_components<G>(y0, cx.q, L.y);
#pragma unroll
                        for (int i = 0; i < 3; ++i) {
                            L.xs[i] = y0[1 + i];
                            L.ps[i] = y0[5 + i];
                        }
                        L.pt = y0[4];
                        L.acc = L.rej = L.flags = L.proj = L.f32 = L.rhs = 0;
                        L.lambda = 0.0;
                        L.g = 0.0;
                        L.cls = RT_CLS_CAPTURED;
                        mlane_finish(L);
                        store_mlane<G>(O, idx, L, T, cx, si);
                    } else if (ok) {
                        mbegin<G>(P, gs, cx, y0, L);
                        active = true;
                    }
                }
            }
        }
        if (__ballot(active) == 0) {
            if (drained) break;
            continue;
        }
        if (active && mstep<G>(P, gs, cx, lane0, L)) {
            mlane_finish(L);
            store_mlane<G>(O, idx, L, T, cx, si);
            if (si == 0) tally.add(L);
            active = false;
        }
    }
    tally.flush(O.counters);
}

// ------------------------------------------------------------ launchers
static bool shape_ok(const LaunchShape& L) { return L.block != 0 && L.block <= 256 && (L.block & 63) == 0; }

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
    if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, reinterpret_cast<const void*>(&k_trace_multi<8, 2>),
                                                         64, 0);
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

// Lanes per superposition ray. Up to eight sources: the smallest power of
// two not below the number of sources (one source per lane; these frames are
// bound by the latency of their slowest rays). From nine sources: 8 lanes
// (two sources per lane, lane_sources), which halves the lanes per ray and
// shares the stage arithmetic and the sums over the group between twice the
// sources (these frames are bound by instruction issue): 16 holes on a 512^2
// frame 55 -> 49 ms, while 8 holes on 4 lanes went from 27 to 31 ms.
uint32_t multi_ray_lanes(uint32_t n_src) {
    uint32_t g = 2;
    while (g < n_src && g < 8) g <<= 1;
    return g;
}

template <int G>
static hipError_t launch_multi_g(const TraceParams& P, const OutPtrs& O, const MultiDev* T, unsigned long long* next,
                                 const LaunchShape& L, hipStream_t stream, uint32_t max_grid, const uint32_t* limit,
                                 unsigned long long* active) {
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const void* fn = (L.waves >= 2) ? reinterpret_cast<const void*>(&k_trace_multi<G, 2>)
                                    : reinterpret_cast<const void*>(&k_trace_multi<G, 1>);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, static_cast<int>(L.block), 0);
    if (e != hipSuccess) return e;
    // no more lanes than G per ray (a small request leaves the rest of the device to others)
    const uint64_t rays = (P.mode == 0) ? static_cast<uint64_t>(P.tiles_x) * ((P.n_rays / P.cols + 7) / 8) * 64
                                        : static_cast<uint64_t>(P.n_rays);
    const uint64_t want = (rays * G + L.block - 1) / L.block;
    uint64_t grid64 = std::min<uint64_t>(want, static_cast<uint64_t>(cus) * static_cast<uint64_t>(per_cu > 0 ? per_cu : 1));
    if (max_grid != 0) grid64 = std::min<uint64_t>(grid64, max_grid);
    const uint32_t grid = static_cast<uint32_t>(std::max<uint64_t>(grid64, 1));
    const uint32_t refill = std::max<uint32_t>(1, std::min<uint32_t>(L.refill_min, 64));
    if (L.waves >= 2)
        hipLaunchKernelGGL((k_trace_multi<G, 2>), dim3(grid), dim3(L.block), 0, stream, P, O, T, next, refill, limit,
                           active);
    else
        hipLaunchKernelGGL((k_trace_multi<G, 1>), dim3(grid), dim3(L.block), 0, stream, P, O, T, next, refill, limit,
                           active);
    return hipGetLastError();
}

hipError_t launch_trace_multi(const TraceParams& P, const OutPtrs& O, const MultiDev* T, unsigned long long* next,
                              const LaunchShape& L, hipStream_t stream, uint32_t max_grid, const uint32_t* limit,
                              unsigned long long* active) {
    if (!shape_ok(L)) return hipErrorInvalidValue;
    if (P.mode == 0 && (P.cols == 0 || P.n_rays % P.cols != 0)) return hipErrorInvalidValue;
    if (P.n_src < 2 || P.n_src > static_cast<uint32_t>(kMaxSources)) return hipErrorInvalidValue;
    switch (multi_ray_lanes(P.n_src)) {
        case 2: return launch_multi_g<2>(P, O, T, next, L, stream, max_grid, limit, active);
        case 4: return launch_multi_g<4>(P, O, T, next, L, stream, max_grid, limit, active);
        case 8: return launch_multi_g<8>(P, O, T, next, L, stream, max_grid, limit, active);
        default: return hipErrorInvalidValue;
    }
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
