// rt_kernels.hip — HIP kernels of the null-geodesic tracer for gfx950.
//
// One lane traces one ray to its end: adaptive DP45 on the Kerr–Schild
// Hamiltonian (first-same-as-last), events located on the dense output
// (horizon, equatorial disk, escape), results written once as SoA. Lanes of a
// wave cover an 8x8 pixel tile so their step counts stay close.
#include <hip/hip_runtime.h>

#include "rt_device.hpp"
#include "rt_kernels.h"

namespace rt {

struct RayState {
    double y[8];
    double xi;  // L_z / E of the ray (disk redshift)
};

struct RayEnd {
    uint32_t cls;
    double y[8];
    double lambda, g;
    uint32_t acc, rej;
    uint32_t rhs, proj, f32;
};

// Initial state of camera pixel (px, py) for a single Kerr hole: pinhole ray
// (camera_projections.hpp:27-41), world direction from the camera basis,
// projection onto e_r/e_theta/e_phi and back (software_compute_engine.hpp:3851-3938,
// 1634-1645), photon momentum -outward in the observer's Gram–Schmidt frame,
// normalised to p_t = -1 (kerr_schild_null_integrator.hpp:179-193). The frame
// itself is identical for every pixel and comes precomputed in P.
__device__ __forceinline__ RayState camera_state_kerr(const TraceParams& P, uint32_t px, uint32_t py) {
    const double v = 1.0 - (static_cast<double>(py) + 0.5) / static_cast<double>(P.height) * 2.0;
    const double u = ((static_cast<double>(px) + 0.5) / static_cast<double>(P.width) * 2.0 - 1.0) * P.aspect;
    const double n1 = 1.0, n2 = v * P.tan_half, n3 = u * P.tan_half;
    const double len = sqrt(n1 * n1 + n2 * n2 + n3 * n3);
    const double inv_len = 1.0 / len;
    const double l0 = n1 * inv_len, l1 = n2 * inv_len, l2 = n3 * inv_len;
    const double d0 = l0 * P.fwd[0] + l2 * P.right[0] + l1 * P.up[0];
    const double d1 = l0 * P.fwd[1] + l2 * P.right[1] + l1 * P.up[1];
    const double d2 = l0 * P.fwd[2] + l2 * P.right[2] + l1 * P.up[2];
    const double st = P.sth, ct = P.cth, sp = P.sph, cp = P.cph;
    const double n_r = d0 * (st * cp) + d1 * (st * sp) + d2 * ct;
    const double n_th = d0 * (ct * cp) + d1 * (ct * sp) + d2 * (-st);
    const double n_ph = d0 * (-sp) + d1 * cp;
    const double o0 = n_r * st * cp + n_th * ct * cp - n_ph * sp;
    const double o1 = n_r * st * sp + n_th * ct * sp + n_ph * cp;
    const double o2 = n_r * ct - n_th * st;
    double vv[4] = {P.u0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        vv[c] -= o0 * P.frame[0][c];
        vv[c] -= o1 * P.frame[1][c];
        vv[c] -= o2 * P.frame[2][c];
    }
    const double two_h = P.two_h_obs;
    const double lv = vv[0] + P.kobs[0] * vv[1] + P.kobs[1] * vv[2] + P.kobs[2] * vv[3];
    const double p_t = -vv[0] + two_h * lv;
    const double inv_e = 1.0 / fmax(-p_t, 1e-12);
    RayState s;
    s.y[0] = 0.0;
    s.y[1] = P.X0[0];
    s.y[2] = P.X0[1];
    s.y[3] = P.X0[2];
    s.y[4] = -1.0;
    s.y[5] = (vv[1] + two_h * lv * P.kobs[0]) * inv_e;
    s.y[6] = (vv[2] + two_h * lv * P.kobs[1]) * inv_e;
    s.y[7] = (vv[3] + two_h * lv * P.kobs[2]) * inv_e;
    s.xi = s.y[1] * s.y[6] - s.y[2] * s.y[5];
    return s;
}

// Render-step heuristic (software_compute_engine.hpp:1677-1683 with the
// GpuCameraPushConstants defaults) as the initial |h|.
__device__ __forceinline__ double initial_step_kerr(const TraceParams& P, double r) {
    const double scale = sqrt(r * fmax(r - P.rh, 0.02 * P.M));
    const double blend = fmin(fmax((r - 20.0 * P.rh) / (80.0 * P.rh), 0.0), 1.0);
    const double far = 1.0 + (7.8 - 1.0) * blend;
    return fmin(fmax(0.5 * scale * far, 0.004), 3.5 * 7.8);
}

// Stages k2..k7 of one DP45 attempt from (y, k1, h): the 5th-order solution
// y5, k7 = f(y5), the error combination ep = sum_{j<7} e_j k_j and, when
// WithDense, dn = sum_{j<7} d_j k_j for the continuous extension. The same
// inputs give bit-identical outputs, so the event path recomputes the stages
// with WithDense instead of keeping dn live on every step.
template <bool WithDense, class R>
__device__ __forceinline__ Field dp45_stages(const double y[8], const double k1[8], double h, double M, double a,
                                             double a2, double y5[8], double ep[8], double k7[8], double dn[8]) {
    double k2[8], k3[8], k4[8], k5[8], k6[8], Y[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) Y[i] = y[i] + h * (dp::a21 * k1[i]);
    rhs_kerr<R>(Y, M, a, a2, k2);
#pragma unroll
    for (int i = 0; i < 8; ++i) Y[i] = y[i] + h * (dp::a31 * k1[i] + dp::a32 * k2[i]);
    rhs_kerr<R>(Y, M, a, a2, k3);
#pragma unroll
    for (int i = 0; i < 8; ++i) Y[i] = y[i] + h * (dp::a41 * k1[i] + dp::a42 * k2[i] + dp::a43 * k3[i]);
    rhs_kerr<R>(Y, M, a, a2, k4);
#pragma unroll
    for (int i = 0; i < 8; ++i)
        Y[i] = y[i] + h * (dp::a51 * k1[i] + dp::a52 * k2[i] + dp::a53 * k3[i] + dp::a54 * k4[i]);
    rhs_kerr<R>(Y, M, a, a2, k5);
#pragma unroll
    for (int i = 0; i < 8; ++i)
        Y[i] = y[i] + h * (dp::a61 * k1[i] + dp::a62 * k2[i] + dp::a63 * k3[i] + dp::a64 * k4[i] + dp::a65 * k5[i]);
    rhs_kerr<R>(Y, M, a, a2, k6);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        y5[i] = y[i] + h * (dp::a71 * k1[i] + dp::a73 * k3[i] + dp::a74 * k4[i] + dp::a75 * k5[i] + dp::a76 * k6[i]);
        ep[i] = dp::e1 * k1[i] + dp::e3 * k3[i] + dp::e4 * k4[i] + dp::e5 * k5[i] + dp::e6 * k6[i];
        if (WithDense) dn[i] = dp::d1 * k1[i] + dp::d3 * k3[i] + dp::d4 * k4[i] + dp::d5 * k5[i] + dp::d6 * k6[i];
    }
    return rhs_kerr<R>(y5, M, a, a2, k7);
}

// Event location inside an accepted step [y, y5] of size h: horizon
// r = r_hz, equatorial disk z = 0 with r in [disk_in, disk_out], escape
// r = r_esc while outgoing; earliest first. Fills the end state and returns
// the class (0: no event).
__device__ __forceinline__ uint32_t locate_event(const TraceParams& P, const double y[8], const double k1[8], double h,
                                              double r0, double r1, bool hz, bool dk, bool es, double xi, bool f32,
                                              double yend[8], double* s_out, double* g_out) {
    const double M = P.M, a = P.a, a2 = P.a2;
    double y5[8], ep[8], k7[8], dn[8];
    if (f32)
        dp45_stages<true, float>(y, k1, h, M, a, a2, y5, ep, k7, dn);
    else
        dp45_stages<true, double>(y, k1, h, M, a, a2, y5, ep, k7, dn);
    Dense q[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        q[i].D = y5[i] - y[i];
        q[i].B = h * k1[i] - q[i].D;
        q[i].C = q[i].D - h * k7[i] - q[i].B;
        q[i].E = h * (dn[i] + dp::d7 * k7[i]);
    }
    auto rad_at = [&](double ss) {
        return ks_radius(dense_at(q[1], y[1], ss), dense_at(q[2], y[2], ss), dense_at(q[3], y[3], ss), a2);
    };
    double best = 2.0, best_g = 0.0;
    uint32_t best_cls = 0;
    if (hz) {
        best = illinois_root([&](double ss) { return rad_at(ss) - P.r_hz; }, r0 - P.r_hz, r1 - P.r_hz);
        best_cls = RT_CLS_CAPTURED;
    }
    if (dk) {
        const double sd = illinois_root([&](double ss) { return dense_at(q[3], y[3], ss); }, y[3], y5[3]);
        const double rc = rad_at(sd);
        if (rc >= P.disk_in && rc <= P.disk_out && rc > P.rh102 && sd < best) {
            best = sd;
            best_cls = RT_CLS_DISK;
            best_g = disk_redshift(M, P.disk_spin, rc, xi, P.clamp_g != 0);
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
        for (int i = 0; i < 8; ++i) yend[i] = dense_at(q[i], y[i], best);
    }
    *s_out = best;
    *g_out = best_g;
    return best_cls;
}

// Step-size controller of the mixed policy: same law, FP32 power.
__device__ __forceinline__ double next_step_fast(double ah, double err, bool accepted, double h_min, double h_max) {
    const float le = __builtin_amdgcn_logf(fmaxf(static_cast<float>(err), 1e-12f));  // log2(err)
    if (accepted) {
        const double factor = (err == 0.0) ? 5.0 : static_cast<double>(__builtin_amdgcn_exp2f(-0.2f * le));
        const double s = fmin(fmax(0.9 * factor, 0.2), 5.0);
        return fmin(fmax(ah * s, h_min), h_max);
    }
    const double factor = static_cast<double>(__builtin_amdgcn_exp2f(-0.25f * le));
    return ah * fmin(fmax(0.9 * factor, 0.1), 0.9);
}

// Trace one ray in a single Kerr spacetime. Mixed: steps that start beyond
// the switch radius evaluate the right-hand side in FP32 (state, stage sums,
// error test, projection and event location stay FP64).
template <bool Mixed>
__device__ RayEnd trace_kerr(const TraceParams& P, const RayState& s0) {
    const double M = P.M, a = P.a, a2 = P.a2;
    RayEnd o;
    o.g = 0.0;
    double y[8], k1[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) y[i] = s0.y[i];
    const Field f0 = rhs_kerr(y, M, a, a2, k1);
    uint32_t rhs_n = 1, proj_n = 0, f32_n = 0;
    double r0 = f0.r;
    double ah = (P.h0 > 0.0) ? P.h0 : initial_step_kerr(P, r0);
    ah = fmin(fmax(ah, P.h_min), P.h_max);
    double lambda = 0.0;
    uint32_t acc = 0, rej = 0, flags = 0, cls = 0;
    for (;;) {
        if (acc + rej >= P.max_steps || ah < P.h_min) {
            flags |= (acc + rej >= P.max_steps) ? RT_FLAG_BUDGET : RT_FLAG_MIN_STEP;
            cls = (r0 < P.near_r) ? RT_CLS_CAPTURED : RT_CLS_ESCAPED;
#pragma unroll
            for (int i = 0; i < 8; ++i) o.y[i] = y[i];
            break;
        }
        const double h = -ah;
        double y5[8], ep[8], k7[8];
        const bool f32 = Mixed && (r0 > P.switch_r);
        const Field f7 = f32 ? dp45_stages<false, float>(y, k1, h, M, a, a2, y5, ep, k7, nullptr)
                             : dp45_stages<false, double>(y, k1, h, M, a, a2, y5, ep, k7, nullptr);
        rhs_n += 6;
        f32_n += f32 ? 1u : 0u;
        double err = 0.0;
        bool finite = true;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            finite = finite && isfinite(y5[i]);
            if (i == 4) continue;  // dp_t/dl = 0: no error in p_t
            const double e = ah * fabs(ep[i] + dp::e7 * k7[i]);
            const double sc = P.rtol * fmax(fabs(y[i]), fabs(y5[i])) + P.atol;
            err = fmax(err, e * rcp64(sc));
        }
        if (!finite || !isfinite(err)) {
            flags |= RT_FLAG_NONFINITE;
            cls = RT_CLS_CAPTURED;
#pragma unroll
            for (int i = 0; i < 8; ++i) o.y[i] = y[i];
            break;
        }
        if (err <= 1.0) {
            ++acc;
            const double r1 = f7.r;
            const double z0 = y[3], z1 = y5[3];
            const bool hz = (r1 <= P.r_hz);
            const bool dk = P.disk_on && ((z0 < 0.0 && z1 >= 0.0) || (z0 > 0.0 && z1 <= 0.0));
            const bool es = (r1 >= P.r_esc) && (-(y5[1] * k7[1] + y5[2] * k7[2] + y5[3] * k7[3]) > 0.0);
            if (hz || dk || es) {
                double s_ev = 1.0, g_ev = 0.0;
                const uint32_t c = locate_event(P, y, k1, h, r0, r1, hz, dk, es, s0.xi, f32, o.y, &s_ev, &g_ev);
                if (c != 0) {
                    rhs_n += 6;
                    const Field fe = ks_field(o.y[1], o.y[2], o.y[3], M, a, a2);
                    if (project_pt(o.y, fe.H, fe.k0 * o.y[5] + fe.k1 * o.y[6] + fe.k2 * o.y[7], P.inv_tol)) {
                        flags |= RT_FLAG_PROJECTED;
                        ++proj_n;
                    }
                    lambda += s_ev * ah;
                    cls = c;
                    o.g = g_ev;
                    break;
                }
            }
#pragma unroll
            for (int i = 0; i < 8; ++i) y[i] = y5[i];
            lambda += ah;
            if (project_pt(y, f7.H, f7.k0 * y[5] + f7.k1 * y[6] + f7.k2 * y[7], P.inv_tol)) {
                flags |= RT_FLAG_PROJECTED;
                ++proj_n;
                if (f32)
                    rhs_kerr<float>(y, M, a, a2, k1);
                else
                    rhs_kerr<double>(y, M, a, a2, k1);
                ++rhs_n;
            } else {
#pragma unroll
                for (int i = 0; i < 8; ++i) k1[i] = k7[i];
            }
            r0 = r1;
            ah = Mixed ? next_step_fast(ah, err, true, P.h_min, P.h_max) : next_step(ah, err, true, P.h_min, P.h_max);
        } else {
            ++rej;
            ah = Mixed ? next_step_fast(ah, err, false, P.h_min, P.h_max)
                       : next_step(ah, err, false, P.h_min, P.h_max);
        }
    }
    const uint32_t c = cls & RT_CLS_MASK;
    if (c == RT_CLS_ESCAPED) o.g = 1.0 / (-o.y[4]);
    if (c == RT_CLS_CAPTURED) o.g = 0.0;
    o.cls = cls | flags;
    o.lambda = lambda;
    o.acc = acc;
    o.rej = rej;
    o.rhs = rhs_n;
    o.proj = proj_n;
    o.f32 = f32_n;
    if (f32_n) o.cls |= RT_FLAG_FP32_STEPS;
    return o;
}

__device__ __forceinline__ void store_end(const TraceParams& P, const OutPtrs& O, uint64_t idx, const RayEnd& e) {
    const uint64_t n = O.n;
    O.cls[idx] = e.cls;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        O.x[c * n + idx] = e.y[c];
        O.p[c * n + idx] = e.y[4 + c];
    }
    double bl[3];
    bl_from_ks(e.y[1], e.y[2], e.y[3], P.a, P.a2, bl);
#pragma unroll
    for (int c = 0; c < 3; ++c) O.bl[c * n + idx] = bl[c];
    O.lambda[idx] = e.lambda;
    O.g[idx] = e.g;
    O.steps[idx] = e.acc;
    O.rejected[idx] = e.rej;
}

__device__ __forceinline__ void add_counters(const OutPtrs& O, bool valid, const RayEnd& e) {
    const unsigned long long acc = wave_sum(valid ? e.acc : 0u);
    const unsigned long long rej = wave_sum(valid ? e.rej : 0u);
    const unsigned long long rhs = wave_sum(valid ? e.rhs : 0u);
    const unsigned long long proj = wave_sum(valid ? e.proj : 0u);
    const unsigned long long f32 = wave_sum(valid ? e.f32 : 0u);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&O.counters[0], acc);
        atomicAdd(&O.counters[1], rej);
        atomicAdd(&O.counters[2], rhs);
        atomicAdd(&O.counters[3], f32);
        atomicAdd(&O.counters[4], proj);
    }
}

// Camera frame, single Kerr. Thread t -> 8x8 tile t/64, lane t%64.
template <bool Mixed, int Waves>
__global__ __launch_bounds__(256, Waves) void k_trace_camera_kerr(TraceParams P, OutPtrs O) {
    const uint64_t t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint32_t tile = static_cast<uint32_t>(t >> 6);
    const uint32_t lane = static_cast<uint32_t>(t & 63);
    const uint32_t px = (tile % P.tiles_x) * 8 + (lane & 7);
    const uint32_t py = (tile / P.tiles_x) * 8 + (lane >> 3);
    const uint32_t rows = P.n_rays / P.width;
    const bool valid = (px < P.width) && (py < rows);
    RayEnd e{};
    if (valid) {
        const RayState s = camera_state_kerr(P, px, P.row_first + py);
        e = trace_kerr<Mixed>(P, s);
        store_end(P, O, static_cast<uint64_t>(py) * P.width + px, e);
    }
    add_counters(O, valid, e);
}

// Explicit initial states, single Kerr.
template <bool Mixed, int Waves>
__global__ __launch_bounds__(256, Waves) void k_trace_rays_kerr(TraceParams P, OutPtrs O) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool valid = i < P.n_rays;
    RayEnd e{};
    if (valid) {
        RayState s;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            s.y[c] = O.in_x[c * O.n + i];
            s.y[4 + c] = O.in_p[c * O.n + i];
        }
        s.xi = (s.y[1] * s.y[6] - s.y[2] * s.y[5]) / (-s.y[4]);
        e = trace_kerr<Mixed>(P, s);
        store_end(P, O, i, e);
    }
    add_counters(O, valid, e);
}

// ------------------------------------------------------------ launchers
template <bool Mixed>
static hipError_t launch_camera_t(const TraceParams& P, const OutPtrs& O, const LaunchShape& L, hipStream_t stream) {
    const uint64_t tiles_y = (P.n_rays / P.width + 7) / 8;
    const uint64_t threads = static_cast<uint64_t>(P.tiles_x) * tiles_y * 64;
    const uint32_t grid = static_cast<uint32_t>((threads + L.block - 1) / L.block);
    if (L.waves >= 2)
        hipLaunchKernelGGL((k_trace_camera_kerr<Mixed, 2>), dim3(grid), dim3(L.block), 0, stream, P, O);
    else
        hipLaunchKernelGGL((k_trace_camera_kerr<Mixed, 1>), dim3(grid), dim3(L.block), 0, stream, P, O);
    return hipGetLastError();
}

hipError_t launch_trace_camera(const TraceParams& P, const OutPtrs& O, const LaunchShape& L, hipStream_t stream) {
    if (L.block == 0 || L.block > 256 || (L.block & 63) != 0) return hipErrorInvalidValue;
    return P.precision == 1 ? launch_camera_t<true>(P, O, L, stream) : launch_camera_t<false>(P, O, L, stream);
}

template <bool Mixed>
static hipError_t launch_rays_t(const TraceParams& P, const OutPtrs& O, const LaunchShape& L, hipStream_t stream) {
    const uint32_t grid = static_cast<uint32_t>((static_cast<uint64_t>(P.n_rays) + L.block - 1) / L.block);
    if (L.waves >= 2)
        hipLaunchKernelGGL((k_trace_rays_kerr<Mixed, 2>), dim3(grid), dim3(L.block), 0, stream, P, O);
    else
        hipLaunchKernelGGL((k_trace_rays_kerr<Mixed, 1>), dim3(grid), dim3(L.block), 0, stream, P, O);
    return hipGetLastError();
}

hipError_t launch_trace_rays(const TraceParams& P, const OutPtrs& O, const LaunchShape& L, hipStream_t stream) {
    if (L.block == 0 || L.block > 256 || (L.block & 63) != 0) return hipErrorInvalidValue;
    return P.precision == 1 ? launch_rays_t<true>(P, O, L, stream) : launch_rays_t<false>(P, O, L, stream);
}

}  // namespace rt
