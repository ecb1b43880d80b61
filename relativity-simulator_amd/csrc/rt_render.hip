// rt_render.hip — RT_METHOD_RENDER_RK4: the reference renderer's own tracer
// on the GPU, for callers that need its pixel records (status_flags,
// iterations_used, redshift, affine_parameter) rather than the DP45
// contract.
//
// Restates trace_kerr_schild_photon (render/software_compute_engine.hpp:1624-1809)
// as dispatch_fp64 runs it for a Kerr frame, without its shading: classic
// RK4 (step_kerr_schild_null_rk4, core/kerr_schild_null_integrator.hpp:116-139)
// on the KS null Hamiltonian (kerr_schild_null_derivatives, :71-114) with the
// render step formula; capture at r <= 1.0001 r_+ or a non-finite state,
// escape at r >= escape_radius; every crossing of theta = pi/2 inside the
// disk sets the disk bit, records g and multiplies the throughput by
// 1 - opacity, the opacity being that of AccretionDiskModel::shade
// (render/accretion_disk/accretion_disk_model.hpp:110-176, :245-314).
//
// The aim is the reference's records bit for bit, so this file is compiled
// without floating-point contraction (-ffp-contract=off, as the reference's
// ISO C++ build) and every expression keeps the reference's operation order,
// with IEEE division and square root. Only atan2 (BL theta and phi) and the
// float transcendentals of the opacity come from the device library; they
// can differ from the host's in the last bit, which changes a record only
// where a decision sits exactly on a threshold.
#include <hip/hip_runtime.h>

#include "rt_camera.hpp"
#include "rt_device.hpp"
#include "rt_kernels.h"

namespace rt {
namespace {

__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double dclamp(double v, double lo, double hi) { return (v < lo) ? lo : ((hi < v) ? hi : v); }
__device__ __forceinline__ float fmaxr(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ float fclampr(float v, float lo, float hi) { return (v < lo) ? lo : ((hi < v) ? hi : v); }

// KS radius of a point (kerr_schild_field's r).
__device__ __forceinline__ double ref_radius(const double x[3], double spin) {
    const double a2 = spin * spin;
    const double z2 = x[2] * x[2];
    const double diff = x[0] * x[0] + x[1] * x[1] + z2 - a2;
    const double r2 = dmax(0.5 * (diff + sqrt(diff * diff + 4.0 * a2 * z2)), 1e-24);
    return sqrt(r2);
}

// Derivative of the KS null state (x, p), p_t = -1, in the reference's
// operation order (kerr_schild_null_derivatives).
__device__ __forceinline__ void ref_rhs(const double x[3], const double p[3], double mass, double spin, double dx[3],
                                        double dp[3]) {
    const double a2 = spin * spin;
    const double z2 = x[2] * x[2];
    const double diff = x[0] * x[0] + x[1] * x[1] + z2 - a2;
    const double fr2 = dmax(0.5 * (diff + sqrt(diff * diff + 4.0 * a2 * z2)), 1e-24);
    const double r = sqrt(fr2);
    const double fden = fr2 * fr2 + a2 * z2;
    const double frpa = fr2 + a2;
    const double h = mass * fr2 * r / fden;
    const double k0 = (r * x[0] + spin * x[1]) / frpa;
    const double k1 = (r * x[1] - spin * x[0]) / frpa;
    const double k2 = x[2] / r;
    const double two = 2.0;
    const double r2 = r * r;
    const double r3 = r2 * r;
    const double r4 = r2 * r2;
    const double den = r4 + a2 * z2;
    const double rpa = r2 + a2;
    const double inv_rpa2 = 1.0 / (rpa * rpa);
    const double inv_den2 = 1.0 / (den * den);
    const double s = k0 * p[0] + k1 * p[1] + k2 * p[2] + 1.0;
    dx[0] = p[0] - two * h * s * k0;
    dx[1] = p[1] - two * h * s * k1;
    dx[2] = p[2] - two * h * s * k2;
    const double dr[3] = {r3 * x[0] / den, r3 * x[1] / den, r * rpa * x[2] / den};
    const double bracket = 3.0 * a2 * z2 - r4;
    const double nx = r * x[0] + spin * x[1];
    const double ny = r * x[1] - spin * x[0];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double dh = mass * dr[i] * r2 * bracket * inv_den2;
        if (i == 2) dh -= two * mass * a2 * r3 * x[2] * inv_den2;
        const double ex = (i == 0) ? 1.0 : 0.0;
        const double ey = (i == 1) ? 1.0 : 0.0;
        const double ez = (i == 2) ? 1.0 : 0.0;
        const double dk1 = ((dr[i] * x[0] + r * ex + spin * ey) * rpa - nx * two * r * dr[i]) * inv_rpa2;
        const double dk2 = ((dr[i] * x[1] + r * ey - spin * ex) * rpa - ny * two * r * dr[i]) * inv_rpa2;
        const double dk3 = (r * ez - x[2] * dr[i]) / r2;
        const double dk_dot_p = dk1 * p[0] + dk2 * p[1] + dk3 * p[2];
        dp[i] = dh * s * s + two * h * s * dk_dot_p;
    }
}

// One classic RK4 step of size h.
__device__ __forceinline__ void ref_rk4(double x[3], double p[3], double h, double mass, double spin) {
    double k1x[3], k1p[3], k2x[3], k2p[3], k3x[3], k3p[3], k4x[3], k4p[3], tx[3], tp[3];
    const double half = 0.5 * h;
    ref_rhs(x, p, mass, spin, k1x, k1p);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        tx[i] = x[i] + half * k1x[i];
        tp[i] = p[i] + half * k1p[i];
    }
    ref_rhs(tx, tp, mass, spin, k2x, k2p);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        tx[i] = x[i] + half * k2x[i];
        tp[i] = p[i] + half * k2p[i];
    }
    ref_rhs(tx, tp, mass, spin, k3x, k3p);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        tx[i] = x[i] + h * k3x[i];
        tp[i] = p[i] + h * k3p[i];
    }
    ref_rhs(tx, tp, mass, spin, k4x, k4p);
    const double sixth = h / 6.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        x[i] += sixth * (k1x[i] + 2.0 * k2x[i] + 2.0 * k3x[i] + k4x[i]);
        p[i] += sixth * (k1p[i] + 2.0 * k2p[i] + 2.0 * k3p[i] + k4p[i]);
    }
}

// BL (r, theta, phi) with the azimuth continued from prev_phi
// (boyer_lindquist_from_kerr_schild).
__device__ __forceinline__ void ref_bl(const double x[3], double spin, double prev_phi, double out[3]) {
    const double r = ref_radius(x, spin);
    const double rho = sqrt(dmax(r * r - x[2] * x[2], 0.0));
    const double two_pi = 2.0 * kPi;
    const double raw_phi = atan2(x[1], x[0]) - atan2(spin, r);
    double delta = raw_phi - prev_phi;
    delta -= two_pi * floor(delta / two_pi + 0.5);
    out[0] = r;
    out[1] = atan2(rho, x[2]);
    out[2] = prev_phi + delta;
}

// ------------------------------------------------------- disk opacity
__device__ __forceinline__ double smooth_unit(double x) {
    const double t = dclamp(x, 0.0, 1.0);
    return t * t * (3.0 - 2.0 * t);
}

__device__ __forceinline__ double reduce_angle(double angle) {
    const double two_pi = 2.0 * kPi;
    return angle - two_pi * floor(angle / two_pi);
}

__device__ __forceinline__ uint32_t hash_u32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

__device__ __forceinline__ float lattice_value(int32_t x, int32_t y, int32_t z, uint32_t seed) {
    const uint32_t h = static_cast<uint32_t>(x) * 374761393U + static_cast<uint32_t>(y) * 668265263U +
                       static_cast<uint32_t>(z) * 2147483647U + seed * 1274126177U;
    return static_cast<float>(hash_u32(h) & 0x00FFFFFFU) * (1.0f / 16777216.0f) * 2.0f - 1.0f;
}

// Trilinear value noise on the integer lattice, smoothstep weights.
__device__ __forceinline__ float value_noise(float x, float y, float z, uint32_t seed) {
    const float fx0 = floorf(x), fy0 = floorf(y), fz0 = floorf(z);
    const int32_t xi = static_cast<int32_t>(fx0), yi = static_cast<int32_t>(fy0), zi = static_cast<int32_t>(fz0);
    const float fx = x - fx0, fy = y - fy0, fz = z - fz0;
    const float ux = fx * fx * (3.0f - 2.0f * fx);
    const float uy = fy * fy * (3.0f - 2.0f * fy);
    const float uz = fz * fz * (3.0f - 2.0f * fz);
    const float c000 = lattice_value(xi, yi, zi, seed), c100 = lattice_value(xi + 1, yi, zi, seed);
    const float c010 = lattice_value(xi, yi + 1, zi, seed), c110 = lattice_value(xi + 1, yi + 1, zi, seed);
    const float c001 = lattice_value(xi, yi, zi + 1, seed), c101 = lattice_value(xi + 1, yi, zi + 1, seed);
    const float c011 = lattice_value(xi, yi + 1, zi + 1, seed), c111 = lattice_value(xi + 1, yi + 1, zi + 1, seed);
    const float c00 = c000 + (c100 - c000) * ux, c10 = c010 + (c110 - c010) * ux;
    const float c01 = c001 + (c101 - c001) * ux, c11 = c011 + (c111 - c011) * ux;
    const float c0 = c00 + (c10 - c00) * uy, c1 = c01 + (c11 - c01) * uy;
    return c0 + (c1 - c0) * uz;
}

__device__ __forceinline__ float fractal_noise(float x, float y, float z, int octaves, uint32_t seed) {
    float total = 0.0f, amplitude = 1.0f, frequency = 1.0f, normalization = 0.0f;
    for (int o = 0; o < octaves; ++o) {
        total += value_noise(x * frequency, y * frequency, z * frequency, seed + static_cast<uint32_t>(o) * 7919U) *
                 amplitude;
        normalization += amplitude;
        amplitude *= 0.5f;
        frequency *= 2.0f;
    }
    return total / fmaxr(normalization, 1e-5f);
}

__device__ __forceinline__ double disk_angular_velocity(double mass, double spin, double radius) {
    const double spin_mag = fabs(spin);
    const double sqrt_m = sqrt(dmax(mass, 1e-12));
    const double sqrt_r = sqrt(dmax(radius, 1e-12));
    const double angular = sqrt_m / (radius * sqrt_r + spin_mag * sqrt_m);
    return (spin < 0.0) ? -angular : angular;
}

// Opacity of the thin procedural disk at a crossing (the part of
// AccretionDiskModel::shade that decides it: envelope, ring, spiral arms,
// turbulence and grain noise).
__device__ double disk_opacity(const TraceParams& P, double radius_in, double azimuth) {
    const double mass = dmax(P.M, 1e-9);
    const double radius = dmax(radius_in, 1e-9);
    const double inner = dmax(P.disk_in, 1e-9);
    const double outer = dmax(P.disk_out, inner * 1.0001);
    const double softness = dmax(static_cast<double>(P.edge_softness), 0.05);
    const double envelope =
        smooth_unit((radius - inner) / (0.8 * softness * mass)) * smooth_unit((outer - radius) / (1.5 * softness * mass));
    const double rotation = static_cast<double>(P.rotation_speed_scale) * P.time;
    const double shear_phase = reduce_angle(azimuth - disk_angular_velocity(mass, P.disk_spin, radius) * rotation);
    const double pattern_phase =
        reduce_angle(azimuth - disk_angular_velocity(mass, P.disk_spin, sqrt(inner * outer)) * rotation);
    const float log_ratio = static_cast<float>(log(dmax(radius / inner, 1.0)));
    const float log_span = fmaxr(static_cast<float>(log(dmax(outer / inner, 1.0001))), 1e-4f);
    const float radial = log_ratio / log_span;
    const float pi_f = 3.14159265358979323846f;
    const float two_pi = 2.0f * pi_f;
    const float ring_wave = 0.5f + 0.5f * cosf(two_pi * P.ring_frequency * radial);
    const float ring_sharpness = fmaxr(P.ring_sharpness, 1.0f);
    const float ring =
        fmaxr(1.0f + P.ring_amplitude * (powf(ring_wave, ring_sharpness) * sqrtf(pi_f * ring_sharpness) - 1.0f), 0.0f);
    const float arms = floorf(P.spiral_arm_count + 0.5f);
    const float spiral =
        (arms > 0.5f)
            ? fmaxr(1.0f + P.spiral_amplitude * cosf(arms * (static_cast<float>(pattern_phase) - P.spiral_pitch * log_ratio)),
                    0.0f)
            : 1.0f;
    const float shear_cos = cosf(static_cast<float>(shear_phase));
    const float shear_sin = sinf(static_cast<float>(shear_phase));
    const int oct = static_cast<int>(P.turbulence_octaves + 0.5f);
    const int octaves = (oct < 1) ? 1 : ((6 < oct) ? 6 : oct);
    const uint32_t seed = static_cast<uint32_t>(fmaxr(P.noise_seed, 0.0f));
    const float detail = static_cast<float>(dclamp(P.detail, 0.0, 1.0));
    const float turbulence_noise = fractal_noise(log_ratio * P.turbulence_scale * P.radial_stretch,
                                                 shear_cos * P.turbulence_scale, shear_sin * P.turbulence_scale,
                                                 octaves, seed);
    const float turbulence = expf(1.5f * P.turbulence_amplitude * detail * turbulence_noise);
    const float grain_noise = value_noise(log_ratio * P.grain_scale * P.radial_stretch, shear_cos * P.grain_scale,
                                          shear_sin * P.grain_scale, seed + 104729U);
    const float grain = fmaxr(1.0f + P.grain_amplitude * detail * grain_noise, 0.0f);
    const float flux = fclampr(ring * spiral * turbulence * grain, 0.0f, 8.0f);
    return dclamp(envelope * 0.95 * static_cast<double>(P.opacity_scale) *
                      dclamp(sqrt(dmin(static_cast<double>(flux), 1.0)), 0.15, 1.0),
                  0.0, 0.98);
}

struct RenderRay {
    double x[3], p[3], point[3];
    double g;
    uint32_t status, iters, rhs;
};

// One pixel through the renderer's tracer from the initial KS state y (p_t = -1).
__device__ __forceinline__ void trace_render(const TraceParams& P, const double y[8], RenderRay& R) {
    const double m = P.M, spin = P.a, rh = P.rh;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        R.x[i] = y[1 + i];
        R.p[i] = y[5 + i];
    }
    const double lz = R.x[0] * R.p[1] - R.x[1] * R.p[0];
    ref_bl(R.x, spin, P.phi_obs, R.point);
    double throughput = 1.0, redshift_rec = 1.0;
    uint32_t status = 0, iters = 0, flags = 0, rk4_steps = 0;
    const double mid_plane = kPi * 0.5;
    bool ended = false;
    for (uint32_t step = 0; step < P.compat_steps && throughput > 0.01; ++step) {
        iters = step + 1;
        if (R.point[0] <= rh * 1.0001) {
            status = RT_CLS_CAPTURED;
            throughput = 0.0;
            ended = true;
            break;
        }
        if (R.point[0] >= P.r_esc) {
            status = RT_CLS_ESCAPED;
            ended = true;
            break;
        }
        const double r_scale = dmax(R.point[0] - rh, 0.02 * m);
        const double far = 1.0 + (P.far_scale - 1.0) * dclamp((R.point[0] - 20.0 * rh) / (80.0 * rh), 0.0, 1.0);
        const double dt_mag = dclamp(P.kappa * sqrt(R.point[0] * r_scale) * far, P.step_min, P.step_max * P.far_scale);
        const double prev[3] = {R.point[0], R.point[1], R.point[2]};
        ref_rk4(R.x, R.p, -dt_mag, m, spin);
        ++rk4_steps;
        if (!isfinite(R.x[0]) || !isfinite(R.x[1]) || !isfinite(R.x[2]) || !isfinite(R.p[0]) || !isfinite(R.p[1]) ||
            !isfinite(R.p[2])) {
            status = RT_CLS_CAPTURED;
            flags |= RT_FLAG_NONFINITE;
            throughput = 0.0;
            ended = true;
            break;
        }
        ref_bl(R.x, spin, prev[2], R.point);
        if (P.disk_on && (prev[1] - mid_plane) * (R.point[1] - mid_plane) <= 0.0) {
            const double span = fabs(R.point[1] - prev[1]);
            const double s_cross = (span > 1e-12) ? dclamp(fabs(prev[1] - mid_plane) / span, 0.0, 1.0) : 0.5;
            const double r_cross = prev[0] + s_cross * (R.point[0] - prev[0]);
            const double phi_cross = prev[2] + s_cross * (R.point[2] - prev[2]);
            if (r_cross >= P.disk_in && r_cross <= P.disk_out && r_cross > rh * 1.02) {
                status |= RT_CLS_DISK;
                redshift_rec = disk_redshift(m, P.disk_spin, r_cross, lz, true);
                throughput *= (1.0 - disk_opacity(P, r_cross, phi_cross));
            }
        }
    }
    if (!ended && throughput > 0.01) flags |= RT_FLAG_BUDGET;
    if ((status & (RT_CLS_CAPTURED | RT_CLS_ESCAPED)) == 0U && R.point[0] < 3.0 * P.rs) throughput = 0.0;
    if ((status & RT_CLS_ESCAPED) != 0U || throughput > 0.01) status |= RT_CLS_ESCAPED;
    R.status = status | flags;
    R.iters = iters;
    R.g = redshift_rec;
    R.rhs = 4u * rk4_steps;
}

// Camera frame: one lane per pixel, the lanes of a wave on an 8x8 tile, the
// tiles around the hole's image first.
template <int Waves>
__global__ __launch_bounds__(256, Waves) void k_trace_render(TraceParams P, OutPtrs O) {
    const uint64_t t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint32_t lane = static_cast<uint32_t>(t & 63);
    uint32_t tx, ty;
    tile_xy(P, static_cast<uint32_t>(t >> 6), tx, ty);
    const uint32_t px = tx * 8 + (lane & 7);
    const uint32_t py = ty * 8 + (lane >> 3);
    const uint32_t rows = P.n_rays / P.cols;
    unsigned long long acc = 0, rhs = 0;
    if (px < P.cols && py < rows) {
        double y0[8], xi;
        camera_state_kerr(P, P.col_first + px, P.row_first + py, y0, &xi);
        RenderRay R;
        trace_render(P, y0, R);
        const uint64_t idx = static_cast<uint64_t>(py) * P.cols + px;
        const uint64_t n = O.n;
        O.cls[idx] = R.status;
        O.x[idx] = 0.0;
        O.p[idx] = -1.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            O.x[(c + 1) * n + idx] = R.x[c];
            O.p[(c + 1) * n + idx] = R.p[c];
            O.bl[c * n + idx] = R.point[c];
        }
        O.lambda[idx] = R.iters * 0.05;
        O.g[idx] = R.g;
        O.steps[idx] = R.iters;
        O.rejected[idx] = 0;
        acc = R.iters;
        rhs = R.rhs;
    }
    const unsigned long long sa = wave_sum(acc), sh = wave_sum(rhs);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&O.counters[0], sa);
        atomicAdd(&O.counters[2], sh);
    }
}

}  // namespace

hipError_t launch_trace_render(const TraceParams& P, const OutPtrs& O, const LaunchShape& L, hipStream_t stream) {
    if (L.block == 0 || L.block > 256 || (L.block & 63) != 0 || P.cols == 0 || P.n_rays % P.cols != 0)
        return hipErrorInvalidValue;
    const uint64_t tiles_y = (P.n_rays / P.cols + 7) / 8;
    const uint64_t threads = static_cast<uint64_t>(P.tiles_x) * tiles_y * 64;
    const uint32_t grid = static_cast<uint32_t>((threads + L.block - 1) / L.block);
    if (L.waves >= 2)
        hipLaunchKernelGGL(k_trace_render<2>, dim3(grid), dim3(L.block), 0, stream, P, O);
    else
        hipLaunchKernelGGL(k_trace_render<1>, dim3(grid), dim3(L.block), 0, stream, P, O);
    return hipGetLastError();
}

}  // namespace rt
