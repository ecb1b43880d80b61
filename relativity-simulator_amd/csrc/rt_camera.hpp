// rt_camera.hpp — pixel -> ray direction on the device: the camera's
// projections and the spherical-basis projection of the world direction.
//
// Follows (paths relative to the Relativity-Simulator tree):
//   CameraProjector::compute_ray_direction and its projections
//     observer/camera_projections.hpp:27-225 (as dispatch_fp64_scalar calls
//     it, with no observer boost: the auto-zoom ray is the pinhole one)
//   pixel -> screen coordinates, world direction, e_r/e_theta/e_phi
//     projection and the `outward` vector
//     render/software_compute_engine.hpp:3843-3938, 1634-1645
// The operations keep the reference's order; the per-frame constants
// (tangents, sines and scales of the field of view) come from the host.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/rt_tracer.h"
#include "rt_params.h"

namespace rt {

__host__ __device__ __forceinline__ void unit3(double n1, double n2, double n3, double nl[3]) {
    const double len = sqrt(n1 * n1 + n2 * n2 + n3 * n3);
    const double inv_len = (len > 0.0) ? (1.0 / len) : 1.0;
    nl[0] = n1 * inv_len;
    nl[1] = n2 * inv_len;
    nl[2] = n3 * inv_len;
}

// (forward, up, right) components of the ray through screen point (u, v).
// P.projection is wave-uniform, so the switch is a scalar branch.
__host__ __device__ __forceinline__ void screen_direction(const TraceParams& P, double u, double v, double nl[3]) {
    constexpr double kPiD = 3.14159265358979323846;
    switch (P.projection) {
        case RT_PROJ_FISHEYE_STEREOGRAPHIC: {  // proj_c = tan(half of the clamped field / 2)
            const double r = sqrt(u * u + v * v);
            if (r < 1e-15) {
                nl[0] = 1.0, nl[1] = 0.0, nl[2] = 0.0;
                return;
            }
            const double theta = 2.0 * atan(r * P.proj_c);
            double sin_t, cos_t;
            sincos(theta, &sin_t, &cos_t);
            unit3(cos_t, v * (sin_t / r), u * (sin_t / r), nl);
            return;
        }
        case RT_PROJ_EQUIRECTANGULAR_360: {  // proj_c = fov / 60 deg
            const double phi = u * kPiD * P.proj_c;
            const double th = kPiD * 0.5 - v * (kPiD * 0.5) * P.proj_c;
            const double theta = (th < 0.0001) ? 0.0001 : ((kPiD - 0.0001) < th ? (kPiD - 0.0001) : th);
            double sin_t, cos_t, sin_p, cos_p;
            sincos(theta, &sin_t, &cos_t);
            sincos(phi, &sin_p, &cos_p);
            unit3(sin_t * cos_p, cos_t, sin_t * sin_p, nl);
            return;
        }
        case RT_PROJ_FISHEYE_EQUIDISTANT: {
            const double r = sqrt(u * u + v * v);
            if (r < 1e-15) {
                nl[0] = 1.0, nl[1] = 0.0, nl[2] = 0.0;
                return;
            }
            const double t = r * P.fov * 0.5;
            const double theta = (kPiD < t) ? kPiD : t;
            double sin_t, cos_t;
            sincos(theta, &sin_t, &cos_t);
            unit3(cos_t, v * (sin_t / r), u * (sin_t / r), nl);
            return;
        }
        case RT_PROJ_FISHEYE_ORTHOGRAPHIC: {  // proj_c = sin(clamped half field)
            const double r = sqrt(u * u + v * v);
            const double s = r * P.proj_c;
            const double sin_t = (s < 0.0) ? 0.0 : ((1.0 < s) ? 1.0 : s);
            const double c2 = 1.0 - sin_t * sin_t;
            const double cos_t = sqrt((c2 < 0.0) ? 0.0 : c2);
            if (r < 1e-15) {
                nl[0] = 1.0, nl[1] = 0.0, nl[2] = 0.0;
                return;
            }
            unit3(cos_t, v * (sin_t / r), u * (sin_t / r), nl);
            return;
        }
        case RT_PROJ_PANINI_CYLINDRICAL: {  // tan_half = tan(fov / 2); d = 1
            const double d = 1.0;
            const double px = u * P.tan_half, py = v * P.tan_half;
            const double q = 1.0 + (1.0 - d * d) * px * px / ((d + 1.0) * (d + 1.0));
            const double s = (d + 1.0) / (d + sqrt((q < 1e-15) ? 1e-15 : q));
            const double sin_phi = (px * s) / (d + 1.0);
            const double c2 = 1.0 - sin_phi * sin_phi;
            const double cos_phi = sqrt((c2 < 0.0) ? 0.0 : c2);
            unit3(cos_phi, py, sin_phi, nl);
            return;
        }
        case RT_PROJ_HAMMER_AITOFF: {  // proj_c = 60 deg / max(fov, 0.01); not renormalised
            const double sqrt2 = 1.4142135623730951;
            const double xs = u * P.proj_c * sqrt2;
            const double x = (xs < -sqrt2) ? -sqrt2 : ((sqrt2 < xs) ? sqrt2 : xs);
            const double ys = v * P.proj_c * (sqrt2 * 0.5);
            const double y = (ys < -(sqrt2 * 0.5)) ? -(sqrt2 * 0.5) : (((sqrt2 * 0.5) < ys) ? (sqrt2 * 0.5) : ys);
            const double z2 = 1.0 - (x * x * 0.25) - (y * y);
            const double z = sqrt((z2 < 1e-15) ? 1e-15 : z2);
            const double phi = 2.0 * atan2(z * x, 2.0 * (2.0 * z * z - 1.0));
            const double as = z * y * sqrt2;
            const double theta = kPiD * 0.5 - asin((as < -1.0) ? -1.0 : ((1.0 < as) ? 1.0 : as));
            double sin_t, cos_t, sin_p, cos_p;
            sincos(theta, &sin_t, &cos_t);
            sincos(phi, &sin_p, &cos_p);
            nl[0] = sin_t * cos_p;
            nl[1] = cos_t;
            nl[2] = sin_t * sin_p;
            return;
        }
        default:  // pinhole, and auto-zoom with no boost
            unit3(1.0, v * P.tan_half, u * P.tan_half, nl);
            return;
    }
}

// Every projection other than the pinhole one, out of line on the device.
// The persistent kernels map pixels inside their ray loop; inlined there, the
// library constants of these projections (sin, cos, atan, asin polynomials)
// would be hoisted out of the loop and held in registers for the kernel's
// whole life.
inline __device__ __noinline__ void screen_direction_out_of_line(const TraceParams& P, double u, double v,
                                                                 double nl[3]) {
    screen_direction(P, u, v, nl);
}

// Camera pixel (px, py) -> the reference's `outward` vector: screen point,
// ray direction, world direction from the camera basis, projection onto
// e_r / e_theta / e_phi at the observer and back to Cartesian.
// InLoop (device code inside a ray loop): the pinhole projection inline, the
// others through screen_direction_out_of_line; same results either way.
template <bool InLoop = false>
__host__ __device__ __forceinline__ void camera_outward(const TraceParams& P, uint32_t px, uint32_t py, double o[3]) {
    const double v = 1.0 - (static_cast<double>(py) + 0.5) / static_cast<double>(P.height) * 2.0;
    const double xs = (static_cast<double>(px) + 0.5) / static_cast<double>(P.width) * 2.0 - 1.0;
    const bool allsky = P.projection == RT_PROJ_EQUIRECTANGULAR_360 || P.projection == RT_PROJ_HAMMER_AITOFF;
    const double u = allsky ? xs : xs * P.aspect;
    double l[3];
    if constexpr (InLoop) {
        if (P.projection == RT_PROJ_PINHOLE || P.projection == RT_PROJ_AUTO_ZOOM)
            unit3(1.0, v * P.tan_half, u * P.tan_half, l);
        else
            screen_direction_out_of_line(P, u, v, l);
    } else {
        screen_direction(P, u, v, l);
    }
    const double d0 = l[0] * P.fwd[0] + l[2] * P.right[0] + l[1] * P.up[0];
    const double d1 = l[0] * P.fwd[1] + l[2] * P.right[1] + l[1] * P.up[1];
    const double d2 = l[0] * P.fwd[2] + l[2] * P.right[2] + l[1] * P.up[2];
    const double st = P.sth, ct = P.cth, sp = P.sph, cp = P.cph;
    const double n_r = d0 * (st * cp) + d1 * (st * sp) + d2 * ct;
    const double n_th = d0 * (ct * cp) + d1 * (ct * sp) + d2 * (-st);
    const double n_ph = d0 * (-sp) + d1 * cp + d2 * 0.0;
    o[0] = n_r * st * cp + n_th * ct * cp - n_ph * sp;
    o[1] = n_r * st * sp + n_th * ct * sp + n_ph * cp;
    o[2] = n_r * ct - n_th * st;
}

// Photon momentum -outward in the observer's Gram–Schmidt frame, normalised
// to p_t = -1 (kerr_schild_null_integrator.hpp:179-193). The frame is the same
// for every pixel and comes precomputed in P.
template <bool InLoop = false>
__host__ __device__ __forceinline__ void camera_state_kerr(const TraceParams& P, uint32_t px, uint32_t py, double y[8],
                                                           double* xi) {
    double o[3];
    camera_outward<InLoop>(P, px, py, o);
    double vv[4] = {P.u0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        vv[c] -= o[0] * P.frame[0][c];
        vv[c] -= o[1] * P.frame[1][c];
        vv[c] -= o[2] * P.frame[2][c];
    }
    const double two_h = P.two_h_obs;
    const double lv = vv[0] + P.kobs[0] * vv[1] + P.kobs[1] * vv[2] + P.kobs[2] * vv[3];
    const double p_t = -vv[0] + two_h * lv;
    const double inv_e = 1.0 / fmax(-p_t, 1e-12);
    y[0] = 0.0;
    y[1] = P.X0[0];
    y[2] = P.X0[1];
    y[3] = P.X0[2];
    y[4] = -1.0;
    y[5] = (vv[1] + two_h * lv * P.kobs[0]) * inv_e;
    y[6] = (vv[2] + two_h * lv * P.kobs[1]) * inv_e;
    y[7] = (vv[3] + two_h * lv * P.kobs[2]) * inv_e;
    *xi = y[1] * y[6] - y[2] * y[5];
}

// Tile t of a frame (0 <= t < tiles) in hot-first order: the hot rectangle
// row by row, then the other tiles row by row.
__device__ __forceinline__ void tile_xy(const TraceParams& P, uint32_t t, uint32_t& tx, uint32_t& ty) {
    const uint32_t TX = P.tiles_x;
    const uint32_t n_hot = P.hot_w * P.hot_h;
    if (t < n_hot) {
        tx = P.hot_x0 + t % P.hot_w;
        ty = P.hot_y0 + t / P.hot_w;
        return;
    }
    uint32_t k = t - n_hot;
    const uint32_t above = P.hot_y0 * TX;
    if (k >= above) {
        k -= above;
        const uint32_t per_row = TX - P.hot_w;
        const uint32_t mid = P.hot_h * per_row;
        if (k < mid) {
            ty = P.hot_y0 + k / per_row;
            tx = k % per_row;
            if (tx >= P.hot_x0) tx += P.hot_w;
            return;
        }
        k = (P.hot_y0 + P.hot_h) * TX + (k - mid);
    }
    tx = k % TX;
    ty = k / TX;
}

}  // namespace rt
