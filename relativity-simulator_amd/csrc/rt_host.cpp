// rt_host.cpp — C++ host runtime of librtsim.so: parameter set-up from the
// C-ABI structs, device buffers, timing, and the single-trace extern "C"
// entry points of include/rt_tracer.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
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

const char* last_error_text() { return g_last_error.c_str(); }

// ------------------------------------------------------ host geometry
namespace {

constexpr double kPiH = 3.14159265358979323846;

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

void to_local_host(const SourceDev& s, const double x[3], double l[3]) {
    const double d[3] = {x[0] - s.c[0], x[1] - s.c[1], x[2] - s.c[2]};
    l[0] = d[0] * s.fx[0] + d[1] * s.fx[1] + d[2] * s.fx[2];
    l[1] = d[0] * s.fy[0] + d[1] * s.fy[1] + d[2] * s.fy[2];
    l[2] = d[0] * s.ax[0] + d[1] * s.ax[1] + d[2] * s.ax[2];
}

}  // namespace

double kerr_isco_host(double m, double a_spin) {
    const double a_star = std::clamp(a_spin / std::max(m, 1e-12), -0.9999999, 0.9999999);
    const double abs_a = std::abs(a_star);
    const double orbit_sign = (a_star >= 0.0) ? -1.0 : 1.0;
    const double cbrt_term = std::cbrt(std::max(1.0 - abs_a * abs_a, 0.0));
    const double z1 = 1.0 + cbrt_term * (std::cbrt(1.0 + abs_a) + std::cbrt(1.0 - abs_a));
    const double z2 = std::sqrt(3.0 * abs_a * abs_a + z1 * z1);
    return (3.0 + z2 + orbit_sign * std::sqrt(std::max((3.0 - z1) * (3.0 + z1 + 2.0 * z2), 0.0))) * m;
}

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
    if (ig->method > RT_METHOD_RENDER_RK4) {
        why = "integrator: unknown method";
        return false;
    }
    if (ig->method == RT_METHOD_RENDER_RK4) {
        if (ig->precision != RT_PRECISION_FP64 || mt->n_sources >= 2) {
            why = "integrator: RT_METHOD_RENDER_RK4 traces single Kerr in FP64";
            return false;
        }
        if (!(ig->step_size_factor > 0.0) || !(ig->min_step > 0.0) || !(ig->far_field_step_scale > 0.0) ||
            !(ig->max_step * ig->far_field_step_scale >= ig->min_step)) {
            why = "integrator: render steps need step_size_factor > 0, min_step > 0, far_field_step_scale > 0 and "
                  "max_step * far_field_step_scale >= min_step";
            return false;
        }
    }
    if (!(tm->escape_radius > 0.0) || !(tm->horizon_eps >= 0.0)) {
        why = "termination: need escape_radius > 0 and horizon_eps >= 0";
        return false;
    }
    if (mt->n_sources > RT_MAX_SOURCES) {
        why = "metric: at most 16 sources";
        return false;
    }
    if (mt->n_sources >= 2) {
        for (uint32_t i = 0; i < mt->n_sources; ++i)
            if (!(mt->src[i].mass > 0.0) || !std::isfinite(mt->src[i].position[0]) ||
                !std::isfinite(mt->src[i].position[1]) || !std::isfinite(mt->src[i].position[2])) {
                why = "metric: every source needs mass > 0 and a finite position";
                return false;
            }
    } else if (!(mt->mass > 0.0)) {
        why = "metric: mass must be > 0";
        return false;
    }
    return true;
}

bool valid_camera(const rt_camera* c, std::string& why) {
    if (!c || c->width == 0 || c->height == 0) {
        why = "camera: width and height must be > 0";
        return false;
    }
    if (c->projection > RT_PROJ_HAMMER_AITOFF) {
        why = "camera: unknown projection";
        return false;
    }
    // the pinhole-type projections take tan(fov / 2); the others any positive angle
    const bool tangent = c->projection == RT_PROJ_PINHOLE || c->projection == RT_PROJ_AUTO_ZOOM ||
                         c->projection == RT_PROJ_PANINI_CYLINDRICAL;
    if (tangent ? !(c->fov_y > 0.0 && c->fov_y < 3.14159) : !(c->fov_y > 0.0 && c->fov_y <= 4.0 * kPiH)) {
        why = tangent ? "camera: fov_y must be in (0, pi) for this projection" : "camera: fov_y must be in (0, 4 pi]";
        return false;
    }
    if (!(c->observer[1] > 0.0)) {
        why = "camera: observer radius must be > 0";
        return false;
    }
    if (static_cast<uint64_t>(c->width) * c->height > (1ull << 31)) {
        why = "camera: more than 2^31 pixels";
        return false;
    }
    return true;
}

// Metric, integrator and termination parts of the parameter block
// (restating dispatch_fp64_scalar's set-up, render/software_compute_engine.hpp:3801-3809).
void fill_params(const rt_metric* mt, const rt_integrator* ig, const rt_termination* tm, TraceParams& P) {
    std::memset(&P, 0, sizeof(P));
    const bool multi = mt->n_sources >= 2;
    if (multi) {
        // central source (src[0]) sets the observer clamp and the budget rule
        const double m0 = std::max(mt->src[0].mass, 1e-4);
        P.M = m0;
        P.a = std::clamp(mt->src[0].spin * m0, -0.999 * m0, 0.999 * m0);
    } else {
        P.M = std::max(mt->mass, 1e-4);
        P.a = std::clamp(mt->spin, -0.999 * P.M, 0.999 * P.M);
    }
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
    P.carter = (!multi && (tm->flags & RT_TERM_CARTER_CAPTURE)) ? 1u : 0u;
    P.carter_gate = P.disk_on ? P.disk_in : 1e300;
    P.n_src = multi ? mt->n_sources : 0u;
    P.rtol = ig->rtol;
    P.atol = ig->atol;
    P.h0 = ig->h0;
    P.h_min = ig->h_min;
    P.h_max = ig->h_max;
    P.inv_tol = ig->invariant_tol;
    P.max_steps = ig->max_steps;
    P.precision = ig->precision;
    P.mixed = (!multi && ig->precision == RT_PRECISION_MIXED) ? 1u : 0u;
    P.switch_r = std::max(ig->mixed_switch_radius, 3.0 * P.rh);
    // RT_METHOD_RENDER_RK4: the render step formula, the exact-path budget
    // rule of dispatch_fp64_scalar (render/software_compute_engine.hpp:3826-3828)
    // and the disk profile's opacity parameters
    P.method = ig->method;
    P.kappa = ig->step_size_factor;
    P.step_min = ig->min_step;
    P.step_max = ig->max_step;
    P.far_scale = ig->far_field_step_scale;
    P.rs = 2.0 * P.M;
    P.compat_steps = std::max(ig->max_steps / 2U, 256U);
    P.time = tm->time;
    const rt_disk_profile& d = tm->disk_profile;
    P.opacity_scale = d.opacity_scale;
    P.edge_softness = d.edge_softness;
    P.rotation_speed_scale = d.rotation_speed_scale;
    P.ring_amplitude = d.ring_amplitude;
    P.ring_frequency = d.ring_frequency;
    P.ring_sharpness = d.ring_sharpness;
    P.spiral_arm_count = d.spiral_arm_count;
    P.spiral_pitch = d.spiral_pitch;
    P.spiral_amplitude = d.spiral_amplitude;
    P.turbulence_amplitude = d.turbulence_amplitude;
    P.turbulence_scale = d.turbulence_scale;
    P.turbulence_octaves = d.turbulence_octaves;
    P.radial_stretch = d.radial_stretch;
    P.grain_amplitude = d.grain_amplitude;
    P.grain_scale = d.grain_scale;
    P.noise_seed = d.noise_seed;
    if (P.method == RT_METHOD_RENDER_RK4) P.carter = 0u;
}

// Source table (restating build_multi_source_data, software_compute_engine.hpp:2738-2798):
// source 0 keeps the world frame when its axis is +z; other sources get the
// frame (helper x axis) x axis, axis x that; spins are a/m clamped to
// +-0.9999 (+-0.999 for source 0).
void fill_sources(const rt_metric* mt, const rt_termination* tm, MultiDev& T) {
    std::memset(&T, 0, sizeof(T));
    T.n = std::min<uint32_t>(mt->n_sources, RT_MAX_SOURCES);
    for (uint32_t i = 0; i < T.n; ++i) {
        const rt_source& b = mt->src[i];
        SourceDev& s = T.src[i];
        const double mb = (i == 0) ? std::max(b.mass, 1e-4) : std::max(b.mass, 1e-9);
        const double ab =
            (i == 0) ? std::clamp(b.spin * mb, -0.999 * mb, 0.999 * mb) : std::clamp(b.spin, -0.9999, 0.9999) * mb;
        s.m = mb;
        s.a = ab;
        s.a2 = ab * ab;
        s.rh = horizon_radius(mb, ab);
        s.r_hz = s.rh + tm->horizon_eps;
        s.rh102 = s.rh * 1.02;
        for (int j = 0; j < 3; ++j) s.c[j] = b.position[j];
        const double al = std::sqrt(b.axis[0] * b.axis[0] + b.axis[1] * b.axis[1] + b.axis[2] * b.axis[2]);
        double ax[3] = {0.0, 0.0, 1.0};
        if (al > 1e-12)
            for (int j = 0; j < 3; ++j) ax[j] = b.axis[j] / al;
        const double helper[3] = {std::abs(ax[2]) < 0.9 ? 0.0 : 1.0, 0.0, std::abs(ax[2]) < 0.9 ? 1.0 : 0.0};
        double fx[3] = {helper[1] * ax[2] - helper[2] * ax[1], helper[2] * ax[0] - helper[0] * ax[2],
                        helper[0] * ax[1] - helper[1] * ax[0]};
        const double fl = std::sqrt(fx[0] * fx[0] + fx[1] * fx[1] + fx[2] * fx[2]);
        for (int j = 0; j < 3; ++j) fx[j] /= fl;
        const double fy[3] = {ax[1] * fx[2] - ax[2] * fx[1], ax[2] * fx[0] - ax[0] * fx[2],
                              ax[0] * fx[1] - ax[1] * fx[0]};
        const bool world = (i == 0 && al > 1e-12 && ax[0] == 0.0 && ax[1] == 0.0 && ax[2] > 0.0);
        for (int j = 0; j < 3; ++j) {
            s.ax[j] = ax[j];
            s.fx[j] = world ? (j == 0 ? 1.0 : 0.0) : fx[j];
            s.fy[j] = world ? (j == 1 ? 1.0 : 0.0) : fy[j];
        }
        s.disk_on = b.disk_enabled ? 1u : 0u;
        const double din = (b.disk_inner > 0.0) ? b.disk_inner : kerr_isco_host(mb, ab);
        s.disk_in = din;
        s.disk_out = std::max((b.disk_outer > 0.0) ? b.disk_outer : 20.0 * mb, din * 1.05);
    }
}

// Camera part: per-frame constants of the pixel loop and the observer frame,
// the same for every pixel, so computed once here. Single Kerr: Gram–Schmidt
// frame of kerr_schild_initial_momentum (core/kerr_schild_null_integrator.hpp:149-177).
// Superposition: flat spherical observer position and the frame of
// ms_initial_momentum (render/software_compute_engine.hpp:2919-2956).
// The rectangle of tiles around the holes' images: rays with impact
// parameter up to 10 m about a hole of mass m (the shadow, at most ~7 m
// across for |a| < m, and the rays that orbit near its edge), where the rays
// that take the most steps are; for a superposition the rectangle holds
// every source's circle. Flat-space estimate from the pinhole mapping of
// camera_outward: a direction with components (f, r, u) on (forward, right,
// up) lands at u = r / (f tan_half), v = u / (f tan_half). A hole behind the
// camera or off the traced window adds nothing.
struct HotHole {
    double c[3];  // world centre
    double m;
};

static void hot_tiles(TraceParams& P, double r_obs, const HotHole* holes, uint32_t n_holes) {
    const double X_obs[3] = {r_obs * P.sth * P.cph, r_obs * P.sth * P.sph, r_obs * P.cth};
    const double W = P.width, H = P.height;
    const double tiles_y = static_cast<double>((P.n_rays / P.cols + 7) / 8);
    double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300;
    for (uint32_t k = 0; k < n_holes; ++k) {
        double d[3], dist2 = 0.0;
        for (int i = 0; i < 3; ++i) {
            d[i] = holes[k].c[i] - X_obs[i];
            dist2 += d[i] * d[i];
        }
        const double dist = std::sqrt(dist2);
        if (!(dist > 0.0) || !(P.tan_half > 0.0)) continue;
        double cf = 0.0, cr = 0.0, cu = 0.0;
        for (int i = 0; i < 3; ++i) {
            cf += P.fwd[i] * d[i] / dist;
            cr += P.right[i] * d[i] / dist;
            cu += P.up[i] * d[i] / dist;
        }
        if (!(cf > 1e-6)) continue;  // behind the camera
        const double u = cr / (cf * P.tan_half), v = cu / (cf * P.tan_half);
        const double px = (u / P.aspect + 1.0) * 0.5 * W - 0.5 - static_cast<double>(P.col_first);
        const double py = (1.0 - v) * 0.5 * H - 0.5 - static_cast<double>(P.row_first);
        const double s = std::min(1.0, 10.0 * holes[k].m / dist);
        const double alpha = std::asin(s);
        const double r_pix = (alpha >= 1.5) ? 1e9 : std::tan(alpha) / P.tan_half * 0.5 * H + 8.0;
        const double hx0 = std::max(0.0, std::floor((px - r_pix) / 8.0));
        const double hx1 = std::min(static_cast<double>(P.tiles_x) - 1.0, std::floor((px + r_pix) / 8.0));
        const double hy0 = std::max(0.0, std::floor((py - r_pix) / 8.0));
        const double hy1 = std::min(tiles_y - 1.0, std::floor((py + r_pix) / 8.0));
        if (!(hx1 >= hx0) || !(hy1 >= hy0)) continue;  // image off the traced window
        x0 = std::min(x0, hx0);
        x1 = std::max(x1, hx1);
        y0 = std::min(y0, hy0);
        y1 = std::max(y1, hy1);
    }
    if (!(x1 >= x0) || !(y1 >= y0)) return;
    P.hot_x0 = static_cast<uint32_t>(x0);
    P.hot_y0 = static_cast<uint32_t>(y0);
    P.hot_w = static_cast<uint32_t>(x1 - x0 + 1.0);
    P.hot_h = static_cast<uint32_t>(y1 - y0 + 1.0);
}

void fill_camera(const rt_camera* c, TraceParams& P, MultiDev* T, const rt_tile* win) {
    // the row window as resolve_row_window takes it (render/software_compute_engine.hpp:1042-1050):
    // count 0 is the whole frame (the offset ignored); otherwise offset + count in 64 bits,
    // clipped at the last row
    uint32_t first = 0, last = c->height;
    if (c->row_count) {
        first = std::min(c->row_offset, c->height);
        last = static_cast<uint32_t>(std::min<uint64_t>(static_cast<uint64_t>(first) + c->row_count, c->height));
    }
    uint32_t col0 = 0, cols = c->width;
    if (win) {  // a rectangle of the frame (checked by valid_window)
        first = win->origin_y;
        last = win->origin_y + win->height;
        col0 = win->origin_x;
        cols = win->width;
    }
    P.width = c->width;
    P.height = c->height;
    P.row_first = first;
    P.col_first = col0;
    P.cols = cols;
    P.n_rays = (last - first) * cols;
    P.tiles_x = (cols + 7) / 8;
    P.mode = 0;
    P.aspect = static_cast<double>(c->width) / static_cast<double>(c->height);
    P.tan_half = std::tan(c->fov_y * 0.5);
    P.fov = c->fov_y;
    P.projection = c->projection;
    // per-frame constants of the projections (observer/camera_projections.hpp:70-196),
    // with the reference's expressions
    switch (c->projection) {
        case RT_PROJ_FISHEYE_STEREOGRAPHIC: {
            const double max_fov = std::clamp(c->fov_y, 0.1, 260.0 * kPiH / 180.0);
            P.proj_c = std::tan(max_fov * 0.5 * 0.5);
            break;
        }
        case RT_PROJ_EQUIRECTANGULAR_360: P.proj_c = c->fov_y / (60.0 * kPiH / 180.0); break;
        case RT_PROJ_FISHEYE_ORTHOGRAPHIC: P.proj_c = std::sin(std::clamp(c->fov_y * 0.5, 0.01, kPiH * 0.499)); break;
        case RT_PROJ_HAMMER_AITOFF: P.proj_c = (60.0 * kPiH / 180.0) / std::max(c->fov_y, 0.01); break;
        default: P.proj_c = 0.0; break;
    }
    for (int i = 0; i < 3; ++i) {
        P.fwd[i] = c->forward[i];
        P.right[i] = c->right[i];
        P.up[i] = c->up[i];
    }
    const double r_obs = std::max(c->observer[1], P.rh * 1.02);
    const double th = std::clamp(c->observer[2], 0.001, kPiH - 0.001);
    const double ph = c->observer[3];
    P.sth = std::sin(th);
    P.cth = std::cos(th);
    P.sph = std::sin(ph);
    P.cph = std::cos(ph);
    P.phi_obs = ph;
    // turbulence anti-aliasing factor of the frame (dispatch_fp64_scalar,
    // render/software_compute_engine.hpp:3834-3836)
    {
        const double r_obs_frame = std::max(c->observer[1], P.rh * 1.02);
        const double angular_pixel_size = c->fov_y / std::max(static_cast<double>(c->width), 1.0);
        const double bh_angular_diameter = (2.0 * P.rh) / r_obs_frame;
        P.detail = std::clamp(bh_angular_diameter / std::max(angular_pixel_size * 24.0, 1e-9), 0.0, 1.0);
    }
    P.hot_x0 = P.hot_y0 = P.hot_w = P.hot_h = 0;
    // hot tiles from the pinhole mapping; other projections trace row by row
    if (c->projection == RT_PROJ_PINHOLE || c->projection == RT_PROJ_AUTO_ZOOM) {
        HotHole holes[kMaxSources];
        uint32_t n_holes = 0;
        if (P.n_src >= 2 && T) {
            for (uint32_t i = 0; i < T->n && i < static_cast<uint32_t>(kMaxSources); ++i) {
                for (int j = 0; j < 3; ++j) holes[n_holes].c[j] = T->src[i].c[j];
                holes[n_holes++].m = T->src[i].m;
            }
        } else if (P.n_src < 2) {
            holes[0] = HotHole{{0.0, 0.0, 0.0}, P.M};
            n_holes = 1;
        }
        hot_tiles(P, r_obs, holes, n_holes);
    }
    if (P.n_src >= 2 && T) {
        P.X0[0] = r_obs * P.sth * P.cph;
        P.X0[1] = r_obs * P.sth * P.sph;
        P.X0[2] = r_obs * P.cth;
        const uint32_t n = T->n;
        T->observer_inside = 0;
        double psum = 0.0;
        for (uint32_t i = 0; i < n; ++i) {
            const SourceDev& s = T->src[i];
            double l[3];
            to_local_host(s, P.X0, l);
            const double diff = (l[0] * l[0] + l[1] * l[1] + l[2] * l[2]) - s.a2;
            const double r2 = std::max(0.5 * (diff + std::sqrt(diff * diff + 4.0 * s.a2 * l[2] * l[2])), 1e-24);
            const double r = std::sqrt(r2);
            if (r <= s.rh * 1.0001 || (r2 - 2.0 * s.m * r + s.a2) <= 0.0) T->observer_inside = 1;
            const HField f = ks_field_host(l, s.m, s.a);
            T->pot[i] = f.H;
            for (int j = 0; j < 3; ++j) T->dir[i][j] = f.k[0] * s.fx[j] + f.k[1] * s.fy[j] + f.k[2] * s.ax[j];
            psum += f.H;
        }
        auto mdot = [&](const double* A, const double* B) {
            double v = -A[0] * B[0] + A[1] * B[1] + A[2] * B[2] + A[3] * B[3];
            for (uint32_t i = 0; i < n; ++i) {
                const double la = A[0] + T->dir[i][0] * A[1] + T->dir[i][1] * A[2] + T->dir[i][2] * A[3];
                const double lb = B[0] + T->dir[i][0] * B[1] + T->dir[i][1] * B[2] + T->dir[i][2] * B[3];
                v += 2.0 * T->pot[i] * la * lb;
            }
            return v;
        };
        const double ils = 1.0 / std::max(1.0 - 2.0 * psum, 1e-6);
        auto lift = [&](double vx, double vy, double vz, double out[4]) {
            double shift = 0.0;
            for (uint32_t i = 0; i < n; ++i)
                shift += T->pot[i] * (T->dir[i][0] * vx + T->dir[i][1] * vy + T->dir[i][2] * vz);
            out[0] = 2.0 * shift * ils;
            out[1] = vx;
            out[2] = vy;
            out[3] = vz;
        };
        double e1[4], e2[4], e3[4];
        lift(1, 0, 0, e1);
        lift(0, 1, 0, e2);
        lift(0, 0, 1, e3);
        const double n1 = std::sqrt(std::max(mdot(e1, e1), 1e-30));
        for (int j = 0; j < 4; ++j) e1[j] /= n1;
        const double e21 = mdot(e2, e1);
        for (int j = 0; j < 4; ++j) e2[j] -= e21 * e1[j];
        const double n2 = std::sqrt(std::max(mdot(e2, e2), 1e-30));
        for (int j = 0; j < 4; ++j) e2[j] /= n2;
        const double e31 = mdot(e3, e1), e32 = mdot(e3, e2);
        for (int j = 0; j < 4; ++j) e3[j] -= e31 * e1[j] + e32 * e2[j];
        const double n3 = std::sqrt(std::max(mdot(e3, e3), 1e-30));
        for (int j = 0; j < 4; ++j) e3[j] /= n3;
        T->u0 = std::sqrt(ils);
        for (int j = 0; j < 4; ++j) {
            T->e[0][j] = e1[j];
            T->e[1][j] = e2[j];
            T->e[2][j] = e3[j];
        }
        return;
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
DeviceArena::~DeviceArena() { release(); }

void DeviceArena::release() {
    if (base) (void)hipFree(base);
    base = nullptr;
    bytes = 0;
}

rt_status DeviceArena::ensure(size_t need) {
    if (need <= bytes && base) return RT_OK;
    release();
    const size_t cap = std::max<size_t>(need, 1 << 16);
    hipError_t e = hipMalloc(&base, cap);
    if (e != hipSuccess) {
        base = nullptr;
        return hip_fail(e, "hipMalloc");
    }
    bytes = cap;
    return RT_OK;
}

static size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

size_t results_bytes(size_t n, bool with_inputs) {
    size_t b = align256(n * 4) + 2 * align256(n * 32) + align256(n * 24) + 2 * align256(n * 8) + 2 * align256(n * 4);
    if (with_inputs) b += 2 * align256(n * 32);
    return b;
}

OutPtrs carve_results(char*& cur, size_t n, bool with_inputs) {
    OutPtrs o{};
    auto take = [&](size_t bytes) {
        char* p = cur;
        cur += align256(bytes);
        return p;
    };
    o.cls = reinterpret_cast<uint32_t*>(take(n * 4));
    o.x = reinterpret_cast<double*>(take(n * 32));
    o.p = reinterpret_cast<double*>(take(n * 32));
    o.bl = reinterpret_cast<double*>(take(n * 24));
    o.lambda = reinterpret_cast<double*>(take(n * 8));
    o.g = reinterpret_cast<double*>(take(n * 8));
    o.steps = reinterpret_cast<uint32_t*>(take(n * 4));
    o.rejected = reinterpret_cast<uint32_t*>(take(n * 4));
    if (with_inputs) {
        o.in_x = reinterpret_cast<const double*>(take(n * 32));
        o.in_p = reinterpret_cast<const double*>(take(n * 32));
    }
    o.n = n;
    return o;
}

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

using namespace rt;

namespace {

// Trace one request (camera or explicit rays) on the context's stream into
// the context arena; times the kernel with HIP events.
// x_soa/p_soa non-null: explicit rays (n of them), else camera cam, over the
// rectangle win of its frame when win is non-null.
rt_status trace_one(rt_context* ctx, const rt_camera* cam, const double* x_soa, const double* p_soa, size_t n_in,
                    const rt_metric* metric, const rt_integrator* integ, const rt_termination* term,
                    const rt_tile* win = nullptr) {
    const auto t0 = std::chrono::steady_clock::now();
    RT_HIP(hipSetDevice(ctx->device));
    TraceParams P;
    fill_params(metric, integ, term, P);
    const bool multi = P.n_src >= 2;
    MultiDev T;
    if (multi) fill_sources(metric, term, T);
    size_t n;
    if (cam) {
        fill_camera(cam, P, multi ? &T : nullptr, win);
        n = P.n_rays;
    } else {
        P.mode = 1;
        P.n_rays = static_cast<uint32_t>(n_in);
        n = n_in;
        if (multi) {
            T.observer_inside = 0;
            T.u0 = 1.0;
        }
    }
    if (n == 0) return fail(RT_ERR_INVALID_ARG, "no rays to trace");
    if (P.method == RT_METHOD_RENDER_RK4 && !cam)
        return fail(RT_ERR_INVALID_ARG, "RT_METHOD_RENDER_RK4 traces cameras, not explicit rays");
    const bool inputs = (cam == nullptr);
    rt_status st = ctx->arena.ensure(results_bytes(n, inputs));
    if (st != RT_OK) return st;
    st = ctx->aux.ensure(sizeof(MultiDev) + 256);
    if (st != RT_OK) return st;
    char* cur = static_cast<char*>(ctx->arena.base);
    OutPtrs O = carve_results(cur, n, inputs);
    unsigned long long* counters = reinterpret_cast<unsigned long long*>(ctx->aux.base);
    MultiDev* Td = reinterpret_cast<MultiDev*>(static_cast<char*>(ctx->aux.base) + 256);
    O.counters = counters;
    if (inputs) {
        for (int c = 0; c < 4; ++c) {
            RT_HIP(hipMemcpyAsync(const_cast<double*>(O.in_x) + c * n, x_soa + c * n, n * 8, hipMemcpyHostToDevice,
                                  ctx->stream));
            RT_HIP(hipMemcpyAsync(const_cast<double*>(O.in_p) + c * n, p_soa + c * n, n * 8, hipMemcpyHostToDevice,
                                  ctx->stream));
        }
    }
    if (multi) RT_HIP(hipMemcpyAsync(Td, &T, sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    RT_HIP(launch_zero_u64(counters, 16, ctx->stream));
    RT_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    if (multi)
        RT_HIP(launch_trace_multi(P, O, Td, counters + 8, kMultiShape, ctx->stream));
    else if (P.method == RT_METHOD_RENDER_RK4)
        RT_HIP(launch_trace_render(P, O, kPlainShape, ctx->stream));
    else if (cam)
        RT_HIP(launch_trace_camera(P, O, ctx->shape, ctx->stream));
    else
        RT_HIP(launch_trace_rays(P, O, ctx->shape, ctx->stream));
    RT_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    RT_HIP(hipEventSynchronize(ctx->ev1));
    float ms = 0.0f;
    RT_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    unsigned long long cnt[5] = {0, 0, 0, 0, 0};
    RT_HIP(hipMemcpy(cnt, counters, sizeof(cnt), hipMemcpyDeviceToHost));
    ctx->last = O;
    ctx->n_last = n;
    rt_stats& s = ctx->stats;
    s.kernel_ms = ms;
    s.rays = n;
    s.accepted_steps = cnt[0];
    s.rejected_steps = cnt[1];
    s.rhs_evals = cnt[2];
    s.fp32_steps = cnt[3];
    s.block_size = multi ? kMultiShape.block : ctx->shape.block;
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
    o->method = RT_METHOD_DP45;
    o->step_size_factor = 0.5;
    o->min_step = 0.004;
    o->max_step = 3.5;
    o->far_field_step_scale = 7.8;
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
    o->time = 0.0;
    // GpuDiskProfile's defaults (render/gpu_types.hpp:42-77)
    o->disk_profile = rt_disk_profile{1.0f,  1.0f,  24.0f, 9500.0f, 1200.0f, 0.75f, 1.0f, 2.0f, 0.45f, 1.0f, 0.0f,
                                      1.0f,  1.0f,  1.0f,  1.0f,    4.0f,    0.35f, 9.0f, 3.0f, 2.0f,  1.2f, 0.25f,
                                      0.45f, 5.0f,  4.0f,  3.0f,    0.2f,    40.0f, 1.0f, 1.0f, 1.0f,  1.0f};
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
const char* rt_version(void) { return "librtsim 0.2 (gfx950, DP45 Kerr-Schild null tracer)"; }
const char* rt_last_error(void) { return rt::last_error_text(); }

rt_status rt_context_create(int device, rt_context** out) {
    if (!out) return fail(RT_ERR_INVALID_ARG, "null out");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(RT_ERR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= n) return fail(RT_ERR_INVALID_ARG, "device index out of range");
    rt_context* ctx = new (std::nothrow) rt_context();
    if (!ctx) return fail(RT_ERR_OOM, "context allocation");
    ctx->device = device;
    ctx->shape = kPlainShape;
    ctx->qshape = kQueueShape;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev0);
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev1);
    for (int i = 0; i < rt_context::kSide && e == hipSuccess; ++i) {
        e = hipStreamCreateWithFlags(&ctx->side[i], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->join[i], hipEventDisableTiming);
    }
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->fork, hipEventDisableTiming);
    if (e != hipSuccess) {
        rt_context_destroy(ctx);
        return hip_fail(e, "context set-up");
    }
    const rt_status ss = rt::scratch_setup(device);
    if (ss != RT_OK) {
        rt_context_destroy(ctx);
        return ss;
    }
    *out = ctx;
    return RT_OK;
}

rt_status rt_context_destroy(rt_context* ctx) {
    if (!ctx) return RT_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    rt::forget_graphs(ctx);
    ctx->arena.release();
    ctx->aux.release();
    ctx->pix.release();
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    for (int i = 0; i < rt_context::kSide; ++i) {
        if (ctx->side[i]) (void)hipStreamDestroy(ctx->side[i]);
        if (ctx->join[i]) (void)hipEventDestroy(ctx->join[i]);
    }
    if (ctx->fork) (void)hipEventDestroy(ctx->fork);
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
    std::lock_guard<std::mutex> lk(ctx->mu);
    rt_status st = trace_one(ctx, cam, nullptr, nullptr, 0, metric, integ, term);
    if (st != RT_OK) return st;
    if (n_rays) *n_rays = ctx->n_last;
    return RT_OK;
}

rt_status rt_trace_tile(rt_context* ctx, const rt_camera* cam, const rt_metric* metric, const rt_integrator* integ,
                        const rt_termination* term, const rt_tile* region, size_t* n_rays) {
    if (!ctx || !region) return fail(RT_ERR_INVALID_ARG, "null context or region");
    std::string why;
    if (!valid_request(metric, integ, term, why) || !valid_camera(cam, why)) return fail(RT_ERR_INVALID_ARG, why);
    if (region->width == 0 || region->height == 0 || region->origin_x >= cam->width ||
        region->origin_y >= cam->height || region->width > cam->width - region->origin_x ||
        region->height > cam->height - region->origin_y)
        return fail(RT_ERR_INVALID_ARG, "region: empty, or not inside the camera's frame");
    std::lock_guard<std::mutex> lk(ctx->mu);
    rt_status st = trace_one(ctx, cam, nullptr, nullptr, 0, metric, integ, term, region);
    if (st != RT_OK) return st;
    if (n_rays) *n_rays = ctx->n_last;
    return RT_OK;
}

rt_status rt_fetch_results(rt_context* ctx, rt_results* out) {
    if (!ctx || !out) return fail(RT_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (ctx->n_last == 0) return fail(RT_ERR_INVALID_ARG, "no resident results");
    RT_HIP(hipSetDevice(ctx->device));
    return copy_results_to_host(ctx->last, ctx->n_last, out, ctx->stream);
}

rt_status rt_fetch_pixels(rt_context* ctx, rt_pixel* out) {
    static_assert(sizeof(rt_pixel) == 32, "rt_pixel is the 32-byte GpuPixelOutput record");
    if (!ctx || !out) return fail(RT_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (ctx->n_last == 0) return fail(RT_ERR_INVALID_ARG, "no resident results");
    RT_HIP(hipSetDevice(ctx->device));
    const size_t bytes = ctx->n_last * sizeof(rt_pixel);
    rt_status st = ctx->pix.ensure(bytes);
    if (st != RT_OK) return st;
    RT_HIP(launch_pack_pixels(ctx->last, ctx->n_last, ctx->pix.base, ctx->stream));
    RT_HIP(hipMemcpyAsync(out, ctx->pix.base, bytes, hipMemcpyDeviceToHost, ctx->stream));
    RT_HIP(hipStreamSynchronize(ctx->stream));
    return RT_OK;
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
    if (n > (1ull << 31)) return fail(RT_ERR_INVALID_ARG, "more than 2^31 rays");
    std::string why;
    if (!valid_request(metric, integ, term, why)) return fail(RT_ERR_INVALID_ARG, why);
    std::lock_guard<std::mutex> lk(ctx->mu);
    rt_status st = trace_one(ctx, nullptr, x_soa, p_soa, n, metric, integ, term);
    if (st != RT_OK) return st;
    return copy_results_to_host(ctx->last, n, out, ctx->stream);
}

}  // extern "C"
