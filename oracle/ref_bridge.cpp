// ref_bridge.cpp — extern "C" shims over the reference's own headers.
//
// TEST INFRASTRUCTURE ONLY. Compiled by oracle/Makefile against the
// Relativity-Simulator headers where they lie (include path given on the
// command line), into oracle/_ref/libref.so. Used by tests/ and by
// tests/golden/make_golden.py to pin the clean-room oracle to the reference's
// own functions; nothing here is reference code, only calls into it.
#include <cmath>
#include <cstddef>
#include <cstring>
#include <cstdint>

#include "relativistic/core/kerr_schild_null_integrator.hpp"
#include "relativistic/integrators/rk45_adaptive.hpp"
#include "relativistic/metrics/kerr_schild.hpp"
#include "relativistic/metrics/bardeen_shadow.hpp"
#include "relativistic/optics/disk_thermal_profile.hpp"
#include "relativistic/render/accretion_disk/accretion_disk_model.hpp"
#include "relativistic/render/software_compute_engine.hpp"
#include "relativistic/distributed/task_payloads.hpp"
#include "relativistic/observer/camera_projections.hpp"

#include <span>
#include <vector>

#include "../include/rt_tracer.h"

using namespace Relativistic;

namespace {

// The renderer's superposition helpers (build_multi_source_data, ms_rhs,
// ms_initial_momentum) are private static members of SoftwareComputeEngine.
// An explicit template instantiation may name private members, and a friend
// defined by that instantiation hands the member pointer out; nothing of the
// reference is restated here.
template <class Tag, auto Member>
struct Expose {
    friend auto exposed(Tag) { return Member; }
};
struct BuildTag {};
struct MsRhsTag {};
struct MsMomentumTag {};
auto exposed(BuildTag);
auto exposed(MsRhsTag);
auto exposed(MsMomentumTag);
template struct Expose<BuildTag, &Render::SoftwareComputeEngine::build_multi_source_data>;
template struct Expose<MsRhsTag, &Render::SoftwareComputeEngine::ms_rhs>;
template struct Expose<MsMomentumTag, &Render::SoftwareComputeEngine::ms_initial_momentum>;

// rt_metric (src[0] = central hole with dimensionless spin, src[1..] = bodies)
// -> the renderer's push constants and GpuBodyData with preset_3d == 8, the
// field mapping of the C ABI. Disk radii map to inner_radius_scale = r_in /
// ISCO (1 when r_in = 0) and outer_radius_mass_units = r_out / m.
auto build_ms(const rt_metric* mt) {
    Render::GpuCameraPushConstants pc{};
    pc.metric_type = 3U;
    const rt_source& c = mt->src[0];
    pc.metric_mass = c.mass;
    pc.metric_spin = c.spin * c.mass;
    auto disk = [](Render::GpuDiskProfile& d, const rt_source& s, double m, double a) {
        d.enabled = s.disk_enabled ? 1.0f : 0.0f;
        const double isco = Optics::DiskThermalProfile::kerr_isco_radius(m, a);
        d.inner_radius_scale = s.disk_inner > 0 ? static_cast<float>(s.disk_inner / isco) : 1.0f;
        d.outer_radius_mass_units = static_cast<float>((s.disk_outer > 0 ? s.disk_outer : 20.0 * m) / m);
    };
    {
        const double m = std::max(c.mass, 1e-4);
        disk(pc.primary_disk, c, m, std::clamp(c.spin * c.mass, -0.999 * m, 0.999 * m));
    }
    std::vector<Render::GpuBodyData> bodies(mt->n_sources > 1 ? mt->n_sources - 1 : 0);
    for (size_t i = 0; i < bodies.size(); ++i) {
        const rt_source& s = mt->src[i + 1];
        Render::GpuBodyData& b = bodies[i];
        b.preset_3d = 8U;
        b.position = {s.position[0], s.position[1], s.position[2], 0.0};
        b.mass = s.mass;
        b.spin_parameter = s.spin;
        b.spin_axis = {s.axis[0], s.axis[1], s.axis[2], 0.0};
        const double m = std::max(s.mass, 1e-9);
        disk(b.disk, s, m, std::clamp(s.spin, -0.9999, 0.9999) * m);
    }
    return exposed(BuildTag{})(pc, std::span<const Render::GpuBodyData>(bodies.data(), bodies.size()));
}

}  // namespace

extern "C" {

// core/kerr_schild_null_integrator.hpp:30-44
void ref_ks_field(const double x[3], double M, double a, double out[5]) {
    const auto f = Core::kerr_schild_field<double>({x[0], x[1], x[2]}, M, a);
    out[0] = f.r;
    out[1] = f.h;
    out[2] = f.k[0];
    out[3] = f.k[1];
    out[4] = f.k[2];
}

// core/kerr_schild_null_integrator.hpp:71-114 (p_t = -1)
void ref_ks_rhs(const double x[3], const double p[3], double M, double a, double dx[3], double dp[3]) {
    Core::KerrSchildNullState<double> s;
    s.position = {x[0], x[1], x[2]};
    s.momentum = {p[0], p[1], p[2]};
    const auto d = Core::kerr_schild_null_derivatives(s, M, a);
    for (int i = 0; i < 3; ++i) {
        dx[i] = d.position[i];
        dp[i] = d.momentum[i];
    }
}

// core/kerr_schild_null_integrator.hpp:116-139
void ref_ks_rk4_step(double x[3], double p[3], double h, double M, double a) {
    Core::KerrSchildNullState<double> s;
    s.position = {x[0], x[1], x[2]};
    s.momentum = {p[0], p[1], p[2]};
    Core::step_kerr_schild_null_rk4(s, h, M, a);
    for (int i = 0; i < 3; ++i) {
        x[i] = s.position[i];
        p[i] = s.momentum[i];
    }
}

// core/kerr_schild_null_integrator.hpp:141-194
void ref_initial_momentum(const double x[3], double M, double a, const double outward[3], double p[3]) {
    const auto q = Core::kerr_schild_initial_momentum<double>({x[0], x[1], x[2]}, M, a,
                                                             {outward[0], outward[1], outward[2]});
    p[0] = q[0];
    p[1] = q[1];
    p[2] = q[2];
}

// core/kerr_schild_null_integrator.hpp:46-54
void ref_ks_from_bl(double r, double th, double ph, double a, double out[3]) {
    const auto q = Core::kerr_schild_position_from_boyer_lindquist<double>(r, th, ph, a);
    out[0] = q[0];
    out[1] = q[1];
    out[2] = q[2];
}

// core/kerr_schild_null_integrator.hpp:56-69
void ref_bl_from_ks(const double x[3], double a, double prev_phi, double out[3]) {
    const auto q = Core::boyer_lindquist_from_kerr_schild<double>({x[0], x[1], x[2]}, a, prev_phi);
    out[0] = q.r;
    out[1] = q.theta;
    out[2] = q.phi;
}

// optics/disk_thermal_profile.hpp:10-20 (same formula as the renderer's
// private kerr_isco_radius, render/software_compute_engine.hpp:1052-1061)
double ref_isco(double M, double a) { return Optics::DiskThermalProfile::kerr_isco_radius(M, a); }

// render/accretion_disk/accretion_disk_model.hpp:205-225
double ref_redshift(double M, double a, double r, double xi) {
    return Render::AccretionDiskModel::redshift_at(M, a, r, xi);
}

// g_{mu nu} and g^{mu nu} of metrics/kerr_schild.hpp:141-189 (row-major 4x4).
void ref_ks_metric(const double x4[4], double M, double a, double g[16], double ginv[16]) {
    const Metrics::KerrSchildMetric<double> m(M, a);
    const Core::FourVector<double> x(x4[0], x4[1], x4[2], x4[3]);
    const auto G = m.metric_tensor(x);
    const auto Gi = m.inverse_metric(x);
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            g[i * 4 + j] = G(i, j);
            ginv[i * 4 + j] = Gi(i, j);
        }
}

// Integrate a null geodesic with the reference's Christoffel-based adaptive
// DP45 (integrators/rk45_adaptive.hpp:158-313) on metrics/kerr_schild.hpp, from
// (x0, u0) forward to affine parameter L exactly (the last step is shortened).
// Returns 0 on success, 1 if the integrator gave up.
int ref_rk45_integrate(const double x0[4], const double u0[4], double M, double a, double L, double rtol,
                       double atol, double h0, double x_out[4], double u_out[4], uint64_t* acc, uint64_t* rej) {
    const Metrics::KerrSchildMetric<double> m(M, a);
    Integrators::RK45Config<double> cfg;
    cfg.rtol = rtol;
    cfg.atol = atol;
    cfg.initial_step = h0;
    cfg.max_step = 1e13;
    Integrators::RK45AdaptiveIntegrator<Metrics::KerrSchildMetric<double>> integ(m, Integrators::GeodesicType::Null,
                                                                               cfg);
    Integrators::GeodesicState<double> s;
    s.x = Core::FourVector<double>(x0[0], x0[1], x0[2], x0[3]);
    s.u = Core::FourVector<double>(u0[0], u0[1], u0[2], u0[3]);
    double lam = 0.0, dt = h0;
    int guard = 0;
    while (lam < L && guard++ < 10000000) {
        double want = dt;
        const bool last = (lam + want >= L);
        if (last) want = L - lam;
        double trial = want;
        const auto used = integ.step(s, trial);
        if (!used) return 1;
        lam += *used;
        // keep the controller's proposal unless this was the shortened last step
        dt = last ? dt : trial;
        if (last && *used == want) break;
    }
    for (int i = 0; i < 4; ++i) {
        x_out[i] = s.x(i);
        u_out[i] = s.u(i);
    }
    *acc = integ.statistics().accepted_steps;
    *rej = integ.statistics().rejected_steps;
    return 0;
}

// render/software_compute_engine.hpp:2738-2798 (build_multi_source_data), in the
// 17-doubles-per-source layout of rto_ms_table. Returns the source count.
int ref_ms_table(const rt_metric* mt, double* out) {
    const auto ms = build_ms(mt);
    for (size_t i = 0; i < ms.count; ++i) {
        double* o = out + 17 * i;
        for (int j = 0; j < 3; ++j) {
            o[j] = ms.position_mass[i][j];
            o[6 + j] = ms.frame_x_horizon[i][j];
            o[9 + j] = ms.frame_y_equatorial[i][j];
            o[12 + j] = ms.axis_spin[i][j];
        }
        o[3] = ms.position_mass[i][3];
        o[4] = ms.axis_spin[i][3];
        o[5] = ms.frame_x_horizon[i][3];
        o[15] = ms.disk_range[i][0];
        o[16] = ms.disk_range[i][1];
    }
    return static_cast<int>(ms.count);
}

// render/software_compute_engine.hpp:2857-2870 (ms_rhs, p_t = -1)
void ref_ms_rhs(const rt_metric* mt, const double x[3], const double p[3], double dx[3], double dp[3]) {
    const auto ms = build_ms(mt);
    std::array<double, 3> ox{}, op{};
    exposed(MsRhsTag{})(ms, {x[0], x[1], x[2]}, {p[0], p[1], p[2]}, ox, op);
    for (int i = 0; i < 3; ++i) {
        dx[i] = ox[i];
        dp[i] = op[i];
    }
}

// render/software_compute_engine.hpp:2919-2964 (ms_initial_momentum)
void ref_ms_initial_momentum(const rt_metric* mt, const double x[3], const double outward[3], double p[3]) {
    const auto ms = build_ms(mt);
    const auto q = exposed(MsMomentumTag{})(ms, {x[0], x[1], x[2]}, {outward[0], outward[1], outward[2]});
    p[0] = q[0];
    p[1] = q[1];
    p[2] = q[2];
}

// ---- 32x32 tile tasks (src/distributed/task_payloads.cpp, linked from
// the reference's own source by oracle/Makefile)

// resolve_tile_geometry (task_payloads.cpp:29-44): 1 and out[4] = origin_x,
// origin_y, width, height, or 0 where it returns nullopt.
int ref_tile_geometry(uint32_t index, uint32_t w, uint32_t h, uint32_t out[4]) {
    const auto g = Distributed::resolve_tile_geometry(index, w, h);
    if (!g) return 0;
    out[0] = g->origin_x;
    out[1] = g->origin_y;
    out[2] = g->width;
    out[3] = g->height;
    return 1;
}

// Sizes and the byte offsets (in the serialized task) of the fields a tile
// worker reads: [0] task bytes, [1] header bytes, [2] push-constant offset,
// [3] push-constant size, then observer_position, tetrad_e1, tetrad_e2,
// tetrad_e3, field_of_view_rad, metric_mass, metric_spin, escape_radius,
// screen_width, screen_height, metric_type, max_integration_steps,
// body_count, primary_disk.enabled, .inner_radius_scale,
// .outer_radius_mass_units, projection_mode, time, step_size_factor,
// min_step_size, max_step_size, far_field_step_scale, render_flags,
// lod_distance_threshold, lod_reduced_steps, observer_boost (offsets from the
// start of the push constants), then sizeof(GpuDiskProfile) and the
// RenderFlags::USE_LOD_SYSTEM bit.
int ref_tile_layout(uint64_t* out) {
    using Render::GpuCameraPushConstants;
    using Render::GpuDiskProfile;
    const uint64_t disk = offsetof(GpuCameraPushConstants, primary_disk);
    const uint64_t v[] = {Distributed::kRenderTileTaskPayloadBytes, sizeof(Distributed::RenderTileTaskHeader),
                          Distributed::kRenderTileTaskPrefixBytes, sizeof(GpuCameraPushConstants),
                          offsetof(GpuCameraPushConstants, observer_position), offsetof(GpuCameraPushConstants, tetrad_e1),
                          offsetof(GpuCameraPushConstants, tetrad_e2), offsetof(GpuCameraPushConstants, tetrad_e3),
                          offsetof(GpuCameraPushConstants, field_of_view_rad), offsetof(GpuCameraPushConstants, metric_mass),
                          offsetof(GpuCameraPushConstants, metric_spin), offsetof(GpuCameraPushConstants, escape_radius),
                          offsetof(GpuCameraPushConstants, screen_width), offsetof(GpuCameraPushConstants, screen_height),
                          offsetof(GpuCameraPushConstants, metric_type),
                          offsetof(GpuCameraPushConstants, max_integration_steps),
                          offsetof(GpuCameraPushConstants, body_count), disk + offsetof(GpuDiskProfile, enabled),
                          disk + offsetof(GpuDiskProfile, inner_radius_scale),
                          disk + offsetof(GpuDiskProfile, outer_radius_mass_units),
                          offsetof(GpuCameraPushConstants, projection_mode), offsetof(GpuCameraPushConstants, time),
                          offsetof(GpuCameraPushConstants, step_size_factor),
                          offsetof(GpuCameraPushConstants, min_step_size),
                          offsetof(GpuCameraPushConstants, max_step_size),
                          offsetof(GpuCameraPushConstants, far_field_step_scale),
                          offsetof(GpuCameraPushConstants, render_flags),
                          offsetof(GpuCameraPushConstants, lod_distance_threshold),
                          offsetof(GpuCameraPushConstants, lod_reduced_steps),
                          offsetof(GpuCameraPushConstants, observer_boost), sizeof(GpuDiskProfile),
                          static_cast<uint64_t>(Render::RenderFlags::USE_LOD_SYSTEM)};
    const int n = static_cast<int>(sizeof(v) / sizeof(v[0]));
    for (int i = 0; i < n; ++i) out[i] = v[i];
    return n;
}

// A task with default push constants except the given fields, header from
// make_render_tile_task_header, written by serialize_render_tile_task
// (task_payloads.cpp:46-112). d: observer[4], e1[4], e2[4], e3[4], fov,
// mass, spin, escape radius, timestamp; u: tile index, width, height,
// lod_reduced_steps, flags, compression mode, max steps, metric type, body
// count; f: disk enabled, inner radius scale, outer radius (M); d2:
// min_step_size, max_step_size, step_size_factor, far_field_step_scale,
// time, lod_distance_threshold, observer_boost[0]; u2: projection_mode,
// render_flags, the push constants' lod_reduced_steps. Returns the bytes
// written (0 where the reference refuses the task).
size_t ref_make_tile_task(const double* d, const uint32_t* u, const float* f, const double* d2, const uint32_t* u2,
                          uint8_t* out, size_t cap) {
    Distributed::RenderTileTaskPayload task{};
    auto& c = task.camera_constants;
    for (int i = 0; i < 4; ++i) {
        c.observer_position[i] = d[i];
        c.tetrad_e1[i] = d[4 + i];
        c.tetrad_e2[i] = d[8 + i];
        c.tetrad_e3[i] = d[12 + i];
    }
    c.field_of_view_rad = d[16];
    c.metric_mass = d[17];
    c.metric_spin = d[18];
    c.escape_radius = d[19];
    task.dispatch_timestamp = d[20];
    c.screen_width = u[1];
    c.screen_height = u[2];
    c.max_integration_steps = u[6];
    c.metric_type = u[7];
    c.body_count = u[8];
    c.primary_disk.enabled = f[0];
    c.primary_disk.inner_radius_scale = f[1];
    c.primary_disk.outer_radius_mass_units = f[2];
    c.min_step_size = d2[0];
    c.max_step_size = d2[1];
    c.step_size_factor = d2[2];
    c.far_field_step_scale = d2[3];
    c.time = d2[4];
    c.lod_distance_threshold = d2[5];
    c.observer_boost[0] = d2[6];
    c.projection_mode = u2[0];
    c.render_flags = u2[1];
    c.lod_reduced_steps = u2[2];
    const auto h = Distributed::make_render_tile_task_header(u[0], u[1], u[2], static_cast<uint16_t>(u[3]),
                                                             static_cast<uint8_t>(u[4]),
                                                             static_cast<Distributed::TileCompressionMode>(u[5]));
    if (!h) return 0;
    task.header = *h;
    return Distributed::serialize_render_tile_task(task, std::span<uint8_t>(out, cap));
}

// CameraProjector<double>::compute_ray_direction (observer/camera_projections.hpp:198-225)
// with the arguments dispatch_fp64_scalar passes (no boost).
void ref_screen_direction(uint32_t mode, double u, double v, double fov, double out[3]) {
    const auto d = Observer::CameraProjector<double>::compute_ray_direction(static_cast<Observer::ProjectionMode>(mode),
                                                                           u, v, fov);
    out[0] = d[0];
    out[1] = d[1];
    out[2] = d[2];
}

// SoftwareComputeEngine::dispatch_fp64 (render/software_compute_engine.hpp:5546-5587)
// on a Kerr frame given in the C ABI's terms: the push constants take the
// camera (observer, tetrad columns, field of view, size, projection, row
// window), the metric (metric_type 2, mass, spin), escape radius, step
// formula, max_integration_steps, time and the disk profile (enabled from
// disk_enabled; inner_radius_scale = disk_inner / ISCO, 1 when disk_inner is
// 0; outer_radius_mass_units = disk_outer / M, as floats). Everything else
// keeps its default. Writes the GpuPixelOutput fields of each traced pixel.
void ref_render_frame(const rt_camera* cam, const rt_metric* mt, const rt_integrator* ig, const rt_termination* tm,
                      uint32_t* status, uint32_t* iters, float* redshift, float* affine, float* rgb) {
    Render::GpuCameraPushConstants pc{};
    for (int i = 0; i < 4; ++i) pc.observer_position[i] = cam->observer[i];
    pc.tetrad_e1 = {0.0, cam->forward[0], cam->forward[1], cam->forward[2]};
    pc.tetrad_e2 = {0.0, cam->right[0], cam->right[1], cam->right[2]};
    pc.tetrad_e3 = {0.0, cam->up[0], cam->up[1], cam->up[2]};
    pc.field_of_view_rad = cam->fov_y;
    pc.screen_width = cam->width;
    pc.screen_height = cam->height;
    pc.projection_mode = cam->projection;
    pc.dispatch_row_offset = cam->row_offset;
    pc.dispatch_row_count = cam->row_count;
    pc.metric_type = 2U;
    pc.metric_mass = mt->mass;
    pc.metric_spin = mt->spin;
    pc.escape_radius = tm->escape_radius;
    pc.max_integration_steps = ig->max_steps;
    pc.step_size_factor = ig->step_size_factor;
    pc.min_step_size = ig->min_step;
    pc.max_step_size = ig->max_step;
    pc.far_field_step_scale = ig->far_field_step_scale;
    pc.time = tm->time;
    static_assert(sizeof(Render::GpuDiskProfile) == sizeof(rt_disk_profile), "disk profile layout");
    std::memcpy(&pc.primary_disk, &tm->disk_profile, sizeof(rt_disk_profile));
    const double m = std::max(mt->mass, 1e-4);
    const double a = std::clamp(mt->spin, -0.999 * m, 0.999 * m);
    const double isco = Optics::DiskThermalProfile::kerr_isco_radius(m, a);
    pc.primary_disk.enabled = tm->disk_enabled ? 1.0f : 0.0f;
    pc.primary_disk.inner_radius_scale = tm->disk_inner > 0.0 ? static_cast<float>(tm->disk_inner / isco) : 1.0f;
    pc.primary_disk.outer_radius_mass_units = static_cast<float>((tm->disk_outer > 0.0 ? tm->disk_outer : 20.0 * m) / m);
    const uint32_t first = cam->row_count ? std::min(cam->row_offset, cam->height) : 0u;
    const uint32_t last =
        cam->row_count ? (uint32_t)std::min<uint64_t>((uint64_t)first + cam->row_count, cam->height) : cam->height;
    const size_t n = static_cast<size_t>(last - first) * cam->width;
    std::vector<Render::GpuPixelOutput> px(n);
    Render::SoftwareComputeEngine::dispatch_fp64(pc, std::span<Render::GpuPixelOutput>(px.data(), n));
    for (size_t i = 0; i < n; ++i) {
        status[i] = px[i].status_flags;
        iters[i] = px[i].iterations_used;
        redshift[i] = px[i].redshift;
        affine[i] = px[i].affine_parameter;
        if (rgb) {
            rgb[3 * i] = px[i].r;
            rgb[3 * i + 1] = px[i].g;
            rgb[3 * i + 2] = px[i].b;
        }
    }
}

// offsetof(GpuPixelOutput, field) for r, g, b, a, redshift, affine_parameter,
// status_flags, iterations_used, then sizeof (render/gpu_types.hpp:1111-1121).
int ref_pixel_layout(uint64_t* out) {
    using Render::GpuPixelOutput;
    const uint64_t v[] = {offsetof(GpuPixelOutput, r), offsetof(GpuPixelOutput, g), offsetof(GpuPixelOutput, b),
                          offsetof(GpuPixelOutput, a), offsetof(GpuPixelOutput, redshift),
                          offsetof(GpuPixelOutput, affine_parameter), offsetof(GpuPixelOutput, status_flags),
                          offsetof(GpuPixelOutput, iterations_used), sizeof(GpuPixelOutput)};
    for (int i = 0; i < 9; ++i) out[i] = v[i];
    return 9;
}

// The reference's tile path on a task's wire bytes: deserialize_render_tile_task
// (task_payloads.cpp:114-129), the push-constant adjustments TileRenderer::render
// makes before tracing (src/distributed/tile_renderer.cpp:150-161: interlace
// and row window off, a non-zero header lod_reduced_steps in, the tile
// origin and extent), then SoftwareComputeEngine::dispatch_tile
// (render/software_compute_engine.hpp:905-1022) without bodies. Writes the
// tile's GpuPixelOutput fields, tile-local row-major; returns the pixel
// count (0 when the reference refuses the task).
size_t ref_render_tile_task(const uint8_t* bytes, size_t n, uint32_t* status, uint32_t* iters, float* redshift,
                            float* affine) {
    Distributed::RenderTileTaskPayload task{};
    if (!Distributed::deserialize_render_tile_task(std::span<const uint8_t>(bytes, n), task)) return 0;
    Render::GpuCameraPushConstants c = task.camera_constants;
    const auto g = Distributed::resolve_tile_geometry(task.header.tile_index, c.screen_width, c.screen_height);
    if (!g) return 0;
    c.interlace_mode = 0U;
    c.interlace_phase = 0U;
    c.dispatch_row_offset = 0U;
    c.dispatch_row_count = 0U;
    if (task.header.lod_reduced_steps != 0U) c.lod_reduced_steps = task.header.lod_reduced_steps;
    c.tile_origin_x = g->origin_x;
    c.tile_origin_y = g->origin_y;
    c.tile_extent_x = g->width;
    c.tile_extent_y = g->height;
    const size_t count = static_cast<size_t>(g->width) * g->height;
    std::vector<Render::GpuPixelOutput> px(count);
    const Render::SoftwareComputeEngine::TileRegion region{g->origin_x, g->origin_y, g->width, g->height};
    if (!Render::SoftwareComputeEngine::dispatch_tile(c, region, std::span<Render::GpuPixelOutput>(px.data(), count),
                                                      {}, nullptr))
        return 0;
    for (size_t i = 0; i < count; ++i) {
        status[i] = px[i].status_flags;
        iters[i] = px[i].iterations_used;
        redshift[i] = px[i].redshift;
        affine[i] = px[i].affine_parameter;
    }
    return count;
}

// BardeenKerrShadow (metrics/bardeen_shadow.hpp:38-162): n boundary points
// (compute_boundary_points) and the enclosed area (compute_shadow_area).
double ref_shadow_boundary(double m, double a, double theta, size_t n, double* alpha, double* beta) {
    const Metrics::BardeenKerrShadow shadow(m, a, theta);
    const auto pts = shadow.compute_boundary_points(n);
    for (size_t i = 0; i < pts.size() && i < n; ++i) {
        alpha[i] = pts[i].alpha;
        beta[i] = pts[i].beta;
    }
    return shadow.compute_shadow_area(n);
}

// deserialize_render_tile_task (task_payloads.cpp:114-129): 1 when it
// accepts the bytes, 0 when it refuses them.
int ref_accepts_tile_task(const uint8_t* bytes, size_t n) {
    Distributed::RenderTileTaskPayload task{};
    return Distributed::deserialize_render_tile_task(std::span<const uint8_t>(bytes, n), task) ? 1 : 0;
}

}  // extern "C"
