// rt_tile.cpp — the cluster's 32x32 tile tasks: tile geometry and the
// decoding of a RenderTileTaskPayload's wire bytes into a trace request.
//
// The wire format is the reference's (include/relativistic/distributed/task_payloads.hpp,
// src/distributed/task_payloads.cpp:114-129): the packed 16-byte
// RenderTileTaskHeader, an f64 dispatch timestamp, then the raw
// GpuCameraPushConstants (render/gpu_types.hpp:966-1109, 11056 bytes,
// little-endian). Only the fields the tracer reads are located, by their
// byte offsets in that struct; the offsets are part of the format and are
// checked against the reference's own struct by the tests.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "rt_host.h"

static_assert(__BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__, "the tile wire format is little-endian");

namespace rt {
namespace {

// RenderTileTaskHeader (distributed/packet_protocol.hpp:158-167, packed).
constexpr size_t kHdrTileIndex = 0, kHdrTileX = 4, kHdrTileY = 6, kHdrTileW = 8, kHdrTileH = 10, kHdrLod = 12,
                 kHdrFlags = 14, kHdrCompression = 15;
constexpr size_t kHeaderBytes = 16;
constexpr size_t kPushOffset = kHeaderBytes + 8;  // after the f64 dispatch timestamp
constexpr size_t kPushBytes = 11056;
static_assert(kPushOffset + kPushBytes == RT_TILE_TASK_BYTES, "task size");

// GpuCameraPushConstants field offsets.
constexpr size_t kObserver = 0, kTetradE1 = 64, kTetradE2 = 96, kTetradE3 = 128, kFov = 160, kMass = 168,
                 kSpin = 176, kMinStep = 216, kMaxStep = 224, kEscape = 240, kLodThreshold = 312, kFarScale = 520,
                 kTime = 528, kStepFactor = 576, kScreenW = 584, kScreenH = 588, kMetricType = 592, kMaxSteps = 604,
                 kRenderFlags = 608, kProjection = 612, kLodSteps = 620, kBodyCount = 660, kPrimaryDisk = 780,
                 kObserverBoost = 9192;
constexpr uint32_t kUseLodSystem = 1u << 9;  // RenderFlags::USE_LOD_SYSTEM (render/gpu_types.hpp:1145)
// GpuDiskProfile (render/gpu_types.hpp:42-77): enabled, inner_radius_scale,
// outer_radius_mass_units are its first three floats.
constexpr size_t kDiskEnabled = 0, kDiskInnerScale = 4, kDiskOuterUnits = 8;
constexpr uint8_t kMaxCompressionMode = 2;  // TileCompressionMode::Lz4Block

template <class T>
T load(const unsigned char* base, size_t off) {
    T v;
    std::memcpy(&v, base + off, sizeof(T));
    return v;
}

bool finite4(const unsigned char* push, size_t off) {
    for (int i = 0; i < 4; ++i)
        if (!std::isfinite(load<double>(push, off + 8 * i))) return false;
    return true;
}

// resolve_tile_geometry (src/distributed/task_payloads.cpp:29-44).
bool tile_geometry(uint32_t index, uint32_t w, uint32_t h, rt_tile& g) {
    if (w == 0 || h == 0 || w > RT_TILE_MAX_SCREEN_WIDTH || h > RT_TILE_MAX_SCREEN_HEIGHT) return false;
    const uint32_t cols = (w + RT_TILE_EDGE - 1) / RT_TILE_EDGE;
    const uint32_t rows = (h + RT_TILE_EDGE - 1) / RT_TILE_EDGE;
    if (static_cast<uint64_t>(index) >= static_cast<uint64_t>(cols) * rows) return false;
    g.origin_x = (index % cols) * RT_TILE_EDGE;
    g.origin_y = (index / cols) * RT_TILE_EDGE;
    g.width = std::min(RT_TILE_EDGE, w - g.origin_x);
    g.height = std::min(RT_TILE_EDGE, h - g.origin_y);
    return true;
}

}  // namespace
}  // namespace rt

extern "C" {

rt_status rt_tile_geometry(uint32_t tile_index, uint32_t screen_width, uint32_t screen_height, rt_tile* out) {
    if (!out) return rt::fail(RT_ERR_INVALID_ARG, "null tile");
    rt_tile g{};
    if (!rt::tile_geometry(tile_index, screen_width, screen_height, g))
        return rt::fail(RT_ERR_INVALID_ARG, "tile index outside the frame, or screen size 0 or above 7680x4320");
    *out = g;
    return RT_OK;
}

rt_status rt_decode_tile_task(const void* bytes, size_t n_bytes, rt_tile_task* out) {
    using namespace rt;
    if (!bytes || !out) return fail(RT_ERR_INVALID_ARG, "null argument");
    if (n_bytes < RT_TILE_TASK_BYTES) return fail(RT_ERR_INVALID_ARG, "tile task shorter than 11080 bytes");
    const unsigned char* b = static_cast<const unsigned char*>(bytes);
    const unsigned char* push = b + kPushOffset;
    rt_tile_task t;
    std::memset(&t, 0, sizeof(t));
    t.tile_index = load<uint32_t>(b, kHdrTileIndex);
    const uint16_t tile_x = load<uint16_t>(b, kHdrTileX), tile_y = load<uint16_t>(b, kHdrTileY);
    const uint16_t tile_w = load<uint16_t>(b, kHdrTileW), tile_h = load<uint16_t>(b, kHdrTileH);
    t.lod_reduced_steps = load<uint16_t>(b, kHdrLod);
    t.flags = load<uint8_t>(b, kHdrFlags);
    t.compression_type = load<uint8_t>(b, kHdrCompression);
    t.dispatch_timestamp = load<double>(b, kHeaderBytes);

    // validate_render_tile_task (src/distributed/task_payloads.cpp:70-101), same order of checks
    const uint32_t W = load<uint32_t>(push, kScreenW), H = load<uint32_t>(push, kScreenH);
    if (W == 0 || H == 0) return fail(RT_ERR_INVALID_ARG, "tile task: screen size 0");
    if (W > RT_TILE_MAX_SCREEN_WIDTH || H > RT_TILE_MAX_SCREEN_HEIGHT)
        return fail(RT_ERR_INVALID_ARG, "tile task: screen larger than 7680x4320");
    const uint32_t max_steps = load<uint32_t>(push, kMaxSteps);
    if (max_steps == 0) return fail(RT_ERR_INVALID_ARG, "tile task: max_integration_steps 0");
    if (!finite4(push, kObserver) || !finite4(push, kTetradE1) || !finite4(push, kTetradE2) ||
        !finite4(push, kTetradE3))
        return fail(RT_ERR_INVALID_ARG, "tile task: non-finite observer position or tetrad");
    const double fov = load<double>(push, kFov), mass = load<double>(push, kMass), spin = load<double>(push, kSpin),
                 escape = load<double>(push, kEscape);
    if (!std::isfinite(fov) || !std::isfinite(mass) || !std::isfinite(spin) || !std::isfinite(escape))
        return fail(RT_ERR_INVALID_ARG, "tile task: non-finite field of view, mass, spin or escape radius");
    if (!std::isfinite(t.dispatch_timestamp)) return fail(RT_ERR_INVALID_ARG, "tile task: non-finite timestamp");
    if (t.compression_type > kMaxCompressionMode)
        return fail(RT_ERR_INVALID_ARG, "tile task: unknown compression mode");
    if (!tile_geometry(t.tile_index, W, H, t.tile)) return fail(RT_ERR_INVALID_ARG, "tile task: tile index outside the frame");
    if (tile_x != t.tile.origin_x / RT_TILE_EDGE || tile_y != t.tile.origin_y / RT_TILE_EDGE || tile_w != t.tile.width ||
        tile_h != t.tile.height)
        return fail(RT_ERR_INVALID_ARG, "tile task: header does not match the tile's geometry");
    t.metric_type = load<uint32_t>(push, kMetricType);
    if (t.metric_type < 1 || t.metric_type > 3)
        return fail(RT_ERR_INVALID_ARG, "tile task: metric_type " + std::to_string(t.metric_type) +
                                            " is not Schwarzschild or Kerr (1, 2, 3)");
    t.body_count = load<uint32_t>(push, kBodyCount);
    // Values the reference accepts but this tracer's contract does not
    // define: refused here, so that every decoded task can be traced.
    if (!(escape > 0.0)) return fail(RT_ERR_INVALID_ARG, "tile task: escape radius not positive");
    double boost[3];
    for (int i = 0; i < 3; ++i) boost[i] = load<double>(push, kObserverBoost + 8 * i);
    if (boost[0] * boost[0] + boost[1] * boost[1] + boost[2] * boost[2] > 1.0e-18)
        return fail(RT_ERR_INVALID_ARG, "tile task: observer boost (aberration) is not traced");

    // push constants -> request, as dispatch_fp64_scalar reads them
    // (render/software_compute_engine.hpp:3794-3813)
    rt_camera& c = t.camera;
    for (int i = 0; i < 4; ++i) c.observer[i] = load<double>(push, kObserver + 8 * i);
    for (int i = 0; i < 3; ++i) {
        c.forward[i] = load<double>(push, kTetradE1 + 8 * (i + 1));
        c.right[i] = load<double>(push, kTetradE2 + 8 * (i + 1));
        c.up[i] = load<double>(push, kTetradE3 + 8 * (i + 1));
    }
    c.fov_y = fov;
    c.width = W;
    c.height = H;
    c.row_offset = 0;
    c.row_count = 0;
    // an unknown projection mode is the pinhole, as compute_ray_direction's default case
    const uint32_t proj = load<uint32_t>(push, kProjection);
    c.projection = (proj <= RT_PROJ_HAMMER_AITOFF) ? proj : RT_PROJ_PINHOLE;
    const double m = std::max(mass, 1e-4);
    const double a = (t.metric_type == 1) ? 0.0 : std::clamp(spin, -0.999 * m, 0.999 * m);
    // step budget under the LOD rule (dispatch_fp64_scalar, software_compute_engine.hpp:3821-3825;
    // the worker puts a non-zero header lod_reduced_steps into the push
    // constants, src/distributed/tile_renderer.cpp:155-157), r_+ from the clamped spin
    uint32_t steps = max_steps;
    {
        const double a_pc = std::clamp(spin, -0.999 * m, 0.999 * m);
        const double rh_pc = (std::abs(a_pc) > 1e-12) ? m + std::sqrt(std::max(m * m - a_pc * a_pc, 0.0)) : 2.0 * m;
        const double r_frame = std::max(c.observer[1], rh_pc * 1.02);
        const uint32_t lod_steps = t.lod_reduced_steps != 0 ? t.lod_reduced_steps : load<uint32_t>(push, kLodSteps);
        const double lod_threshold = load<double>(push, kLodThreshold);
        if ((load<uint32_t>(push, kRenderFlags) & kUseLodSystem) && lod_threshold > 0.0 && r_frame > lod_threshold)
            steps = std::min(steps, lod_steps);
    }
    // observer radius clamped to 1.02 r_+ (dispatch_fp64_scalar, software_compute_engine.hpp:3916)
    c.observer[1] = std::max(c.observer[1], 1.02 * (m + std::sqrt(std::max(m * m - a * a, 0.0))));
    std::string why;
    if (!valid_camera(&c, why)) return fail(RT_ERR_INVALID_ARG, "tile task: " + why);
    rt_kerr_metric(&t.metric, m, a);
    rt_default_integrator(&t.integrator);
    t.integrator.max_steps = std::max(steps, 1u);
    t.integrator.min_step = load<double>(push, kMinStep);
    t.integrator.max_step = load<double>(push, kMaxStep);
    t.integrator.step_size_factor = load<double>(push, kStepFactor);
    t.integrator.far_field_step_scale = load<double>(push, kFarScale);
    rt_default_termination(&t.termination);
    t.termination.escape_radius = escape;
    t.termination.time = load<double>(push, kTime);
    static_assert(sizeof(rt_disk_profile) == 128, "rt_disk_profile is GpuDiskProfile's 128 bytes");
    std::memcpy(&t.termination.disk_profile, push + kPrimaryDisk, sizeof(rt_disk_profile));
    const unsigned char* disk = push + kPrimaryDisk;
    const float enabled = load<float>(disk, kDiskEnabled);
    const double inner_scale = std::max(load<float>(disk, kDiskInnerScale), 1.0f);
    const double outer_units = load<float>(disk, kDiskOuterUnits);
    const double inner = kerr_isco_host(m, a) * inner_scale;
    // a NaN radius makes the reference's range test fail for every crossing:
    // the disk is never hit, which is the disk switched off
    const bool radii_defined = !std::isnan(inner_scale) && !std::isnan(outer_units);
    t.termination.disk_enabled = (enabled > 0.5f && radii_defined) ? 1u : 0u;
    t.termination.disk_inner = (inner_scale > 1.0) ? inner : 0.0;  // 0: the ISCO itself
    t.termination.disk_outer = std::max(outer_units * m, inner * 1.05);
    *out = t;
    return RT_OK;
}

}  // extern "C"
