/*
 * rt_tracer.h — C ABI of the MI355X null-geodesic tracer (librtsim.so).
 *
 * One backward null geodesic per ray in Kerr / Kerr–Schild spacetime,
 * integrated with adaptive Dormand–Prince 5(4) on the Kerr–Schild
 * Hamiltonian, on one AMD Instinct MI355X (gfx950).
 *
 * Plain C: every type is a POD, every buffer is a pointer plus a count,
 * no entry point throws, every entry point returns rt_status.
 *
 * Which reference interface each entry point replaces (paths relative to
 * the Relativity-Simulator tree):
 *   rt_trace_camera      SoftwareComputeEngine::dispatch_fp64
 *                        (include/relativistic/render/software_compute_engine.hpp:5546-5587)
 *                        and VulkanComputeExecutor::dispatch_and_readback
 *                        (include/relativistic/render/vulkan_compute_executor.hpp:1399)
 *   rt_trace_rays        per-ray tracing from an explicit (x, p):
 *                        GeodesicRayProbe::trace
 *                        (include/relativistic/optics/geodesic_ray_probe.hpp:452)
 *   rt_trace_batch       many frames in one dispatch: the tile/band loop of
 *                        TileRenderer::render (src/distributed/tile_renderer.cpp:133-182)
 *                        and VulkanComputeExecutor banding (vulkan_compute_executor.hpp:1394-1465)
 *   rt_decode_tile_task  deserialize_render_tile_task + validate_render_tile_task
 *                        (src/distributed/task_payloads.cpp:70-129)
 *   rt_trace_tile        SoftwareComputeEngine::dispatch_tile
 *                        (include/relativistic/render/software_compute_engine.hpp:905-1022)
 *   rt_trace_tiles       many tile tasks in one dispatch: the task loop of
 *                        TileRenderer::render (src/distributed/tile_renderer.cpp:133-182)
 *   RT_METHOD_RENDER_RK4 the renderer's own tracer, trace_kerr_schild_photon
 *                        (include/relativistic/render/software_compute_engine.hpp:1624-1809)
 *                        as dispatch_fp64 runs it for a Kerr frame
 *   rt_server_*          the windowed worker of GeodesicComputePipeline
 *                        (include/relativistic/render/geodesic_compute_pipeline.hpp:529-728):
 *                        requests, cancellation, completion polling;
 *                        rt_server_submit_ex deadlines follow the scheduler's
 *                        in-flight expiry, DistributionScheduler::collect_expired
 *                        (src/distributed/distribution_scheduler.cpp:389-405)
 */
#ifndef RT_TRACER_H
#define RT_TRACER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RT_ABI_VERSION 3u
#define RT_MAX_SOURCES 16u

/* ------------------------------------------------------------------ status */
typedef enum rt_status {
    RT_OK = 0,
    RT_ERR_INVALID_ARG = 1, /* null pointer, zero size, out-of-range parameter */
    RT_ERR_OOM = 2,         /* device or host allocation failed */
    RT_ERR_DEVICE = 3,      /* HIP runtime error (message in rt_last_error) */
    RT_ERR_TIMEOUT = 4,     /* rt_server_wait deadline passed */
    RT_ERR_CANCELLED = 5,   /* request cancelled before completion */
    RT_ERR_NO_DEVICE = 6,   /* no gfx950 device visible */
    RT_ERR_BUSY = 7,        /* admission refused: footprint exceeds the budget */
    RT_ERR_NOT_FOUND = 8,   /* unknown request id */
    RT_ERR_EXPIRED = 9      /* request deadline passed before it finished (rt_server_submit_ex) */
} rt_status;

/* --------------------------------------------- per-ray class and flag bits
 * The class bits are the reference's PixelFlags values
 * (include/relativistic/render/gpu_types.hpp:1123-1130). */
#define RT_CLS_CAPTURED 1u       /* HORIZON_ABSORBED: r <= r_+ + horizon_eps     */
#define RT_CLS_ESCAPED 2u        /* CELESTIAL_HIT: r >= escape_radius, outgoing  */
#define RT_CLS_DISK 4u           /* ACCRETION_DISK_HIT: z = 0 within the disk    */
#define RT_CLS_MASK 7u
#define RT_FLAG_BUDGET (1u << 8)     /* max_steps attempts used up            */
#define RT_FLAG_MIN_STEP (1u << 9)   /* |h| fell below h_min                  */
#define RT_FLAG_NONFINITE (1u << 10) /* state became non-finite               */
#define RT_FLAG_PROJECTED (1u << 11) /* null-constraint projection fired once */
#define RT_FLAG_FP32_STEPS (1u << 12)/* mixed precision: some steps ran FP32  */
#define RT_FLAG_CARTER (1u << 13)    /* captured by the Carter test (RT_TERM_CARTER_CAPTURE),
                                        end state where the test fired         */
#define RT_FLAG_INVALID (1u << 15)   /* input ray rejected (not traced)       */

/* ------------------------------------------------------------ parameters */
typedef enum rt_precision {
    RT_PRECISION_FP64 = 0,  /* every operation in FP64                          */
    RT_PRECISION_MIXED = 1  /* FP32 right-hand side beyond mixed_switch_radius;
                               FP64 state, error test, projection, root finding */
} rt_precision;

typedef struct rt_source {
    /* One Kerr–Schild source of a superposition. Source 0 is the central hole
     * (reference: build_multi_source_data, software_compute_engine.hpp:2738-2798). */
    double position[3];   /* world Cartesian centre                               */
    double mass;          /* m > 0                                                */
    double spin;          /* a/m in [-0.9999, 0.9999] (dimensionless, like
                             GpuBodyData::spin_parameter)                          */
    double axis[3];       /* spin axis (normalised by the library)                */
    double disk_inner;    /* <= 0: the source's prograde ISCO                     */
    double disk_outer;    /* <= 0: 20 m;  < 0 on both: disk disabled            */
    uint32_t disk_enabled;
    uint32_t reserved;
} rt_source;

typedef struct rt_metric {
    double mass;          /* M of the central hole (clamped to >= 1e-4)           */
    double spin;          /* a in length units, clamped to [-0.999 M, 0.999 M]    */
    uint32_t n_sources;   /* 0 or 1: single Kerr (mass, spin); 2..16: superposition
                             described by src[0..n_sources-1], src[0] central     */
    uint32_t reserved;
    rt_source src[RT_MAX_SOURCES];
} rt_metric;

/* Integration method. */
typedef enum rt_method {
    RT_METHOD_DP45 = 0,       /* adaptive Dormand–Prince 5(4) (the default)       */
    /* The reference renderer's tracer (trace_kerr_schild_photon,
     * render/software_compute_engine.hpp:1624-1809, as dispatch_fp64 runs it
     * for a Kerr frame): classic RK4 with the render step
     *   |h| = clamp(step_size_factor sqrt(r max(r - r_+, 0.02 M)) F(r),
     *               min_step, max_step far_field_step_scale),
     *   F(r) = 1 + (far_field_step_scale - 1) clamp((r - 20 r_+) / (80 r_+), 0, 1)
     * on BL r; capture at r <= 1.0001 r_+ (or a non-finite state), escape at
     * r >= escape_radius; a translucent disk: every crossing of theta = pi/2
     * with r in [disk_inner, disk_outer] and r > 1.02 r_+ sets the disk bit,
     * records g and multiplies the throughput by 1 - opacity (the disk
     * profile's), and the ray ends once the throughput is <= 0.01; after the
     * loop, a ray neither captured nor escaped inside 3 r_s has throughput 0,
     * and the escaped bit is set when escaped or throughput > 0.01.
     * max_steps is the reference's max_integration_steps, halved with a floor
     * of 256 as dispatch_fp64_scalar does for the exact Kerr path. Results:
     * cls = the reference's status_flags (any combination of 1, 2, 4), steps
     * = iterations_used (loop entries), lambda = iterations_used x 0.05 (its
     * affine_parameter), g = its redshift (1 unless a disk crossing set it),
     * x and p the end state (p_t = -1), bl the end point with the azimuth
     * continued from the observer's. FP64 only; single Kerr only; cameras
     * only (not explicit rays). */
    RT_METHOD_RENDER_RK4 = 1
} rt_method;

typedef struct rt_integrator {
    double rtol;          /* DP45 relative tolerance (1e-10)                      */
    double atol;          /* DP45 absolute tolerance (1e-12)                      */
    double h0;            /* initial |h|; 0: the reference render-step heuristic  */
    double h_min;         /* |h| below this ends the ray with RT_FLAG_MIN_STEP    */
    double h_max;         /* cap on |h|                                           */
    double invariant_tol; /* null-constraint projection threshold (1e-12)         */
    double mixed_switch_radius; /* MIXED: FP64 RHS inside this radius; the library
                                   raises it to at least 3 r_+                    */
    uint32_t max_steps;   /* cap on attempted (accepted + rejected) steps         */
    uint32_t precision;   /* rt_precision                                         */
    uint32_t method;      /* rt_method                                            */
    uint32_t reserved;
    /* RT_METHOD_RENDER_RK4 step formula (GpuCameraPushConstants fields of the
     * same names, render/gpu_types.hpp:966-1109; defaults 0.5, 0.004, 3.5, 7.8) */
    double step_size_factor;
    double min_step;
    double max_step;
    double far_field_step_scale;
} rt_integrator;

#define RT_TERM_NO_G_CLAMP 1u /* do not clamp disk g to [0.05, 4]                */
/* Single Kerr: end a ray as captured, without following it to r_+ + eps, as
 * soon as it moves inward inside the disk's inner radius (anywhere when the
 * disk is off), when Carter's radial potential R(r) of its constants (taken
 * from the initial state) has no root outside r_+: the certain-absorption
 * verdict of CarterPhaseClassifier (optics/carter_ray_classifier.hpp:116-164),
 * with a margin that leaves rays near the critical curve to the integrator.
 * Such rays move inward from there on, cannot meet the disk again, and are
 * captured. Their end state is where the test fired (flag RT_FLAG_CARTER).
 * Off by default. */
#define RT_TERM_CARTER_CAPTURE 2u

/* The disk's appearance parameters: the float fields of the reference's
 * GpuDiskProfile (render/gpu_types.hpp:42-77) in its order, 128 bytes, so a
 * caller can copy that struct in. RT_METHOD_RENDER_RK4 reads the ones that
 * decide the disk's opacity (AccretionDiskModel::shade,
 * render/accretion_disk/accretion_disk_model.hpp:245-314): opacity_scale,
 * edge_softness, rotation_speed_scale, the ring, spiral, turbulence and
 * grain fields and noise_seed. The disk's on/off switch and radii are
 * rt_termination's disk_enabled, disk_inner and disk_outer. */
typedef struct rt_disk_profile {
    float enabled, inner_radius_scale, outer_radius_mass_units, peak_temperature_k;
    float floor_temperature_k, temperature_exponent, zero_torque_strength, extra_beaming_exponent;
    float brightness, color_saturation, tint_strength, opacity_scale;
    float tint_r, tint_g, tint_b, rotation_speed_scale;
    float ring_amplitude, ring_frequency, ring_sharpness, spiral_arm_count;
    float spiral_pitch, spiral_amplitude, turbulence_amplitude, turbulence_scale;
    float turbulence_octaves, radial_stretch, grain_amplitude, grain_scale;
    float reference_luminance, noise_seed, temperature_normalization, edge_softness;
} rt_disk_profile;

typedef struct rt_termination {
    double horizon_eps;   /* captured when r <= r_+ + horizon_eps (1e-6)           */
    double escape_radius; /* escaped when r >= escape_radius while outgoing (1000) */
    double disk_inner;    /* <= 0: prograde ISCO of (M, a)                        */
    double disk_outer;    /* equatorial disk outer radius (20 M)                  */
    uint32_t disk_enabled;
    uint32_t flags;       /* RT_TERM_* bits                                        */
    double time;          /* RT_METHOD_RENDER_RK4: the disk pattern's time (push
                             constant `time`; rotates the opacity pattern)       */
    rt_disk_profile disk_profile; /* RT_METHOD_RENDER_RK4: opacity parameters     */
} rt_termination;

/* Camera projections: the reference's ProjectionMode values
 * (observer/camera_projections.hpp:13-22, CameraProjector::compute_ray_direction
 * :198-225, as dispatch_fp64_scalar calls it with no boost). The all-sky
 * modes (3, 7) take the horizontal screen coordinate without the aspect
 * ratio (software_compute_engine.hpp:3855-3856). */
#define RT_PROJ_PINHOLE 0u
#define RT_PROJ_AUTO_ZOOM 1u          /* with no observer boost: the pinhole   */
#define RT_PROJ_FISHEYE_STEREOGRAPHIC 2u
#define RT_PROJ_EQUIRECTANGULAR_360 3u
#define RT_PROJ_FISHEYE_EQUIDISTANT 4u
#define RT_PROJ_FISHEYE_ORTHOGRAPHIC 5u
#define RT_PROJ_PANINI_CYLINDRICAL 6u
#define RT_PROJ_HAMMER_AITOFF 7u

typedef struct rt_camera {
    /* Camera, as GpuCameraPushConstants (render/gpu_types.hpp:966-1109):
     * observer_position (t, r, theta, phi) and the spatial parts of
     * tetrad_e1 (forward), tetrad_e2 (right), tetrad_e3 (up), world Cartesian. */
    double observer[4];
    double forward[3];
    double right[3];
    double up[3];
    double fov_y;         /* field of view, radians (the pinhole's vertical full
                             angle; each projection reads it as the reference does) */
    uint32_t width;
    uint32_t height;
    uint32_t row_offset;  /* first traced row (dispatch_row_offset); read only
                             when row_count is not 0                           */
    uint32_t row_count;   /* traced rows, clipped at the last row; 0 = the whole
                             frame (dispatch_row_count)                        */
    uint32_t projection;  /* RT_PROJ_* (projection_mode); 0 = pinhole           */
    uint32_t reserved;
} rt_camera;

typedef struct rt_results {
    /* Caller-owned host arrays, SoA, one entry per traced ray (row-major pixel
     * order for cameras). Any pointer may be NULL to skip that output. */
    uint32_t* cls;        /* class bits | flag bits                               */
    double* x;            /* 4*n: t[n] | x[n] | y[n] | z[n]  (Kerr–Schild)        */
    double* p;            /* 4*n: p_t[n] | p_x[n] | p_y[n] | p_z[n] (covariant)   */
    double* bl;           /* 3*n: r[n] | theta[n] | phi[n] (spheroidal, as
                             boyer_lindquist_from_kerr_schild; phi in (-pi, pi])  */
    double* lambda;       /* affine length sum |h| up to the end point            */
    double* g;            /* disk: redshift_at; escaped: 1/E; captured: 0         */
    uint32_t* steps;      /* accepted DP45 steps                                  */
    uint32_t* rejected;   /* rejected DP45 steps                                  */
} rt_results;

typedef struct rt_stats {
    double kernel_ms;          /* device time of the last trace (HIP events)      */
    double total_ms;           /* host wall time of the last call                 */
    uint64_t rays;             /* rays traced by the last call                    */
    uint64_t accepted_steps;   /* summed over rays                                */
    uint64_t rejected_steps;
    uint64_t rhs_evals;        /* right-hand-side evaluations actually executed   */
    uint64_t fp32_steps;       /* MIXED: steps that ran the FP32 right-hand side  */
    uint32_t block_size;       /* launch shape used                               */
    uint32_t grid_size;
} rt_stats;

typedef struct rt_request {
    /* One camera request for batched tracing / serving. */
    rt_camera camera;
    rt_metric metric;
    rt_integrator integrator;
    rt_termination termination;
} rt_request;

typedef struct rt_context rt_context;
typedef struct rt_server rt_server;

/* ------------------------------------------------------------- defaults */
/* Fill with the documented defaults (rtol 1e-10, atol 1e-12, ...). */
void rt_default_integrator(rt_integrator* out);
void rt_default_termination(rt_termination* out);
/* Observer at (r, theta, phi) looking at the hole: forward = -e_r,
 * right = e_phi, up = -e_theta. */
void rt_default_camera(rt_camera* out, double r_obs, double theta_obs, double phi_obs,
                       double fov_y, uint32_t width, uint32_t height);
/* Kerr metric with mass M and spin a (length units). */
void rt_kerr_metric(rt_metric* out, double mass, double spin);

/* ------------------------------------------------------------- lifetime */
uint32_t rt_abi_version(void);
const char* rt_version(void);
/* Message of the last failed call on this thread ("" if none). */
const char* rt_last_error(void);

rt_status rt_context_create(int device, rt_context** out);
rt_status rt_context_destroy(rt_context* ctx);
rt_status rt_get_stats(const rt_context* ctx, rt_stats* out);

/* ------------------------------------------------------------ one frame */
/* Trace every pixel of a camera; results are copied into host arrays. */
rt_status rt_trace_camera(rt_context* ctx, const rt_camera* cam, const rt_metric* metric,
                          const rt_integrator* integ, const rt_termination* term,
                          rt_results* out);

/* Trace n rays from explicit initial states (x^mu, p_mu) given as SoA host
 * arrays of 4*n doubles each (t|x|y|z and p_t|p_x|p_y|p_z). */
rt_status rt_trace_rays(rt_context* ctx, const double* x_soa, const double* p_soa, size_t n,
                        const rt_metric* metric, const rt_integrator* integ,
                        const rt_termination* term, rt_results* out);

/* Device-resident variant of rt_trace_camera: results stay in device
 * buffers owned by the context (valid until the next trace call on it);
 * returns after the kernels finished. *n_rays receives the ray count. */
rt_status rt_trace_camera_resident(rt_context* ctx, const rt_camera* cam, const rt_metric* metric,
                                   const rt_integrator* integ, const rt_termination* term,
                                   size_t* n_rays);
/* Copy the device-resident results of the last resident trace to the host. */
rt_status rt_fetch_results(rt_context* ctx, rt_results* out);

/* The reference's per-pixel output record (GpuPixelOutput,
 * render/gpu_types.hpp:1111-1121): same 32-byte layout and field meaning.
 * r, g, b, a are the shaded colour, which this library does not compute
 * (0, 0, 0, 1); the caller's shading fills them from the class and redshift. */
typedef struct rt_pixel {
    float r, g, b, a;
    float redshift;          /* g as a float                                   */
    float affine_parameter;  /* lambda as a float                              */
    uint32_t status_flags;   /* PixelFlags bits: 1 captured, 2 escaped, 4 disk */
    uint32_t iterations_used;/* accepted DP45 steps                            */
} rt_pixel;

/* Pack the device-resident results of the last resident trace into rt_pixel
 * records on the device and copy those (32 B per ray instead of 116) to the
 * caller's array of n_rays records. Replaces the readback loop of
 * VulkanComputeExecutor::dispatch_and_readback (render/vulkan_compute_executor.hpp:1467-1600),
 * which fills a std::vector<GpuPixelOutput>. */
rt_status rt_fetch_pixels(rt_context* ctx, rt_pixel* out);

/* ------------------------------------------------------ 32x32 tile tasks
 * The cluster's tile protocol (include/relativistic/distributed/task_payloads.hpp):
 * a frame is cut into 32x32 tiles numbered row-major, the last column and row
 * of tiles ragged; a worker receives a RenderTileTaskPayload and returns the
 * tile's GpuPixelOutput records, tile-local row-major. */
#define RT_TILE_EDGE 32u               /* kTileEdgePixels                      */
#define RT_TILE_MAX_SCREEN_WIDTH 7680u /* kMaximumScreenWidth                  */
#define RT_TILE_MAX_SCREEN_HEIGHT 4320u/* kMaximumScreenHeight                 */
/* Wire size of one task (kRenderTileTaskPayloadBytes): the 16-byte
 * RenderTileTaskHeader, the f64 dispatch timestamp, then the 11056-byte
 * GpuCameraPushConstants (render/gpu_types.hpp:966-1109), little-endian. */
#define RT_TILE_TASK_BYTES 11080u

/* A rectangle of a frame (TileGeometry, task_payloads.hpp:42-51). */
typedef struct rt_tile {
    uint32_t origin_x, origin_y, width, height;
} rt_tile;

/* Geometry of tile tile_index of a screen_width x screen_height frame.
 * Replaces resolve_tile_geometry (src/distributed/task_payloads.cpp:29-44):
 * RT_ERR_INVALID_ARG where it returns nullopt (zero or oversized screen,
 * index past the last tile). */
rt_status rt_tile_geometry(uint32_t tile_index, uint32_t screen_width, uint32_t screen_height, rt_tile* out);

/* A decoded tile task: the header fields, the resolved tile, and the push
 * constants mapped onto this library's request (single Kerr; see
 * rt_decode_tile_task). */
typedef struct rt_tile_task {
    uint32_t tile_index;          /* RenderTileTaskHeader fields                 */
    uint16_t lod_reduced_steps;
    uint8_t flags;                /* kTileTaskFlag* bits, passed through         */
    uint8_t compression_type;     /* result TileCompressionMode, passed through  */
    double dispatch_timestamp;
    rt_tile tile;                 /* resolve_tile_geometry of the task           */
    uint32_t metric_type;         /* 1 Schwarzschild, 2 or 3 Kerr                */
    uint32_t body_count;          /* GpuBodyData records sent beside the task    */
    rt_camera camera;             /* whole frame; the tile is the traced window  */
    rt_metric metric;
    rt_integrator integrator;
    rt_termination termination;
} rt_tile_task;

/* Decode and validate n_bytes of wire data into *out. Replaces
 * deserialize_render_tile_task + validate_render_tile_task
 * (src/distributed/task_payloads.cpp:70-129): RT_ERR_INVALID_ARG exactly where
 * they return false (short buffer, zero or oversized screen, zero step
 * budget, non-finite observer / tetrad / field of view / mass / spin /
 * escape radius / timestamp, unknown compression mode, header that does not
 * match the tile's geometry), and also for values the reference accepts but
 * this tracer does not define: a metric_type outside the Kerr family
 * (1, 2, 3), a field of view outside (0, pi), an escape radius that is not
 * positive. Every task it accepts can be traced. The push constants map as
 * dispatch_fp64_scalar reads them (render/software_compute_engine.hpp:3794-3813):
 * observer_position (radius raised to 1.02 r_+, as the reference clamps it)
 * and the spatial parts of tetrad_e1/e2/e3 -> camera;
 * metric_mass and metric_spin (0 for metric_type 1, clamped to +-0.999 M)
 * -> metric; escape_radius -> termination.escape_radius;
 * max_integration_steps -> integrator.max_steps (other integrator fields:
 * rt_default_integrator); disk on when primary_disk.enabled > 0.5 and
 * neither radius field is NaN (a NaN radius fails the reference's range test
 * for every crossing), inner radius ISCO x max(inner_radius_scale, 1), outer
 * radius max(outer_radius_mass_units x M, 1.05 x inner). */
rt_status rt_decode_tile_task(const void* bytes, size_t n_bytes, rt_tile_task* out);

/* Trace the rectangle *region of the camera's frame (rows and columns of the
 * full width x height frame; the camera's row window is ignored) and keep
 * the results device-resident, tile-local row-major, for rt_fetch_results /
 * rt_fetch_pixels; *n_rays receives width * height. Every pixel's result is
 * the one rt_trace_camera gives for it in the whole frame. Replaces
 * SoftwareComputeEngine::dispatch_tile (render/software_compute_engine.hpp:905-1022)
 * as called by TileRenderer::render (src/distributed/tile_renderer.cpp:133-182);
 * any rectangle inside the frame is accepted, 32x32 tiles being one case. */
rt_status rt_trace_tile(rt_context* ctx, const rt_camera* cam, const rt_metric* metric,
                        const rt_integrator* integ, const rt_termination* term, const rt_tile* region,
                        size_t* n_rays);

/* Trace n_tasks decoded tile tasks (any mix of frames, metrics and
 * tolerances) as one batch: a persistent kernel over a shared work queue of
 * 8x8 pixel blocks (queue_kind 0 at two waves per SIMD when the context's
 * shape asks for per-request kernels; a tile is too small to fill the device
 * alone),
 * captured in a hipGraph that is replayed while the batch layout stays the
 * same. Results stay device-resident, task after task, each tile-local
 * row-major, for rt_fetch_results / rt_fetch_pixels; *n_rays receives their
 * total. Each task's results equal rt_trace_tile's to 1e-12 relative, with
 * class, flags and step counts identical: the same formulas in another
 * kernel, which the compiler contracts into FMAs differently, so the last
 * bits can differ.
 * Replaces the per-task loop of a tile worker, TileRenderer::render called
 * once per RenderTileTaskPayload (src/distributed/tile_renderer.cpp:133-182),
 * whose GPU dispatches gpu_mutex_ serialises. */
rt_status rt_trace_tiles(rt_context* ctx, const rt_tile_task* tasks, size_t n_tasks, size_t* n_rays);

/* ------------------------------------------- allocation-free tile batch
 * No counterpart in the reference. For callers that cannot allocate device
 * memory: code under HIP graph capture, serving loops that must not call
 * hipMalloc, harnesses that forbid it. Traces n_tasks decoded tile tasks
 * (single Kerr) like rt_trace_tiles, on the same work-queue kernel at two
 * waves per SIMD, but allocates nothing: the per-task tables, counters and
 * queue head live in a fixed device array of the library (64 KiB per GPU,
 * held until the process ends), and the results go to the caller's device
 * arrays. dev_out holds DEVICE pointers, every one non-NULL and naturally
 * aligned, each array sized for `capacity` rays (x and p: 4 * capacity,
 * bl: 3 * capacity, laid out component after component with leading
 * dimension capacity). Results are written task after task, each tile-local
 * row-major; *n_rays receives their total. RT_ERR_INVALID_ARG for more than
 * RT_NOALLOC_MAX_TASKS tasks, for more rays than capacity, and for a
 * superposition task. Calls on one GPU share its array: the library
 * serialises them; work launched by other means that overlaps a call on
 * the same GPU and uses that array would race. Returns after the kernel
 * finished; the context's resident results are left as they were. */
#define RT_NOALLOC_MAX_TASKS 64u
rt_status rt_trace_tiles_noalloc(rt_context* ctx, const rt_tile_task* tasks, size_t n_tasks,
                                 const rt_results* dev_out, size_t capacity, size_t* n_rays);

/* ------------------------------------------------------- critical curve
 * The Kerr critical curve (the shadow's edge) in the impact-parameter plane
 * (alpha, beta) of an observer at inclination theta_obs, following
 * BardeenKerrShadow (include/relativistic/metrics/bardeen_shadow.hpp:38-162):
 * rays whose constants (xi = L_z / E, eta = Q / E^2) lie inside it are
 * captured, outside it escape. Host computations, no device needed. */

/* Each traced pixel's distance (units of M) from its ray's
 * (alpha, beta) = (-xi / sin theta_obs, sqrt(eta + a^2 cos^2 theta_obs -
 * xi^2 cot^2 theta_obs)) to the critical curve; out holds n_rays values in
 * the camera's pixel order. Rays closer than about 1e-6 M may change class
 * with the rounding of the integration. Single Kerr only. */
rt_status rt_critical_distance(const rt_camera* cam, const rt_metric* metric, double* out);
/* n >= 4 points of the curve, as BardeenKerrShadow::compute_boundary_points:
 * the upper branch over the visible photon-orbit radii, then the lower branch
 * back; the circle of radius sqrt(27) M when a = 0 or the observer is polar. */
rt_status rt_shadow_boundary(double mass, double spin, double theta_obs, size_t n, double* alpha, double* beta);
/* Distance (units of M) of the point (alpha, beta) to the curve; -1 on invalid input. */
double rt_impact_distance(double mass, double spin, double theta_obs, double alpha, double beta);

/* ------------------------------------------------------- launch shapes */
typedef struct rt_launch_shape {
    uint32_t block;       /* threads per workgroup: 64, 128, 192 or 256          */
    uint32_t waves;       /* occupancy the kernel is built for: 1 or 2 waves/SIMD */
    uint32_t refill_min;  /* lane-refill queue: idle lanes that trigger a refill (1..64) */
    uint32_t queue_kind;  /* batches: 0 = one persistent kernel, a wave takes
                             whole 8x8 tiles from a shared queue;
                             1 = one persistent kernel, idle lanes refilled ray
                             by ray; 2 = one per-frame kernel per request, the
                             requests spread over concurrent streams          */
} rt_launch_shape;

/* Shapes used by the single-Kerr per-frame kernels (plain) and work-queue
 * kernels (queue). The superposition kernel has a fixed shape. */
rt_status rt_get_launch_shape(const rt_context* ctx, rt_launch_shape* plain, rt_launch_shape* queue);
rt_status rt_set_launch_shape(rt_context* ctx, const rt_launch_shape* plain, const rt_launch_shape* queue);
/* Time the work-queue kernel over candidate shapes on the given sample
 * requests (single-Kerr requests only), keep the fastest in the context and
 * return it in *best. */
rt_status rt_autotune(rt_context* ctx, const rt_request* sample, size_t n_sample, rt_launch_shape* best);

/* ------------------------------------------------------- batched frames */
/* Trace n_req camera requests of different sizes and tolerances in one
 * persistent kernel fed by a device work queue (superposition requests run
 * as their own kernels in the same graph). The launch sequence is captured
 * into a hipGraph that is replayed while the batch layout stays the same.
 * outs[i] receives request i; outs may be NULL to keep the results resident
 * (then only the statistics are updated). */
rt_status rt_trace_batch(rt_context* ctx, const rt_request* reqs, size_t n_req, rt_results* outs);

/* ------------------------------------------------------------- serving */
/* A request loop on one device. Requests are admitted in arrival order as
 * device memory (hbm_budget_bytes), pinned staging (staging_mib), a stream
 * (four, one per hardware queue) and the device's wave slots allow; a request
 * that does not fit is passed by later ones (until it has waited 20 ms, when
 * memory is what it lacks). Each admitted request runs at once, next to those
 * already on the device, as a persistent kernel sized to the wave slots it was
 * granted: single-Kerr requests up to their need, superpositions at most three
 * quarters of the device and never the last quarter, so that requests
 * arriving meanwhile find room. The kernel writes the outputs asked for straight into pinned
 * staging. Every request must be taken (rt_server_take) or cancelled to free
 * its staging. */
typedef struct rt_server_config {
    uint64_t hbm_budget_bytes;  /* device memory for results and tables           */
    uint32_t max_inflight;      /* requests on the device at once                 */
    uint32_t staging_mib;       /* pinned host staging for finished results, MiB
                                   (0: 1024), at most hbm_budget_bytes            */
} rt_server_config;

/* Outputs of a request (rt_submit_options.outputs; rt_results fields). */
#define RT_OUT_CLS (1u << 0)
#define RT_OUT_X (1u << 1)
#define RT_OUT_P (1u << 2)
#define RT_OUT_BL (1u << 3)
#define RT_OUT_LAMBDA (1u << 4)
#define RT_OUT_G (1u << 5)
#define RT_OUT_STEPS (1u << 6)
#define RT_OUT_REJECTED (1u << 7)
#define RT_OUT_ALL 0xFFu
/* The reference's pixel record (rt_pixel) needs class, steps, g and lambda. */
#define RT_OUT_PIXEL (RT_OUT_CLS | RT_OUT_LAMBDA | RT_OUT_G | RT_OUT_STEPS)

typedef struct rt_submit_options {
    uint32_t outputs;     /* RT_OUT_* bits: only these are copied back and taken */
    uint32_t reserved;    /* 0 */
    int64_t deadline_ms;  /* > 0: the request ends RT_ERR_EXPIRED if it has not
                             finished this many ms after submit (a running kernel
                             completes and its results are dropped); 0: none   */
} rt_submit_options;

rt_status rt_server_create(int device, const rt_server_config* cfg, rt_server** out);
/* Cancels pending requests, waits for those on the device, frees the pools. */
rt_status rt_server_destroy(rt_server* srv);
/* Queue a request (copied). Refused with RT_ERR_BUSY if its footprint can
 * never fit the budgets; otherwise queued until memory frees up. Same as
 * rt_server_submit_ex with every output and no deadline. */
rt_status rt_server_submit(rt_server* srv, const rt_request* req, uint64_t* id);
rt_status rt_server_submit_ex(rt_server* srv, const rt_request* req, const rt_submit_options* opt, uint64_t* id);
/* 1 in *done when request id has finished (done, cancelled, expired or failed). */
rt_status rt_server_poll(rt_server* srv, uint64_t id, int* done);
/* Block until the request has finished or timeout_ms passed (< 0: no limit;
 * RT_ERR_TIMEOUT). Returns the request's status: RT_OK, RT_ERR_CANCELLED,
 * RT_ERR_EXPIRED or a device error. */
rt_status rt_server_wait(rt_server* srv, uint64_t id, int64_t timeout_ms);
/* Finished results into host arrays (NULL fields are skipped; asking for an
 * output not requested at submit is RT_ERR_INVALID_ARG and leaves the
 * request in place), then forget the request. A cancelled, expired or failed
 * request returns its status and is forgotten. RT_ERR_BUSY while unfinished. */
rt_status rt_server_take(rt_server* srv, uint64_t id, rt_results* out);
/* Pending: removed at once. Running: reported cancelled at once; the kernel
 * completes and its results are dropped. Done: results dropped. */
rt_status rt_server_cancel(rt_server* srv, uint64_t id);

#ifdef __cplusplus
}
#endif
#endif /* RT_TRACER_H */
