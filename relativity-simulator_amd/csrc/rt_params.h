// rt_params.h — parameter blocks shared by the C++ host and the HIP kernels.
// Built on the host (rt_host.cpp) from the C-ABI structs. A single-request
// launch passes TraceParams by value (scalar registers); the work-queue
// kernel reads one TraceParams per request from a device table.
#pragma once
#include <stdint.h>

namespace rt {

constexpr int kMaxSources = 16;

struct TraceParams {
    // ---- metric (single Kerr, or the central source of a superposition)
    double M, a, a2, rh;
    double r_hz;        // rh + horizon_eps: capture radius
    double r_esc;       // escape radius (KS r, or |x| for superpositions)
    double disk_in, disk_out, disk_spin, rh102;
    double near_r;      // 3 r_s: budget/step-underflow capture radius
    double carter_gate; // Carter capture test runs inside this radius (disk_in, or everywhere)
    uint32_t carter;    // 1: RT_TERM_CARTER_CAPTURE (single Kerr)
    uint32_t disk_on, clamp_g;
    uint32_t n_src;     // 0: single Kerr; >= 2: superposition (MultiDev)
    uint32_t mixed;     // 1: mixed precision policy
    // ---- integrator
    double rtol, atol, h0, h_min, h_max, inv_tol;
    double switch_r;    // mixed precision: FP32 RHS for steps starting at r > switch_r
    uint32_t max_steps;
    uint32_t precision; // rt_precision
    // ---- camera (per-frame constants; identical for every pixel)
    uint32_t width, height;    // frame size (the pixel -> screen mapping)
    // Traced window: rows row_first.., columns col_first..col_first+cols-1;
    // n_rays = rows * cols, results row-major over the window.
    uint32_t row_first, col_first, cols, n_rays;
    uint32_t tiles_x;   // 8x8 pixel tiles per row of tiles of the window
    uint32_t mode;      // 0 camera, 1 explicit rays
    // Tile order: the rectangle of tiles around the hole's image (the
    // captured and near-critical rays, which take up to ~100x the steps of
    // the rest) is traced first, then the other tiles row by row, so that
    // the slow rays overlap with the bulk of the frame. Tile units, relative
    // to the traced row window; hot_w * hot_h == 0: plain row-major order.
    uint32_t hot_x0, hot_y0, hot_w, hot_h;
    uint32_t projection;             // RT_PROJ_* mode of the pixel -> ray mapping
    uint32_t method;                 // rt_method
    double aspect, tan_half;
    double fov;                      // field of view as given
    double proj_c;                   // per-frame constant of the projection (see fill_camera)
    double fwd[3], right[3], up[3];
    double sth, cth, sph, cph;       // observer angles
    double X0[3];                    // observer position (KS, or flat for superpositions)
    double frame[3][4];              // Gram–Schmidt spatial frame e_1..e_3 at the observer
    double u0;                       // observer e_0 time component
    double kobs[3], two_h_obs;       // KS null vector and 2H at the observer (single Kerr)
    // ---- RT_METHOD_RENDER_RK4 (the reference renderer's tracer)
    double kappa, step_min, step_max, far_scale;  // render step formula
    double rs;                       // 2 M
    double phi_obs;                  // observer azimuth (start of the continued BL azimuth)
    double time;                     // disk pattern time
    double detail;                   // turbulence anti-aliasing factor of the frame
    uint32_t compat_steps;           // loop budget: max(max_integration_steps / 2, 256)
    uint32_t pad_compat;
    // opacity parameters of the disk profile (GpuDiskProfile fields)
    float opacity_scale, edge_softness, rotation_speed_scale, ring_amplitude;
    float ring_frequency, ring_sharpness, spiral_arm_count, spiral_pitch;
    float spiral_amplitude, turbulence_amplitude, turbulence_scale, turbulence_octaves;
    float radial_stretch, grain_amplitude, grain_scale, noise_seed;
};

// One Kerr–Schild source of a superposition, in its local frame.
struct SourceDev {
    double c[3];      // world centre
    double fx[3];     // local x axis (world components)
    double fy[3];     // local y axis
    double ax[3];     // spin axis = local z
    double m, a, a2, rh;
    double r_hz;      // rh + horizon_eps
    double disk_in, disk_out, rh102;
    uint32_t disk_on, pad;
};

// Superposition table of one request: staged into LDS by every workgroup.
struct MultiDev {
    uint32_t n;                      // number of sources (2..16)
    uint32_t observer_inside;        // observer inside a horizon: every ray captured at once
    double u0;                       // observer frame (superposed metric)
    double e[3][4];
    double pot[kMaxSources];         // H_i at the observer
    double dir[kMaxSources][3];      // world k_i at the observer
    SourceDev src[kMaxSources];
};

// Device output pointers (SoA) of one trace.
struct OutPtrs {
    uint32_t* cls;
    double* x;        // 4*n
    double* p;        // 4*n
    double* bl;       // 3*n
    double* lambda;
    double* g;
    uint32_t* steps;
    uint32_t* rejected;
    unsigned long long* counters; // [0] accepted [1] rejected [2] rhs evals [3] fp32 steps [4] projections
    const double* in_x;           // explicit rays: 4*n input positions
    const double* in_p;           // explicit rays: 4*n input momenta
    uint64_t n;                   // leading dimension of the SoA arrays
};

// Work queue of the persistent kernels: the hot tiles of every request
// first, then the remaining tiles of every request. Segment s < n_req holds
// request s's hot tiles, segment n_req + s its other tiles; req_offset[s] is
// the first queue index of segment s (2 n_req + 1 entries; queue indices
// count tiles, or rays for the lane-refill kernel).
struct QueueDev {
    const TraceParams* reqs;
    const OutPtrs* outs;
    const uint64_t* req_offset;   // 2 n_req + 1 entries
    unsigned long long* next;     // queue head (atomic)
    unsigned long long* counters; // summed statistics, as OutPtrs::counters
    uint64_t total;               // queue length (req_offset[2 n_req])
    uint32_t n_req;
    uint32_t refill_min;          // refill a wave when at least this many lanes are idle
};

}  // namespace rt
