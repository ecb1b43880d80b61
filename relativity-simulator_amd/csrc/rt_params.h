// rt_params.h — per-launch parameter block shared by the C++ host and the
// HIP kernels. Built on the host (rt_setup.cpp) from the C-ABI structs, passed
// by value as a kernel argument so every field lands in scalar registers.
#pragma once
#include <stdint.h>

namespace rt {

constexpr int kMaxSources = 16;

// One Kerr–Schild source of a superposition in its local frame.
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

struct TraceParams {
    // ---- metric (single Kerr)
    double M, a, a2, rh;
    double r_hz;        // rh + horizon_eps: capture radius
    double r_esc;       // escape radius (KS r, or |x| for superpositions)
    double disk_in, disk_out, disk_spin, rh102;
    double near_r;      // 3 r_s: budget/step-underflow capture radius
    uint32_t disk_on, clamp_g;
    uint32_t n_src;     // 0: single Kerr; >= 2: superposition (sources in srcs)
    uint32_t pad0;
    // ---- integrator
    double rtol, atol, h0, h_min, h_max, inv_tol;
    double switch_r;    // mixed precision: FP32 RHS for steps starting at r > switch_r
    uint32_t max_steps;
    uint32_t precision; // rt_precision
    // ---- camera (per-frame constants; identical for every pixel)
    uint32_t width, height, row_first, n_rays;
    uint32_t tiles_x;   // 8x8 pixel tiles per row of tiles
    uint32_t mode;      // 0 camera, 1 explicit rays
    double inv_w, inv_h, aspect, tan_half;
    double fwd[3], right[3], up[3];
    double sth, cth, sph, cph;       // observer angles
    double X0[3];                    // observer position (KS or flat for superpositions)
    double frame[3][4];              // Gram–Schmidt spatial frame e_1..e_3 at the observer
    double u0;                       // observer e_0 time component
    double kobs[3], two_h_obs;       // KS null vector and 2H at the observer (single Kerr)
    // superposition observer data (multi): per-source potential and world k at X0
    double pot0[kMaxSources];
    double dir0[kMaxSources][3];
    uint32_t observer_inside;        // superposition: observer inside a horizon
    uint32_t pad1;
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

}  // namespace rt
