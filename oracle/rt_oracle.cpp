// rt_oracle.cpp — FP64 CPU oracle: ray tracer, frame driver and C API.
//
// TEST INFRASTRUCTURE ONLY (see rt_oracle.hpp). Loaded by tests/ through
// ctypes, by __graft_entry__.smoke() as the checker and by bench.py as the
// CPU baseline; the product library never links it.
//
// Tracing semantics (the contract the HIP kernels also implement):
//  * camera -> initial state restates dispatch_fp64_scalar's pixel loop and
//    trace_kerr_schild_photon's set-up (render/software_compute_engine.hpp:3843-3938,
//    1634-1652); multi-source set-up restates trace_multi_source
//    (software_compute_engine.hpp:3054-3066) and ms_initial_momentum (:2919-2964);
//  * stepping is adaptive DP45 (integrators/rk45_adaptive.hpp:196-311) on the
//    Hamiltonian right-hand side, first-same-as-last when the state is not
//    projected, with the null-constraint projection of rk45_adaptive.hpp:91-139
//    restated for p_t;
//  * events are located inside an accepted step on the DP45 dense output
//    with an Illinois root finder: horizon r = r_+ + eps, equatorial disk z = 0
//    with r in [r_in, r_out], escape r = r_esc while outgoing; the earliest wins;
//  * exhausted budget / step underflow classify as captured when r < 3 r_s, as
//    escaped otherwise (software_compute_engine.hpp:1767-1769).
#include <atomic>
#include <cstring>
#include <thread>
#include <vector>

#include "rt_oracle.hpp"

namespace rto {

// Diagnostic counters (summed over all rays traced since the last reset).
std::atomic<uint64_t> g_rhs_evals{0}, g_projections{0}, g_accepted{0}, g_rejected{0};

struct Source {
    double c[3];
    double m, a, rh;
    double fx[3], fy[3], ax[3];
    double disk_in, disk_out;
    bool disk_on;
};

struct Spacetime {
    bool multi = false;
    int n = 1;
    double M = 1, a = 0, rh = 2;      // central hole (single-Kerr path)
    double disk_in = 6, disk_out = 20, disk_spin = 0;
    bool disk_on = true, clamp_g = true;
    double eps_h = 1e-6, r_esc = 1000;
    bool carter = false;              // RT_TERM_CARTER_CAPTURE (single Kerr)
    double carter_gate = 1e300;       // the test runs inside this radius
    Source s[RT_MAX_SOURCES];
};

// Margin of the Carter test, relative to r^4 (R(r) / r^4 is dimensionless).
constexpr double kCarterMargin = 1e-6;

struct Integ {
    double rtol, atol, h0, h_min, h_max, inv_tol;
    uint32_t max_steps;
};

static double horizon_radius(double m, double a) { return m + sqrt(max(m * m - a * a, 0.0)); }

// Restates the parameter set-up of dispatch_fp64_scalar (software_compute_engine.hpp:3801-3809)
// and build_multi_source_data (:2738-2798).
static Spacetime make_spacetime(const rt_metric* mt, const rt_termination* tm) {
    Spacetime st;
    st.eps_h = tm->horizon_eps;
    st.r_esc = tm->escape_radius;
    st.clamp_g = (tm->flags & RT_TERM_NO_G_CLAMP) == 0;
    if (mt->n_sources >= 2) {
        st.multi = true;
        st.n = (int)std::min<uint32_t>(mt->n_sources, RT_MAX_SOURCES);
        for (int i = 0; i < st.n; ++i) {
            const rt_source& b = mt->src[i];
            Source& s = st.s[i];
            const double mb = (i == 0) ? max(b.mass, 1e-4) : max(b.mass, 1e-9);
            const double ab = (i == 0) ? std::clamp(b.spin * mb, -0.999 * mb, 0.999 * mb)
                                       : std::clamp(b.spin, -0.9999, 0.9999) * mb;
            s.m = mb;
            s.a = ab;
            s.rh = horizon_radius(mb, ab);
            for (int j = 0; j < 3; ++j) s.c[j] = b.position[j];
            const double al = sqrt(b.axis[0] * b.axis[0] + b.axis[1] * b.axis[1] + b.axis[2] * b.axis[2]);
            double ax[3] = {0, 0, 1};
            if (al > 1e-12) {
                ax[0] = b.axis[0] / al;
                ax[1] = b.axis[1] / al;
                ax[2] = b.axis[2] / al;
            }
            const double helper[3] = {abs(ax[2]) < 0.9 ? 0.0 : 1.0, 0.0, abs(ax[2]) < 0.9 ? 1.0 : 0.0};
            double fx[3] = {helper[1] * ax[2] - helper[2] * ax[1], helper[2] * ax[0] - helper[0] * ax[2],
                            helper[0] * ax[1] - helper[1] * ax[0]};
            const double fl = sqrt(fx[0] * fx[0] + fx[1] * fx[1] + fx[2] * fx[2]);
            fx[0] /= fl;
            fx[1] /= fl;
            fx[2] /= fl;
            const double fy[3] = {ax[1] * fx[2] - ax[2] * fx[1], ax[2] * fx[0] - ax[0] * fx[2],
                                  ax[0] * fx[1] - ax[1] * fx[0]};
            // The central source keeps the world frame exactly, as build_multi_source_data does.
            const bool world = (i == 0 && al > 1e-12 && ax[0] == 0.0 && ax[1] == 0.0 && ax[2] > 0.0);
            for (int j = 0; j < 3; ++j) {
                s.ax[j] = ax[j];
                s.fx[j] = world ? (j == 0 ? 1.0 : 0.0) : fx[j];
                s.fy[j] = world ? (j == 1 ? 1.0 : 0.0) : fy[j];
            }
            s.disk_on = b.disk_enabled != 0;
            const double isco = kerr_isco(mb, ab);
            const double din = (b.disk_inner > 0) ? b.disk_inner : isco;
            s.disk_in = din;
            s.disk_out = max((b.disk_outer > 0) ? b.disk_outer : 20.0 * mb, din * 1.05);
        }
        st.M = st.s[0].m;
        st.a = st.s[0].a;
        st.rh = st.s[0].rh;
        st.disk_on = false;
        return st;
    }
    st.M = max(mt->mass, 1e-4);
    st.a = std::clamp(mt->spin, -0.999 * st.M, 0.999 * st.M);
    st.rh = horizon_radius(st.M, st.a);
    st.disk_spin = std::clamp(st.a, -0.999 * st.M, 0.999 * st.M);
    st.disk_on = tm->disk_enabled != 0;
    const double isco = kerr_isco(st.M, st.a);
    st.disk_in = (tm->disk_inner > 0) ? tm->disk_inner : isco;
    st.disk_out = max((tm->disk_outer > 0) ? tm->disk_outer : 20.0 * st.M, st.disk_in * 1.05);
    st.carter = (tm->flags & RT_TERM_CARTER_CAPTURE) != 0;
    st.carter_gate = st.disk_on ? st.disk_in : 1e300;
    return st;
}

static Integ make_integ(const rt_integrator* ig) {
    Integ o;
    o.rtol = ig->rtol;
    o.atol = ig->atol;
    o.h0 = ig->h0;
    o.h_min = ig->h_min;
    o.h_max = ig->h_max;
    o.inv_tol = ig->invariant_tol;
    o.max_steps = ig->max_steps;
    return o;
}

// ---------------------------------------------------- multi-source pieces
static inline void to_local(const Source& s, const double d[3], double l[3]) {
    l[0] = d[0] * s.fx[0] + d[1] * s.fx[1] + d[2] * s.fx[2];
    l[1] = d[0] * s.fy[0] + d[1] * s.fy[1] + d[2] * s.fy[2];
    l[2] = d[0] * s.ax[0] + d[1] * s.ax[1] + d[2] * s.ax[2];
}
static inline void to_world(const Source& s, const double l[3], double d[3]) {
    d[0] = l[0] * s.fx[0] + l[1] * s.fy[0] + l[2] * s.ax[0];
    d[1] = l[0] * s.fx[1] + l[1] * s.fy[1] + l[2] * s.ax[1];
    d[2] = l[0] * s.fx[2] + l[1] * s.fy[2] + l[2] * s.ax[2];
}

// Restates ms_rhs (software_compute_engine.hpp:2857-2870): sum of the rotated
// per-source KS terms on top of the flat part, with the carried p_t.
static void ms_rhs(const Spacetime& st, const double y[8], double d[8]) {
    d[0] = -y[4];
    d[1] = y[5];
    d[2] = y[6];
    d[3] = y[7];
    d[4] = 0.0;
    d[5] = d[6] = d[7] = 0.0;
    for (int i = 0; i < st.n; ++i) {
        const Source& s = st.s[i];
        const double dx[3] = {y[1] - s.c[0], y[2] - s.c[1], y[3] - s.c[2]};
        double lx[3], lp[3];
        to_local(s, dx, lx);
        to_local(s, y + 5, lp);
        double two_hs, corr[3], dp[3], wc[3], wp[3];
        ks_terms(lx[0], lx[1], lx[2], y[4], lp[0], lp[1], lp[2], s.m, s.a, two_hs, corr, dp);
        to_world(s, corr, wc);
        to_world(s, dp, wp);
        d[0] += two_hs;
        d[1] += wc[0];
        d[2] += wc[1];
        d[3] += wc[2];
        d[5] += wp[0];
        d[6] += wp[1];
        d[7] += wp[2];
    }
}

static void rhs(const Spacetime& st, const double y[8], double d[8]) {
    if (st.multi)
        ms_rhs(st, y, d);
    else
        ks_rhs(y, st.M, st.a, d);
}

// Coefficients of the null constraint as a quadratic in p_t:
//   g^{mu nu} p p = A p_t^2 + B p_t + C0
static void constraint_quadratic(const Spacetime& st, const double y[8], double& A, double& B, double& C0) {
    const double P2 = y[5] * y[5] + y[6] * y[6] + y[7] * y[7];
    double sum_h = 0.0, sum_hq = 0.0, sum_hq2 = 0.0;
    auto add = [&](double H, double q) {
        sum_h += H;
        sum_hq += H * q;
        sum_hq2 += H * q * q;
    };
    if (!st.multi) {
        const KSField<double> f = ks_field(y[1], y[2], y[3], st.M, st.a);
        add(f.H, f.k0 * y[5] + f.k1 * y[6] + f.k2 * y[7]);
    } else {
        for (int i = 0; i < st.n; ++i) {
            const Source& s = st.s[i];
            const double dx[3] = {y[1] - s.c[0], y[2] - s.c[1], y[3] - s.c[2]};
            double lx[3], lp[3];
            to_local(s, dx, lx);
            to_local(s, y + 5, lp);
            const KSField<double> f = ks_field(lx[0], lx[1], lx[2], s.m, s.a);
            add(f.H, f.k0 * lp[0] + f.k1 * lp[1] + f.k2 * lp[2]);
        }
    }
    A = -(1.0 + 2.0 * sum_h);
    B = 4.0 * sum_hq;
    C0 = P2 - 2.0 * sum_hq2;
}

double constraint(const Spacetime& st, const double y[8]) {
    double A, B, C0;
    constraint_quadratic(st, y, A, B, C0);
    return (A * y[4] + B) * y[4] + C0;
}

// Null-constraint projection, restating project_invariant (rk45_adaptive.hpp:91-139)
// for the Hamiltonian state: when |g^{mu nu} p p| >= tol, p_t is replaced by the
// root of the quadratic closest to its current value. Returns true if it fired.
static bool project(const Spacetime& st, double y[8], double tol) {
    double A, B, C0;
    constraint_quadratic(st, y, A, B, C0);
    const double res = (A * y[4] + B) * y[4] + C0;
    // Tolerance relative to the momentum scale: near r_+ the covariant components
    // of a backward ray grow like 1/(r - r_+), and an absolute 1e-12 on a sum of
    // terms of size |p|^2 would sit below the rounding of those terms.
    const double scale = max(1.0, y[4] * y[4] + y[5] * y[5] + y[6] * y[6] + y[7] * y[7]);
    if (!(abs(res) >= tol * scale)) return false;
    if (abs(A) > 1e-14) {
        const double disc = B * B - 4.0 * A * C0;
        if (disc >= 0.0) {
            const double sd = sqrt(disc);
            const double r1 = (-B + sd) / (2.0 * A);
            const double r2 = (-B - sd) / (2.0 * A);
            y[4] = (abs(r1 - y[4]) <= abs(r2 - y[4])) ? r1 : r2;
            return true;
        }
        return false;
    }
    if (abs(B) > 1e-14) {
        y[4] = -C0 / B;
        return true;
    }
    return false;
}

// Reference render step heuristic (software_compute_engine.hpp:1677-1683 with the
// GpuCameraPushConstants defaults of render/gpu_types.hpp:981-1034) used as the
// initial |h| when the caller gives none.
static double initial_step(const Spacetime& st, const double y[8]) {
    const double kappa = 0.5, hmin = 0.004, hmax = 3.5, sff = 7.8;
    double scale = 1e300, blend = 1.0;
    auto one = [&](double r, double rh, double m) {
        scale = min(scale, sqrt(r * max(r - rh, 0.02 * m)));
        blend = min(blend, std::clamp((r - 20.0 * rh) / (80.0 * rh), 0.0, 1.0));
    };
    if (!st.multi) {
        one(ks_radius(y[1], y[2], y[3], st.a), st.rh, st.M);
    } else {
        for (int i = 0; i < st.n; ++i) {
            const Source& s = st.s[i];
            const double dx[3] = {y[1] - s.c[0], y[2] - s.c[1], y[3] - s.c[2]};
            double l[3];
            to_local(s, dx, l);
            one(ks_radius(l[0], l[1], l[2], s.a), s.rh, s.m);
        }
    }
    const double far = 1.0 + (sff - 1.0) * blend;
    return std::clamp(kappa * scale * far, hmin, hmax * sff);
}

// ------------------------------------------------------------------ rays
struct RayOut {
    uint32_t cls = 0;
    double y[8] = {0};
    double bl[3] = {0};
    double lambda = 0, g = 0;
    uint32_t acc = 0, rej = 0;
};

static void finish_bl(const Spacetime& st, RayOut& o) {
    if (!st.multi) {
        bl_from_ks(o.y + 1, st.a, o.bl);
    } else {
        const Source& s = st.s[0];
        const double dx[3] = {o.y[1] - s.c[0], o.y[2] - s.c[1], o.y[3] - s.c[2]};
        double l[3];
        to_local(s, dx, l);
        bl_from_ks(l, s.a, o.bl);
    }
}

// Position of the ray relative to the escape/budget rules.
static bool near_any_hole(const Spacetime& st, const double y[8]) {
    if (!st.multi) return ks_radius(y[1], y[2], y[3], st.a) < 3.0 * (2.0 * st.M);
    for (int i = 0; i < st.n; ++i) {
        const Source& s = st.s[i];
        const double dx = y[1] - s.c[0], dy = y[2] - s.c[1], dz = y[3] - s.c[2];
        if (sqrt(dx * dx + dy * dy + dz * dz) < 6.0 * s.m) return true;
    }
    return false;
}

static double escape_radius_of(const Spacetime& st, const double* x) {
    if (!st.multi) return ks_radius(x[0], x[1], x[2], st.a);
    return sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
}

// Trace one ray from y_init (t, x, y, z, p_t, p_x, p_y, p_z). xi is L_z/E of
// the ray for the single-hole disk redshift (reference: angular_momentum_z of the
// initial state, software_compute_engine.hpp:1650).
// Optional record of every crossing of the equatorial plane z = 0 along a
// single-Kerr ray (test infrastructure: rto_z_crossings).
struct CrossingLog {
    static constexpr int kCap = 512;
    double r[kCap], lambda[kCap];
    int n = 0;
};
static thread_local CrossingLog* g_crossings = nullptr;

RayOut trace_ray(const Spacetime& st, const Integ& ig, const double y_init[8], double xi) {
    RayOut o;
    double y[8], y5[8], k[7][8];
    std::memcpy(y, y_init, sizeof(y));
    auto f = [&](const double* yy, double* dd) { rhs(st, yy, dd); };
    f(y, k[0]);
    double ah = (ig.h0 > 0) ? ig.h0 : initial_step(st, y);
    ah = std::clamp(ah, ig.h_min, ig.h_max);
    const double sgn = -1.0;  // backward in affine parameter
    double lambda = 0.0;
    uint32_t acc = 0, rej = 0, flags = 0;
    uint32_t cls = 0;
    bool done = false;
    // RT_TERM_CARTER_CAPTURE: the test runs once, on the initial state (the
    // constants hold along the ray and are best conditioned far from the hole)
    double r_prev = st.multi ? 0.0 : ks_radius(y[1], y[2], y[3], st.a);
    bool carter_no_turn = false;
    if (st.carter && !st.multi) {
        double xi_c, eta_c;
        carter_no_turn = carter_constants(y, st.a, r_prev, &xi_c, &eta_c) &&
                         carter_no_turning_point(st.M, st.a, st.rh, xi_c, eta_c, kCarterMargin) == 1;
    }
    while (!done) {
        if (acc + rej >= ig.max_steps || ah < ig.h_min) {
            flags |= (acc + rej >= ig.max_steps) ? RT_FLAG_BUDGET : RT_FLAG_MIN_STEP;
            cls = near_any_hole(st, y) ? RT_CLS_CAPTURED : RT_CLS_ESCAPED;
            std::memcpy(o.y, y, sizeof(y));
            break;
        }
        const double h = sgn * ah;
        const double err = dp45_attempt<double, 8>(f, y, h, ig.rtol, ig.atol, k, y5);
        bool finite = std::isfinite(err);
        for (int i = 0; i < 8; ++i) finite = finite && std::isfinite(y5[i]);
        if (!finite) {
            flags |= RT_FLAG_NONFINITE;
            cls = RT_CLS_CAPTURED;
            std::memcpy(o.y, y, sizeof(y));
            break;
        }
        if (err <= 1.0) {
            // ---- events inside the accepted step, earliest first
            double best = 2.0;
            uint32_t best_cls = 0;
            double best_r = 0.0, best_g = 0.0;
            auto dens = [&](double s, int i) { return dense_eval<8>(s, i, y, y5, h, k); };
            auto dens_x = [&](double s, double out[3]) {
                out[0] = dens(s, 1);
                out[1] = dens(s, 2);
                out[2] = dens(s, 3);
            };
            if (!st.multi) {
                const double r0 = ks_radius(y[1], y[2], y[3], st.a);
                const double r1 = ks_radius(y5[1], y5[2], y5[3], st.a);
                const double r_hz = st.rh + st.eps_h;
                if (r1 <= r_hz) {
                    const double s = illinois_root(
                        [&](double ss) {
                            double x[3];
                            dens_x(ss, x);
                            return ks_radius(x[0], x[1], x[2], st.a) - r_hz;
                        },
                        r0 - r_hz, r1 - r_hz);
                    if (s < best) {
                        best = s;
                        best_cls = RT_CLS_CAPTURED;
                    }
                }
                const double z0 = y[3], z1 = y5[3];
                if (g_crossings && g_crossings->n < CrossingLog::kCap &&
                    ((z0 < 0.0 && z1 >= 0.0) || (z0 > 0.0 && z1 <= 0.0))) {
                    const double s = illinois_root([&](double ss) { return dens(ss, 3); }, z0, z1);
                    double x[3];
                    dens_x(s, x);
                    g_crossings->r[g_crossings->n] = ks_radius(x[0], x[1], x[2], st.a);
                    g_crossings->lambda[g_crossings->n] = lambda + s * ah;
                    ++g_crossings->n;
                }
                if (st.disk_on && ((z0 < 0.0 && z1 >= 0.0) || (z0 > 0.0 && z1 <= 0.0))) {
                    const double s = illinois_root([&](double ss) { return dens(ss, 3); }, z0, z1);
                    double x[3];
                    dens_x(s, x);
                    const double rc = ks_radius(x[0], x[1], x[2], st.a);
                    if (rc >= st.disk_in && rc <= st.disk_out && rc > st.rh * 1.02 && s < best) {
                        best = s;
                        best_cls = RT_CLS_DISK;
                        best_r = rc;
                        best_g = disk_redshift(st.M, st.disk_spin, rc, xi, st.clamp_g);
                    }
                }
                if (r1 >= st.r_esc && sgn * (y5[1] * k[6][1] + y5[2] * k[6][2] + y5[3] * k[6][3]) > 0.0) {
                    double s = 1.0;
                    if (r0 < st.r_esc) {
                        s = illinois_root(
                            [&](double ss) {
                                double x[3];
                                dens_x(ss, x);
                                return ks_radius(x[0], x[1], x[2], st.a) - st.r_esc;
                            },
                            r0 - st.r_esc, r1 - st.r_esc);
                    }
                    if (s < best) {
                        best = s;
                        best_cls = RT_CLS_ESCAPED;
                    }
                }
            } else {
                for (int i = 0; i < st.n; ++i) {
                    const Source& src = st.s[i];
                    auto local = [&](const double* x, double l[3]) {
                        const double d[3] = {x[0] - src.c[0], x[1] - src.c[1], x[2] - src.c[2]};
                        to_local(src, d, l);
                    };
                    double l0[3], l1[3];
                    local(y + 1, l0);
                    local(y5 + 1, l1);
                    const double r0 = ks_radius(l0[0], l0[1], l0[2], src.a);
                    const double r1 = ks_radius(l1[0], l1[1], l1[2], src.a);
                    const double r_hz = src.rh + st.eps_h;
                    if (r1 <= r_hz && r0 > r_hz) {
                        const double s = illinois_root(
                            [&](double ss) {
                                double x[3], l[3];
                                dens_x(ss, x);
                                local(x, l);
                                return ks_radius(l[0], l[1], l[2], src.a) - r_hz;
                            },
                            r0 - r_hz, r1 - r_hz);
                        if (s < best) {
                            best = s;
                            best_cls = RT_CLS_CAPTURED;
                        }
                    }
                    const double z0 = l0[2], z1 = l1[2];
                    if (src.disk_on && ((z0 < 0.0 && z1 >= 0.0) || (z0 > 0.0 && z1 <= 0.0))) {
                        const double s = illinois_root(
                            [&](double ss) {
                                double x[3], l[3];
                                dens_x(ss, x);
                                local(x, l);
                                return l[2];
                            },
                            z0, z1);
                        double x[3], l[3];
                        dens_x(s, x);
                        local(x, l);
                        const double rc = ks_radius(l[0], l[1], l[2], src.a);
                        if (rc >= src.disk_in && rc <= src.disk_out && rc > src.rh * 1.02 && s < best) {
                            // xi = local L_z over |p|, as ms_disk_crossing (software_compute_engine.hpp:3021-3042).
                            const double pw[3] = {dens(s, 5), dens(s, 6), dens(s, 7)};
                            double pl[3];
                            to_local(src, pw, pl);
                            const double pm = sqrt(pl[0] * pl[0] + pl[1] * pl[1] + pl[2] * pl[2]);
                            const double lz = (l[0] * pl[1] - l[1] * pl[0]) / max(pm, 1e-30);
                            best = s;
                            best_cls = RT_CLS_DISK;
                            best_r = rc;
                            best_g = disk_redshift(src.m, src.a, rc, lz, st.clamp_g);
                        }
                    }
                }
                const double R0 = escape_radius_of(st, y + 1), R1 = escape_radius_of(st, y5 + 1);
                if (R1 >= st.r_esc && sgn * (y5[1] * k[6][1] + y5[2] * k[6][2] + y5[3] * k[6][3]) > 0.0) {
                    double s = 1.0;
                    if (R0 < st.r_esc) {
                        s = illinois_root(
                            [&](double ss) {
                                double x[3];
                                dens_x(ss, x);
                                return sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]) - st.r_esc;
                            },
                            R0 - st.r_esc, R1 - st.r_esc);
                    }
                    if (s < best) {
                        best = s;
                        best_cls = RT_CLS_ESCAPED;
                    }
                }
            }
            ++acc;
            if (best_cls != 0) {
                if (best >= 1.0) {
                    std::memcpy(o.y, y5, sizeof(y5));
                } else {
                    for (int i = 0; i < 8; ++i) o.y[i] = dens(best, i);
                }
                if (project(st, o.y, ig.inv_tol)) flags |= RT_FLAG_PROJECTED;
                lambda += best * ah;
                cls = best_cls;
                if (cls == RT_CLS_DISK) {
                    o.g = best_g;
                    (void)best_r;
                }
                done = true;
                break;
            }
            std::memcpy(y, y5, sizeof(y));
            lambda += ah;
            if (project(st, y, ig.inv_tol)) {
                flags |= RT_FLAG_PROJECTED;
                g_projections.fetch_add(1, std::memory_order_relaxed);
                f(y, k[0]);
            } else {
                std::memcpy(k[0], k[6], sizeof(k[0]));
            }
            if (carter_no_turn) {
                // no turning point: captured once moving inward inside the gate
                const double r1 = ks_radius(y[1], y[2], y[3], st.a);
                if (r1 < r_prev && r1 < st.carter_gate) {
                    cls = RT_CLS_CAPTURED;
                    flags |= RT_FLAG_CARTER;
                    std::memcpy(o.y, y, sizeof(y));
                    done = true;
                    break;
                }
                r_prev = r1;
            }
            ah = dp45_next_step(ah, err, true, ig.h_min, ig.h_max);
        } else {
            ++rej;
            ah = dp45_next_step(ah, err, false, ig.h_min, ig.h_max);
        }
    }
    g_accepted.fetch_add(acc, std::memory_order_relaxed);
    g_rejected.fetch_add(rej, std::memory_order_relaxed);
    o.cls = cls | flags;
    o.lambda = lambda;
    o.acc = acc;
    o.rej = rej;
    if ((cls & RT_CLS_MASK) == RT_CLS_ESCAPED) o.g = 1.0 / (-o.y[4]);
    if ((cls & RT_CLS_MASK) == RT_CLS_CAPTURED) o.g = 0.0;
    finish_bl(st, o);
    return o;
}

// Restates ms_initial_momentum (software_compute_engine.hpp:2919-2964): the
// superposed metric's Gram-Schmidt frame at x, and the covariant spatial
// momentum of the backward ray along -outward, normalised to p_t = -1.
static void ms_initial_momentum(const Spacetime& st, const double x[3], const double outward[3], double p[3]) {
    double pot[RT_MAX_SOURCES], dir[RT_MAX_SOURCES][3];
    double psum = 0.0;
    for (int i = 0; i < st.n; ++i) {
        const Source& s = st.s[i];
        const double dx[3] = {x[0] - s.c[0], x[1] - s.c[1], x[2] - s.c[2]};
        double l[3];
        to_local(s, dx, l);
        const KSField<double> f = ks_field(l[0], l[1], l[2], s.m, s.a);
        const double kl[3] = {f.k0, f.k1, f.k2};
        pot[i] = f.H;
        to_world(s, kl, dir[i]);
        psum += f.H;
    }
    const int n = st.n;
    auto mdot = [&](const double* A, const double* B) {
        double v = -A[0] * B[0] + A[1] * B[1] + A[2] * B[2] + A[3] * B[3];
        for (int i = 0; i < n; ++i) {
            const double la = A[0] + dir[i][0] * A[1] + dir[i][1] * A[2] + dir[i][2] * A[3];
            const double lb = B[0] + dir[i][0] * B[1] + dir[i][1] * B[2] + dir[i][2] * B[3];
            v += 2.0 * pot[i] * la * lb;
        }
        return v;
    };
    const double ils = 1.0 / max(1.0 - 2.0 * psum, 1e-6);
    const double u0 = sqrt(ils);
    auto lift = [&](double vx, double vy, double vz, double out[4]) {
        double shift = 0.0;
        for (int i = 0; i < n; ++i) shift += pot[i] * (dir[i][0] * vx + dir[i][1] * vy + dir[i][2] * vz);
        out[0] = 2.0 * shift * ils;
        out[1] = vx;
        out[2] = vy;
        out[3] = vz;
    };
    double e1[4], e2[4], e3[4];
    lift(1, 0, 0, e1);
    lift(0, 1, 0, e2);
    lift(0, 0, 1, e3);
    const double n1_ = sqrt(max(mdot(e1, e1), 1e-30));
    for (int j = 0; j < 4; ++j) e1[j] /= n1_;
    const double e21 = mdot(e2, e1);
    for (int j = 0; j < 4; ++j) e2[j] -= e21 * e1[j];
    const double n2_ = sqrt(max(mdot(e2, e2), 1e-30));
    for (int j = 0; j < 4; ++j) e2[j] /= n2_;
    const double e31 = mdot(e3, e1), e32 = mdot(e3, e2);
    for (int j = 0; j < 4; ++j) e3[j] -= e31 * e1[j] + e32 * e2[j];
    const double n3_ = sqrt(max(mdot(e3, e3), 1e-30));
    for (int j = 0; j < 4; ++j) e3[j] /= n3_;
    const double vv[4] = {u0 - outward[0] * e1[0] - outward[1] * e2[0] - outward[2] * e3[0],
                          -outward[0] * e1[1] - outward[1] * e2[1] - outward[2] * e3[1],
                          -outward[0] * e1[2] - outward[1] * e2[2] - outward[2] * e3[2],
                          -outward[0] * e1[3] - outward[1] * e2[3] - outward[2] * e3[3]};
    double mom[3] = {vv[1], vv[2], vv[3]};
    double tc = -vv[0];
    for (int i = 0; i < n; ++i) {
        const double lv = vv[0] + dir[i][0] * vv[1] + dir[i][1] * vv[2] + dir[i][2] * vv[3];
        tc += 2.0 * pot[i] * lv;
        mom[0] += 2.0 * pot[i] * lv * dir[i][0];
        mom[1] += 2.0 * pot[i] * lv * dir[i][1];
        mom[2] += 2.0 * pot[i] * lv * dir[i][2];
    }
    const double inv_pt = 1.0 / max(-tc, 1e-12);
    p[0] = mom[0] * inv_pt;
    p[1] = mom[1] * inv_pt;
    p[2] = mom[2] * inv_pt;
}

// -------------------------------------------------------------- camera
// Unit vector (n1, n2, n3) = (forward, up, right) components of the ray
// through screen point (u, v), restating CameraProjector::compute_ray_direction
// (observer/camera_projections.hpp:198-225) and its projections (:27-196)
// for the call dispatch_fp64_scalar makes (no boost: the auto-zoom ray is the
// pinhole one). Unknown modes fall back to the pinhole, as the switch does.
static void normalized(double n1, double n2, double n3, double nl[3]) {
    const double len = sqrt(n1 * n1 + n2 * n2 + n3 * n3);
    const double inv_len = (len > 0.0) ? (1.0 / len) : 1.0;
    nl[0] = n1 * inv_len;
    nl[1] = n2 * inv_len;
    nl[2] = n3 * inv_len;
}

void screen_direction(uint32_t mode, double u, double v, double fov, double nl[3]) {
    switch (mode) {
        case RT_PROJ_FISHEYE_STEREOGRAPHIC: {  // :70-94
            const double max_fov = std::clamp(fov, 0.1, 260.0 * kPi / 180.0);
            const double r = sqrt(u * u + v * v);
            if (r < 1e-15) {
                nl[0] = 1.0, nl[1] = 0.0, nl[2] = 0.0;
                return;
            }
            const double half_max_fov = max_fov * 0.5;
            const double theta = 2.0 * std::atan(r * std::tan(half_max_fov * 0.5));
            const double sin_t = std::sin(theta), cos_t = std::cos(theta);
            normalized(cos_t, v * (sin_t / r), u * (sin_t / r), nl);
            return;
        }
        case RT_PROJ_EQUIRECTANGULAR_360: {  // :174-196
            const double fov_scale = fov / (60.0 * kPi / 180.0);
            const double phi = u * kPi * fov_scale;
            const double theta = std::clamp(kPi * 0.5 - v * (kPi * 0.5) * fov_scale, 0.0001, kPi - 0.0001);
            const double sin_t = std::sin(theta), cos_t = std::cos(theta);
            const double sin_p = std::sin(phi), cos_p = std::cos(phi);
            normalized(sin_t * cos_p, cos_t, sin_t * sin_p, nl);
            return;
        }
        case RT_PROJ_FISHEYE_EQUIDISTANT: {  // :96-113
            const double r = sqrt(u * u + v * v);
            if (r < 1e-15) {
                nl[0] = 1.0, nl[1] = 0.0, nl[2] = 0.0;
                return;
            }
            const double theta = min(r * fov * 0.5, kPi);
            const double sin_t = std::sin(theta), cos_t = std::cos(theta);
            normalized(cos_t, v * (sin_t / r), u * (sin_t / r), nl);
            return;
        }
        case RT_PROJ_FISHEYE_ORTHOGRAPHIC: {  // :115-132
            const double r = sqrt(u * u + v * v);
            const double sin_half_fov = std::sin(std::clamp(fov * 0.5, 0.01, kPi * 0.499));
            const double sin_t = std::clamp(r * sin_half_fov, 0.0, 1.0);
            const double cos_t = sqrt(max(1.0 - sin_t * sin_t, 0.0));
            if (r < 1e-15) {
                nl[0] = 1.0, nl[1] = 0.0, nl[2] = 0.0;
                return;
            }
            normalized(cos_t, v * (sin_t / r), u * (sin_t / r), nl);
            return;
        }
        case RT_PROJ_PANINI_CYLINDRICAL: {  // :134-152
            const double d = 1.0;
            const double tan_h = std::tan(fov * 0.5);
            const double px = u * tan_h, py = v * tan_h;
            const double s = (d + 1.0) / (d + sqrt(max(1.0 + (1.0 - d * d) * px * px / ((d + 1.0) * (d + 1.0)), 1e-15)));
            const double sin_phi = (px * s) / (d + 1.0);
            const double cos_phi = sqrt(max(1.0 - sin_phi * sin_phi, 0.0));
            normalized(cos_phi, py, sin_phi, nl);
            return;
        }
        case RT_PROJ_HAMMER_AITOFF: {  // :154-172 (not renormalised)
            const double sqrt2 = 1.4142135623730951;
            const double zoom = (60.0 * kPi / 180.0) / max(fov, 0.01);
            const double x = std::clamp(u * zoom * sqrt2, -sqrt2, sqrt2);
            const double y = std::clamp(v * zoom * (sqrt2 * 0.5), -(sqrt2 * 0.5), sqrt2 * 0.5);
            const double z = sqrt(max(1.0 - (x * x * 0.25) - (y * y), 1e-15));
            const double phi = 2.0 * std::atan2(z * x, 2.0 * (2.0 * z * z - 1.0));
            const double theta = kPi * 0.5 - std::asin(std::clamp(z * y * sqrt2, -1.0, 1.0));
            const double sin_t = std::sin(theta), cos_t = std::cos(theta);
            const double sin_p = std::sin(phi), cos_p = std::cos(phi);
            nl[0] = sin_t * cos_p;
            nl[1] = cos_t;
            nl[2] = sin_t * sin_p;
            return;
        }
        default: {  // pinhole (:27-41), auto-zoom with no boost (:43-68)
            const double tan_half = std::tan(fov * 0.5);
            normalized(1.0, v * tan_half, u * tan_half, nl);
            return;
        }
    }
}

// Pixel -> initial state, restating dispatch_fp64_scalar (software_compute_engine.hpp:3843-3938)
// with the projections above and the set-up of trace_kerr_schild_photon (:1634-1652) /
// trace_multi_source (:3054-3066). The all-sky modes take u without the
// aspect ratio (:3855-3856). Returns false when the observer sits inside a
// horizon (captured at once).
void camera_outward(const rt_camera* cam, double rh, uint32_t px, uint32_t py, double outward[3]) {
    const double W = cam->width, Hh = cam->height;
    const double aspect = W / Hh;
    const double v = 1.0 - (static_cast<double>(py) + 0.5) / Hh * 2.0;
    const double u_norm = ((static_cast<double>(px) + 0.5) / W * 2.0 - 1.0) * aspect;
    const bool allsky = cam->projection == RT_PROJ_EQUIRECTANGULAR_360 || cam->projection == RT_PROJ_HAMMER_AITOFF;
    const double u = allsky ? (((static_cast<double>(px) + 0.5) / W) * 2.0 - 1.0) : u_norm;
    double nl[3];
    screen_direction(cam->projection, u, v, cam->fov_y, nl);
    const double d[3] = {nl[0] * cam->forward[0] + nl[2] * cam->right[0] + nl[1] * cam->up[0],
                         nl[0] * cam->forward[1] + nl[2] * cam->right[1] + nl[1] * cam->up[1],
                         nl[0] * cam->forward[2] + nl[2] * cam->right[2] + nl[1] * cam->up[2]};
    (void)rh;
    const double th = std::clamp(cam->observer[2], 0.001, kPi - 0.001);
    const double ph = cam->observer[3];
    const double st_ = std::sin(th), ct = std::cos(th), sp = std::sin(ph), cp = std::cos(ph);
    const double n_r = d[0] * (st_ * cp) + d[1] * (st_ * sp) + d[2] * ct;
    const double n_th = d[0] * (ct * cp) + d[1] * (ct * sp) + d[2] * (-st_);
    const double n_ph = d[0] * (-sp) + d[1] * cp + d[2] * 0.0;
    outward[0] = n_r * st_ * cp + n_th * ct * cp - n_ph * sp;
    outward[1] = n_r * st_ * sp + n_th * ct * sp + n_ph * cp;
    outward[2] = n_r * ct - n_th * st_;
}

bool camera_ray(const Spacetime& st, const rt_camera* cam, uint32_t px, uint32_t py, double y[8], double* xi) {
    double outward[3];
    camera_outward(cam, st.rh, px, py, outward);
    const double r_obs = max(cam->observer[1], st.rh * 1.02);
    const double th = std::clamp(cam->observer[2], 0.001, kPi - 0.001);
    const double ph = cam->observer[3];
    const double st_ = std::sin(th), ct = std::cos(th), sp = std::sin(ph), cp = std::cos(ph);
    y[0] = 0.0;
    y[4] = -1.0;
    if (!st.multi) {
        ks_position_from_bl(r_obs, th, ph, st.a, y + 1);
        ks_initial_momentum(y + 1, st.M, st.a, outward, y + 5);
        *xi = y[1] * y[6] - y[2] * y[5];
        return true;
    }
    // multi-source: flat spherical position, superposed-metric frame.
    y[1] = r_obs * st_ * cp;
    y[2] = r_obs * st_ * sp;
    y[3] = r_obs * ct;
    *xi = 0.0;
    for (int i = 0; i < st.n; ++i) {
        const Source& s = st.s[i];
        const double dx[3] = {y[1] - s.c[0], y[2] - s.c[1], y[3] - s.c[2]};
        double l[3];
        to_local(s, dx, l);
        const double a2 = s.a * s.a;
        const double diff = (l[0] * l[0] + l[1] * l[1] + l[2] * l[2]) - a2;
        const double r2 = max(0.5 * (diff + sqrt(diff * diff + 4.0 * a2 * l[2] * l[2])), 1e-24);
        const double r = sqrt(r2);
        if (r <= s.rh * 1.0001 || (r2 - 2.0 * s.m * r + a2) <= 0.0) return false;
    }
    ms_initial_momentum(st, y + 1, outward, y + 5);
    return true;
}

static void store(rt_results* out, size_t n, size_t idx, const RayOut& o) {
    if (out->cls) out->cls[idx] = o.cls;
    if (out->x)
        for (int c = 0; c < 4; ++c) out->x[c * n + idx] = o.y[c];
    if (out->p)
        for (int c = 0; c < 4; ++c) out->p[c * n + idx] = o.y[4 + c];
    if (out->bl)
        for (int c = 0; c < 3; ++c) out->bl[c * n + idx] = o.bl[c];
    if (out->lambda) out->lambda[idx] = o.lambda;
    if (out->g) out->g[idx] = o.g;
    if (out->steps) out->steps[idx] = o.acc;
    if (out->rejected) out->rejected[idx] = o.rej;
}

template <class Fn>
static void parallel_rows(size_t rows, int threads, Fn&& fn) {
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    std::atomic<size_t> next{0};
    auto worker = [&]() {
        for (;;) {
            const size_t r = next.fetch_add(1);
            if (r >= rows) return;
            fn(r);
        }
    };
    if (threads == 1) {
        worker();
        return;
    }
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) pool.emplace_back(worker);
    for (auto& t : pool) t.join();
}


// ------------------------------------------------ the renderer's tracer
// RT_METHOD_RENDER_RK4: restates trace_kerr_schild_photon
// (render/software_compute_engine.hpp:1624-1809) without its shading (sky,
// colour, tonemapping), keeping everything that decides the pixel record's
// status_flags, iterations_used and redshift: RK4 steps of the render step
// formula, relative horizon, escape radius, the translucent disk whose
// opacity (AccretionDiskModel::shade, render/accretion_disk/accretion_disk_model.hpp:245-314)
// multiplies the throughput, and the after-loop rules. Set-up as
// dispatch_fp64_scalar (:3794-3832): budget halved with a floor of 256,
// turbulence anti-aliasing factor of the frame.

struct RenderCfg {
    double m, a, rs, rh, isco, r_esc, disk_in, disk_out, disk_spin;
    bool disk_on;
    uint32_t max_steps;
    double kappa, h_min, h_max, far_scale, time, detail;
    rt_disk_profile prof;
};

static RenderCfg make_render(const rt_camera* cam, const rt_metric* mt, const rt_integrator* ig,
                             const rt_termination* tm) {
    RenderCfg c;
    c.m = max(mt->mass, 1e-4);
    c.a = std::clamp(mt->spin, -0.999 * c.m, 0.999 * c.m);
    c.rs = 2.0 * c.m;
    c.rh = c.m + sqrt(max(c.m * c.m - c.a * c.a, 0.0));
    c.isco = kerr_isco(c.m, c.a);
    c.r_esc = tm->escape_radius;
    c.disk_on = tm->disk_enabled != 0;
    c.disk_in = (tm->disk_inner > 0.0) ? tm->disk_inner : c.isco;
    c.disk_out = max((tm->disk_outer > 0.0) ? tm->disk_outer : 20.0 * c.m, c.disk_in * 1.05);
    c.disk_spin = std::clamp(c.a, -0.999 * c.m, 0.999 * c.m);
    c.max_steps = std::max(ig->max_steps / 2U, 256U);
    c.kappa = ig->step_size_factor;
    c.h_min = ig->min_step;
    c.h_max = ig->max_step;
    c.far_scale = ig->far_field_step_scale;
    c.time = tm->time;
    c.prof = tm->disk_profile;
    // turbulence anti-aliasing factor (dispatch_fp64_scalar :3834-3836); its r_h
    // is 2 M for |a| <= 1e-12, the same number
    const double rh_frame = (abs(c.a) > 1e-12) ? c.rh : c.rs;
    const double r_obs_frame = max(cam->observer[1], rh_frame * 1.02);
    const double angular_pixel_size = cam->fov_y / max(static_cast<double>(cam->width), 1.0);
    const double bh_angular_diameter = (2.0 * rh_frame) / r_obs_frame;
    c.detail = std::clamp(bh_angular_diameter / max(angular_pixel_size * 24.0, 1e-9), 0.0, 1.0);
    return c;
}

// BL point with the azimuth continued from prev_phi (boyer_lindquist_from_kerr_schild,
// core/kerr_schild_null_integrator.hpp:56-69).
static void bl_point(const double x[3], double a, double prev_phi, double out[3]) {
    const double r = ks_radius(x[0], x[1], x[2], a);
    const double rho = sqrt(max(r * r - x[2] * x[2], 0.0));
    const double two_pi = 2.0 * kPi;
    const double raw_phi = atan2(x[1], x[0]) - atan2(a, r);
    double delta = raw_phi - prev_phi;
    delta -= two_pi * floor(delta / two_pi + 0.5);
    out[0] = r;
    out[1] = atan2(rho, x[2]);
    out[2] = prev_phi + delta;
}

// One classic RK4 step (step_kerr_schild_null_rk4, core/kerr_schild_null_integrator.hpp:116-139)
// of the KS null state, p_t = -1; y[1..3] position, y[5..7] momentum.
static void render_rk4(double y[8], double h, double M, double a) {
    double k1[8], k2[8], k3[8], k4[8], t[8];
    const double half = 0.5 * h;
    auto advance = [&](const double* slope, double scale) {
        t[0] = y[0];
        t[4] = y[4];
        for (int i = 1; i <= 3; ++i) {
            t[i] = y[i] + scale * slope[i];
            t[4 + i] = y[4 + i] + scale * slope[4 + i];
        }
    };
    ks_rhs(y, M, a, k1);
    advance(k1, half);
    ks_rhs(t, M, a, k2);
    advance(k2, half);
    ks_rhs(t, M, a, k3);
    advance(k3, h);
    ks_rhs(t, M, a, k4);
    const double sixth = h / 6.0;
    for (int i = 1; i <= 3; ++i) {
        y[i] += sixth * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]);
        y[4 + i] += sixth * (k1[4 + i] + 2.0 * k2[4 + i] + 2.0 * k3[4 + i] + k4[4 + i]);
    }
}

// The disk's opacity at a crossing: the opacity path of AccretionDiskModel::shade
// (render/accretion_disk/accretion_disk_model.hpp:110-176 helpers, :245-314).
static double smooth_unit(double x) {
    const double t = std::clamp(x, 0.0, 1.0);
    return t * t * (3.0 - 2.0 * t);
}

static double reduce_angle(double angle) {
    const double two_pi = 2.0 * kPi;
    return angle - two_pi * floor(angle / two_pi);
}

static uint32_t hash_u32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

static float lattice_value(int32_t x, int32_t y, int32_t z, uint32_t seed) {
    const uint32_t h = static_cast<uint32_t>(x) * 374761393U + static_cast<uint32_t>(y) * 668265263U +
                       static_cast<uint32_t>(z) * 2147483647U + seed * 1274126177U;
    return static_cast<float>(hash_u32(h) & 0x00FFFFFFU) * (1.0f / 16777216.0f) * 2.0f - 1.0f;
}

static float value_noise(float x, float y, float z, uint32_t seed) {
    const float fx0 = std::floor(x), fy0 = std::floor(y), fz0 = std::floor(z);
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

static float fractal_noise(float x, float y, float z, int octaves, uint32_t seed) {
    float total = 0.0f, amplitude = 1.0f, frequency = 1.0f, normalization = 0.0f;
    for (int o = 0; o < octaves; ++o) {
        total += value_noise(x * frequency, y * frequency, z * frequency, seed + static_cast<uint32_t>(o) * 7919U) *
                 amplitude;
        normalization += amplitude;
        amplitude *= 0.5f;
        frequency *= 2.0f;
    }
    return total / std::max(normalization, 1e-5f);
}

// Keplerian angular velocity of AccretionDiskModel::orbital_frame (:205-216).
static double disk_angular_velocity(double mass, double spin, double radius) {
    const double spin_mag = abs(spin);
    const double sqrt_m = sqrt(max(mass, 1e-12));
    const double sqrt_r = sqrt(max(radius, 1e-12));
    const double angular = sqrt_m / (radius * sqrt_r + spin_mag * sqrt_m);
    return (spin < 0.0) ? -angular : angular;
}

static double disk_opacity(const rt_disk_profile& pf, double m, double spin, double inner_r, double outer_r,
                           double radius_in, double azimuth, double time, double detail_scale) {
    const double mass = max(m, 1e-9);
    const double radius = max(radius_in, 1e-9);
    const double inner = max(inner_r, 1e-9);
    const double outer = max(outer_r, inner * 1.0001);
    const double softness = max(static_cast<double>(pf.edge_softness), 0.05);
    const double envelope =
        smooth_unit((radius - inner) / (0.8 * softness * mass)) * smooth_unit((outer - radius) / (1.5 * softness * mass));
    const double rotation = static_cast<double>(pf.rotation_speed_scale) * time;
    const double shear_phase = reduce_angle(azimuth - disk_angular_velocity(mass, spin, radius) * rotation);
    const double pattern_phase = reduce_angle(azimuth - disk_angular_velocity(mass, spin, sqrt(inner * outer)) * rotation);
    const float log_ratio = static_cast<float>(std::log(max(radius / inner, 1.0)));
    const float log_span = std::max(static_cast<float>(std::log(max(outer / inner, 1.0001))), 1e-4f);
    const float radial = log_ratio / log_span;
    const float pi_f = 3.14159265358979323846f;
    const float two_pi = 2.0f * pi_f;
    const float ring_wave = 0.5f + 0.5f * std::cos(two_pi * pf.ring_frequency * radial);
    const float ring_sharpness = std::max(pf.ring_sharpness, 1.0f);
    const float ring = std::max(
        1.0f + pf.ring_amplitude * (std::pow(ring_wave, ring_sharpness) * std::sqrt(pi_f * ring_sharpness) - 1.0f), 0.0f);
    const float arms = std::floor(pf.spiral_arm_count + 0.5f);
    const float spiral =
        (arms > 0.5f)
            ? std::max(1.0f + pf.spiral_amplitude * std::cos(arms * (static_cast<float>(pattern_phase) - pf.spiral_pitch * log_ratio)),
                       0.0f)
            : 1.0f;
    const float shear_cos = std::cos(static_cast<float>(shear_phase));
    const float shear_sin = std::sin(static_cast<float>(shear_phase));
    const int octaves = std::clamp(static_cast<int>(pf.turbulence_octaves + 0.5f), 1, 6);
    const uint32_t seed = static_cast<uint32_t>(std::max(pf.noise_seed, 0.0f));
    const float detail = static_cast<float>(std::clamp(detail_scale, 0.0, 1.0));
    const float turbulence_noise = fractal_noise(log_ratio * pf.turbulence_scale * pf.radial_stretch,
                                                 shear_cos * pf.turbulence_scale, shear_sin * pf.turbulence_scale,
                                                 octaves, seed);
    const float turbulence = std::exp(1.5f * pf.turbulence_amplitude * detail * turbulence_noise);
    const float grain_noise = value_noise(log_ratio * pf.grain_scale * pf.radial_stretch, shear_cos * pf.grain_scale,
                                          shear_sin * pf.grain_scale, seed + 104729U);
    const float grain = std::max(1.0f + pf.grain_amplitude * detail * grain_noise, 0.0f);
    const float flux = std::clamp(ring * spiral * turbulence * grain, 0.0f, 8.0f);
    return std::clamp(envelope * 0.95 * static_cast<double>(pf.opacity_scale) *
                          std::clamp(sqrt(min(static_cast<double>(flux), 1.0)), 0.15, 1.0),
                      0.0, 0.98);
}

// One pixel through the renderer's tracer. y holds the initial KS state (p_t = -1).
static RayOut trace_render(const RenderCfg& c, double phi_obs, const double y0[8]) {
    RayOut o;
    double y[8];
    std::memcpy(y, y0, sizeof(y));
    const double lz = y[1] * y[6] - y[2] * y[5];
    double point[3];
    bl_point(y + 1, c.a, phi_obs, point);
    double throughput = 1.0, redshift_rec = 1.0;
    uint32_t status = 0, iters = 0, flags = 0;
    const double mid_plane = kPi * 0.5;
    bool ended = false;
    for (uint32_t step = 0; step < c.max_steps && throughput > 0.01; ++step) {
        iters = step + 1;
        if (point[0] <= c.rh * 1.0001) {
            status = RT_CLS_CAPTURED;
            throughput = 0.0;
            ended = true;
            break;
        }
        if (point[0] >= c.r_esc) {
            status = RT_CLS_ESCAPED;
            ended = true;
            break;
        }
        const double r_scale = max(point[0] - c.rh, 0.02 * c.m);
        const double far = 1.0 + (c.far_scale - 1.0) * std::clamp((point[0] - 20.0 * c.rh) / (80.0 * c.rh), 0.0, 1.0);
        const double dt_mag = std::clamp(c.kappa * sqrt(point[0] * r_scale) * far, c.h_min, c.h_max * c.far_scale);
        const double prev[3] = {point[0], point[1], point[2]};
        render_rk4(y, -dt_mag, c.m, c.a);
        if (!std::isfinite(y[1]) || !std::isfinite(y[2]) || !std::isfinite(y[3]) || !std::isfinite(y[5]) ||
            !std::isfinite(y[6]) || !std::isfinite(y[7])) {
            status = RT_CLS_CAPTURED;
            flags |= RT_FLAG_NONFINITE;
            throughput = 0.0;
            ended = true;
            break;
        }
        bl_point(y + 1, c.a, prev[2], point);
        if (c.disk_on && (prev[1] - mid_plane) * (point[1] - mid_plane) <= 0.0) {
            const double span = abs(point[1] - prev[1]);
            const double s_cross = (span > 1e-12) ? std::clamp(abs(prev[1] - mid_plane) / span, 0.0, 1.0) : 0.5;
            const double r_cross = prev[0] + s_cross * (point[0] - prev[0]);
            const double phi_cross = prev[2] + s_cross * (point[2] - prev[2]);
            if (r_cross >= c.disk_in && r_cross <= c.disk_out && r_cross > c.rh * 1.02) {
                status |= RT_CLS_DISK;
                redshift_rec = disk_redshift(c.m, c.disk_spin, r_cross, lz, true);
                const double opacity = disk_opacity(c.prof, c.m, c.disk_spin, c.disk_in, c.disk_out, r_cross, phi_cross,
                                                    c.time, c.detail);
                throughput *= (1.0 - opacity);
            }
        }
    }
    if (!ended && throughput > 0.01) flags |= RT_FLAG_BUDGET;
    if ((status & (RT_CLS_CAPTURED | RT_CLS_ESCAPED)) == 0U && point[0] < 3.0 * c.rs) throughput = 0.0;
    if ((status & RT_CLS_ESCAPED) != 0U || throughput > 0.01) status |= RT_CLS_ESCAPED;
    o.cls = status | flags;
    std::memcpy(o.y, y, sizeof(y));
    o.bl[0] = point[0];
    o.bl[1] = point[1];
    o.bl[2] = point[2];
    o.lambda = iters * 0.05;
    o.g = redshift_rec;
    o.acc = iters;
    o.rej = 0;
    return o;
}

}  // namespace rto

// ===================================================================== C API
using namespace rto;

extern "C" {


// The camera's row window [first, last), as resolve_row_window computes it
// (render/software_compute_engine.hpp:1042-1050): count 0 is the whole frame,
// whatever the offset; otherwise offset + count in 64 bits, clipped.
static void row_window(const rt_camera* cam, uint32_t& first, uint32_t& last) {
    first = 0;
    last = cam->height;
    if (cam->row_count) {
        first = std::min(cam->row_offset, cam->height);
        last = (uint32_t)std::min<uint64_t>((uint64_t)first + cam->row_count, cam->height);
    }
}

int rto_trace_camera(const rt_camera* cam, const rt_metric* mt, const rt_integrator* ig, const rt_termination* tm,
                     rt_results* out, int threads) {
    if (!cam || !mt || !ig || !tm || !out || cam->width == 0 || cam->height == 0) return RT_ERR_INVALID_ARG;
    const Spacetime st = make_spacetime(mt, tm);
    const Integ in = make_integ(ig);
    uint32_t first, last;
    row_window(cam, first, last);
    const size_t W = cam->width;
    const size_t n = (size_t)(last - first) * W;
    parallel_rows(last - first, threads, [&](size_t row) {
        const uint32_t py = first + (uint32_t)row;
        for (uint32_t px = 0; px < cam->width; ++px) {
            double y[8], xi;
            RayOut o;
            if (camera_ray(st, cam, px, py, y, &xi)) {
                o = trace_ray(st, in, y, xi);
            } else {
                std::memcpy(o.y, y, sizeof(y));
                o.cls = RT_CLS_CAPTURED;
                finish_bl(st, o);
            }
            store(out, n, row * W + px, o);
        }
    });
    return RT_OK;
}

// Same as rto_trace_camera on a subset of rows: rows first + j*stride, j < count.
// Output index is j*W + x (dense in the subset).
int rto_trace_camera_rows(const rt_camera* cam, const rt_metric* mt, const rt_integrator* ig,
                          const rt_termination* tm, uint32_t first, uint32_t stride, uint32_t count,
                          rt_results* out, int threads) {
    if (!cam || !mt || !ig || !tm || !out || stride == 0) return RT_ERR_INVALID_ARG;
    const Spacetime st = make_spacetime(mt, tm);
    const Integ in = make_integ(ig);
    const size_t W = cam->width;
    const size_t n = (size_t)count * W;
    parallel_rows(count, threads, [&](size_t j) {
        const uint32_t py = first + (uint32_t)j * stride;
        for (uint32_t px = 0; px < cam->width; ++px) {
            double y[8], xi;
            RayOut o;
            if (py < cam->height && camera_ray(st, cam, px, py, y, &xi)) {
                o = trace_ray(st, in, y, xi);
            } else {
                o.cls = RT_CLS_CAPTURED;
            }
            store(out, n, j * W + px, o);
        }
    });
    return RT_OK;
}

int rto_trace_rays(const double* x_soa, const double* p_soa, size_t n, const rt_metric* mt, const rt_integrator* ig,
                   const rt_termination* tm, rt_results* out, int threads) {
    if (!x_soa || !p_soa || !mt || !ig || !tm || !out) return RT_ERR_INVALID_ARG;
    const Spacetime st = make_spacetime(mt, tm);
    const Integ in = make_integ(ig);
    const size_t chunk = 64;
    parallel_rows((n + chunk - 1) / chunk, threads, [&](size_t c) {
        for (size_t i = c * chunk; i < std::min(n, (c + 1) * chunk); ++i) {
            double y[8];
            for (int j = 0; j < 4; ++j) {
                y[j] = x_soa[j * n + i];
                y[4 + j] = p_soa[j * n + i];
            }
            const double xi = (y[1] * y[6] - y[2] * y[5]) / (-y[4]);
            store(out, n, i, trace_ray(st, in, y, xi));
        }
    });
    return RT_OK;
}

// RT_METHOD_RENDER_RK4 camera frame (single Kerr): the renderer's tracer on
// every pixel of the camera's row window.
int rto_trace_render_camera(const rt_camera* cam, const rt_metric* mt, const rt_integrator* ig,
                            const rt_termination* tm, rt_results* out, int threads) {
    if (!cam || !mt || !ig || !tm || !out || cam->width == 0 || cam->height == 0 || mt->n_sources >= 2)
        return RT_ERR_INVALID_ARG;
    const Spacetime st = make_spacetime(mt, tm);
    const RenderCfg rc = make_render(cam, mt, ig, tm);
    uint32_t first, last;
    row_window(cam, first, last);
    const size_t W = cam->width;
    const size_t n = (size_t)(last - first) * W;
    parallel_rows(last - first, threads, [&](size_t row) {
        const uint32_t py = first + (uint32_t)row;
        for (uint32_t px = 0; px < cam->width; ++px) {
            double y[8], xi;
            camera_ray(st, cam, px, py, y, &xi);
            store(out, n, row * W + px, trace_render(rc, cam->observer[3], y));
        }
    });
    return RT_OK;
}

// The equatorial crossings (KS radius, affine length) of a single-Kerr ray
// traced with the disk switched off, in order along the ray; returns their
// number (at most cap). Test infrastructure for the disk-event checks.
int rto_z_crossings(const double y0[8], const rt_metric* mt, const rt_integrator* ig, const rt_termination* tm,
                    double* r_out, double* lambda_out, int cap) {
    rt_termination off = *tm;
    off.disk_enabled = 0;
    const Spacetime st = make_spacetime(mt, &off);
    const Integ in = make_integ(ig);
    CrossingLog log;
    g_crossings = &log;
    double y[8];
    std::memcpy(y, y0, sizeof(y));
    const double xi = (y[1] * y[6] - y[2] * y[5]) / (-y[4]);
    trace_ray(st, in, y, xi);
    g_crossings = nullptr;
    const int n = std::min(log.n, cap);
    for (int i = 0; i < n; ++i) {
        r_out[i] = log.r[i];
        lambda_out[i] = log.lambda[i];
    }
    return n;
}

// Initial state of one pixel (ok = 0 if the observer is inside a horizon).
int rto_camera_state(const rt_camera* cam, const rt_metric* mt, const rt_termination* tm, uint32_t px, uint32_t py,
                     double y[8], double* xi) {
    const Spacetime st = make_spacetime(mt, tm);
    return camera_ray(st, cam, px, py, y, xi) ? 1 : 0;
}

// Ray direction (forward, up, right components) of screen point (u, v) for a
// projection mode (CameraProjector::compute_ray_direction as dispatch_fp64_scalar calls it).
void rto_screen_direction(uint32_t mode, double u, double v, double fov, double out[3]) {
    screen_direction(mode, u, v, fov, out);
}

// Diagnostic counters: accepted, rejected, projections; reset when reset != 0.
void rto_counters(uint64_t out[3], int reset) {
    out[0] = g_accepted.load();
    out[1] = g_rejected.load();
    out[2] = g_projections.load();
    if (reset) {
        g_accepted = 0;
        g_rejected = 0;
        g_projections = 0;
    }
}

void rto_ks_field(const double x[3], double M, double a, double out[5]) {
    const KSField<double> f = ks_field(x[0], x[1], x[2], M, a);
    out[0] = f.r;
    out[1] = f.H;
    out[2] = f.k0;
    out[3] = f.k1;
    out[4] = f.k2;
}

void rto_rhs(const double y[8], const rt_metric* mt, double d[8]) {
    rt_termination tm{};
    tm.horizon_eps = 1e-6;
    tm.escape_radius = 1000;
    tm.disk_outer = 20;
    const Spacetime st = make_spacetime(mt, &tm);
    rhs(st, y, d);
}

double rto_constraint(const double y[8], const rt_metric* mt) {
    rt_termination tm{};
    tm.horizon_eps = 1e-6;
    tm.escape_radius = 1000;
    const Spacetime st = make_spacetime(mt, &tm);
    return constraint(st, y);
}

int rto_project(double y[8], const rt_metric* mt, double tol) {
    rt_termination tm{};
    tm.horizon_eps = 1e-6;
    tm.escape_radius = 1000;
    const Spacetime st = make_spacetime(mt, &tm);
    return project(st, y, tol) ? 1 : 0;
}

void rto_initial_momentum(const double x[3], double M, double a, const double outward[3], double p[3]) {
    ks_initial_momentum(x, M, a, outward, p);
}

// Superposition source table (build_multi_source_data) as 17 doubles per
// source: centre[3], m, a, r_+, frame x[3], frame y[3], axis[3], disk in, disk
// out (-1, -1 when the source's disk is off). Returns the source count.
int rto_ms_table(const rt_metric* mt, const rt_termination* tm, double* out) {
    const Spacetime st = make_spacetime(mt, tm);
    if (!st.multi) return 0;
    for (int i = 0; i < st.n; ++i) {
        const Source& s = st.s[i];
        double* o = out + 17 * i;
        for (int j = 0; j < 3; ++j) {
            o[j] = s.c[j];
            o[6 + j] = s.fx[j];
            o[9 + j] = s.fy[j];
            o[12 + j] = s.ax[j];
        }
        o[3] = s.m;
        o[4] = s.a;
        o[5] = s.rh;
        o[15] = s.disk_on ? s.disk_in : -1.0;
        o[16] = s.disk_on ? s.disk_out : -1.0;
    }
    return st.n;
}

void rto_ms_initial_momentum(const rt_metric* mt, const double x[3], const double outward[3], double p[3]) {
    rt_termination tm{};
    tm.horizon_eps = 1e-6;
    tm.escape_radius = 1000;
    const Spacetime st = make_spacetime(mt, &tm);
    ms_initial_momentum(st, x, outward, p);
}

void rto_ks_from_bl(double r, double th, double ph, double a, double out[3]) { ks_position_from_bl(r, th, ph, a, out); }
void rto_bl_from_ks(const double x[3], double a, double out[3]) { bl_from_ks(x, a, out); }
double rto_isco(double M, double a) { return kerr_isco(M, a); }
int rto_carter_verdict(double M, double a, double xi, double eta, double margin) {
    return carter_no_turning_point(M, a, horizon_radius(M, a), xi, eta, margin);
}
double rto_redshift(double M, double a, double r, double xi, int clamp) { return disk_redshift(M, a, r, xi, clamp != 0); }

// One DP45 attempt on the geodesic system from y with step h; fills y5, the
// seven stage derivatives k (7x8, row-major) and returns the error norm.
double rto_dp45_attempt(const double y[8], double h, const rt_metric* mt, double rtol, double atol, double y5[8],
                        double kout[56]) {
    rt_termination tm{};
    tm.horizon_eps = 1e-6;
    tm.escape_radius = 1000;
    const Spacetime st = make_spacetime(mt, &tm);
    double k[7][8];
    rhs(st, y, k[0]);
    auto f = [&](const double* yy, double* dd) { rhs(st, yy, dd); };
    const double err = dp45_attempt<double, 8>(f, y, h, rtol, atol, k, y5);
    for (int s = 0; s < 7; ++s)
        for (int i = 0; i < 8; ++i) kout[s * 8 + i] = k[s][i];
    return err;
}

// Fixed-step DP45 on test problems with closed forms, for pinning the tableau
// and the dense output: kind 0: y' = lam*y; kind 1: y' = 5 t^4 (as an autonomous
// 2-system (t, y)); kind 2: harmonic oscillator (x' = v, v' = -x).
// Integrates n_steps of size h from (t0, y0[0..1]); writes y_end[0..1], and the
// dense value at fraction s of the LAST step in dense_out[0..1]; returns max err norm.
double rto_test_ode(int kind, double lam, const double y0in[2], double t0, double h, int n_steps, double s,
                    double y_end[2], double dense_out[2]) {
    constexpr int N = 3;  // (t, y0, y1)
    auto f = [&](const double* yy, double* dd) {
        dd[0] = 1.0;
        if (kind == 0) {
            dd[1] = lam * yy[1];
            dd[2] = lam * yy[2];
        } else if (kind == 1) {
            const double t = yy[0];
            dd[1] = 5.0 * t * t * t * t;
            dd[2] = 4.0 * t * t * t;
        } else {
            dd[1] = yy[2];
            dd[2] = -yy[1];
        }
    };
    double y[N] = {t0, y0in[0], y0in[1]}, y5[N], k[7][N];
    double emax = 0.0;
    for (int i = 0; i < n_steps; ++i) {
        f(y, k[0]);
        const double e = dp45_attempt<double, N>(f, y, h, 1e-300, 1e-300, k, y5);
        emax = max(emax, e);
        if (i == n_steps - 1) {
            dense_out[0] = dense_eval<N>(s, 1, y, y5, h, k);
            dense_out[1] = dense_eval<N>(s, 2, y, y5, h, k);
        }
        for (int j = 0; j < N; ++j) y[j] = y5[j];
    }
    y_end[0] = y[1];
    y_end[1] = y[2];
    return emax;
}

// Illinois root of a polynomial c0 + c1 s + c2 s^2 + c3 s^3 on [0,1].
double rto_test_root(const double c[4]) {
    auto g = [&](double s) { return ((c[3] * s + c[2]) * s + c[1]) * s + c[0]; };
    return illinois_root(g, g(0.0), g(1.0));
}

// Trace a single ray given its initial state (for small-case tests).
void rto_trace_one(const double y0[8], const rt_metric* mt, const rt_integrator* ig, const rt_termination* tm,
                   uint32_t* cls, double y_end[8], double bl[3], double* lambda, double* g, uint32_t* acc,
                   uint32_t* rej) {
    const Spacetime st = make_spacetime(mt, tm);
    const Integ in = make_integ(ig);
    const double xi = (y0[1] * y0[6] - y0[2] * y0[5]) / (-y0[4]);
    const RayOut o = trace_ray(st, in, y0, xi);
    *cls = o.cls;
    for (int i = 0; i < 8; ++i) y_end[i] = o.y[i];
    for (int i = 0; i < 3; ++i) bl[i] = o.bl[i];
    *lambda = o.lambda;
    *g = o.g;
    *acc = o.acc;
    *rej = o.rej;
}

}  // extern "C"

// ======================================================= FLOP counting
// Instantiates the templated RHS and DP45 attempt with a scalar that counts
// + - * / and sqrt (1 each; pow counts 1), to fix the algorithmic FLOP
// constants used for roofline reporting.
namespace rto_count {
static thread_local uint64_t g_ops = 0;
struct Cnt {
    double v;
    Cnt() : v(0) {}
    Cnt(double x) : v(x) {}
};
inline Cnt operator+(Cnt a, Cnt b) { ++g_ops; return Cnt(a.v + b.v); }
inline Cnt operator-(Cnt a, Cnt b) { ++g_ops; return Cnt(a.v - b.v); }
inline Cnt operator*(Cnt a, Cnt b) { ++g_ops; return Cnt(a.v * b.v); }
inline Cnt operator/(Cnt a, Cnt b) { ++g_ops; return Cnt(a.v / b.v); }
inline Cnt operator-(Cnt a) { return Cnt(-a.v); }
inline bool operator<(Cnt a, Cnt b) { return a.v < b.v; }
inline bool operator>(Cnt a, Cnt b) { return a.v > b.v; }
inline bool operator<=(Cnt a, Cnt b) { return a.v <= b.v; }
inline bool operator==(Cnt a, Cnt b) { return a.v == b.v; }
inline Cnt sqrt(Cnt a) { ++g_ops; return Cnt(std::sqrt(a.v)); }
inline Cnt abs(Cnt a) { return Cnt(std::fabs(a.v)); }
inline Cnt pow(Cnt a, Cnt b) { ++g_ops; return Cnt(std::pow(a.v, b.v)); }
inline Cnt max(Cnt a, Cnt b) { return a.v < b.v ? b : a; }
inline Cnt min(Cnt a, Cnt b) { return b.v < a.v ? b : a; }
}  // namespace rto_count

namespace rto {
using rto_count::Cnt;
inline Cnt sqrt(Cnt a) { return rto_count::sqrt(a); }
inline Cnt abs(Cnt a) { return rto_count::abs(a); }
inline Cnt pow(Cnt a, Cnt b) { return rto_count::pow(a, b); }
inline Cnt max(Cnt a, Cnt b) { return rto_count::max(a, b); }
}  // namespace rto

extern "C" void rto_flop_counts(double out[4]) {
    using rto_count::Cnt;
    using rto_count::g_ops;
    Cnt y[8] = {0.0, 700.0, 200.0, 300.0, -1.0, -0.7, -0.2, -0.3}, d[8];
    g_ops = 0;
    rto::ks_rhs<Cnt>(y, Cnt(1.0), Cnt(0.9), d);
    out[0] = static_cast<double>(g_ops);  // one right-hand-side evaluation
    Cnt k[7][8], y5[8];
    for (int i = 0; i < 8; ++i) k[0][i] = d[i];
    uint64_t rhs_ops = 0;
    auto f = [&](const Cnt* yy, Cnt* dd) {
        const uint64_t before = g_ops;
        rto::ks_rhs<Cnt>(yy, Cnt(1.0), Cnt(0.9), dd);
        rhs_ops += g_ops - before;
    };
    g_ops = 0;
    const Cnt err = rto::dp45_attempt<Cnt, 8>(f, y, Cnt(-20.0), Cnt(1e-10), Cnt(1e-12), k, y5);
    (void)rto::dp45_next_step<Cnt>(Cnt(20.0), err, true, Cnt(1e-12), Cnt(1e13));
    out[1] = static_cast<double>(g_ops - rhs_ops);  // stage sums + error norm + controller per attempt
    out[2] = static_cast<double>(rhs_ops);          // six RHS evaluations inside an attempt
    // one source's share of the superposed RHS (ms_rhs): offset from its
    // centre, position and momentum into its frame, its KS terms, the
    // correction and momentum derivative back to the world frame, the sums
    Cnt c[3] = {3.0, -2.0, 1.0}, fr[3][3] = {{0.6, 0.8, 0.0}, {-0.8, 0.6, 0.0}, {0.0, 0.0, 1.0}};
    auto rot = [&](const Cnt v[3], Cnt o[3], bool inverse) {
        for (int i = 0; i < 3; ++i) {
            o[i] = inverse ? v[0] * fr[0][i] + v[1] * fr[1][i] + v[2] * fr[2][i]
                           : v[0] * fr[i][0] + v[1] * fr[i][1] + v[2] * fr[i][2];
        }
    };
    g_ops = 0;
    const Cnt dx[3] = {y[1] - c[0], y[2] - c[1], y[3] - c[2]};
    Cnt lx[3], lp[3], two_hs, corr[3], dp[3], wc[3], wp[3];
    rot(dx, lx, false);
    rot(y + 5, lp, false);
    rto::ks_terms<Cnt>(lx[0], lx[1], lx[2], y[4], lp[0], lp[1], lp[2], Cnt(0.5), Cnt(0.3), two_hs, corr, dp);
    rot(corr, wc, true);
    rot(dp, wp, true);
    Cnt acc[7];
    acc[0] = d[0] + two_hs;
    for (int i = 0; i < 3; ++i) {
        acc[1 + i] = d[1 + i] + wc[i];
        acc[4 + i] = d[5 + i] + wp[i];
    }
    out[3] = static_cast<double>(g_ops);
}

// Integrate y0 over an affine length L (sign of L = direction) with the same
// DP45 steps, projection and controller as trace_ray but no events; the last
// step is shortened to land on L exactly. Single hole only. Returns accepted
// steps (negative if the step size fell below h_min).
extern "C" int rto_integrate_to(const double y0[8], double M, double a, double L, double rtol, double atol,
                                double h0, double y_out[8]) {
    rt_metric mt{};
    mt.mass = M;
    mt.spin = a;
    rt_termination tm{};
    tm.horizon_eps = 1e-6;
    tm.escape_radius = 1e30;
    const rto::Spacetime st = rto::make_spacetime(&mt, &tm);
    double y[8], y5[8], k[7][8];
    std::memcpy(y, y0, sizeof(y));
    auto f = [&](const double* yy, double* dd) { rto::rhs(st, yy, dd); };
    f(y, k[0]);
    const double sgn = (L < 0.0) ? -1.0 : 1.0;
    const double Labs = std::fabs(L);
    double lam = 0.0, ah = h0;
    int acc = 0;
    while (lam < Labs) {
        if (ah < 1e-14) return -acc;
        const bool last = (lam + ah >= Labs);
        const double step = last ? (Labs - lam) : ah;
        const double err = rto::dp45_attempt<double, 8>(f, y, sgn * step, rtol, atol, k, y5);
        if (err <= 1.0) {
            std::memcpy(y, y5, sizeof(y));
            ++acc;
            lam = last ? Labs : lam + step;
            if (rto::project(st, y, 1e-12))
                f(y, k[0]);
            else
                std::memcpy(k[0], k[6], sizeof(k[0]));
            if (!last) ah = rto::dp45_next_step(step, err, true, 1e-14, 1e13);
        } else {
            ah = rto::dp45_next_step(step, err, false, 1e-14, 1e13);
        }
    }
    std::memcpy(y_out, y, sizeof(y));
    return acc;
}
