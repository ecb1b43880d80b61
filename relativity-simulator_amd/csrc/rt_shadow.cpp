// rt_shadow.cpp — the Kerr critical curve (the shadow's edge) and each
// camera pixel's distance to it, on the host.
//
// A null ray of a Kerr hole has the constants xi = L_z / E and
// eta = Q / E^2. Seen from inclination theta_o, its impact parameters are
// alpha = -xi / sin theta_o and beta = +-sqrt(eta + a^2 cos^2 theta_o -
// xi^2 cot^2 theta_o). The rays that stay on spherical photon orbits,
// r in [r_ph1, r_ph2], have
//   xi(r)  = (-r^3 + 3 M r^2 - a^2 r - a^2 M) / (a (r - M)),
//   eta(r) = r^3 (4 M a^2 - r (r - 3M)^2) / (a^2 (r - M)^2),
// and their (alpha, beta) trace the critical curve: rays inside it are
// captured, rays outside escape. The curve follows BardeenKerrShadow
// (metrics/bardeen_shadow.hpp:38-162): photon-orbit radii, xi(r), eta(r),
// the visible arc where beta^2 >= 0, and the circle of radius sqrt(27) M for
// a = 0 or a polar observer. The distance of a ray to the curve, in the
// (alpha, beta) plane and in units of M, decides which rays lie "within
// 1e-6 of the critical curve", where the class of a ray may depend on the
// rounding of the integration.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <thread>
#include <vector>

#include "../../include/rt_tracer.h"
#include "rt_camera.hpp"
#include "rt_host.h"
#include "rt_params.h"

namespace rt {
namespace {

constexpr double kPiS = 3.14159265358979323846;

struct CriticalCurve {
    double m, a, sin_t, cos_t, cot_t;
    bool circle = false;
    double r0 = 0.0, r1 = 0.0;  // visible arc of photon-orbit radii

    void xi_eta(double r, double& xi, double& eta) const {
        const double u = r - m;
        const double us = (std::abs(u) < 1e-15) ? ((u >= 0.0) ? 1e-15 : -1e-15) : u;
        xi = (-r * r * r + 3.0 * m * r * r - a * a * r - a * a * m) / (a * us);
        const double bracket = 4.0 * m * a * a - r * (r - 3.0 * m) * (r - 3.0 * m);
        eta = (r * r * r * bracket) / (a * a * us * us);
    }
    double beta_sq(double r) const {
        double xi, eta;
        xi_eta(r, xi, eta);
        return eta + a * a * cos_t * cos_t - xi * xi * cot_t * cot_t;
    }
    // (alpha, beta >= 0) of the upper branch at photon-orbit radius r
    void point(double r, double& al, double& be) const {
        double xi, eta;
        xi_eta(r, xi, eta);
        al = -xi / sin_t;
        be = std::sqrt(std::max(eta + a * a * cos_t * cos_t - xi * xi * cot_t * cot_t, 0.0));
    }

    CriticalCurve(double mass, double spin, double theta_o) : m(mass), a(spin) {
        sin_t = std::sin(theta_o);
        cos_t = std::cos(theta_o);
        cot_t = cos_t / sin_t;
        if (std::abs(a) < 1e-12 || std::abs(sin_t) < 1e-12) {
            circle = true;
            return;
        }
        const double am = std::clamp(a / m, -0.999999999, 0.999999999);
        const double ra = 2.0 * m * (1.0 + std::cos(2.0 / 3.0 * std::acos(-am)));
        const double rb = 2.0 * m * (1.0 + std::cos(2.0 / 3.0 * std::acos(am)));
        r0 = std::min(ra, rb);
        r1 = std::max(ra, rb);
        if (beta_sq(r0) < 0.0 || beta_sq(r1) < 0.0) {
            // only an arc of the orbits reaches theta_o: its ends where beta^2 = 0
            double peak = 0.5 * (r0 + r1), best = -1e300;
            for (int k = 0; k <= 400; ++k) {
                const double r = r0 + (r1 - r0) * k / 400.0;
                const double b = beta_sq(r);
                if (b > best) {
                    best = b;
                    peak = r;
                }
            }
            if (best > 0.0) {
                double lo = r0, hi = peak;
                for (int i = 0; i < 80; ++i) {
                    const double mid = 0.5 * (lo + hi);
                    (beta_sq(mid) >= 0.0 ? hi : lo) = mid;
                }
                const double start = hi;
                lo = peak;
                hi = r1;
                for (int i = 0; i < 80; ++i) {
                    const double mid = 0.5 * (lo + hi);
                    (beta_sq(mid) >= 0.0 ? lo : hi) = mid;
                }
                r0 = start;
                r1 = lo;
            }
        }
    }

    // Distance from (alpha, beta) to the curve (both branches: beta is taken
    // by its magnitude, the curve being symmetric).
    double distance(double al, double be) const {
        if (circle) return std::abs(std::hypot(al, be) - std::sqrt(27.0) * m);
        be = std::abs(be);
        auto d2 = [&](double r) {
            double ca, cb;
            point(r, ca, cb);
            return (al - ca) * (al - ca) + (be - cb) * (be - cb);
        };
        constexpr int kSamples = 96;
        int best_k = 0;
        double best = 1e300;
        for (int k = 0; k <= kSamples; ++k) {
            const double v = d2(r0 + (r1 - r0) * k / kSamples);
            if (v < best) {
                best = v;
                best_k = k;
            }
        }
        // golden-section refinement on the bracket around the best sample
        double lo = r0 + (r1 - r0) * std::max(best_k - 1, 0) / kSamples;
        double hi = r0 + (r1 - r0) * std::min(best_k + 1, kSamples) / kSamples;
        const double g = 0.6180339887498949;
        double c = hi - g * (hi - lo), d = lo + g * (hi - lo);
        double fc = d2(c), fd = d2(d);
        for (int i = 0; i < 100 && hi - lo > 1e-15 * std::max(1.0, hi); ++i) {
            if (fc < fd) {
                hi = d;
                d = c;
                fd = fc;
                c = hi - g * (hi - lo);
                fc = d2(c);
            } else {
                lo = c;
                c = d;
                fc = fd;
                d = lo + g * (hi - lo);
                fd = d2(d);
            }
        }
        best = std::min({best, fc, fd, d2(lo), d2(hi)});
        return std::sqrt(best);
    }
};

// The ray's Carter constants from its KS null state y (E = -p_t):
// xi = L_z / E, eta = Q / E^2 with Q = p_theta^2 + cos^2 theta (L_z^2 / sin^2 theta - a^2 E^2).
bool ray_constants(const double y[8], double a, double& xi, double& eta, double& p_theta) {
    const double x = y[1], yy = y[2], z = y[3];
    const double a2 = a * a;
    const double diff = x * x + yy * yy + z * z - a2;
    const double r = std::sqrt(std::max(0.5 * (diff + std::sqrt(diff * diff + 4.0 * a2 * z * z)), 1e-24));
    const double E = -y[4];
    const double cth = z / r;
    const double s2 = 1.0 - cth * cth;
    if (!(s2 > 1e-24) || !(E > 0.0)) return false;
    const double sth = std::sqrt(s2);
    const double lz = x * y[6] - yy * y[5];
    p_theta = (cth / sth) * (x * y[5] + yy * y[6]) - r * sth * y[7];
    const double q = p_theta * p_theta + cth * cth * (lz * lz / s2 - a2 * E * E);
    xi = lz / E;
    eta = q / (E * E);
    return true;
}

}  // namespace
}  // namespace rt

using namespace rt;

extern "C" rt_status rt_critical_distance(const rt_camera* cam, const rt_metric* metric, double* out) {
    if (!cam || !metric || !out) return fail(RT_ERR_INVALID_ARG, "null argument");
    std::string why;
    if (!valid_camera(cam, why)) return fail(RT_ERR_INVALID_ARG, why);
    if (metric->n_sources >= 2) return fail(RT_ERR_INVALID_ARG, "the critical curve is a single hole's");
    if (!(metric->mass > 0.0)) return fail(RT_ERR_INVALID_ARG, "metric: mass must be > 0");
    rt_integrator ig;
    rt_termination tm;
    rt_default_integrator(&ig);
    rt_default_termination(&tm);
    TraceParams P;
    fill_params(metric, &ig, &tm, P);
    fill_camera(cam, P, nullptr);
    const double theta_o = std::clamp(cam->observer[2], 0.001, kPiS - 0.001);
    const CriticalCurve curve(P.M, P.a, theta_o);
    const uint32_t rows = P.n_rays / P.cols;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const uint32_t n_threads = std::min<uint32_t>(hw, std::max<uint32_t>(1, rows));
    auto work = [&](uint32_t t) {
        for (uint32_t row = t; row < rows; row += n_threads)
            for (uint32_t col = 0; col < P.cols; ++col) {
                double y[8], lz;
                camera_state_kerr(P, P.col_first + col, P.row_first + row, y, &lz);
                double xi, eta, pth;
                double d = 0.0;
                if (ray_constants(y, P.a, xi, eta, pth)) {
                    const double st = std::sin(theta_o), ct = std::cos(theta_o);
                    const double b2 = eta + P.a2 * ct * ct - xi * xi * (ct / st) * (ct / st);
                    d = curve.distance(-xi / st, std::sqrt(std::max(b2, 0.0)));
                }
                out[static_cast<size_t>(row) * P.cols + col] = d;
            }
    };
    std::vector<std::thread> pool;
    for (uint32_t t = 1; t < n_threads; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto& th : pool) th.join();
    return RT_OK;
}

// n points of the critical curve: the upper branch (beta >= 0) over the
// visible arc of photon-orbit radii, then the lower branch back (as
// BardeenKerrShadow::compute_boundary_points orders them); the circle of
// radius sqrt(27) M for a = 0 or a polar observer.
extern "C" rt_status rt_shadow_boundary(double mass, double spin, double theta_obs, size_t n, double* alpha,
                                        double* beta) {
    if (!alpha || !beta || n < 4 || !(mass > 0.0) || !std::isfinite(spin) || !std::isfinite(theta_obs))
        return fail(RT_ERR_INVALID_ARG, "rt_shadow_boundary: need mass > 0, finite spin and angle, n >= 4");
    const CriticalCurve c(mass, spin, theta_obs);
    if (c.circle) {
        const double rs = std::sqrt(27.0) * mass;
        for (size_t i = 0; i < n; ++i) {
            const double ph = 2.0 * kPiS * static_cast<double>(i) / static_cast<double>(n);
            alpha[i] = rs * std::cos(ph);
            beta[i] = rs * std::sin(ph);
        }
        return RT_OK;
    }
    const size_t half = n / 2;
    for (size_t i = 0; i < half; ++i) {
        const double r = c.r0 + (c.r1 - c.r0) * static_cast<double>(i) / static_cast<double>(half - 1);
        c.point(r, alpha[i], beta[i]);
        alpha[n - 1 - i] = alpha[i];
        beta[n - 1 - i] = -beta[i];
    }
    if (n % 2) {  // odd n: the middle point closes the curve on the far end
        alpha[half] = alpha[half - 1];
        beta[half] = 0.0;
    }
    return RT_OK;
}

// Distance from the point (alpha, beta) of the impact-parameter plane to the
// critical curve, in units of M (negative on error).
extern "C" double rt_impact_distance(double mass, double spin, double theta_obs, double alpha, double beta) {
    if (!(mass > 0.0) || !std::isfinite(spin) || !std::isfinite(theta_obs)) return -1.0;
    return CriticalCurve(mass, spin, theta_obs).distance(alpha, beta);
}
