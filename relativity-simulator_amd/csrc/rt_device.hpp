// rt_device.hpp — device-side building blocks of the null-geodesic tracer
// (CDNA4 / gfx950). One lane owns one ray; the whole DP45 step, the
// Hamiltonian right-hand side and the event location live in registers.
//
// The mathematics follows the reference (paths relative to the
// Relativity-Simulator tree):
//   KS field / Hamiltonian RHS   core/kerr_schild_null_integrator.hpp:30-114
//   DP45 tableau, error, control integrators/rk45_adaptive.hpp:196-311,
//                                integrators/step_controller.hpp:44-78
//   constraint projection        integrators/rk45_adaptive.hpp:91-139 (for p_t)
//   disk redshift                render/accretion_disk/accretion_disk_model.hpp:205-225
//   camera / initial momentum    render/software_compute_engine.hpp:3843-3938, 1634-1652,
//                                core/kerr_schild_null_integrator.hpp:141-194
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/rt_tracer.h"
#include "rt_params.h"

namespace rt {

// ------------------------------------------------------------- tableau
namespace dp {
constexpr double a21 = 1.0 / 5.0;
constexpr double a31 = 3.0 / 40.0, a32 = 9.0 / 40.0;
constexpr double a41 = 44.0 / 45.0, a42 = -56.0 / 15.0, a43 = 32.0 / 9.0;
constexpr double a51 = 19372.0 / 6561.0, a52 = -25360.0 / 2187.0, a53 = 64448.0 / 6561.0, a54 = -212.0 / 729.0;
constexpr double a61 = 9017.0 / 3168.0, a62 = -355.0 / 33.0, a63 = 46732.0 / 5247.0, a64 = 49.0 / 176.0,
                 a65 = -5103.0 / 18656.0;
constexpr double a71 = 35.0 / 384.0, a73 = 500.0 / 1113.0, a74 = 125.0 / 192.0, a75 = -2187.0 / 6784.0,
                 a76 = 11.0 / 84.0;
constexpr double e1 = 71.0 / 57600.0, e3 = -71.0 / 16695.0, e4 = 71.0 / 1920.0, e5 = -17253.0 / 339200.0,
                 e6 = 22.0 / 525.0, e7 = -1.0 / 40.0;
constexpr double d1 = -12715105075.0 / 11282082432.0, d3 = 87487479700.0 / 32700410799.0,
                 d4 = -10690763975.0 / 1880347072.0, d5 = 701980252875.0 / 199316789632.0,
                 d6 = -1453857185.0 / 822651844.0, d7 = 69997945.0 / 29380423.0;
}  // namespace dp

constexpr double kPi = 3.14159265358979323846;

struct Field {
    double r, H, k0, k1, k2;
};

// KS radius: positive root of r^4 - (R^2 - a^2) r^2 - a^2 z^2 = 0, floored.
__device__ __forceinline__ double ks_radius(double x, double y, double z, double a2) {
    const double z2 = z * z;
    const double diff = x * x + y * y + z2 - a2;
    const double r2 = fmax(0.5 * (diff + sqrt(diff * diff + 4.0 * a2 * z2)), 1e-24);
    return sqrt(r2);
}

__device__ __forceinline__ Field ks_field(double x, double y, double z, double M, double a, double a2) {
    const double z2 = z * z;
    const double diff = x * x + y * y + z2 - a2;
    const double r2 = fmax(0.5 * (diff + sqrt(diff * diff + 4.0 * a2 * z2)), 1e-24);
    const double r = sqrt(r2);
    const double den = r2 * r2 + a2 * z2;
    const double rpa = r2 + a2;
    Field f;
    f.r = r;
    f.H = M * r2 * r / den;
    f.k0 = (r * x + a * y) / rpa;
    f.k1 = (r * y - a * x) / rpa;
    f.k2 = z / r;
    return f;
}

// Hamiltonian right-hand side of one source in its own frame:
// two_hs = 2 H s, corr = -2 H s k, dp = s^2 grad H + 2 H s grad(k.p), s = k.p - p_t.
__device__ __forceinline__ Field ks_terms(double X, double Y, double Z, double pt, double px, double py, double pz,
                                          double M, double a, double a2, double& two_hs, double corr[3],
                                          double dp[3]) {
    const double z2 = Z * Z;
    const double diff = X * X + Y * Y + z2 - a2;
    const double r2 = fmax(0.5 * (diff + sqrt(diff * diff + 4.0 * a2 * z2)), 1e-24);
    const double r = sqrt(r2);
    const double r3 = r2 * r;
    const double r4 = r2 * r2;
    const double den = r4 + a2 * z2;
    const double rpa = r2 + a2;
    // three reciprocals replace the reference's eleven divisions
    const double inv_den = 1.0 / den;
    const double inv_rpa = 1.0 / rpa;
    const double inv_r = 1.0 / r;
    const double H = M * r3 * inv_den;
    const double k0 = (r * X + a * Y) * inv_rpa;
    const double k1 = (r * Y - a * X) * inv_rpa;
    const double k2 = Z * inv_r;
    const double s = k0 * px + k1 * py + k2 * pz - pt;
    two_hs = 2.0 * H * s;
    corr[0] = -(two_hs * k0);
    corr[1] = -(two_hs * k1);
    corr[2] = -(two_hs * k2);

    const double inv_den2 = inv_den * inv_den;
    const double inv_rpa2 = inv_rpa * inv_rpa;
    const double inv_r2 = inv_r * inv_r;
    const double r3d = r3 * inv_den;
    const double dr0 = r3d * X, dr1 = r3d * Y, dr2 = r * rpa * Z * inv_den;
    const double bracket = 3.0 * a2 * z2 - r4;
    const double nx = r * X + a * Y;
    const double ny = r * Y - a * X;
    const double hb = M * r2 * bracket * inv_den2;  // dH = dr_i * hb (+ z term)
    const double ss = s * s;
    const double two_h_s = 2.0 * H * s;
    const double two_r = 2.0 * r;
    // grad(k.p) per component, written out (ex/ey/ez unit terms folded in)
    const double kx0 = ((dr0 * X + r) * rpa - nx * two_r * dr0) * inv_rpa2;
    const double ky0 = ((dr0 * Y - a) * rpa - ny * two_r * dr0) * inv_rpa2;
    const double kz0 = (-Z * dr0) * inv_r2;
    const double kx1 = ((dr1 * X + a) * rpa - nx * two_r * dr1) * inv_rpa2;
    const double ky1 = ((dr1 * Y + r) * rpa - ny * two_r * dr1) * inv_rpa2;
    const double kz1 = (-Z * dr1) * inv_r2;
    const double kx2 = ((dr2 * X) * rpa - nx * two_r * dr2) * inv_rpa2;
    const double ky2 = ((dr2 * Y) * rpa - ny * two_r * dr2) * inv_rpa2;
    const double kz2 = (r - Z * dr2) * inv_r2;
    const double dh0 = dr0 * hb;
    const double dh1 = dr1 * hb;
    const double dh2 = dr2 * hb - 2.0 * M * a2 * r3 * Z * inv_den2;
    dp[0] = dh0 * ss + two_h_s * (kx0 * px + ky0 * py + kz0 * pz);
    dp[1] = dh1 * ss + two_h_s * (kx1 * px + ky1 * py + kz1 * pz);
    dp[2] = dh2 * ss + two_h_s * (kx2 * px + ky2 * py + kz2 * pz);
    Field f;
    f.r = r;
    f.H = H;
    f.k0 = k0;
    f.k1 = k1;
    f.k2 = k2;
    return f;
}

// Single-hole RHS for y = (t, x, y, z, p_t, p_x, p_y, p_z); d[4] = 0.
__device__ __forceinline__ Field rhs_kerr(const double y[8], double M, double a, double a2, double d[8]) {
    double two_hs, corr[3], dpp[3];
    const Field f = ks_terms(y[1], y[2], y[3], y[4], y[5], y[6], y[7], M, a, a2, two_hs, corr, dpp);
    d[0] = -y[4] + two_hs;
    d[1] = y[5] + corr[0];
    d[2] = y[6] + corr[1];
    d[3] = y[7] + corr[2];
    d[4] = 0.0;
    d[5] = dpp[0];
    d[6] = dpp[1];
    d[7] = dpp[2];
    return f;
}

// Null constraint g^{mu nu} p p = A p_t^2 + B p_t + C0 from the field at the
// current position; projection picks the root for p_t closest to the current
// value when |residual| >= tol * max(1, |p|^2). Returns true if it fired.
__device__ __forceinline__ bool project_pt(double y[8], double H, double q, double tol) {
    const double P2 = y[5] * y[5] + y[6] * y[6] + y[7] * y[7];
    const double A = -(1.0 + 2.0 * H);
    const double B = 4.0 * (H * q);
    const double C0 = P2 - 2.0 * (H * q * q);
    const double res = (A * y[4] + B) * y[4] + C0;
    const double scale = fmax(1.0, y[4] * y[4] + P2);
    if (!(fabs(res) >= tol * scale)) return false;
    const double disc = B * B - 4.0 * A * C0;
    if (disc < 0.0) return false;
    const double sd = sqrt(disc);
    const double r1 = (-B + sd) / (2.0 * A);
    const double r2 = (-B - sd) / (2.0 * A);
    y[4] = (fabs(r1 - y[4]) <= fabs(r2 - y[4])) ? r1 : r2;
    return true;
}

// Prograde ISCO (Bardeen–Press–Teukolsky), host and device.
__host__ __device__ inline double kerr_isco(double m, double a_spin) {
    double a_star = a_spin / (m > 1e-12 ? m : 1e-12);
    a_star = a_star < -0.9999999 ? -0.9999999 : (a_star > 0.9999999 ? 0.9999999 : a_star);
    const double abs_a = fabs(a_star);
    const double orbit_sign = (a_star >= 0.0) ? -1.0 : 1.0;
    const double cbrt_term = cbrt(fmax(1.0 - abs_a * abs_a, 0.0));
    const double z1 = 1.0 + cbrt_term * (cbrt(1.0 + abs_a) + cbrt(1.0 - abs_a));
    const double z2 = sqrt(3.0 * abs_a * abs_a + z1 * z1);
    return (3.0 + z2 + orbit_sign * sqrt(fmax((3.0 - z1) * (3.0 + z1 + 2.0 * z2), 0.0))) * m;
}

// Circular-orbit disk redshift g for a photon with xi = L_z/E.
__device__ __forceinline__ double disk_redshift(double mass, double spin, double radius, double xi, bool clamp_g) {
    const double spin_mag = fabs(spin);
    const double sqrt_m = sqrt(fmax(mass, 1e-12));
    const double sqrt_r = sqrt(fmax(radius, 1e-12));
    const double r15 = radius * sqrt_r;
    const double ang = sqrt_m / (r15 + spin_mag * sqrt_m);
    const double bracket = fmax(r15 - 3.0 * mass * sqrt_r + 2.0 * spin_mag * sqrt_m, 1e-9);
    const double omega = (spin < 0.0) ? -ang : ang;
    const double ut = (r15 + spin_mag * sqrt_m) / (sqrt(r15) * sqrt(bracket));
    const double den = ut * (1.0 - omega * xi);
    if (!(fabs(den) > 1e-9)) return 1.0;
    const double g = 1.0 / den;
    return clamp_g ? fmin(fmax(g, 0.05), 4.0) : g;
}

// Standard integral controller with safety 0.9; returns the new |h|.
__device__ __forceinline__ double next_step(double ah, double err, bool accepted, double h_min, double h_max) {
    if (accepted) {
        const double factor = (err == 0.0) ? 5.0 : pow(1.0 / fmax(err, 1e-12), 0.2);
        const double s = fmin(fmax(0.9 * factor, 0.2), 5.0);
        return fmin(fmax(ah * s, h_min), h_max);
    }
    const double factor = pow(1.0 / fmax(err, 1e-12), 0.25);
    return ah * fmin(fmax(0.9 * factor, 0.1), 0.9);
}

// DP45 dense output (order-4 continuous extension) of one component.
struct Dense {
    double D, B, C, E;
};
__device__ __forceinline__ double dense_at(const Dense& q, double y0, double s) {
    const double s1 = 1.0 - s;
    return y0 + s * (q.D + s1 * (q.B + s * (q.C + s1 * q.E)));
}

// Illinois root on [0,1] (same iteration as the CPU oracle).
template <class G>
__device__ __forceinline__ double illinois_root(G&& g, double g0, double g1) {
    if (g1 == 0.0) return 1.0;
    double a = 0.0, fa = g0, b = 1.0, fb = g1;
    double c = 1.0, c_prev = -1.0;
    for (int it = 0; it < 64; ++it) {
        c = (a * fb - b * fa) / (fb - fa);
        if (!(c > 0.0)) c = 0.0;
        if (c > 1.0) c = 1.0;
        if (fabs(c - c_prev) <= 1e-15) break;
        const double fc = g(c);
        if (fc == 0.0) break;
        if ((fc < 0.0) != (fb < 0.0)) {
            a = b;
            fa = fb;
        } else {
            fa *= 0.5;
        }
        b = c;
        fb = fc;
        c_prev = c;
    }
    return c;
}

// Spheroidal (r, theta, phi) of a KS point, phi wrapped to [-pi, pi).
__device__ __forceinline__ void bl_from_ks(double x, double y, double z, double a, double a2, double out[3]) {
    const double r = ks_radius(x, y, z, a2);
    const double rho = sqrt(fmax(r * r - z * z, 0.0));
    const double two_pi = 2.0 * kPi;
    double d = atan2(y, x) - atan2(a, r);
    d -= two_pi * floor(d / two_pi + 0.5);
    out[0] = r;
    out[1] = atan2(rho, z);
    out[2] = d;
}

// Wave-wide sum of a 64-bit counter (wave64).
__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

}  // namespace rt
