// rt_oracle.hpp — FP64 CPU oracle of the Kerr null-geodesic tracer.
//
// TEST INFRASTRUCTURE ONLY. This is a clean-room restatement of the reference
// algorithm, used by tests/, by __graft_entry__.smoke() and by the cpu_baseline
// leg of bench.py as the checker and the CPU baseline. The product
// (librtsim.so) never links, loads or calls it.
//
// Every function cites the reference file:line it restates (paths are relative
// to the Relativity-Simulator tree). Templated on the scalar type so the same
// code can be instantiated with an operation-counting type to fix the
// algorithmic FLOP count per right-hand-side evaluation and per DP45 attempt.
//
// State of one ray: y = (t, x, y, z, p_t, p_x, p_y, p_z) in Kerr–Schild
// Cartesian coordinates, covariant momentum, future-directed (p_t = -E < 0);
// the ray is traced backward with negative affine steps, as
// trace_kerr_schild_photon does (software_compute_engine.hpp:1677-1683).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>

#include "../include/rt_tracer.h"

namespace rto {

using std::abs;
using std::atan2;
using std::cbrt;
using std::floor;
using std::max;
using std::min;
using std::pow;
using std::sqrt;

constexpr double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------- KS field
// Restates kerr_schild_field, core/kerr_schild_null_integrator.hpp:30-44
// (same expression order, so FP64 results are bit-identical).
template <class S>
struct KSField {
    S r, H, k0, k1, k2;
};

template <class S>
inline KSField<S> ks_field(S x, S y, S z, S M, S a) {
    const S a2 = a * a;
    const S z2 = z * z;
    const S diff = x * x + y * y + z2 - a2;
    const S r2 = max(S(0.5) * (diff + sqrt(diff * diff + S(4) * a2 * z2)), S(1e-24));
    const S r = sqrt(r2);
    const S den = r2 * r2 + a2 * z2;
    const S rpa = r2 + a2;
    KSField<S> f;
    f.r = r;
    f.H = M * r2 * r / den;
    f.k0 = (r * x + a * y) / rpa;
    f.k1 = (r * y - a * x) / rpa;
    f.k2 = z / r;
    return f;
}

// KS radius only (kerr_schild_field with mass 0, as boyer_lindquist_from_kerr_schild
// does at core/kerr_schild_null_integrator.hpp:58).
template <class S>
inline S ks_radius(S x, S y, S z, S a) {
    const S a2 = a * a;
    const S z2 = z * z;
    const S diff = x * x + y * y + z2 - a2;
    const S r2 = max(S(0.5) * (diff + sqrt(diff * diff + S(4) * a2 * z2)), S(1e-24));
    return sqrt(r2);
}

// ------------------------------------------------------ Hamiltonian RHS
// Restates kerr_schild_null_derivatives, core/kerr_schild_null_integrator.hpp:71-114
// (and its twin ks_source_terms, render/software_compute_engine.hpp:2812-2855),
// generalised from p_t = -1 to a carried p_t:
//   Ham = 1/2 (-p_t^2 + |p|^2 - 2 H s^2),  s = k.p - p_t
//   dt/dl = -p_t + 2 H s,  dx/dl = p - 2 H s k,
//   dp/dl = s^2 grad H + 2 H s grad(k.p),  dp_t/dl = 0.
// With p_t = -1, -p_t is exactly 1 and every term is the reference's
// expression evaluated in the reference's order (bit-identical in FP64).
// ks_terms returns the source's share: two_hs = 2 H s, corr = -2 H s k and dp.
template <class S>
inline KSField<S> ks_terms(S X, S Y, S Z, S pt, S px, S py, S pz, S M, S a, S& two_hs, S corr[3], S dp[3]) {
    const KSField<S> f = ks_field(X, Y, Z, M, a);
    const S two = S(2);
    const S r = f.r;
    const S r2 = r * r;
    const S r3 = r2 * r;
    const S r4 = r2 * r2;
    const S a2 = a * a;
    const S z2 = Z * Z;
    const S den = r4 + a2 * z2;
    const S rpa = r2 + a2;
    const S inv_rpa2 = S(1) / (rpa * rpa);
    const S inv_den2 = S(1) / (den * den);
    const S s = f.k0 * px + f.k1 * py + f.k2 * pz - pt;

    two_hs = two * f.H * s;
    corr[0] = -(two_hs * f.k0);
    corr[1] = -(two_hs * f.k1);
    corr[2] = -(two_hs * f.k2);

    const S dr[3] = {r3 * X / den, r3 * Y / den, r * rpa * Z / den};
    const S bracket = S(3) * a2 * z2 - r4;
    const S nx = r * X + a * Y;
    const S ny = r * Y - a * X;
    // d/dx_i of H and k. The reference multiplies unit-vector components
    // (r*ex + a*ey, r*ez); they are written out per component here, which
    // leaves every non-zero term and its rounding unchanged.
    const S dh0 = M * dr[0] * r2 * bracket * inv_den2;
    const S dh1 = M * dr[1] * r2 * bracket * inv_den2;
    const S dh2 = M * dr[2] * r2 * bracket * inv_den2 - two * M * a2 * r3 * Z * inv_den2;
    const S dk1_0 = ((dr[0] * X + r) * rpa - nx * two * r * dr[0]) * inv_rpa2;
    const S dk2_0 = ((dr[0] * Y - a) * rpa - ny * two * r * dr[0]) * inv_rpa2;
    const S dk3_0 = (-(Z * dr[0])) / r2;
    const S dk1_1 = ((dr[1] * X + a) * rpa - nx * two * r * dr[1]) * inv_rpa2;
    const S dk2_1 = ((dr[1] * Y + r) * rpa - ny * two * r * dr[1]) * inv_rpa2;
    const S dk3_1 = (-(Z * dr[1])) / r2;
    const S dk1_2 = ((dr[2] * X) * rpa - nx * two * r * dr[2]) * inv_rpa2;
    const S dk2_2 = ((dr[2] * Y) * rpa - ny * two * r * dr[2]) * inv_rpa2;
    const S dk3_2 = (r - Z * dr[2]) / r2;
    dp[0] = dh0 * s * s + two * f.H * s * (dk1_0 * px + dk2_0 * py + dk3_0 * pz);
    dp[1] = dh1 * s * s + two * f.H * s * (dk1_1 * px + dk2_1 * py + dk3_1 * pz);
    dp[2] = dh2 * s * s + two * f.H * s * (dk1_2 * px + dk2_2 * py + dk3_2 * pz);
    return f;
}

// Full single-hole right-hand side d = dy/dl for y = (t, x, y, z, p_t, p_x, p_y, p_z).
template <class S>
inline KSField<S> ks_rhs(const S y[8], S M, S a, S d[8]) {
    S two_hs, corr[3], dp[3];
    const KSField<S> f = ks_terms(y[1], y[2], y[3], y[4], y[5], y[6], y[7], M, a, two_hs, corr, dp);
    d[0] = -y[4] + two_hs;
    d[1] = y[5] + corr[0];
    d[2] = y[6] + corr[1];
    d[3] = y[7] + corr[2];
    d[4] = S(0);
    d[5] = dp[0];
    d[6] = dp[1];
    d[7] = dp[2];
    return f;
}

// Twice the Hamiltonian, g^{mu nu} p_mu p_nu, for a single Kerr hole.
template <class S>
inline S ks_constraint(const S y[8], S M, S a) {
    const KSField<S> f = ks_field(y[1], y[2], y[3], M, a);
    const S s = f.k0 * y[5] + f.k1 * y[6] + f.k2 * y[7] - y[4];
    return -y[4] * y[4] + y[5] * y[5] + y[6] * y[6] + y[7] * y[7] - S(2) * f.H * s * s;
}

// ------------------------------------------------- BL <-> KS conversions
// Restates kerr_schild_position_from_boyer_lindquist, core/kerr_schild_null_integrator.hpp:46-54.
inline void ks_position_from_bl(double r, double th, double ph, double a, double out[3]) {
    const double st = std::sin(th);
    out[0] = (r * std::cos(ph) - a * std::sin(ph)) * st;
    out[1] = (r * std::sin(ph) + a * std::cos(ph)) * st;
    out[2] = r * std::cos(th);
}

// Restates boyer_lindquist_from_kerr_schild, core/kerr_schild_null_integrator.hpp:56-69,
// with previous_phi = 0, so phi is wrapped into [-pi, pi).
inline void bl_from_ks(const double x[3], double a, double out[3]) {
    const double r = ks_radius(x[0], x[1], x[2], a);
    const double rho = sqrt(max(r * r - x[2] * x[2], 0.0));
    const double two_pi = 2.0 * kPi;
    const double raw_phi = atan2(x[1], x[0]) - atan2(a, r);
    double delta = raw_phi;
    delta -= two_pi * floor(delta / two_pi + 0.5);
    out[0] = r;
    out[1] = atan2(rho, x[2]);
    out[2] = delta;
}

// --------------------------------------------------- initial momentum
// Restates kerr_schild_initial_momentum, core/kerr_schild_null_integrator.hpp:141-194:
// Gram–Schmidt frame of the KS metric at x, photon direction -outward in
// that frame, normalised to p_t = -1.
inline void ks_initial_momentum(const double x[3], double M, double a, const double outward[3], double p_out[3]) {
    const KSField<double> f = ks_field(x[0], x[1], x[2], M, a);
    const double h = f.H;
    const double k[3] = {f.k0, f.k1, f.k2};
    const double two_h = 2.0 * h;
    auto inner = [&](const double* A, const double* B) {
        const double la = A[0] + k[0] * A[1] + k[1] * A[2] + k[2] * A[3];
        const double lb = B[0] + k[0] * B[1] + k[1] * B[2] + k[2] * B[3];
        return -A[0] * B[0] + A[1] * B[1] + A[2] * B[2] + A[3] * B[3] + two_h * la * lb;
    };
    const double u0 = 1.0 / sqrt(max(1.0 - two_h, 1e-6));
    double e[3][4];
    for (int i = 0; i < 3; ++i) {
        e[i][0] = two_h * k[i] * u0 * u0;
        e[i][1] = e[i][2] = e[i][3] = 0.0;
        e[i][1 + i] = 1.0;
    }
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < i; ++j) {
            const double proj = inner(e[i], e[j]);
            for (int c = 0; c < 4; ++c) e[i][c] -= proj * e[j][c];
        }
        const double inv_norm = 1.0 / sqrt(max(inner(e[i], e[i]), 1e-30));
        for (int c = 0; c < 4; ++c) e[i][c] *= inv_norm;
    }
    double v[4] = {u0, 0.0, 0.0, 0.0};
    for (int i = 0; i < 3; ++i)
        for (int c = 0; c < 4; ++c) v[c] -= outward[i] * e[i][c];
    const double lv = v[0] + k[0] * v[1] + k[1] * v[2] + k[2] * v[3];
    const double p_t = -v[0] + two_h * lv;
    const double inv_e = 1.0 / max(-p_t, 1e-12);
    p_out[0] = (v[1] + two_h * lv * k[0]) * inv_e;
    p_out[1] = (v[2] + two_h * lv * k[1]) * inv_e;
    p_out[2] = (v[3] + two_h * lv * k[2]) * inv_e;
}

// ------------------------------------------------------------ disk
// Carter's radial potential of a null ray with E = 1, xi = L_z/E, eta = Q/E^2:
//   R(r) = r^4 + (a^2 - xi^2 - eta) r^2 + 2M (eta + (xi - a)^2) r - a^2 eta
// (the polynomial CarterPhaseClassifier::classify builds,
// optics/carter_ray_classifier.hpp:130-134). A ray moving inward is certainly
// absorbed when R has no root outside r_+ (:149-153). Here: R(r_+) >= 0
// always (Delta(r_+) = 0), R -> +inf, so R has no root outside r_+ exactly
// when R > 0 at r_+ and at every critical point (R' = 0) outside r_+. The
// critical points solve the depressed cubic c^3 + (P/2) c + Qc/4 = 0 (closed
// form: Cardano, or the trigonometric form for three real roots). Unlike the
// reference classifier, which leaves eta < 0 to integration, rays with a
// negative Carter constant go through the same test. Returns 1
// (no root, every value above margin * r^4), 0 (a root: some critical value
// below -margin * c^4), -1 (undecided: within the margin, near the critical
// curve).
inline int carter_no_turning_point(double M, double a, double rp, double xi, double eta, double margin) {
    const double e = eta;  // eta < 0 (vortical rays) uses the same polynomial
    const double P = a * a - xi * xi - e;
    const double Qc = 2.0 * M * (e + (xi - a) * (xi - a));
    const double R0 = -a * a * e;
    auto R = [&](double r) { return ((r * r + P) * r + Qc) * r + R0; };
    bool undecided = !(R(rp) > margin * rp * rp * rp * rp);
    const double p = 0.5 * P, q = 0.25 * Qc;
    double c[3];
    int nc;
    const double disc = 0.25 * q * q + p * p * p / 27.0;
    if (disc > 0.0) {
        const double sd = std::sqrt(disc);
        c[0] = std::cbrt(-0.5 * q + sd) + std::cbrt(-0.5 * q - sd);
        nc = 1;
    } else {
        const double m = 2.0 * std::sqrt(-p / 3.0);
        const double arg = std::clamp(3.0 * q / (p * m), -1.0, 1.0);
        const double th = std::acos(arg) / 3.0;
        for (int k = 0; k < 3; ++k) c[k] = m * std::cos(th - 2.0943951023931957 * k);
        nc = 3;
    }
    for (int k = 0; k < nc; ++k) {
        if (!(c[k] > rp)) continue;
        const double c4 = c[k] * c[k] * c[k] * c[k];
        const double v = R(c[k]);
        if (v < -margin * c4) return 0;
        if (!(v > margin * c4)) undecided = true;
    }
    return undecided ? -1 : 1;
}

// Constants of a KS null state y = (t, x, y, z, p_t, p_x, p_y, p_z) about a
// Kerr hole of spin a: xi = L_z / E and eta = Q / E^2, with E = -p_t, L_z =
// x p_y - y p_x and p_theta of the spheroidal angle (KS and BL share theta).
// Returns false near the spin axis, where Q is ill-conditioned.
inline bool carter_constants(const double y[8], double a, double r, double* xi, double* eta) {
    const double E = -y[4];
    const double cth = y[3] / r;
    const double s2 = 1.0 - cth * cth;
    if (!(s2 > 1e-12) || !(E > 0.0)) return false;
    const double sth = std::sqrt(s2);
    const double Lz = y[1] * y[6] - y[2] * y[5];
    const double pth = (cth / sth) * (y[1] * y[5] + y[2] * y[6]) - r * sth * y[7];
    const double Q = pth * pth + cth * cth * (Lz * Lz / s2 - a * a * E * E);
    *xi = Lz / E;
    *eta = Q / (E * E);
    return true;
}

// Restates kerr_isco_radius, render/software_compute_engine.hpp:1052-1061
// (Bardeen–Press–Teukolsky prograde ISCO).
inline double kerr_isco(double m, double a_spin) {
    const double a_star = std::clamp(a_spin / max(m, 1e-12), -0.9999999, 0.9999999);
    const double abs_a = abs(a_star);
    const double orbit_sign = (a_star >= 0.0) ? -1.0 : 1.0;
    const double cbrt_term = cbrt(max(1.0 - abs_a * abs_a, 0.0));
    const double z1 = 1.0 + cbrt_term * (cbrt(1.0 + abs_a) + cbrt(1.0 - abs_a));
    const double z2 = sqrt(3.0 * abs_a * abs_a + z1 * z1);
    const double r_over_m = 3.0 + z2 + orbit_sign * sqrt(max((3.0 - z1) * (3.0 + z1 + 2.0 * z2), 0.0));
    return r_over_m * m;
}

// Restates AccretionDiskModel::orbital_frame / redshift_factor / redshift_at,
// render/accretion_disk/accretion_disk_model.hpp:205-225. xi = L_z / E.
inline double disk_redshift(double mass, double spin, double radius, double xi, bool clamp_g) {
    const double spin_mag = abs(spin);
    const double sqrt_m = sqrt(max(mass, 1e-12));
    const double sqrt_r = sqrt(max(radius, 1e-12));
    const double r15 = radius * sqrt_r;
    const double ang = sqrt_m / (r15 + spin_mag * sqrt_m);
    const double bracket = max(r15 - 3.0 * mass * sqrt_r + 2.0 * spin_mag * sqrt_m, 1e-9);
    const double omega = (spin < 0.0) ? -ang : ang;
    const double ut = (r15 + spin_mag * sqrt_m) / (sqrt(r15) * sqrt(bracket));
    const double den = ut * (1.0 - omega * xi);
    if (!(abs(den) > 1e-9)) return 1.0;
    const double g = 1.0 / den;
    return clamp_g ? std::clamp(g, 0.05, 4.0) : g;
}

// ---------------------------------------------------------- DP45 tableau
// Dormand & Prince (1980) 5(4) pair, as rk45_adaptive.hpp:196-203 and
// docs/MATHEMATICAL_FORMULATION.md:2050-2072.
struct DP {
    static constexpr double a21 = 1.0 / 5.0;
    static constexpr double a31 = 3.0 / 40.0, a32 = 9.0 / 40.0;
    static constexpr double a41 = 44.0 / 45.0, a42 = -56.0 / 15.0, a43 = 32.0 / 9.0;
    static constexpr double a51 = 19372.0 / 6561.0, a52 = -25360.0 / 2187.0, a53 = 64448.0 / 6561.0,
                            a54 = -212.0 / 729.0;
    static constexpr double a61 = 9017.0 / 3168.0, a62 = -355.0 / 33.0, a63 = 46732.0 / 5247.0,
                            a64 = 49.0 / 176.0, a65 = -5103.0 / 18656.0;
    static constexpr double a71 = 35.0 / 384.0, a73 = 500.0 / 1113.0, a74 = 125.0 / 192.0,
                            a75 = -2187.0 / 6784.0, a76 = 11.0 / 84.0;
    static constexpr double e1 = 71.0 / 57600.0, e3 = -71.0 / 16695.0, e4 = 71.0 / 1920.0,
                            e5 = -17253.0 / 339200.0, e6 = 22.0 / 525.0, e7 = -1.0 / 40.0;
    // Continuous extension of order 4 (Hairer, Norsett & Wanner, Solving ODE I,
    // sect. II.6, DOPRI5 dense output), written in the Horner form
    //   y(s) = y0 + s (D + (1-s) (B + s (C + (1-s) E)))
    // with D = y1 - y0, B = h k1 - D, C = D - h k7 - B, E = h sum d_j k_j.
    static constexpr double d1 = -12715105075.0 / 11282082432.0, d3 = 87487479700.0 / 32700410799.0,
                            d4 = -10690763975.0 / 1880347072.0, d5 = 701980252875.0 / 199316789632.0,
                            d6 = -1453857185.0 / 822651844.0, d7 = 69997945.0 / 29380423.0;
};

// One DP45 attempt of an N-component system. k[0] holds f(y0) on entry
// (first same as last); fills k[1..6], y5 = 5th-order solution, and returns the
// max-norm error of rk45_adaptive.hpp:254-261 measured with |h| (the reference
// multiplies by the signed step, which makes every backward step's error
// non-positive; the magnitude is what the norm needs).
template <class S, int N, class F>
inline S dp45_attempt(F&& f, const S* y0, S h, S rtol, S atol, S k[7][N], S* y5) {
    S ys[N];
    for (int i = 0; i < N; ++i) ys[i] = y0[i] + h * (S(DP::a21) * k[0][i]);
    f(ys, k[1]);
    for (int i = 0; i < N; ++i) ys[i] = y0[i] + h * (S(DP::a31) * k[0][i] + S(DP::a32) * k[1][i]);
    f(ys, k[2]);
    for (int i = 0; i < N; ++i)
        ys[i] = y0[i] + h * (S(DP::a41) * k[0][i] + S(DP::a42) * k[1][i] + S(DP::a43) * k[2][i]);
    f(ys, k[3]);
    for (int i = 0; i < N; ++i)
        ys[i] = y0[i] + h * (S(DP::a51) * k[0][i] + S(DP::a52) * k[1][i] + S(DP::a53) * k[2][i] +
                             S(DP::a54) * k[3][i]);
    f(ys, k[4]);
    for (int i = 0; i < N; ++i)
        ys[i] = y0[i] + h * (S(DP::a61) * k[0][i] + S(DP::a62) * k[1][i] + S(DP::a63) * k[2][i] +
                             S(DP::a64) * k[3][i] + S(DP::a65) * k[4][i]);
    f(ys, k[5]);
    for (int i = 0; i < N; ++i)
        y5[i] = y0[i] + h * (S(DP::a71) * k[0][i] + S(DP::a73) * k[2][i] + S(DP::a74) * k[3][i] +
                             S(DP::a75) * k[4][i] + S(DP::a76) * k[5][i]);
    f(y5, k[6]);
    const S ah = abs(h);
    S err = S(0);
    for (int i = 0; i < N; ++i) {
        const S e = ah * abs(S(DP::e1) * k[0][i] + S(DP::e3) * k[2][i] + S(DP::e4) * k[3][i] +
                             S(DP::e5) * k[4][i] + S(DP::e6) * k[5][i] + S(DP::e7) * k[6][i]);
        const S sc = rtol * max(abs(y0[i]), abs(y5[i])) + atol;
        err = max(err, e / sc);
    }
    return err;
}

// Standard (integral) step controller, rk45_adaptive.hpp:293-311 with
// step_controller.hpp:44-78 (mode Standard), safety factor 0.9.
// Returns the new |h|.
template <class S>
inline S dp45_next_step(S abs_h, S err, bool accepted, S h_min, S h_max) {
    if (accepted) {
        const S factor = (err == S(0)) ? S(5) : pow(S(1) / max(err, S(1e-12)), S(1) / S(5));
        S ah = abs_h * std::clamp(S(0.9) * factor, S(0.2), S(5));
        return std::clamp(ah, h_min, h_max);
    }
    const S factor = pow(S(1) / max(err, S(1e-12)), S(1) / S(4));
    return abs_h * std::clamp(S(0.9) * factor, S(0.1), S(0.9));
}

// Dense output at s in [0,1] for component i.
template <int N>
inline double dense_eval(double s, int i, const double* y0, const double* y1, double h, const double k[7][N]) {
    const double D = y1[i] - y0[i];
    const double B = h * k[0][i] - D;
    const double C = D - h * k[6][i] - B;
    const double E = h * (DP::d1 * k[0][i] + DP::d3 * k[2][i] + DP::d4 * k[3][i] + DP::d5 * k[4][i] +
                          DP::d6 * k[5][i] + DP::d7 * k[6][i]);
    const double s1 = 1.0 - s;
    return y0[i] + s * (D + s1 * (B + s * (C + s1 * E)));
}

// Illinois (modified regula falsi) root of g on [0,1] given g(0)=g0, g(1)=g1
// with g0*g1 <= 0. Deterministic: at most 64 iterations, stops when the
// iterate no longer moves or g vanishes.
template <class G>
inline double illinois_root(G&& g, double g0, double g1) {
    if (g1 == 0.0) return 1.0;
    double a = 0.0, fa = g0, b = 1.0, fb = g1;
    double c = 1.0, c_prev = -1.0;
    for (int it = 0; it < 64; ++it) {
        c = (a * fb - b * fa) / (fb - fa);
        if (!(c > 0.0)) c = 0.0;
        if (c > 1.0) c = 1.0;
        if (abs(c - c_prev) <= 1e-15) break;
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

}  // namespace rto
