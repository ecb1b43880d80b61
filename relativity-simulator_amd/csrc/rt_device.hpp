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

// ---------------------------------------------------- fast FP64 math
// Reciprocal of a positive normal x: v_rcp_f64 (about 2^-29 relative) and one
// Newton step (error squared, then the final rounding): within about 1 ulp.
__device__ __forceinline__ double rcp64(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
}

// sqrt(x) and 1/sqrt(x) of a positive normal x: v_rsq_f64 (about 2^-23
// relative), one Goldschmidt iteration (2^-46) and a residual correction of
// the root; within 1 ulp.
__device__ __forceinline__ double sqrt_rsqrt64(double x, double& rs) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y;
    double h = 0.5 * y;
    const double r = fma(-g, h, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    const double d = fma(-g, g, x);
    rs = h + h;
    return fma(d, h, g);
}

// sqrt of x >= 0 (0 allowed), same sequence.
__device__ __forceinline__ double sqrt64(double x) {
    double rs;
    const double v = sqrt_rsqrt64(x, rs);
    return (x > 0.0) ? v : 0.0;
}

struct Field {
    double r, H, k0, k1, k2;
};

// KS radius: positive root of r^4 - (R^2 - a^2) r^2 - a^2 z^2 = 0, floored.
__device__ __forceinline__ double ks_radius(double x, double y, double z, double a2) {
    const double z2 = z * z;
    const double diff = x * x + y * y + z2 - a2;
    const double r2 = fmax(0.5 * (diff + sqrt64(diff * diff + 4.0 * a2 * z2)), 1e-24);
    double rs;
    return sqrt_rsqrt64(r2, rs);
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
// Generic over the arithmetic type R (double for FP64 steps, float for the
// far-field steps of the mixed-precision policy); inputs and outputs are FP64.
template <class R>
__device__ __forceinline__ R rcpR(R x);
template <>
__device__ __forceinline__ double rcpR<double>(double x) {
    return rcp64(x);
}
template <>
__device__ __forceinline__ float rcpR<float>(float x) {
    return __builtin_amdgcn_rcpf(x);
}
template <class R>
__device__ __forceinline__ R sqrt_rsqrtR(R x, R& rs);
template <>
__device__ __forceinline__ double sqrt_rsqrtR<double>(double x, double& rs) {
    return sqrt_rsqrt64(x, rs);
}
template <>
__device__ __forceinline__ float sqrt_rsqrtR<float>(float x, float& rs) {
    rs = __builtin_amdgcn_rsqf(x);
    return x * rs;
}
template <class R>
__device__ __forceinline__ R sqrtR(R x) {
    R rs;
    const R v = sqrt_rsqrtR<R>(x, rs);
    return (x > R(0)) ? v : R(0);
}

// Per-hole constants of the right-hand side.
struct KerrConst {
    double M, a, a2, four_a2, two_M, three_M, m_four_M, two_Ma2;
};
__device__ __forceinline__ KerrConst kerr_const(double M, double a, double a2) {
    KerrConst K;
    K.M = M;
    K.a = a;
    K.a2 = a2;
    K.four_a2 = 4.0 * a2;
    K.two_M = 2.0 * M;
    K.three_M = 3.0 * M;
    K.m_four_M = -4.0 * M;
    K.two_Ma2 = 2.0 * M * a2;
    return K;
}

// Hamiltonian right-hand side of one hole in its own frame:
// ths = 2 H s, k (the KS null vector), dp = s^2 grad H + 2 H s grad(k.p), s = k.p - p_t;
// the position rate is p - 2 H s k.
//
// With diff = x^2 + y^2 + z^2 - a^2 and sq = sqrt(diff^2 + 4 a^2 z^2) the KS
// radius is r^2 = (diff + sq) / 2, and r^4 + a^2 z^2 = r^2 sq. Hence
//   H = M r^3 / (r^4 + a^2 z^2) = M r / sq,
//   grad r = (x r / sq, y r / sq, z rho / (r sq)),   rho = r^2 + a^2,
//   grad H = hb grad r - (0, 0, 2 M a^2 z / (r sq^2)),
//   hb = M (3 a^2 z^2 - r^4) / (r^2 sq^2) = M (3 sq - 4 r^2) / sq^2,
// so the field takes the two reciprocal square roots (of sq^2 and of r^2) and
// one reciprocal (of rho). The gradient of k.p is taken in closed form instead
// of through the nine-entry Jacobian of k: with u = x p_x + y p_y,
// q1 = (r u + a (y p_x - x p_y)) / rho = k_x p_x + k_y p_y and q2 = z p_z / r,
//   grad(k.p) = A + C grad r,  A = ((r p_x - a p_y)/rho, (r p_y + a p_x)/rho, p_z/r),
//   C = (u - 2 r q1)/rho - q2/r.
// Same function as the reference's expression (core/kerr_schild_null_integrator.hpp:30-114).
// Generic over the arithmetic type R (double for FP64 steps, float for the
// far-field steps of the mixed-precision policy); inputs and outputs are FP64.
// The returned Field carries H = (2 H) / 2 (exact), used only where a caller
// reads it.
template <class R>
__device__ __forceinline__ Field ks_core(double Xd, double Yd, double Zd, double ptd, double pxd, double pyd,
                                         double pzd, const KerrConst& K, double& ths_out, double dp[3]) {
    const R X = R(Xd), Y = R(Yd), Z = R(Zd), pt = R(ptd), px = R(pxd), py = R(pyd), pz = R(pzd);
    const R a = R(K.a), a2 = R(K.a2);
    const R z2 = Z * Z;
    const R diff = fma(X, X, fma(Y, Y, z2 - a2));
    R inv_sq;
    const R sq = sqrt_rsqrtR<R>(fma(diff, diff, R(K.four_a2) * z2), inv_sq);
    const R r2 = fmax(R(0.5) * (diff + sq), R(1e-24));
    R inv_r;
    const R r = sqrt_rsqrtR<R>(r2, inv_r);
    const R rho = r2 + a2;
    const R inv_rho = rcpR<R>(rho);
    const R drc = r * inv_sq;  // (d r / d x) / x
    const R H2 = R(K.two_M) * drc;
    const R k0 = fma(r, X, a * Y) * inv_rho;
    const R k1 = fma(r, Y, -(a * X)) * inv_rho;
    const R k2 = Z * inv_r;
    const R q1 = fma(k0, px, k1 * py);
    const R q2 = k2 * pz;
    const R s = (q1 + q2) - pt;
    const R ths = H2 * s;
    ths_out = double(ths);
    const R u = fma(X, px, Y * py);
    const R C = fma(fma(-(r + r), q1, u), inv_rho, -(q2 * inv_r));
    const R hb = fma(R(K.m_four_M), r2, R(K.three_M) * sq) * (inv_sq * inv_sq);
    const R ss = s * s;
    const R G = fma(ss, hb, ths * C);
    const R Gd = G * drc;
    const R tr = ths * inv_rho;
    dp[0] = double(fma(Gd, X, tr * fma(r, px, -(a * py))));
    dp[1] = double(fma(Gd, Y, tr * fma(r, py, a * px)));
    dp[2] = double(inv_r * fma(Z * inv_sq, fma(-(ss * R(K.two_Ma2)), inv_sq, G * rho), ths * pz));
    Field f;
    f.r = double(r);
    f.H = double(R(0.5) * H2);
    f.k0 = double(k0);
    f.k1 = double(k1);
    f.k2 = double(k2);
    return f;
}

// The same terms as two_hs = 2 H s, corr = -2 H s k (the superposition sums them over holes).
template <class R>
__device__ __forceinline__ Field ks_terms(double Xd, double Yd, double Zd, double ptd, double pxd, double pyd,
                                          double pzd, const KerrConst& K, double& two_hs, double corr[3],
                                          double dp[3]) {
    const Field f = ks_core<R>(Xd, Yd, Zd, ptd, pxd, pyd, pzd, K, two_hs, dp);
    corr[0] = -(two_hs * f.k0);
    corr[1] = -(two_hs * f.k1);
    corr[2] = -(two_hs * f.k2);
    return f;
}

// Single-hole RHS for y = (t, x, y, z, p_t, p_x, p_y, p_z); d[4] = 0.
template <class R = double>
__device__ __forceinline__ Field rhs_kerr(const double y[8], const KerrConst& K, double d[8]) {
    double ths, dpp[3];
    const Field f = ks_core<R>(y[1], y[2], y[3], y[4], y[5], y[6], y[7], K, ths, dpp);
    d[0] = ths - y[4];
    d[1] = fma(-ths, f.k0, y[5]);
    d[2] = fma(-ths, f.k1, y[6]);
    d[3] = fma(-ths, f.k2, y[7]);
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

// e^(-1/5) and e^(-1/4) of e in [1e-12, ...): FP32 seed (v_log_f32 /
// v_exp_f32, about 3e-7 relative) and one Newton step on y^n e = 1 in FP64
// (about 6e-13 relative), in place of a full FP64 pow.
__device__ __forceinline__ double inv_root5(double e) {
    const double y = static_cast<double>(__builtin_amdgcn_exp2f(-0.2f * __builtin_amdgcn_logf(static_cast<float>(e))));
    const double y2 = y * y;
    return y * fma(1.0 - y2 * y2 * y * e, 0.2, 1.0);
}
__device__ __forceinline__ double inv_root4(double e) {
    const double y = static_cast<double>(__builtin_amdgcn_exp2f(-0.25f * __builtin_amdgcn_logf(static_cast<float>(e))));
    const double y2 = y * y;
    return y * fma(1.0 - y2 * y2 * e, 0.25, 1.0);
}

// Standard integral controller with safety 0.9; returns the new |h|.
__device__ __forceinline__ double next_step(double ah, double err, bool accepted, double h_min, double h_max) {
    if (accepted) {
        const double factor = (err == 0.0) ? 5.0 : inv_root5(fmax(err, 1e-12));
        const double s = fmin(fmax(0.9 * factor, 0.2), 5.0);
        return fmin(fmax(ah * s, h_min), h_max);
    }
    const double factor = inv_root4(fmax(err, 1e-12));
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
// Carter capture test (RT_TERM_CARTER_CAPTURE). Constants of the KS null
// state y about a hole of spin a at KS radius r: xi = L_z / E, eta = Q / E^2
// (E = -p_t, L_z = x p_y - y p_x, p_theta of the spheroidal angle). Then
// Carter's radial potential R(r) = r^4 + (a^2 - xi^2 - eta) r^2 +
// 2M (eta + (xi - a)^2) r - a^2 eta has no root outside r_+ exactly when it
// is positive at r_+ and at its critical points outside r_+ (R -> +inf, and
// R(r_+) >= 0 because Delta(r_+) = 0); the critical points are the real roots
// of c^3 + P/2 c + Qc/4 = 0. true: no root, every value above margin * r^4
// (certain capture for a ray moving inward); false: a root, or within the
// margin (near the critical curve), or near the spin axis.
__device__ __forceinline__ bool carter_certain_capture(const double y[8], double r, double M, double a, double rp,
                                                       double margin) {
    const double E = -y[4];
    const double cth = y[3] / r;
    const double s2 = 1.0 - cth * cth;
    if (!(s2 > 1e-12) || !(E > 0.0)) return false;
    const double sth = sqrt(s2);
    const double Lz = y[1] * y[6] - y[2] * y[5];
    const double pth = (cth / sth) * (y[1] * y[5] + y[2] * y[6]) - r * sth * y[7];
    const double Q = pth * pth + cth * cth * (Lz * Lz / s2 - a * a * E * E);
    const double xi = Lz / E, e = Q / (E * E);  // e < 0 (vortical rays): same polynomial
    const double P = a * a - xi * xi - e;
    const double Qc = 2.0 * M * (e + (xi - a) * (xi - a));
    const double R0 = -a * a * e;
    auto R = [&](double x) { return ((x * x + P) * x + Qc) * x + R0; };
    if (!(R(rp) > margin * rp * rp * rp * rp)) return false;
    const double p = 0.5 * P, q = 0.25 * Qc;
    const double disc = 0.25 * q * q + p * p * p / 27.0;
    double c[3];
    int nc;
    if (disc > 0.0) {
        const double sd = sqrt(disc);
        c[0] = cbrt(-0.5 * q + sd) + cbrt(-0.5 * q - sd);
        nc = 1;
    } else {
        const double m = 2.0 * sqrt(-p / 3.0);
        const double arg = fmin(fmax(3.0 * q / (p * m), -1.0), 1.0);
        const double th = acos(arg) / 3.0;
        c[0] = m * cos(th);
        c[1] = m * cos(th - 2.0943951023931957);
        c[2] = m * cos(th - 4.1887902047863905);
        nc = 3;
    }
    for (int k = 0; k < nc; ++k) {
        if (!(c[k] > rp)) continue;
        const double c4 = c[k] * c[k] * c[k] * c[k];
        if (!(R(c[k]) > margin * c4)) return false;
    }
    return true;
}

// Margin of the Carter test, relative to r^4.
constexpr double kCarterMargin = 1e-6;
// Lane-internal flag bit (not exported): the Carter test, made once on the
// ray's initial state (its constants are those of the whole ray, and they are
// best conditioned far from the hole), found no turning point.
constexpr uint32_t kCarterNoTurn = 1u << 30;

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}


// Lane-internal flag bit (not exported): the event test found a sign change
// in the last accepted attempt but no event inside it (a crossing of z = 0
// off the disk); the attempt is repeated and taken as a plain step.
constexpr uint32_t kNoEventRetry = 1u << 29;

// Per-wave totals of the step counters, added to the launch's counters once
// per wave (counters: [0] accepted [1] rejected [2] rhs evals [3] fp32 steps
// [4] projections).
struct Tally {
    unsigned long long acc = 0, rej = 0, rhs = 0, f32 = 0, proj = 0;
    template <class LaneT>
    __device__ void add(const LaneT& L) {
        acc += L.acc;
        rej += L.rej;
        rhs += L.rhs;
        f32 += L.f32;
        proj += L.proj;
    }
    __device__ void flush(unsigned long long* counters) {
        const unsigned long long sa = wave_sum(acc), sr = wave_sum(rej), sh = wave_sum(rhs), sf = wave_sum(f32),
                                 sp = wave_sum(proj);
        if ((threadIdx.x & 63) == 0) {
            atomicAdd(&counters[0], sa);
            atomicAdd(&counters[1], sr);
            atomicAdd(&counters[2], sh);
            atomicAdd(&counters[3], sf);
            atomicAdd(&counters[4], sp);
        }
    }
};

// Wave-level yield of the serving loop's persistent kernels. *active counts
// the request's running waves; a wave leaves when more are running than the
// host's current limit for the request (a word in mapped pinned host memory,
// read at system scope, lowered by the host to make room for other requests).
// The host keeps the limit at 1 or more, so the request's last waves always
// drain its queue. limit == nullptr: never yield.
__device__ __forceinline__ bool wave_yields(const uint32_t* limit, unsigned long long* active) {
    if (limit == nullptr) return false;
    int go = 0;
    if ((threadIdx.x & 63) == 0) {
        const uint32_t lim = __hip_atomic_load(limit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        const unsigned long long a = __hip_atomic_load(active, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (a > lim && atomicCAS(active, a, a - 1) == a) go = 1;
    }
    return __shfl(go, 0, 64) != 0;
}

__device__ __forceinline__ void wave_joins(const uint32_t* limit, unsigned long long* active) {
    if (limit != nullptr && (threadIdx.x & 63) == 0) atomicAdd(active, 1ull);
}


}  // namespace rt
