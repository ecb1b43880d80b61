// rt_multi.hip — the superposition kernel (multi-source Kerr–Schild, 2 to
// 16 sources) of the null-geodesic tracer for gfx950, and its launchers.
//
//   k_trace_multi        G lanes per ray (2, 4 or 8), each with one or two
//                        sources and a share of the state components;
//                        position and momentum gathered per stage, source
//                        sums reduce-scattered by DPP; persistent, idle
//                        groups refilled ray by ray
//
// Follows (paths relative to the Relativity-Simulator tree):
//   ms_rhs, ms_initial_momentum, trace_multi_source
//     render/software_compute_engine.hpp:2857-2870, 2919-2964, 3044-3252
//   DP45 tableau, error, control as for one hole (rt_device.hpp)
#include <hip/hip_runtime.h>

#include <algorithm>

#include "rt_camera.hpp"
#include "rt_device.hpp"
#include "rt_kernels.h"

namespace rt {

// ------------------------------------------------------- superposition
// Component-distributed layout. The G lanes of a group (G = 2, 4 or 8; see
// multi_ray_lanes) trace one ray together. Lane l of the group evaluates
// source l (and source l + 8 in groups of 8), and owns the state components
// c = q + Gc m (m < C) of the DP45 step, with q = l, Gc = G and C = 8 / G: it
// keeps their values, their seven stages and their stage sums, so no lane
// holds the whole step. Per right-hand side the group gathers the position
// and momentum from their owners (every lane needs them for its sources),
// each lane adds its sources' terms, and the sums over the sources are
// reduced onto the owners of the components (reduce-scatter by DPP
// butterflies, no memory). Values every lane needs
// (p_t, the current position and momentum, the step size, counters) are kept
// in every lane. Floating-point addition is commutative bit for bit, so both
// partners of a butterfly level hold identical bits: every group-wide value
// is the same in all lanes of the group, and the lanes take the same
// accept/reject and event decisions in lock step. A ray's latency per step is
// that of its lane's sources and one share of the stage arithmetic, which
// matters because a frame with few sources is bound by its slowest rays (rays
// that graze a horizon take thousands of steps).

// DPP lane moves inside rows of 16: xor 1 and xor 2 (quad permutes), mirror
// inside 8 lanes and inside 16 lanes. Each pairs every lane with a partner in
// the same aligned group of 2, 4, 8 or 16 lanes.
constexpr int kDppXor1 = 0xB1;        // quad_perm [1,0,3,2]
constexpr int kDppXor2 = 0x4E;        // quad_perm [2,3,0,1]
constexpr int kDppHalfMirror = 0x141; // row_half_mirror
constexpr int kDppMirror = 0x140;     // row_mirror

template <int Ctrl>
__device__ __forceinline__ double dpp_f64(double v) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(static_cast<uint32_t>(b)), Ctrl, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(static_cast<uint32_t>(b >> 32)), Ctrl, 0xF, 0xF,
                                               false);
    return __builtin_bit_cast(double, (static_cast<uint64_t>(static_cast<uint32_t>(hi)) << 32) |
                                          static_cast<uint32_t>(lo));
}

template <int Ctrl>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v) {
    return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), Ctrl, 0xF, 0xF, false));
}

// The value of v in the lane whose byte address (4 x lane) is addr.
__device__ __forceinline__ double bperm_f64(int addr, double v) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const int lo = __builtin_amdgcn_ds_bpermute(addr, static_cast<int>(static_cast<uint32_t>(b)));
    const int hi = __builtin_amdgcn_ds_bpermute(addr, static_cast<int>(static_cast<uint32_t>(b >> 32)));
    return __builtin_bit_cast(double, (static_cast<uint64_t>(static_cast<uint32_t>(hi)) << 32) |
                                          static_cast<uint32_t>(lo));
}

// Sum over the group (identical bits in every lane of the group).
template <int G>
__device__ __forceinline__ double group_sum(double v) {
    if constexpr (G == 1) return v;
    v = v + dpp_f64<kDppXor1>(v);
    if constexpr (G >= 4) v = v + dpp_f64<kDppXor2>(v);
    if constexpr (G >= 8) v = v + dpp_f64<kDppHalfMirror>(v);
    if constexpr (G >= 16) v = v + dpp_f64<kDppMirror>(v);
    return v;
}

template <int G>
__device__ __forceinline__ double group_min(double v) {
    if constexpr (G == 1) return v;
    v = fmin(v, dpp_f64<kDppXor1>(v));
    if constexpr (G >= 4) v = fmin(v, dpp_f64<kDppXor2>(v));
    if constexpr (G >= 8) v = fmin(v, dpp_f64<kDppHalfMirror>(v));
    if constexpr (G >= 16) v = fmin(v, dpp_f64<kDppMirror>(v));
    return v;
}

// Earliest event of the group: smallest s, ties to the lower source index
// (the order in which a sequential scan over the sources would keep them).
template <int Ctrl>
__device__ __forceinline__ void argmin_step(double& s, uint32_t& i) {
    const double so = dpp_f64<Ctrl>(s);
    const uint32_t io = dpp_u32<Ctrl>(i);
    if (so < s || (so == s && io < i)) {
        s = so;
        i = io;
    }
}

template <int G>
__device__ __forceinline__ void group_argmin(double& s, uint32_t& i) {
    if constexpr (G == 1) return;
    argmin_step<kDppXor1>(s, i);
    if constexpr (G >= 4) argmin_step<kDppXor2>(s, i);
    if constexpr (G >= 8) argmin_step<kDppHalfMirror>(s, i);
    if constexpr (G >= 16) argmin_step<kDppMirror>(s, i);
}

// Layout constants of a group of G lanes.
template <int G>
struct MG {
    static_assert(G == 2 || G == 4 || G == 8, "groups of 2, 4 or 8 lanes");
    static constexpr int Gc = G;      // lanes owning components
    static constexpr int C = 8 / G;   // components per lane
};

// Per-lane constants of the group layout.
struct MCtx {
    uint32_t q;   // component slot: the lane owns components q + Gc m
    int bp;       // 4 x (first lane of the group): ds_bpermute address of the group's first lane
};

// Component c (compile time) of a group-distributed vector v, in every lane of
// the group: read from its owner (lane q_c = c mod Gc of the group, register
// m = c / Gc). Pairs and quads broadcast by a quad permute; groups of 8 read
// the owner through the LDS crossbar.
template <int G, int c>
__device__ __forceinline__ double from_owner(const double* v, const MCtx& cx) {
    constexpr int Gc = MG<G>::Gc;
    constexpr int qc = c % Gc, m = c / Gc;
    if constexpr (Gc == 2)
        return dpp_f64<(qc == 0) ? 0xA0 : 0xF5>(v[m]);  // quad_perm [0,0,2,2] / [1,1,3,3]
    else if constexpr (Gc == 4)
        return dpp_f64<qc * 0x55>(v[m]);  // quad_perm [qc,qc,qc,qc]
    else
        return bperm_f64(cx.bp + 4 * qc, v[m]);
}

// Sum over the group of the 8-slot vectors v (slot = component), delivered to
// the owners: S[m] = sum over the group's lanes of v[q + Gc m]. Each level
// halves the slots a lane keeps: it sends its partner the half the partner
// keeps and adds the half it receives to its own.
template <int G>
__device__ __forceinline__ void reduce_scatter(double v[8], uint32_t q, double* S) {
    constexpr int Gc = MG<G>::Gc;
    if constexpr (Gc == 8) {
        // lanes 0-3 keep components 0-3, lanes 4-7 components 4-7 (partner 7 - q)
        const bool lo = q < 4;
        double w[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const double own = lo ? v[k] : v[4 + k];
            const double send = lo ? v[4 + k] : v[k];
            w[k] = own + dpp_f64<kDppHalfMirror>(send);
        }
        // partner q ^ 2: keep w[0..1] or w[2..3]
        const bool b1 = (q & 2) == 0;
        double u[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const double own = b1 ? w[k] : w[2 + k];
            const double send = b1 ? w[2 + k] : w[k];
            u[k] = own + dpp_f64<kDppXor2>(send);
        }
        // partner q ^ 1: component q
        const bool b0 = (q & 1) == 0;
        S[0] = (b0 ? u[0] : u[1]) + dpp_f64<kDppXor1>(b0 ? u[1] : u[0]);
    } else if constexpr (Gc == 4) {
        // partner q ^ 2: lanes 0, 1 keep components {0, 1, 4, 5}, lanes 2, 3 {2, 3, 6, 7}
        const bool b1 = (q & 2) == 0;
        double w[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int ca = (k & 1) + 4 * (k >> 1);
            const double own = b1 ? v[ca] : v[ca + 2];
            const double send = b1 ? v[ca + 2] : v[ca];
            w[k] = own + dpp_f64<kDppXor2>(send);
        }
        // partner q ^ 1: components q and q + 4
        const bool b0 = (q & 1) == 0;
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            const double own = b0 ? w[2 * m] : w[2 * m + 1];
            const double send = b0 ? w[2 * m + 1] : w[2 * m];
            S[m] = own + dpp_f64<kDppXor1>(send);
        }
    } else {
        // partner q ^ 1: lane 0 keeps the even components, lane 1 the odd ones
        const bool b0 = (q & 1) == 0;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const double own = b0 ? v[2 * m] : v[2 * m + 1];
            const double send = b0 ? v[2 * m + 1] : v[2 * m];
            S[m] = own + dpp_f64<kDppXor1>(send);
        }
    }
}

// Maximum over the lanes of the group.
template <int G>
__device__ __forceinline__ double owners_max(double v) {
    constexpr int Gc = MG<G>::Gc;
    v = fmax(v, dpp_f64<kDppXor1>(v));
    if constexpr (Gc >= 4) v = fmax(v, dpp_f64<kDppXor2>(v));
    if constexpr (Gc >= 8) v = fmax(v, dpp_f64<kDppHalfMirror>(v));
    return v;
}

// Component q + Gc m of a whole state vector y (select chain: the index
// depends on the lane).
template <int G, int m>
__device__ __forceinline__ double own_component(const double y[8], uint32_t q) {
    constexpr int Gc = MG<G>::Gc;
    double v = y[Gc * m];
#pragma unroll
    for (int j = 1; j < Gc; ++j) v = (q == static_cast<uint32_t>(j)) ? y[Gc * m + j] : v;
    return v;
}

// The components q + Gc m (m < C) of a whole state vector y.
template <int G>
__device__ __forceinline__ void own_components(const double y[8], uint32_t q, double* out) {
    constexpr int C = MG<G>::C;
    out[0] = own_component<G, 0>(y, q);
    if constexpr (C > 1) out[1] = own_component<G, 1>(y, q);
    if constexpr (C > 2) {
        out[2] = own_component<G, 2>(y, q);
        out[3] = own_component<G, 3>(y, q);
    }
}

__device__ __forceinline__ void to_local(const SourceDev& s, double x, double y, double z, double l[3]) {
    const double dx = x - s.c[0], dy = y - s.c[1], dz = z - s.c[2];
    l[0] = dx * s.fx[0] + dy * s.fx[1] + dz * s.fx[2];
    l[1] = dx * s.fy[0] + dy * s.fy[1] + dz * s.fy[2];
    l[2] = dx * s.ax[0] + dy * s.ax[1] + dz * s.ax[2];
}

// The sources of one lane: sources si and, in groups of 8 lanes, si + 8 of
// the table in LDS, those that exist (si: the lane's index in its group of G).
struct GroupSrc {
    const SourceDev* S;
    uint32_t first, n;
};

// Sources per lane in a group of G lanes (multi_ray_lanes picks G).
template <int G>
constexpr int lane_sources() {
    return (G == 8) ? 2 : 1;
}

// f(V, i) for each source of the lane, in order.
template <int G, class F>
__device__ __forceinline__ void for_sources(const GroupSrc& gs, F&& f) {
#pragma unroll
    for (int k = 0; k < lane_sources<G>(); ++k) {
        const uint32_t i = gs.first + static_cast<uint32_t>(k * G);
        if (i < gs.n) f(gs.S + i, i);
    }
}

// One source's terms in the superposed Hamiltonian (ms_rhs,
// software_compute_engine.hpp:2857-2870, with a carried p_t) at world
// position X, p_t and world momentum Pm, added to v: (2 H s, -2 H s k, 0, dp)
// with k and dp rotated to world axes.
__device__ __forceinline__ void add_source_terms(const SourceDev* V, const double X[3], double pt, const double Pm[3],
                                                 double v[8]) {
    double ths, c0, c1, c2, q0, q1, q2;
    {
        const double dx = X[0] - V->c[0], dy = X[1] - V->c[1], dz = X[2] - V->c[2];
        double lx, ly, lz, px, py, pz;
        {
            const double f0 = V->fx[0], f1 = V->fx[1], f2 = V->fx[2];
            lx = dx * f0 + dy * f1 + dz * f2;
            px = Pm[0] * f0 + Pm[1] * f1 + Pm[2] * f2;
        }
        {
            const double f0 = V->fy[0], f1 = V->fy[1], f2 = V->fy[2];
            ly = dx * f0 + dy * f1 + dz * f2;
            py = Pm[0] * f0 + Pm[1] * f1 + Pm[2] * f2;
        }
        {
            const double f0 = V->ax[0], f1 = V->ax[1], f2 = V->ax[2];
            lz = dx * f0 + dy * f1 + dz * f2;
            pz = Pm[0] * f0 + Pm[1] * f1 + Pm[2] * f2;
        }
        double t, c[3], q[3];
        ks_terms<double>(lx, ly, lz, pt, px, py, pz, kerr_const(V->m, V->a, V->a2), t, c, q);
        ths = t;
        {
            const double f0 = V->fx[0], f1 = V->fx[1], f2 = V->fx[2];
            c0 = c[0] * f0;
            c1 = c[0] * f1;
            c2 = c[0] * f2;
            q0 = q[0] * f0;
            q1 = q[0] * f1;
            q2 = q[0] * f2;
        }
        {
            const double f0 = V->fy[0], f1 = V->fy[1], f2 = V->fy[2];
            c0 += c[1] * f0;
            c1 += c[1] * f1;
            c2 += c[1] * f2;
            q0 += q[1] * f0;
            q1 += q[1] * f1;
            q2 += q[1] * f2;
        }
        {
            const double f0 = V->ax[0], f1 = V->ax[1], f2 = V->ax[2];
            c0 += c[2] * f0;
            c1 += c[2] * f1;
            c2 += c[2] * f2;
            q0 += q[2] * f0;
            q1 += q[2] * f1;
            q2 += q[2] * f2;
        }
    }
    v[0] += ths;
    v[1] += c0;
    v[2] += c1;
    v[3] += c2;
    v[5] += q0;
    v[6] += q1;
    v[7] += q2;
}

// Superposed right-hand side of the group-distributed state Y (p_t given):
// gathers the position X and momentum Pm (returned; the same in every lane),
// adds the lane's source and reduces onto the owners, d[m] being the rate of
// component q + Gc m: t' = -p_t + sum 2Hs, x' = p - sum 2Hs k, p_t' = 0,
// p' = sum dp.
template <int G>
__device__ __forceinline__ void mrhs(const GroupSrc& gs, const MCtx& cx, const double* Y, double pt, double* d,
                                     double X[3], double Pm[3]) {
    constexpr int Gc = MG<G>::Gc, C = MG<G>::C;
    X[0] = from_owner<G, 1>(Y, cx);
    X[1] = from_owner<G, 2>(Y, cx);
    X[2] = from_owner<G, 3>(Y, cx);
    Pm[0] = from_owner<G, 5>(Y, cx);
    Pm[1] = from_owner<G, 6>(Y, cx);
    Pm[2] = from_owner<G, 7>(Y, cx);
    double v[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for_sources<G>(gs, [&](const SourceDev* V, uint32_t) { add_source_terms(V, X, pt, Pm, v); });
    double S[C];
    reduce_scatter<G>(v, cx.q, S);
#pragma unroll
    for (int m = 0; m < C; ++m) {
        if (Gc * m < 4) {
            const uint32_t c = cx.q + Gc * m;
            double base = 0.0;
            base = (c == 1) ? Pm[0] : base;
            base = (c == 2) ? Pm[1] : base;
            base = (c == 3) ? Pm[2] : base;
            base = (c == 0) ? -pt : base;
            d[m] = base + S[m];
        } else {
            d[m] = S[m];
        }
    }
}

// The seven stages of one DP45 attempt on the group's components (as
// dp45_stages for one hole; WithDense adds the dense-output combination).
// X5, P5: position and momentum of y5 in every lane.
template <bool WithDense, int G>
__device__ __forceinline__ void mstages(const GroupSrc& gs, const MCtx& cx, const double* y, const double* k1,
                                        double pt, double h, double* y5, double* ep, double* k7, double* dn,
                                        double X5[3], double P5[3]) {
    constexpr int C = MG<G>::C;
    double k2[C], k3[C], k4[C], k5[C], k6[C], Y[C], X[3], Pm[3];
#pragma unroll
    for (int i = 0; i < C; ++i) Y[i] = y[i] + h * (dp::a21 * k1[i]);
    mrhs<G>(gs, cx, Y, pt, k2, X, Pm);
#pragma unroll
    for (int i = 0; i < C; ++i) Y[i] = y[i] + h * (dp::a31 * k1[i] + dp::a32 * k2[i]);
    mrhs<G>(gs, cx, Y, pt, k3, X, Pm);
#pragma unroll
    for (int i = 0; i < C; ++i) Y[i] = y[i] + h * (dp::a41 * k1[i] + dp::a42 * k2[i] + dp::a43 * k3[i]);
    mrhs<G>(gs, cx, Y, pt, k4, X, Pm);
#pragma unroll
    for (int i = 0; i < C; ++i)
        Y[i] = y[i] + h * (dp::a51 * k1[i] + dp::a52 * k2[i] + dp::a53 * k3[i] + dp::a54 * k4[i]);
    mrhs<G>(gs, cx, Y, pt, k5, X, Pm);
#pragma unroll
    for (int i = 0; i < C; ++i)
        Y[i] = y[i] + h * (dp::a61 * k1[i] + dp::a62 * k2[i] + dp::a63 * k3[i] + dp::a64 * k4[i] + dp::a65 * k5[i]);
    mrhs<G>(gs, cx, Y, pt, k6, X, Pm);
#pragma unroll
    for (int i = 0; i < C; ++i) {
        y5[i] = y[i] + h * (dp::a71 * k1[i] + dp::a73 * k3[i] + dp::a74 * k4[i] + dp::a75 * k5[i] + dp::a76 * k6[i]);
        ep[i] = dp::e1 * k1[i] + dp::e3 * k3[i] + dp::e4 * k4[i] + dp::e5 * k5[i] + dp::e6 * k6[i];
        if (WithDense) dn[i] = dp::d1 * k1[i] + dp::d3 * k3[i] + dp::d4 * k4[i] + dp::d5 * k5[i] + dp::d6 * k6[i];
    }
    mrhs<G>(gs, cx, y5, pt, k7, X5, P5);
}

// Local coordinates of a world point in the frame of source V.
__device__ __forceinline__ void to_local_v(const SourceDev* V, double x, double y, double z, double l[3]) {
    const double dx = x - V->c[0], dy = y - V->c[1], dz = z - V->c[2];
    l[0] = dx * V->fx[0] + dy * V->fx[1] + dz * V->fx[2];
    l[1] = dx * V->fy[0] + dy * V->fy[1] + dz * V->fy[2];
    l[2] = dx * V->ax[0] + dy * V->ax[1] + dz * V->ax[2];
}

// A superposition ray in its group: own components of the state and of its
// derivative (first-same-as-last), and what every lane keeps.
template <int C>
struct MLane {
    double y[C], k1[C];
    double xs[3], ps[3], pt;  // position, momentum and p_t of the state
    double ah, lambda, g;
    uint32_t acc, rej, flags, cls;
    uint32_t rhs, proj, f32;
};

// Superposed null constraint as a quadratic in p_t at (X, pt, Pm); projection
// as for one hole. Updates pt; true if it fired.
template <int G>
__device__ __forceinline__ bool mproject(const GroupSrc& gs, const double X[3], const double Pm[3], double& pt,
                                         double tol) {
    double h = 0.0, hq = 0.0, hq2 = 0.0;
    for_sources<G>(gs, [&](const SourceDev* V, uint32_t) {
        double l[3];
        to_local_v(V, X[0], X[1], X[2], l);
        const double px = Pm[0] * V->fx[0] + Pm[1] * V->fx[1] + Pm[2] * V->fx[2];
        const double py = Pm[0] * V->fy[0] + Pm[1] * V->fy[1] + Pm[2] * V->fy[2];
        const double pz = Pm[0] * V->ax[0] + Pm[1] * V->ax[1] + Pm[2] * V->ax[2];
        const Field f = ks_field(l[0], l[1], l[2], V->m, V->a, V->a2);
        const double q = f.k0 * px + f.k1 * py + f.k2 * pz;
        h += f.H;
        hq += f.H * q;
        hq2 += f.H * q * q;
    });
    const double sh = group_sum<G>(h), shq = group_sum<G>(hq), shq2 = group_sum<G>(hq2);
    const double P2 = Pm[0] * Pm[0] + Pm[1] * Pm[1] + Pm[2] * Pm[2];
    const double A = -(1.0 + 2.0 * sh);
    const double B = 4.0 * shq;
    const double C0 = P2 - 2.0 * shq2;
    const double res = (A * pt + B) * pt + C0;
    const double scale = fmax(1.0, pt * pt + P2);
    if (!(fabs(res) >= tol * scale)) return false;
    const double disc = B * B - 4.0 * A * C0;
    if (disc < 0.0) return false;
    const double sd = sqrt(disc);
    const double r1 = (-B + sd) / (2.0 * A);
    const double r2 = (-B - sd) / (2.0 * A);
    pt = (fabs(r1 - pt) <= fabs(r2 - pt)) ? r1 : r2;
    return true;
}

// Keep the owner's copy of component 4 equal to p_t after a projection.
template <int G>
__device__ __forceinline__ void sync_pt(const MCtx& cx, double* y, double pt) {
    constexpr int Gc = MG<G>::Gc;
    constexpr int m = 4 / Gc;
    if (cx.q == static_cast<uint32_t>(4 % Gc)) y[m] = pt;
}

// Camera pixel -> initial state in a superposition (trace_multi_source set-up,
// software_compute_engine.hpp:3054-3066, and ms_initial_momentum :2919-2964):
// flat spherical observer position, superposed-metric frame precomputed.
__device__ __forceinline__ void camera_state_multi(const TraceParams& P, const MultiDev& T, uint32_t px, uint32_t py,
                                                   double y[8]) {
    double o[3];
    camera_outward<true>(P, px, py, o);
    double vv[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) vv[c] = (c == 0 ? T.u0 : 0.0) - o[0] * T.e[0][c] - o[1] * T.e[1][c] - o[2] * T.e[2][c];
    double mom0 = vv[1], mom1 = vv[2], mom2 = vv[3];
    double tc = -vv[0];
    for (uint32_t i = 0; i < T.n; ++i) {
        const double lv = vv[0] + T.dir[i][0] * vv[1] + T.dir[i][1] * vv[2] + T.dir[i][2] * vv[3];
        const double w = 2.0 * T.pot[i] * lv;
        tc += w;
        mom0 += w * T.dir[i][0];
        mom1 += w * T.dir[i][1];
        mom2 += w * T.dir[i][2];
    }
    const double inv_pt = 1.0 / fmax(-tc, 1e-12);
    y[0] = 0.0;
    y[1] = P.X0[0];
    y[2] = P.X0[1];
    y[3] = P.X0[2];
    y[4] = -1.0;
    y[5] = mom0 * inv_pt;
    y[6] = mom1 * inv_pt;
    y[7] = mom2 * inv_pt;
}

// A ray through a superposition (FP64): DP45 as for one hole; events per
// source (horizon of its local KS radius, crossing of its local equatorial
// plane inside its disk), escape on the Euclidean radius while outgoing.
// mbegin sets a group up from the whole initial state y0 (every lane), mstep
// makes one attempt and returns true (in every lane of the group) once the
// ray has terminated.
template <int G>
__device__ __forceinline__ void mbegin(const TraceParams& P, const GroupSrc& gs, const MCtx& cx, const double y0[8],
                                       MLane<MG<G>::C>& L) {
    own_components<G>(y0, cx.q, L.y);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        L.xs[i] = y0[1 + i];
        L.ps[i] = y0[5 + i];
    }
    L.pt = y0[4];
    double X[3], Pm[3];
    mrhs<G>(gs, cx, L.y, L.pt, L.k1, X, Pm);
    L.rhs = 1;
    L.proj = L.f32 = 0;
    L.acc = L.rej = L.flags = L.cls = 0;
    L.lambda = 0.0;
    L.g = 0.0;
    double scale = 1e300, blend = 1.0;
    for_sources<G>(gs, [&](const SourceDev* V, uint32_t) {
        double l[3];
        to_local_v(V, L.xs[0], L.xs[1], L.xs[2], l);
        const double r = ks_radius(l[0], l[1], l[2], V->a2);
        const double rh = V->rh;
        scale = fmin(scale, sqrt(r * fmax(r - rh, 0.02 * V->m)));
        blend = fmin(blend, fmin(fmax((r - 20.0 * rh) / (80.0 * rh), 0.0), 1.0));
    });
    scale = group_min<G>(scale);
    blend = group_min<G>(blend);
    const double ah = (P.h0 > 0.0) ? P.h0 : fmin(fmax(0.5 * scale * (1.0 + 6.8 * blend), 0.004), 3.5 * 7.8);
    L.ah = fmin(fmax(ah, P.h_min), P.h_max);
}

// Dense-output coefficients of component c (compile time), in every lane.
template <int G, int c>
__device__ __forceinline__ Dense dense_from_owner(const double* D, const double* B, const double* Cc, const double* E,
                                                  const MCtx& cx) {
    Dense q;
    q.D = from_owner<G, c>(D, cx);
    q.B = from_owner<G, c>(B, cx);
    q.C = from_owner<G, c>(Cc, cx);
    q.E = from_owner<G, c>(E, cx);
    return q;
}

// Event location inside an accepted step [y, y5] of size h: every lane the
// horizons and disks of its own sources, then the group's earliest (ties to
// the lower source index, as a scan over the sources in order keeps them),
// then escape on the Euclidean radius while outgoing. Recomputes the step
// with the dense-output combination; the dense coefficients of the position
// and momentum are gathered into every lane. Fills the end state (own
// components ye; position Xe and momentum Pe in every lane); returns the
// class (0: no event), the same in every lane of the group.
template <int G>
__device__ __forceinline__ uint32_t mlocate_event(const TraceParams& P, const GroupSrc& gs, const MCtx& cx,
                                                  uint32_t lane0, const MLane<MG<G>::C>& L, double h,
                                                  bool es, double* ye, double Xe[3], double Pe[3], double* s_out,
                                                  double* g_out) {
    constexpr int C = MG<G>::C;
    double y5[C], ep[C], k7[C], dn[C], X5[3], P5[3];
    mstages<true, G>(gs, cx, L.y, L.k1, L.pt, h, y5, ep, k7, dn, X5, P5);
    double qD[C], qB[C], qC[C], qE[C];
#pragma unroll
    for (int i = 0; i < C; ++i) {
        qD[i] = y5[i] - L.y[i];
        qB[i] = h * L.k1[i] - qD[i];
        qC[i] = qD[i] - h * k7[i] - qB[i];
        qE[i] = h * (dn[i] + dp::d7 * k7[i]);
    }
    const Dense qx = dense_from_owner<G, 1>(qD, qB, qC, qE, cx);
    const Dense qy = dense_from_owner<G, 2>(qD, qB, qC, qE, cx);
    const Dense qz = dense_from_owner<G, 3>(qD, qB, qC, qE, cx);
    const Dense q5 = dense_from_owner<G, 5>(qD, qB, qC, qE, cx);
    const Dense q6 = dense_from_owner<G, 6>(qD, qB, qC, qE, cx);
    const Dense q7 = dense_from_owner<G, 7>(qD, qB, qC, qE, cx);
    const double* y = L.xs;
    double best = 2.0, best_g = 0.0;
    uint32_t best_cls = 0, best_src = 0xFFFFFFFFu;
    for_sources<G>(gs, [&](const SourceDev* V, uint32_t i) {
        auto local_at = [&](double ss, double l[3]) {
            to_local_v(V, dense_at(qx, y[0], ss), dense_at(qy, y[1], ss), dense_at(qz, y[2], ss), l);
        };
        double l0[3], l1[3];
        to_local_v(V, y[0], y[1], y[2], l0);
        to_local_v(V, X5[0], X5[1], X5[2], l1);
        const double a2 = V->a2, r_hz = V->r_hz;
        const double r0 = ks_radius(l0[0], l0[1], l0[2], a2);
        const double r1 = ks_radius(l1[0], l1[1], l1[2], a2);
        if (r1 <= r_hz && r0 > r_hz) {
            const double sh = illinois_root(
                [&](double ss) {
                    double l[3];
                    local_at(ss, l);
                    return ks_radius(l[0], l[1], l[2], a2) - r_hz;
                },
                r0 - r_hz, r1 - r_hz);
            if (sh < best) {
                best = sh;
                best_cls = RT_CLS_CAPTURED;
                best_src = i;
            }
        }
        const double z0 = l0[2], z1 = l1[2];
        if (V->disk_on && ((z0 < 0.0 && z1 >= 0.0) || (z0 > 0.0 && z1 <= 0.0))) {
            const double sd = illinois_root(
                [&](double ss) {
                    double l[3];
                    local_at(ss, l);
                    return l[2];
                },
                z0, z1);
            double l[3];
            local_at(sd, l);
            const double rc = ks_radius(l[0], l[1], l[2], a2);
            if (rc >= V->disk_in && rc <= V->disk_out && rc > V->rh102 && sd < best) {
                const double pw0 = dense_at(q5, L.ps[0], sd), pw1 = dense_at(q6, L.ps[1], sd),
                             pw2 = dense_at(q7, L.ps[2], sd);
                const double pl0 = pw0 * V->fx[0] + pw1 * V->fx[1] + pw2 * V->fx[2];
                const double pl1 = pw0 * V->fy[0] + pw1 * V->fy[1] + pw2 * V->fy[2];
                const double pl2 = pw0 * V->ax[0] + pw1 * V->ax[1] + pw2 * V->ax[2];
                const double pm = sqrt(pl0 * pl0 + pl1 * pl1 + pl2 * pl2);
                const double lzv = (l[0] * pl1 - l[1] * pl0) / fmax(pm, 1e-30);
                best = sd;
                best_cls = RT_CLS_DISK;
                best_src = i;
                best_g = disk_redshift(V->m, V->a, rc, lzv, P.clamp_g != 0);
            }
        }
    });
    {
        uint32_t win = best_src;
        group_argmin<G>(best, win);
        if (best < 2.0) {
            const int from = static_cast<int>(lane0 + (win & (G - 1)));
            best_cls = static_cast<uint32_t>(__shfl(static_cast<int>(best_cls), from, 64));
            best_g = __shfl(best_g, from, 64);
        }
    }
    if (!(best < 2.0)) best_cls = 0;
    if (es) {
        const double R0 = sqrt(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
        const double R1 = sqrt(X5[0] * X5[0] + X5[1] * X5[1] + X5[2] * X5[2]);
        double se = 1.0;
        if (R0 < P.r_esc)
            se = illinois_root(
                [&](double ss) {
                    const double x = dense_at(qx, y[0], ss), yy = dense_at(qy, y[1], ss), z = dense_at(qz, y[2], ss);
                    return sqrt(x * x + yy * yy + z * z) - P.r_esc;
                },
                R0 - P.r_esc, R1 - P.r_esc);
        if (se < best) {
            best = se;
            best_cls = RT_CLS_ESCAPED;
        }
    }
    if (best_cls == 0) return 0;
    if (best >= 1.0) {
#pragma unroll
        for (int i = 0; i < C; ++i) ye[i] = y5[i];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            Xe[i] = X5[i];
            Pe[i] = P5[i];
        }
    } else {
#pragma unroll
        for (int i = 0; i < C; ++i) {
            Dense q;
            q.D = qD[i];
            q.B = qB[i];
            q.C = qC[i];
            q.E = qE[i];
            ye[i] = dense_at(q, L.y[i], best);
        }
        Xe[0] = dense_at(qx, y[0], best);
        Xe[1] = dense_at(qy, y[1], best);
        Xe[2] = dense_at(qz, y[2], best);
        Pe[0] = dense_at(q5, L.ps[0], best);
        Pe[1] = dense_at(q6, L.ps[1], best);
        Pe[2] = dense_at(q7, L.ps[2], best);
    }
    *s_out = best;
    *g_out = best_g;
    return best_cls;
}

template <int G>
__device__ __forceinline__ bool mstep(const TraceParams& P, const GroupSrc& gs, const MCtx& cx, uint32_t lane0,
                                      MLane<MG<G>::C>& L) {
    constexpr int Gc = MG<G>::Gc, C = MG<G>::C;
    if (L.acc + L.rej >= P.max_steps || L.ah < P.h_min) {
        L.flags |= (L.acc + L.rej >= P.max_steps) ? RT_FLAG_BUDGET : RT_FLAG_MIN_STEP;
        double near = 0.0;
        for_sources<G>(gs, [&](const SourceDev* V, uint32_t) {
            const double dx = L.xs[0] - V->c[0], dy = L.xs[1] - V->c[1], dz = L.xs[2] - V->c[2];
            if (sqrt(dx * dx + dy * dy + dz * dz) < 6.0 * V->m) near = 1.0;
        });
        L.cls = (group_min<G>(-near) < 0.0) ? RT_CLS_CAPTURED : RT_CLS_ESCAPED;
        return true;
    }
    const double ah = L.ah, h = -ah;
    double y5[C], ep[C], k7[C], X5[3], P5[3];
    mstages<false, G>(gs, cx, L.y, L.k1, L.pt, h, y5, ep, k7, nullptr, X5, P5);
    L.rhs += 6;
    // err = |h| max_i |sum_j e_j k_j,i| / (rtol max(|y_i|, |y5_i|) + atol) over the
    // components other than p_t; infinite when a component is not finite
    double err = 0.0;
    bool ok = true;
#pragma unroll
    for (int m = 0; m < C; ++m) {
        const uint32_t c = cx.q + Gc * m;
        const double e = ah * fabs(ep[m] + dp::e7 * k7[m]);
        const double sc = P.rtol * fmax(fabs(L.y[m]), fabs(y5[m])) + P.atol;
        const double r = e * rcp64(sc);
        ok = ok && isfinite(y5[m]);
        if (c != 4) {
            ok = ok && isfinite(r);
            err = fmax(err, r);
        }
    }
    err = owners_max<G>(ok ? err : __builtin_huge_val());
    if (!(err < __builtin_huge_val())) {
        L.flags |= RT_FLAG_NONFINITE;
        L.cls = RT_CLS_CAPTURED;
        return true;
    }
    if (!(err <= 1.0)) {
        ++L.rej;
        L.ah = next_step(ah, err, false, P.h_min, P.h_max);
        return false;
    }
    ++L.acc;
    // any event in this step? (a horizon reached or a disk plane crossed by one
    // of the group's sources, escape)
    double mine = 0.0;
    for_sources<G>(gs, [&](const SourceDev* V, uint32_t) {
        double l0[3], l1[3];
        to_local_v(V, L.xs[0], L.xs[1], L.xs[2], l0);
        to_local_v(V, X5[0], X5[1], X5[2], l1);
        const bool hz = ks_radius(l1[0], l1[1], l1[2], V->a2) <= V->r_hz;
        const bool dk = V->disk_on && ((l0[2] < 0.0 && l1[2] >= 0.0) || (l0[2] > 0.0 && l1[2] <= 0.0));
        if (hz || dk) mine = 1.0;
    });
    const bool src_ev = group_min<G>(-mine) < 0.0;
    const double R1 = sqrt(X5[0] * X5[0] + X5[1] * X5[1] + X5[2] * X5[2]);
    bool es = false;
    if (R1 >= P.r_esc) {
        const double kx = from_owner<G, 1>(k7, cx), ky = from_owner<G, 2>(k7, cx), kz = from_owner<G, 3>(k7, cx);
        es = -(X5[0] * kx + X5[1] * ky + X5[2] * kz) > 0.0;
    }
    if (src_ev || es) {
        double s_ev = 1.0, g_ev = 0.0, ye[C], Xe[3], Pe[3];
        const uint32_t c = mlocate_event<G>(P, gs, cx, lane0, L, h, es, ye, Xe, Pe, &s_ev, &g_ev);
        if (c != 0) {
            L.rhs += 6;
            if (mproject<G>(gs, Xe, Pe, L.pt, P.inv_tol)) {
                L.flags |= RT_FLAG_PROJECTED;
                ++L.proj;
            }
#pragma unroll
            for (int i = 0; i < C; ++i) L.y[i] = ye[i];
            sync_pt<G>(cx, L.y, L.pt);
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                L.xs[i] = Xe[i];
                L.ps[i] = Pe[i];
            }
            L.lambda += s_ev * ah;
            L.cls = c;
            L.g = g_ev;
            return true;
        }
    }
#pragma unroll
    for (int i = 0; i < C; ++i) L.y[i] = y5[i];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        L.xs[i] = X5[i];
        L.ps[i] = P5[i];
    }
    L.lambda += ah;
    if (mproject<G>(gs, L.xs, L.ps, L.pt, P.inv_tol)) {
        L.flags |= RT_FLAG_PROJECTED;
        ++L.proj;
        sync_pt<G>(cx, L.y, L.pt);
        double X[3], Pm[3];
        mrhs<G>(gs, cx, L.y, L.pt, L.k1, X, Pm);
        ++L.rhs;
    } else {
#pragma unroll
        for (int i = 0; i < C; ++i) L.k1[i] = k7[i];
    }
    L.ah = next_step(ah, err, true, P.h_min, P.h_max);
    return false;
}

// Redshift per class and the final flag word of a superposition ray.
template <int C>
__device__ __forceinline__ void mlane_finish(MLane<C>& L) {
    const uint32_t c = L.cls & RT_CLS_MASK;
    if (c == RT_CLS_ESCAPED) L.g = 1.0 / (-L.pt);
    if (c == RT_CLS_CAPTURED) L.g = 0.0;
    L.cls |= L.flags & ~(kCarterNoTurn | kNoEventRetry);
}

// The per-ray outputs other than the state (not inlined, for the same reason
// as camera_state_multi: atan2's constants).
__device__ __noinline__ void store_multi_rest(const OutPtrs& O, uint64_t idx, uint32_t cls, double x, double y,
                                              double z, double lambda, double g, uint32_t acc, uint32_t rej,
                                              const SourceDev& s0) {
    const uint64_t n = O.n;
    O.cls[idx] = cls;
    double lc[3], bl[3];
    to_local(s0, x, y, z, lc);
    bl_from_ks(lc[0], lc[1], lc[2], s0.a, s0.a2, bl);
#pragma unroll
    for (int c = 0; c < 3; ++c) O.bl[c * n + idx] = bl[c];
    O.lambda[idx] = lambda;
    O.g[idx] = g;
    O.steps[idx] = acc;
    O.rejected[idx] = rej;
}

// Store: every owner its components, lane 0 of the group the rest, with BL coordinates taken in the frame of source 0 (the
// central hole).
template <int G>
__device__ __forceinline__ void store_mlane(const OutPtrs& O, uint64_t idx, const MLane<MG<G>::C>& L,
                                            const MultiDev& T, const MCtx& cx, uint32_t l) {
    constexpr int Gc = MG<G>::Gc, C = MG<G>::C;
    const uint64_t n = O.n;
    if (l < static_cast<uint32_t>(Gc)) {
#pragma unroll
        for (int m = 0; m < C; ++m) {
            const uint32_t c = cx.q + Gc * m;
            double* dst = (c < 4) ? O.x : O.p;
            dst[(c & 3) * n + idx] = L.y[m];
        }
    }
    if (l == 0) store_multi_rest(O, idx, L.cls, L.xs[0], L.xs[1], L.xs[2], L.lambda, L.g, L.acc, L.rej, T.src[0]);
}

// Superposition request: a persistent kernel, G lanes per ray (G = 2, 4 or
// 8; lane s of a group takes its sources and a share of the components).
// Idle groups are refilled from a device counter (ballot over the group
// leaders, one atomicAdd per refill, rank by popcount), so a group whose ray
// terminates takes the next ray while the slow rays of the wave continue.
// Camera rays are handed out in 8x8-tile order (ray g: tile g / 64, pixel
// g % 64), the tiles around the holes' images first (tile_xy): the rays that
// take the most steps start at once instead of forming the frame's tail. The
// source table is staged into LDS. *next starts at zero.
template <int G, int Waves>
__global__ __launch_bounds__(256, Waves) void k_trace_multi(TraceParams P, OutPtrs O,
                                                            const MultiDev* __restrict__ Tg,
                                                            unsigned long long* __restrict__ next,
                                                            uint32_t refill_min, const uint32_t* limit,
                                                            unsigned long long* active_waves) {
    constexpr int C = MG<G>::C;
    __shared__ MultiDev sm;
    {
        const uint32_t words = sizeof(MultiDev) / 8;
        const double* src = reinterpret_cast<const double*>(Tg);
        double* dst = reinterpret_cast<double*>(&sm);
        for (uint32_t w = threadIdx.x; w < words; w += blockDim.x) dst[w] = src[w];
        __syncthreads();
    }
    const MultiDev& T = sm;
    const uint32_t lane_id = threadIdx.x & 63;
    const uint32_t si = lane_id & (G - 1);
    const uint32_t lane0 = lane_id & ~static_cast<uint32_t>(G - 1);
    const uint64_t below = (lane0 == 0) ? 0ull : (~0ull >> (64 - lane0));
    const GroupSrc gs{sm.src, si, T.n};
    const MCtx cx{si & static_cast<uint32_t>(MG<G>::Gc - 1), static_cast<int>(4 * lane0)};
    const uint64_t total = (P.mode == 0) ? static_cast<uint64_t>(P.tiles_x) * ((P.n_rays / P.cols + 7) / 8) * 64
                                         : static_cast<uint64_t>(P.n_rays);
    MLane<C> L;
    bool active = false, drained = false;
    uint64_t idx = 0;
    Tally tally;
    wave_joins(limit, active_waves);
    for (;;) {
        const uint64_t idle = __ballot(!active && si == 0);
        const uint32_t n_idle = __popcll(idle);
        // the host asks for fewer waves: take no more rays, leave once idle
        if (!drained && n_idle != 0 && (n_idle * G >= refill_min || n_idle * G == 64) &&
            wave_yields(limit, active_waves))
            drained = true;
        if (!drained && n_idle != 0 && (n_idle * G >= refill_min || n_idle * G == 64)) {
            const uint32_t first = __ffsll(static_cast<unsigned long long>(idle)) - 1;
            unsigned long long base = 0;
            if (lane_id == first) base = atomicAdd(next, static_cast<unsigned long long>(n_idle));
            base = __shfl(base, first, 64);
            if (base + n_idle >= total) drained = true;
            if (!active) {
                const uint64_t g = base + __popcll(idle & below);
                if (g < total) {
                    double y0[8];
                    bool ok = true;
                    if (P.mode == 0) {
                        const uint32_t l = static_cast<uint32_t>(g & 63);
                        uint32_t tx, ty;
                        tile_xy(P, static_cast<uint32_t>(g >> 6), tx, ty);
                        const uint32_t px = tx * 8 + (l & 7);
                        const uint32_t py = ty * 8 + (l >> 3);
                        ok = (px < P.cols) && (py < P.n_rays / P.cols);
                        if (ok) {
                            camera_state_multi(P, T, P.col_first + px, P.row_first + py, y0);
                            idx = static_cast<uint64_t>(py) * P.cols + px;
                        }
                    } else {
                        idx = g;
#pragma unroll
                        for (int c = 0; c < 4; ++c) {
                            y0[c] = O.in_x[c * O.n + g];
                            y0[4 + c] = O.in_p[c * O.n + g];
                        }
                    }
                    if (ok && P.mode == 0 && T.observer_inside) {
                        // observer inside a horizon: captured where it stands
                        own_components<G>(y0, cx.q, L.y);
#pragma unroll
                        for (int i = 0; i < 3; ++i) {
                            L.xs[i] = y0[1 + i];
                            L.ps[i] = y0[5 + i];
                        }
                        L.pt = y0[4];
                        L.acc = L.rej = L.flags = L.proj = L.f32 = L.rhs = 0;
                        L.lambda = 0.0;
                        L.g = 0.0;
                        L.cls = RT_CLS_CAPTURED;
                        mlane_finish(L);
                        store_mlane<G>(O, idx, L, T, cx, si);
                    } else if (ok) {
                        mbegin<G>(P, gs, cx, y0, L);
                        active = true;
                    }
                }
            }
        }
        if (__ballot(active) == 0) {
            if (drained) break;
            continue;
        }
        // A wave holding a ray past 300 attempts issues first on its SIMD: the
        // frame ends with those rays, and a wave beside them on the SIMD would
        // otherwise take half the issue slots while the frame's bulk runs
        // (8 and 16 holes on 512^2: 23.3 -> 22.2 and 47.4 -> 43.9 ms).
        if (__ballot(active && (L.acc + L.rej) > 300u) != 0)
            __builtin_amdgcn_s_setprio(3);
        else
            __builtin_amdgcn_s_setprio(0);
        if (active && mstep<G>(P, gs, cx, lane0, L)) {
            mlane_finish(L);
            store_mlane<G>(O, idx, L, T, cx, si);
            if (si == 0) tally.add(L);
            active = false;
        }
    }
    tally.flush(O.counters);
}

// ------------------------------------------------------------ launchers
// Lanes per superposition ray. Up to eight sources: the smallest power of
// two not below the number of sources (one source per lane; these frames are
// bound by the latency of their slowest rays). From nine sources: 8 lanes
// (two sources per lane, lane_sources), which halves the lanes per ray and
// shares the stage arithmetic and the sums over the group between twice the
// sources (these frames are bound by instruction issue): 16 holes on a 512^2
// frame 55 -> 49 ms, while 8 holes on 4 lanes went from 27 to 31 ms.
uint32_t multi_ray_lanes(uint32_t n_src) {
    uint32_t g = 2;
    while (g < n_src && g < 8) g <<= 1;
    return g;
}

template <int G>
static hipError_t launch_multi_g(const TraceParams& P, const OutPtrs& O, const MultiDev* T, unsigned long long* next,
                                 const LaunchShape& L, hipStream_t stream, uint32_t max_grid, const uint32_t* limit,
                                 unsigned long long* active) {
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const void* fn = (L.waves >= 2) ? reinterpret_cast<const void*>(&k_trace_multi<G, 2>)
                                    : reinterpret_cast<const void*>(&k_trace_multi<G, 1>);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, static_cast<int>(L.block), 0);
    if (e != hipSuccess) return e;
    // no more lanes than G per ray (a small request leaves the rest of the device to others)
    const uint64_t rays = (P.mode == 0) ? static_cast<uint64_t>(P.tiles_x) * ((P.n_rays / P.cols + 7) / 8) * 64
                                        : static_cast<uint64_t>(P.n_rays);
    const uint64_t want = (rays * G + L.block - 1) / L.block;
    uint64_t grid64 = std::min<uint64_t>(want, static_cast<uint64_t>(cus) * static_cast<uint64_t>(per_cu > 0 ? per_cu : 1));
    if (max_grid != 0) grid64 = std::min<uint64_t>(grid64, max_grid);
    const uint32_t grid = static_cast<uint32_t>(std::max<uint64_t>(grid64, 1));
    const uint32_t refill = std::max<uint32_t>(1, std::min<uint32_t>(L.refill_min, 64));
    if (L.waves >= 2)
        hipLaunchKernelGGL((k_trace_multi<G, 2>), dim3(grid), dim3(L.block), 0, stream, P, O, T, next, refill, limit,
                           active);
    else
        hipLaunchKernelGGL((k_trace_multi<G, 1>), dim3(grid), dim3(L.block), 0, stream, P, O, T, next, refill, limit,
                           active);
    return hipGetLastError();
}

hipError_t launch_trace_multi(const TraceParams& P, const OutPtrs& O, const MultiDev* T, unsigned long long* next,
                              const LaunchShape& L, hipStream_t stream, uint32_t max_grid, const uint32_t* limit,
                              unsigned long long* active) {
    if (!shape_ok(L)) return hipErrorInvalidValue;
    if (P.mode == 0 && (P.cols == 0 || P.n_rays % P.cols != 0)) return hipErrorInvalidValue;
    if (P.n_src < 2 || P.n_src > static_cast<uint32_t>(kMaxSources)) return hipErrorInvalidValue;
    switch (multi_ray_lanes(P.n_src)) {
        case 2: return launch_multi_g<2>(P, O, T, next, L, stream, max_grid, limit, active);
        case 4: return launch_multi_g<4>(P, O, T, next, L, stream, max_grid, limit, active);
        case 8: return launch_multi_g<8>(P, O, T, next, L, stream, max_grid, limit, active);
        default: return hipErrorInvalidValue;
    }
}

// Workgroups of the superposition kernel (groups of 8 lanes, two waves per
// SIMD) resident on one compute unit.
hipError_t multi_blocks_per_cu(int* blocks) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, reinterpret_cast<const void*>(&k_trace_multi<8, 2>),
                                                        64, 0);
}

}  // namespace rt
