"""Closed-form and invariant checks that pin the oracle without the reference:
ISCO and horizon known answers (tests/test_novikov_thorne_disk.cpp:38-44,
tests/test_kerr_metrics.cpp of the reference), the Kerr-Schild field (null k,
det g = -1), Hamilton's equations by finite differences, the DP45 tableau and
its continuous extension on problems with closed forms, the Illinois root,
conservation of E, L_z and the Carter constant along traced rays, and the
critical impact parameters that separate captured from escaping rays."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib as O
import rtsim

PD = C.POINTER(C.c_double)


def ptr(a):
    return np.ascontiguousarray(a, np.float64).ctypes.data_as(PD)


# ------------------------------------------------------------- ISCO
def test_isco_known_answers():
    isco = O.oracle().rto_isco
    assert isco(1.0, 0.0) == pytest.approx(6.0, abs=1e-12)
    assert 1.0 < isco(1.0, 0.998) < 1.3
    assert isco(1.0, 0.998) == pytest.approx(1.2369, abs=1e-3)
    assert isco(2.0, 0.0) == pytest.approx(12.0, abs=1e-12)
    assert isco(1.0, -0.9999999) == pytest.approx(9.0, abs=1e-3)  # retrograde limit
    # monotone in prograde spin
    v = [isco(1.0, a) for a in np.linspace(0, 0.999, 30)]
    assert all(x > y for x, y in zip(v, v[1:]))


# ------------------------------------------------------------- KS field
def field(x, M, a):
    o = np.zeros(5)
    O.oracle().rto_ks_field(ptr(x), M, a, o.ctypes.data_as(PD))
    return o


def test_ks_field_identities():
    rng = np.random.default_rng(11)
    for _ in range(300):
        a = rng.uniform(-0.99, 0.99)
        x = rng.normal(size=3) * rng.uniform(2, 300)
        r, H, k0, k1, k2 = field(x, 1.0, a)
        # r solves r^4 - (R^2 - a^2) r^2 - a^2 z^2 = 0
        R2 = x @ x
        assert abs(r ** 4 - (R2 - a * a) * r * r - a * a * x[2] ** 2) < 1e-10 * r ** 4
        # k is a unit vector (l = (1, k) null for eta)
        assert abs(k0 * k0 + k1 * k1 + k2 * k2 - 1.0) < 1e-13
        assert H == pytest.approx(r ** 3 / (r ** 4 + a * a * x[2] ** 2), rel=1e-14)
        # det g = -1 for g = eta + 2H l l
        l = np.array([1.0, k0, k1, k2])
        g = np.diag([-1.0, 1, 1, 1]) + 2 * H * np.outer(l, l)
        assert np.linalg.det(g) == pytest.approx(-1.0, abs=1e-11)


# ------------------------------------------------------- Hamilton's equations
def test_rhs_is_hamiltonian_flow():
    """dx/dl = dHam/dp, dp/dl = -dHam/dx with 2 Ham = g^{mu nu} p p (the
    oracle's constraint), by 4th-order central differences."""
    rng = np.random.default_rng(12)
    for a in (0.0, 0.6, -0.95):
        met = rtsim.kerr_metric(1.0, a)
        for _ in range(20):
            x = rng.normal(size=3) * rng.uniform(3, 40)
            y = np.array([0.0, *x, -rng.uniform(0.5, 2), *rng.normal(size=3)])
            d = O.rhs(y, met)

            def grad(i):
                h = 1e-4 * max(1.0, abs(y[i]))
                f = []
                for s in (-2, -1, 1, 2):
                    z = y.copy()
                    z[i] += s * h
                    f.append(O.constraint(z, met))
                return (f[0] - 8 * f[1] + 8 * f[2] - f[3]) / (12 * h) / 2.0

            for i in range(1, 4):  # dp_i/dl = -dHam/dx_i
                assert d[4 + i] == pytest.approx(-grad(i), rel=1e-6, abs=1e-9)
            for i in range(4, 8):  # dx/dl = dHam/dp (index 4 -> t)
                assert d[i - 4] == pytest.approx(grad(i), rel=1e-7, abs=1e-9)


# ------------------------------------------------------------- DP45
def ode(kind, lam, y0, t0, h, n, s):
    ye, dn = np.zeros(2), np.zeros(2)
    e = O.oracle().rto_test_ode(kind, lam, ptr(np.array(y0, float)), t0, h, n, s, ye.ctypes.data_as(PD),
                                dn.ctypes.data_as(PD))
    return ye, dn, e


def test_dp45_exact_on_quartic_quadrature():
    """y' = 5 t^4 (y = t^5 needs order 5: exact); dense output is order 4, so
    exact for y' = 4 t^3 (y = t^4) at interior points."""
    ye, dn, _ = ode(1, 0.0, [0.0, 0.0], 0.0, 0.1, 10, 0.37)
    assert ye[0] == pytest.approx(1.0, abs=1e-14)   # t^5 at t = 1
    assert ye[1] == pytest.approx(1.0, abs=1e-14)   # t^4 at t = 1
    t = 0.9 + 0.37 * 0.1
    assert dn[1] == pytest.approx(t ** 4, abs=1e-14)


def test_dp45_order_five():
    errs = []
    for n in (8, 16, 32):
        ye, _, _ = ode(0, -1.3, [1.0, 2.0], 0.0, 1.0 / n, n, 0.5)
        errs.append(abs(ye[0] - math.exp(-1.3)))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 26 < r1 < 38 and 26 < r2 < 38  # 2^5 = 32


def test_dense_output_order_and_endpoints():
    errs = []
    for n in (8, 16, 32):
        _, dn, _ = ode(2, 0.0, [1.0, 0.0], 0.0, 1.0 / n, n, 0.5)
        t = 1.0 - 0.5 / n
        errs.append(abs(dn[0] - math.cos(t)))
    assert errs[0] / errs[1] > 14 and errs[1] / errs[2] > 14  # at least order 4 locally
    ye, d1, _ = ode(2, 0.0, [1.0, 0.0], 0.0, 0.1, 5, 1.0)
    assert d1[0] == pytest.approx(ye[0], abs=1e-15)


def test_embedded_error_estimate_order_four():
    e = [ode(0, -2.0, [1.0, 1.0], 0.0, h, 1, 0.5)[2] for h in (0.2, 0.1)]
    # rtol = atol = 1e-300, so err ~ |b5 - b4| part ~ h^5
    assert 20 < e[0] / e[1] < 45


def test_illinois_root():
    rng = np.random.default_rng(13)
    for _ in range(200):
        root = rng.uniform(0.01, 0.99)
        c3, c2 = rng.normal(), rng.normal()
        # p(s) = (s - root) * (c3 s^2 + c2 s + 5): one root in [0, 1]
        poly = np.polymul([1.0, -root], [c3 * 0.1, c2 * 0.1, 5.0])
        s = O.oracle().rto_test_root(ptr(poly[::-1].copy()))
        assert s == pytest.approx(root, abs=1e-13)


# ---------------------------------------------------- initial momentum
def test_camera_state_is_null_with_unit_energy():
    for a in (0.0, 0.9, 0.998):
        met = rtsim.kerr_metric(1.0, a)
        cam = rtsim.default_camera(30.0, math.radians(70), 0.2, math.radians(60), 5, 5)
        for px in range(5):
            ok, y, xi = O.camera_state(cam, met, rtsim.default_termination(), px, (px * 3) % 5)
            assert ok and y[4] == -1.0
            assert abs(O.constraint(y, met)) < 1e-13
            assert xi == pytest.approx(y[1] * y[6] - y[2] * y[5])


def test_center_pixel_points_at_the_hole():
    """Centre of an odd frame: direction = forward = -e_r. For a = 0 and an
    observer on the x axis (the first axis of the Gram-Schmidt frame, so the
    frame's first leg is radial) the photon has zero angular momentum."""
    met = rtsim.kerr_metric(1.0, 0.0)
    cam = rtsim.default_camera(500.0, math.radians(90), 0.0, math.radians(10), 9, 9)
    ok, y, xi = O.camera_state(cam, met, rtsim.default_termination(), 4, 4)
    x, p = y[1:4], y[5:8]
    assert np.linalg.norm(np.cross(x, p)) < 1e-12 * np.linalg.norm(x) * np.linalg.norm(p)


# ---------------------------------------------------- conserved quantities
def carter(y, a):
    """E, L_z and Carter's constant of a null ray from its KS state."""
    x, p = y[1:4], y[5:8]
    r = field(x, 0.0, a)[0]
    cth = x[2] / r
    sth2 = max(1.0 - cth * cth, 1e-30)
    p_th = (cth / math.sqrt(sth2)) * (x[0] * p[0] + x[1] * p[1]) - r * math.sqrt(sth2) * p[2]
    E, Lz = -y[4], x[0] * p[1] - x[1] * p[0]
    return E, Lz, p_th * p_th + cth * cth * (Lz * Lz / sth2 - a * a * E * E)


@pytest.mark.parametrize("a", [0.0, 0.7, 0.998])
def test_invariants_conserved_along_rays(a):
    met = rtsim.kerr_metric(1.0, a)
    cam = rtsim.default_camera(200.0, math.radians(66), 0.1, math.radians(12), 7, 7)
    tm = rtsim.default_termination(escape_radius=500.0, disk_enabled=False)
    for px, py in [(0, 0), (3, 2), (6, 5), (2, 6)]:
        ok, y0, _ = O.camera_state(cam, met, tm, px, py)
        out = O.trace_one(y0, met, rtsim.default_integrator(), tm)
        if out["cls"] & 7 == 1:
            continue
        E0, L0, Q0 = carter(y0, a)
        E1, L1, Q1 = carter(out["y"], a)
        scale = L0 * L0 + abs(Q0) + 1.0
        assert abs(E1 - E0) < 1e-9
        assert abs(L1 - L0) < 1e-8 * max(1.0, abs(L0))
        assert abs(Q1 - Q0) < 1e-7 * scale


# ------------------------------------------------- critical impact parameters
def equatorial_state(a, xi_target):
    """Equatorial ray from r = 1000 M whose L_z/E equals xi_target."""
    M = 1.0
    x = np.zeros(3)
    O.oracle().rto_ks_from_bl(1000.0, math.pi / 2, 0.0, a, x.ctypes.data_as(PD))

    def xi_of(alpha):
        out = np.array([-math.cos(alpha), math.sin(alpha), 0.0])
        p = np.zeros(3)
        O.oracle().rto_initial_momentum(ptr(x), M, a, ptr(out), p.ctypes.data_as(PD))
        return (x[0] * p[1] - x[1] * p[0]), p

    lo, hi = -0.02, 0.02
    flo = xi_of(lo)[0] - xi_target
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        fm = xi_of(mid)[0] - xi_target
        if (fm < 0) == (flo < 0):
            lo, flo = mid, fm
        else:
            hi = mid
    xi, p = xi_of(0.5 * (lo + hi))
    return np.array([0.0, *x, -1.0, *p]), xi


@pytest.mark.parametrize("a", [0.0, 0.5, 0.9])
def test_equatorial_critical_impact_parameters(a):
    """Circular photon orbits: |xi_c| = -+a + 6 M cos(arccos(-+a/M)/3)
    (prograde upper sign); rays 1e-3 inside are captured, 1e-3 outside escape."""
    met = rtsim.kerr_metric(1.0, a)
    tm = rtsim.default_termination(escape_radius=1000.0, disk_enabled=False)
    ig = rtsim.default_integrator()
    for sign in (+1, -1):
        b = -sign * a + 6.0 * math.cos(math.acos(-sign * a) / 3.0)
        for fac, want in ((1 - 1e-3, 1), (1 + 1e-3, 2)):
            y0, xi = equatorial_state(a, sign * b * fac)
            out = O.trace_one(y0, met, ig, tm)
            assert out["cls"] & 7 == want, f"a={a} sign={sign} xi={xi}: class {out['cls'] & 7}"


def test_schwarzschild_shadow_edge():
    """Observer at r = 50 M, a = 0: the shadow's angular radius is
    asin(3 sqrt(3) M sqrt(1 - 2M/r) / r) (tests/test_schwarzschild_shadow_render.cpp:30-31);
    pixels well inside are captured, well outside escape."""
    r = 50.0
    alpha = math.asin(3 * math.sqrt(3) * math.sqrt(1 - 2 / r) / r)
    n = 41
    fov = 4 * alpha
    cam = rtsim.default_camera(r, math.radians(90), 0.0, fov, n, 1)
    res = O.trace_camera(cam, rtsim.kerr_metric(1.0, 0.0), rtsim.default_integrator(),
                         rtsim.default_termination(escape_radius=1000.0, disk_enabled=False))
    tan_half = math.tan(fov / 2)
    for px in range(n):
        u = ((px + 0.5) / n * 2 - 1) * n  # aspect = n
        ang = math.atan(abs(u) * tan_half)
        if ang < 0.97 * alpha:
            assert res.klass[px] == 1, (px, ang, alpha)
        elif ang > 1.03 * alpha:
            assert res.klass[px] == 2, (px, ang, alpha)


# ------------------------------------------------- Carter radial classifier
def carter_captured(xi, eta, a, M=1.0):
    """Integration-free class of a ray from afar with constants (xi, eta)
    (CarterPhaseClassifier, optics/carter_ray_classifier.hpp:116-164): captured
    iff R(r) = (r^2 + a^2 - a xi)^2 - Delta (eta + (xi - a)^2) has no real root
    outside r_+. Returns (captured, margin) where margin = min R(r)/r^4 over
    r in [r_+, 50 M] (0 on the critical curve)."""
    K = eta + (xi - a) ** 2
    coeffs = [1.0, 0.0, 2 * a * a - 2 * a * xi - K, 2 * M * K, (a * a - a * xi) ** 2 - a * a * K]
    rp = M + math.sqrt(M * M - a * a)
    roots = np.roots(coeffs)
    real = roots[np.abs(roots.imag) < 1e-9].real
    captured = not np.any(real > rp)
    rr = np.linspace(rp, 50.0, 4000)
    R = np.polyval(coeffs, rr) / rr ** 4
    return captured, float(np.min(np.abs(R)))


@pytest.mark.parametrize("a,incl", [(0.0, 60.0), (0.9, 60.0), (0.998, 85.0), (0.6, 17.0)])
def test_classes_match_carter_classifier(a, incl):
    met = rtsim.kerr_metric(1.0, a)
    tm = rtsim.default_termination(escape_radius=1000.0, disk_enabled=False)
    n = 24
    cam = rtsim.default_camera(1000.0, math.radians(incl), 0.0, math.radians(1.6), n, n)
    res = O.trace_camera(cam, met, rtsim.default_integrator(), tm)
    checked = 0
    for i in range(n * n):
        ok, y0, _ = O.camera_state(cam, met, tm, i % n, i // n)
        E, Lz, Q = carter(y0, a)
        cap, margin = carter_captured(Lz / E, Q / (E * E), a)
        if margin < 1e-4:
            continue  # within reach of the critical curve
        checked += 1
        assert (res.klass[i] == 1) == cap, (i, Lz, Q, res.klass[i])
    assert checked > n * n * 0.8
    assert (res.klass == 1).sum() > 20 and (res.klass == 2).sum() > 20


# ------------------------------------------------------------ superposition
def rotation(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q @ np.diag(np.sign(np.diag(r)))
    return q if np.linalg.det(q) > 0 else -q


def sources_config(rng, k):
    src = [{"mass": float(rng.uniform(0.5, 1.5)), "spin": float(rng.uniform(-0.9, 0.9)), "axis": (0.0, 0.0, 1.0)}]
    for _ in range(k - 1):
        ax = rng.normal(size=3)
        src.append({"mass": float(rng.uniform(0.1, 1.0)), "spin": float(rng.uniform(-0.95, 0.95)),
                    "position": tuple(rng.uniform(-20, 20, size=3)), "axis": tuple(ax / np.linalg.norm(ax))})
    return src


def test_superposition_rotation_covariance():
    """Rotating every source centre and spin axis and the state by R rotates
    the right-hand side by R and leaves p_t's derivative and dt/dl unchanged:
    pins the per-source frames and the local <-> world maps (a transposed or
    mis-signed frame breaks it for general axes)."""
    rng = np.random.default_rng(21)
    for _ in range(60):
        src = sources_config(rng, int(rng.integers(2, 9)))
        R = rotation(rng)
        rot = [dict(s, position=tuple(R @ np.array(s.get("position", (0.0, 0.0, 0.0)))),
                    axis=tuple(R @ np.array(s["axis"]))) for s in src]
        j = int(rng.integers(0, len(src)))
        x = np.array(src[j].get("position", (0.0, 0.0, 0.0))) + rng.normal(size=3) * 4.0
        p = rng.normal(size=3)
        y = np.array([0.0, *x, -1.0, *p])
        yr = np.array([0.0, *(R @ x), -1.0, *(R @ p)])
        d = O.rhs(y, rtsim.multi_metric(src))
        dr = O.rhs(yr, rtsim.multi_metric(rot))
        scale = np.abs(d).max()
        assert abs(dr[0] - d[0]) <= 1e-12 * scale and dr[4] == d[4] == 0.0
        assert np.abs(dr[1:4] - R @ d[1:4]).max() <= 1e-12 * scale
        assert np.abs(dr[5:8] - R @ d[5:8]).max() <= 1e-12 * scale
        assert abs(O.constraint(yr, rtsim.multi_metric(rot)) - O.constraint(y, rtsim.multi_metric(src))) \
            <= 1e-12 * (1.0 + p @ p)


def test_superposition_spin_axis_flip():
    """A body of spin s about n is the same hole as spin -s about -n."""
    rng = np.random.default_rng(22)
    for _ in range(60):
        src = sources_config(rng, int(rng.integers(2, 6)))
        flip = [src[0]] + [dict(s, spin=-s["spin"], axis=tuple(-np.array(s["axis"]))) for s in src[1:]]
        x = np.array(src[-1]["position"]) + rng.normal(size=3) * 3.0
        y = np.array([0.0, *x, -1.0, *rng.normal(size=3)])
        d, df = O.rhs(y, rtsim.multi_metric(src)), O.rhs(y, rtsim.multi_metric(flip))
        assert np.abs(df - d).max() <= 1e-12 * np.abs(d).max()


@pytest.mark.parametrize("a", [0.0, 0.9])
def test_superposition_reduces_to_kerr(a):
    """A central hole plus a body of negligible mass far away: the same
    right-hand side as the single Kerr hole (to the body's ~1e-9 M / 1e6 M
    field), and the same rays: identical classes, disk hits at the same
    points, escapes within the escape tests' own difference (Euclidean
    |x| = r_esc against KS r = r_esc: |x| - r = a^2 sin^2(theta) / 2r).
    Redshifts are not compared: the superposition path takes xi = local L_z
    over |p| at the crossing (ms_disk_crossing, software_compute_engine.hpp:3021-3042),
    the single-hole path L_z / E of the ray (:1650)."""
    single = rtsim.kerr_metric(1.0, a)
    multi = rtsim.multi_metric([{"mass": 1.0, "spin": a},
                                {"mass": 1e-9, "spin": 0.0, "position": (1e6, 0.0, 0.0), "disk_enabled": False}])
    rng = np.random.default_rng(23)
    for _ in range(100):
        x = rng.normal(size=3)
        x *= rng.uniform(2.5, 300.0) / np.linalg.norm(x)
        y = np.array([0.0, *x, -1.0, *rng.normal(size=3)])
        d1, d2 = O.rhs(y, single), O.rhs(y, multi)
        assert np.abs(d2 - d1).max() <= 1e-13 * np.abs(d1).max() + 1e-14
    cam = rtsim.default_camera(60.0, math.radians(75), 0.0, math.radians(30), 9, 9)
    r_esc = 200.0
    tm = rtsim.default_termination(escape_radius=r_esc)
    xs, ps = [], []
    for i in range(81):
        ok, y0, _ = O.camera_state(cam, single, tm, i % 9, i // 9)
        xs.append(y0[:4])
        ps.append(y0[4:])
    xs, ps = np.array(xs).T.copy(), np.array(ps).T.copy()
    r1 = O.trace_rays(xs, ps, single, rtsim.default_integrator(), tm)
    r2 = O.trace_rays(xs, ps, multi, rtsim.default_integrator(), tm)
    assert np.array_equal(r1.klass, r2.klass)
    assert {1, 2, 4} <= set(r1.klass.tolist())
    disk, esc = r1.klass == 4, r1.klass == 2
    assert np.abs(r2.x[1:, disk] - r1.x[1:, disk]).max() <= 1e-9 * 20.0
    assert np.abs(r2.x[1:, esc] - r1.x[1:, esc]).max() <= a * a / (2 * r_esc) * 1.01 + 1e-9 * r_esc


def test_superposition_mirror_camera():
    """Two holes placed as mirror images through the plane x = 0: centres
    (+-9, 0, 0), spin 0.7 about n and about its mirror image (n_x, -n_y, -n_z)
    (spin is an axial vector), a light central hole without disk, and an
    observer on the y axis whose right vector is -x: pixel (W-1-i, j) sees the
    mirror image of pixel (i, j), with the same class and redshift. (The tilt
    keeps the two disks out of one plane: coplanar disks would tie and go to
    the lower source index.)"""
    n1 = np.array([0.3, 0.2, 1.0]) / math.sqrt(1.13)
    met = rtsim.multi_metric([
        {"mass": 0.2, "spin": 0.0, "disk_enabled": False},
        {"mass": 0.8, "spin": 0.7, "position": (9.0, 0.0, 0.0), "axis": tuple(n1)},
        {"mass": 0.8, "spin": 0.7, "position": (-9.0, 0.0, 0.0), "axis": (n1[0], -n1[1], -n1[2])}])
    n = 16
    cam = rtsim.default_camera(80.0, math.radians(70), math.radians(90), math.radians(40), n, n)
    assert np.allclose(cam.right[:], [-1.0, 0.0, 0.0], atol=1e-15)
    res = O.trace_camera(cam, met, rtsim.default_integrator(), rtsim.default_termination(escape_radius=300.0))
    k = res.klass.reshape(n, n)
    assert np.array_equal(k, k[:, ::-1])
    assert {1, 2, 4} <= set(res.klass.tolist())
    X = res.x.reshape(4, n, n)
    ok = k != 1
    assert np.abs(X[1][ok] + X[1][:, ::-1][ok]).max() < 1e-6
    assert np.abs(X[2][ok] - X[2][:, ::-1][ok]).max() < 1e-6
    assert np.abs(X[3][ok] - X[3][:, ::-1][ok]).max() < 1e-6
    g = res.g.reshape(n, n)
    assert np.abs(g - g[:, ::-1]).max() < 1e-7


# ------------------------------------------- Carter test (RT_TERM_CARTER_CAPTURE)
def test_carter_verdict_matches_sampled_potential():
    """The closed-form test (critical points of R by the cubic formula) against
    R sampled on a fine grid outside r_+: no root <=> verdict 1, a root <=> 0."""
    rng = np.random.default_rng(31)
    decided = 0
    for _ in range(3000):
        a = float(rng.uniform(-0.998, 0.998))
        xi = float(rng.uniform(-9, 9))
        eta = float(rng.uniform(-1, 40))
        v = O.oracle().rto_carter_verdict(1.0, a, xi, eta, 1e-6)
        cap, margin = carter_captured(xi, eta, a)
        if v == -1:
            assert margin < 1e-3, (a, xi, eta, margin)
            continue
        decided += 1
        assert (v == 1) == cap, (a, xi, eta, v, cap)
    assert decided > 2900


@pytest.mark.parametrize("a,incl,disk", [(0.998, 60.0, True), (0.5, 80.0, False), (0.0, 30.0, True)])
def test_carter_capture_only_ends_captured_rays(a, incl, disk):
    """With RT_TERM_CARTER_CAPTURE every ray the test ends is one that the full
    integration captures; every other ray's result is unchanged."""
    met = rtsim.kerr_metric(1.0, a)
    tm = rtsim.default_termination(disk_enabled=disk)
    tmc = rtsim.default_termination(disk_enabled=disk)
    tmc.flags |= rtsim.RT_TERM_CARTER_CAPTURE
    n = 24
    cam = rtsim.default_camera(1000.0, math.radians(incl), 0.0, math.radians(1.6), n, n)
    full = O.trace_camera(cam, met, rtsim.default_integrator(), tm)
    fast = O.trace_camera(cam, met, rtsim.default_integrator(), tmc)
    early = (fast.cls & rtsim.RT_FLAG_CARTER) != 0
    assert early.sum() >= max(5, 0.5 * (full.klass == 1).sum()), (early.sum(), (full.klass == 1).sum())
    assert np.all(full.klass[early] == 1) and np.all(fast.klass[early] == 1)
    assert np.array_equal(fast.klass, full.klass)
    same = ~early
    assert np.array_equal(fast.steps[same], full.steps[same]) and np.array_equal(fast.x[:, same], full.x[:, same])
    assert fast.steps[early].sum() < 0.2 * full.steps[early].sum()


def test_event_end_points_with_the_disk_on():
    """Where each class of ray ends, with the disk on, checked against the same
    ray traced with the disk off (which records every crossing of z = 0):
    - a disk hit ends on z = 0 (BL theta = pi/2) at the ray's FIRST crossing
      whose radius lies in [r_isco, r_out] and above 1.02 r_+, at that
      crossing's radius and affine length;
    - an escaped ray never crossed the disk, ends on r = r_esc, moving outward
      along the integration direction;
    - a captured ray never crossed the disk and ends on r = r_+ + eps.
    A disk taken at a later crossing, a crossing outside the radial range, an
    escape taken at the wrong radius or while moving inward, or a capture
    short of the horizon all fail these checks."""
    import ctypes as C

    lib = O.oracle()
    lib.rto_z_crossings.argtypes = [O.PD, C.POINTER(rtsim.rt_metric), C.POINTER(rtsim.rt_integrator),
                                    C.POINTER(rtsim.rt_termination), O.PD, O.PD, C.c_int]
    lib.rto_z_crossings.restype = C.c_int
    a = 0.9
    met = rtsim.kerr_metric(1.0, a)
    ig = rtsim.default_integrator()
    tm = rtsim.default_termination(escape_radius=400.0, disk_outer=15.0)
    rh = 1.0 + math.sqrt(1.0 - a * a)
    r_in, r_out = O.oracle().rto_isco(1.0, a), 15.0
    cam = rtsim.default_camera(200.0, math.radians(76.0), 0.4, math.radians(9.0), 28, 28)
    res = O.trace_camera(cam, met, ig, tm)
    counts = {1: 0, 2: 0, 4: 0}
    multi_crossing = 0
    rbuf, lbuf = np.zeros(512), np.zeros(512)
    for i in range(res.n):
        ok, y0, _ = O.camera_state(cam, met, tm, i % cam.width, i // cam.width)
        assert ok
        k = lib.rto_z_crossings(y0.ctypes.data_as(O.PD), C.byref(met), C.byref(ig), C.byref(tm),
                                rbuf.ctypes.data_as(O.PD), lbuf.ctypes.data_as(O.PD), 512)
        inside = [(r, l) for r, l in zip(rbuf[:k], lbuf[:k]) if r_in <= r <= r_out and r > 1.02 * rh]
        cls = int(res.klass[i])
        counts[cls] += 1
        r_end, th_end, lam = res.bl[0, i], res.bl[1, i], res.lam[i]
        if cls == rtsim.RT_CLS_DISK:
            assert inside, f"ray {i}: disk hit without a crossing in range"
            r_c, l_c = inside[0]
            assert abs(res.x[3, i]) <= 1e-9 * r_end and abs(th_end - math.pi / 2) <= 1e-9
            assert abs(r_end - r_c) <= 1e-7 * r_c and abs(lam - l_c) <= 1e-7 * l_c, (i, r_end, r_c, lam, l_c)
            assert r_in <= r_end <= r_out and r_end > 1.02 * rh
            multi_crossing += len(inside) > 1
        else:
            assert not [1 for r, l in inside if l <= lam * (1 + 1e-12)], f"ray {i}: class {cls} crossed the disk"
            if cls == rtsim.RT_CLS_ESCAPED:
                assert abs(r_end - 400.0) <= 1e-10 * 400.0
                d = O.rhs(np.concatenate([res.x[:, i], res.p[:, i]]), met)
                assert -(res.x[1:, i] @ d[1:4]) > 0.0, f"ray {i}: escaped moving inward"
            else:
                assert abs(r_end - (rh + 1e-6)) <= 1e-9, (i, r_end - rh)
    print(f"classes {counts}, disk rays with more than one in-range crossing: {multi_crossing}")
    assert min(counts.values()) >= 20 and multi_crossing >= 5
