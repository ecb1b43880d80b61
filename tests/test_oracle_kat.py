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
