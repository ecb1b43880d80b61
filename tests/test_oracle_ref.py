"""The CPU oracle against the reference's own functions, compiled from the
reference headers into oracle/_ref/libref.so (skipped where that library is
absent). Bit-identical where the oracle restates the same expression; the
whole-ray integration is compared with the reference's Christoffel-based
RK45AdaptiveIntegrator (integrators/rk45_adaptive.hpp) on KerrSchildMetric."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib as O
import rtsim

pytestmark = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref/libref.so not built")
PD = C.POINTER(C.c_double)


def ptr(a):
    return a.ctypes.data_as(PD)


def rand_states(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        a = rng.uniform(-0.999, 0.999)
        R = 10 ** rng.uniform(0.2, 3.5)
        x = rng.normal(size=3)
        x *= R / np.linalg.norm(x)
        out.append((a, x, rng.normal(size=3)))
    return out


def test_ks_field_bitwise():
    for a, x, _ in rand_states(500, 1):
        o, r = np.zeros(5), np.zeros(5)
        O.oracle().rto_ks_field(ptr(x), 1.3, a, ptr(o))
        O.ref().ref_ks_field(ptr(x), 1.3, a, ptr(r))
        assert np.array_equal(o, r)


def test_rhs_bitwise():
    for a, x, p in rand_states(500, 2):
        dx, dp = np.zeros(3), np.zeros(3)
        O.ref().ref_ks_rhs(ptr(x), ptr(p), 1.0, a, ptr(dx), ptr(dp))
        d = O.rhs(np.array([0.0, *x, -1.0, *p]), rtsim.kerr_metric(1.0, a))
        assert np.array_equal(d[1:4], dx) and np.array_equal(d[5:8], dp)


def test_initial_momentum_bitwise():
    rng = np.random.default_rng(3)
    for a, x, _ in rand_states(300, 3):
        out = rng.normal(size=3)
        out /= np.linalg.norm(out)
        po, pr = np.zeros(3), np.zeros(3)
        O.oracle().rto_initial_momentum(ptr(x), 1.0, a, ptr(out), ptr(po))
        O.ref().ref_initial_momentum(ptr(x), 1.0, a, ptr(out), ptr(pr))
        assert np.array_equal(po, pr)


def test_coordinate_conversions_bitwise():
    rng = np.random.default_rng(4)
    for _ in range(300):
        a = rng.uniform(-0.99, 0.99)
        r, th, ph = rng.uniform(1.5, 500), rng.uniform(0.01, 3.13), rng.uniform(-3.1, 3.1)
        o, f = np.zeros(3), np.zeros(3)
        O.oracle().rto_ks_from_bl(r, th, ph, a, ptr(o))
        O.ref().ref_ks_from_bl(r, th, ph, a, ptr(f))
        assert np.array_equal(o, f)
        b1, b2 = np.zeros(3), np.zeros(3)
        O.oracle().rto_bl_from_ks(ptr(o), a, ptr(b1))
        O.ref().ref_bl_from_ks(ptr(o), a, 0.0, ptr(b2))
        assert np.array_equal(b1, b2)


def test_isco_and_redshift_bitwise():
    rng = np.random.default_rng(5)
    for _ in range(300):
        m = rng.uniform(0.1, 3)
        a = rng.uniform(-0.999, 0.999) * m
        assert O.oracle().rto_isco(m, a) == O.ref().ref_isco(m, a)
        r = rng.uniform(1.0, 30.0) * m
        xi = rng.uniform(-8, 8) * m
        assert O.oracle().rto_redshift(m, a, r, xi, 1) == O.ref().ref_redshift(m, a, r, xi)


@pytest.mark.parametrize("a", [0.0, 0.5, 0.9, 0.998])
def test_integration_matches_reference_rk45(a):
    """Backward ray from a camera at r = 60 M, 40 M of affine length, against
    the reference's own adaptive DP45 on the Christoffel form (forward on the
    time-reversed momentum, which traces the same curve)."""
    lib = O.oracle()
    lib.rto_integrate_to.argtypes = [PD, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.c_double, PD]
    lib.rto_integrate_to.restype = C.c_int
    cam = rtsim.default_camera(60.0, math.radians(63.0), 0.3, math.radians(40.0), 9, 9)
    for px, py in [(4, 4), (1, 6), (7, 2)]:
        ok, y, _ = O.camera_state(cam, rtsim.kerr_metric(1.0, a), rtsim.default_termination(), px, py)
        yf = y.copy()
        yf[4:] = -yf[4:]
        out = np.zeros(8)
        assert lib.rto_integrate_to(ptr(yf), 1.0, a, 40.0, 1e-12, 1e-14, 0.5, ptr(out)) > 0
        g, gi = np.zeros(16), np.zeros(16)
        O.ref().ref_ks_metric(ptr(yf[:4].copy()), 1.0, a, ptr(g), ptr(gi))
        u = gi.reshape(4, 4) @ yf[4:]
        xo, uo = np.zeros(4), np.zeros(4)
        acc, rej = C.c_uint64(), C.c_uint64()
        assert O.ref().ref_rk45_integrate(ptr(yf[:4].copy()), ptr(u), 1.0, a, 40.0, 1e-12, 1e-14, 0.5, ptr(xo),
                                          ptr(uo), C.byref(acc), C.byref(rej)) == 0
        O.ref().ref_ks_metric(ptr(out[:4].copy()), 1.0, a, ptr(g), ptr(gi))
        u_end = gi.reshape(4, 4) @ out[4:]
        assert np.abs(out[:4] - xo).max() < 1e-10 * max(1.0, np.abs(xo).max())
        assert np.abs(u_end - uo).max() < 1e-10
