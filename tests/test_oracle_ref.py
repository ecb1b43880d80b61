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


def rand_multi(rng, n_src=None):
    """A superposition like the serving workload's: central hole plus bodies
    with masses U[0.1, 1], positions U[-25, 25]^3, spins U[-0.95, 0.95] on
    random axes, some disks off."""
    k = int(rng.integers(2, 17)) if n_src is None else n_src
    src = [{"mass": float(rng.uniform(0.5, 2.0)), "spin": float(rng.uniform(-0.99, 0.99)),
            "disk_enabled": bool(rng.random() < 0.7)}]
    for _ in range(k - 1):
        src.append({"mass": float(rng.uniform(0.1, 1.0)), "spin": float(rng.uniform(-0.95, 0.95)),
                    "position": tuple(float(v) for v in rng.uniform(-25, 25, size=3)),
                    "axis": tuple(float(v) for v in rng.normal(size=3)),
                    "disk_enabled": bool(rng.random() < 0.7)})
    return rtsim.multi_metric(src)


def test_ms_source_table_bitwise():
    """build_multi_source_data (render/software_compute_engine.hpp:2738-2798):
    centres, masses, clamped spins, horizons, rotated frames and disk ranges."""
    rng = np.random.default_rng(11)
    tm = rtsim.default_termination()
    for _ in range(200):
        met = rand_multi(rng)
        o, r = np.zeros(17 * 16), np.zeros(17 * 16)
        n = O.oracle().rto_ms_table(C.byref(met), C.byref(tm), ptr(o))
        assert n == met.n_sources == O.ref().ref_ms_table(C.byref(met), ptr(r))
        assert np.array_equal(o[:17 * n], r[:17 * n])


def test_ms_rhs_matches_reference():
    """ms_rhs (software_compute_engine.hpp:2857-2870) at p_t = -1, at points
    within a few horizon radii of the sources and far away. The oracle sums
    the same per-source terms in its own algebraic order (the core KS form of
    kerr_schild_null_derivatives, not the renderer's ks_source_terms), so the
    two agree to rounding, not bit for bit."""
    rng = np.random.default_rng(12)
    worst = 0.0
    for _ in range(400):
        met = rand_multi(rng)
        j = int(rng.integers(0, met.n_sources))
        c = np.array(met.src[j].position[:])
        x = c + rng.normal(size=3) * rng.choice([3.0, 10.0, 100.0])
        p = rng.normal(size=3)
        dx, dp = np.zeros(3), np.zeros(3)
        O.ref().ref_ms_rhs(C.byref(met), ptr(x), ptr(p), ptr(dx), ptr(dp))
        d = O.rhs(np.array([0.0, *x, -1.0, *p]), met)
        ref = np.concatenate([dx, dp])
        got = np.concatenate([d[1:4], d[5:8]])
        worst = max(worst, float(np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300 + np.abs(ref).max()))))
    assert worst < 1e-14, worst


def test_ms_initial_momentum_matches_reference():
    """ms_initial_momentum (software_compute_engine.hpp:2919-2964) at observer
    positions 40-1000 M out, random outward directions: bit-identical."""
    rng = np.random.default_rng(13)
    for _ in range(300):
        met = rand_multi(rng)
        x = rng.normal(size=3)
        x *= rng.uniform(40.0, 1000.0) / np.linalg.norm(x)
        out = rng.normal(size=3)
        out /= np.linalg.norm(out)
        po, pr = np.zeros(3), np.zeros(3)
        O.oracle().rto_ms_initial_momentum(C.byref(met), ptr(x), ptr(out), ptr(po))
        O.ref().ref_ms_initial_momentum(C.byref(met), ptr(x), ptr(out), ptr(pr))
        assert np.array_equal(po, pr)
