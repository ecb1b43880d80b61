"""The Kerr critical curve and the distance of rays to it (rt_shadow_boundary,
rt_impact_distance, rt_critical_distance), host-side.

Pinned by: the Schwarzschild circle of radius sqrt(27) M (distance
|b - sqrt(27) M| for impact parameter b, area 27 pi M^2, as the reference's
test_kerr_shadow_eht.cpp); points of the curve at distance 0 and points moved
along the curve's normal at the distance moved; the reference's own
BardeenKerrShadow boundary and area where oracle/_ref is built; and the
oracle's traced classes: on a frame across the shadow edge, captured and
escaped neighbours meet only where the distance is within a pixel's extent,
and the captured pixels cover the curve's area."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib as O
import rtsim


def polygon_area(al, be):
    return 0.5 * abs(np.dot(al, np.roll(be, -1)) - np.dot(be, np.roll(al, -1)))


def test_schwarzschild_circle():
    rs = math.sqrt(27.0)
    for al, be in [(0.0, 0.0), (3.0, 4.0), (rs, 0.0), (-10.0, 2.5), (0.3, -rs)]:
        assert rtsim.impact_distance(1.0, 0.0, 1.1, al, be) == pytest.approx(abs(math.hypot(al, be) - rs), abs=1e-14)
    al, be = rtsim.shadow_boundary(1.0, 0.0, 1.1, 4000)
    assert polygon_area(al, be) == pytest.approx(27.0 * math.pi, rel=1e-5)
    al, be = rtsim.shadow_boundary(2.0, 0.0, 0.4, 4000)
    assert polygon_area(al, be) == pytest.approx(4 * 27.0 * math.pi, rel=1e-5)


@pytest.mark.parametrize("a,theta", [(0.5, math.pi / 2), (0.9, 1.0), (0.998, 1.48), (-0.7, 0.6), (0.9, 0.3)])
def test_curve_points_and_normal_offsets(a, theta):
    al, be = rtsim.shadow_boundary(1.0, a, theta, 2001)
    for i in range(5, 990, 37):
        assert rtsim.impact_distance(1.0, a, theta, al[i], be[i]) < 1e-9
        # unit normal from the neighbours
        ta, tb = al[i + 1] - al[i - 1], be[i + 1] - be[i - 1]
        nrm = math.hypot(ta, tb)
        na, nb = -tb / nrm, ta / nrm
        for d in (1e-3, 1e-2):
            got = rtsim.impact_distance(1.0, a, theta, al[i] + d * na, be[i] + d * nb)
            assert got == pytest.approx(d, rel=2e-2), (i, d, got)


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference tree absent)")
@pytest.mark.parametrize("a,theta", [(0.5, math.pi / 2), (0.9, math.radians(17.0)), (0.9999, math.pi / 2),
                                     (0.998, 1.48), (-0.6, 1.2)])
def test_boundary_against_reference(a, theta):
    """The reference's BardeenKerrShadow points all lie on this curve; the
    areas agree (the reference's test_kerr_shadow_eht.cpp compares areas to 1e-5)."""
    ref = O.ref()
    ref.ref_shadow_boundary.argtypes = [C.c_double, C.c_double, C.c_double, C.c_size_t, O.PD, O.PD]
    ref.ref_shadow_boundary.restype = C.c_double
    n = 2000
    ra, rb = np.zeros(n), np.zeros(n)
    area = ref.ref_shadow_boundary(1.0, a, theta, n, ra.ctypes.data_as(O.PD), rb.ctypes.data_as(O.PD))
    d = [rtsim.impact_distance(1.0, a, theta, x, y) for x, y in zip(ra, rb)]
    assert max(d) < 1e-6, max(d)
    al, be = rtsim.shadow_boundary(1.0, a, theta, n)
    assert polygon_area(al, be) == pytest.approx(area, rel=1e-5)


def test_distance_marks_the_class_boundary():
    """A 96 x 96 frame across the shadow of a = 0.9 at 70 deg (disk off):
    class changes between 4-neighbours only where the distance is below the
    pixel's extent in the impact-parameter plane, and the captured pixels'
    area matches the curve's."""
    a, theta = 0.9, math.radians(70.0)
    r_obs, fov, side = 1000.0, math.radians(1.4), 96
    cam = rtsim.default_camera(r_obs, theta, 0.0, fov, side, side)
    met = rtsim.kerr_metric(1.0, a)
    tm = rtsim.default_termination(disk_enabled=False)
    cls = O.trace_camera(cam, met, rtsim.default_integrator(), tm).klass.reshape(side, side)
    dist = rtsim.critical_distance(cam, met).reshape(side, side)
    pix = r_obs * 2.0 * math.tan(fov / 2) / side  # pixel extent in units of M at the observer
    edges = np.zeros_like(cls, bool)
    edges[:, 1:] |= cls[:, 1:] != cls[:, :-1]
    edges[:, :-1] |= cls[:, 1:] != cls[:, :-1]
    edges[1:, :] |= cls[1:, :] != cls[:-1, :]
    edges[:-1, :] |= cls[1:, :] != cls[:-1, :]
    assert edges.sum() > 50
    assert dist[edges].max() < 1.5 * pix, (dist[edges].max(), pix)
    assert dist.min() < 0.05 * pix
    al, be = rtsim.shadow_boundary(1.0, a, theta, 4000)
    captured_area = (cls == rtsim.RT_CLS_CAPTURED).sum() * pix * pix
    assert captured_area == pytest.approx(polygon_area(al, be), rel=0.03)


def test_row_window_count_to_the_last_row():
    """A row count that runs past the last row (up to 2^32 - 1) means "to the
    last row", as the reference's row window takes offset + count in 64 bits:
    the same distances as the exact count, and the oracle's frame the same.
    A count of 0 is the whole frame, whatever the offset."""
    met = rtsim.kerr_metric(1.0, 0.9)
    for count in (0xFFFFFFFF, 0xFFFFFFF0, 100):
        far = rtsim.default_camera(30.0, 1.1, 0.2, 0.9, 12, 20, row_offset=5, row_count=count)
        exact = rtsim.default_camera(30.0, 1.1, 0.2, 0.9, 12, 20, row_offset=5, row_count=15)
        assert rtsim.n_rays_of(far) == 15 * 12
        assert np.array_equal(rtsim.critical_distance(far, met), rtsim.critical_distance(exact, met))
        ig, tm = rtsim.default_integrator(), rtsim.default_termination(escape_radius=100.0)
        a, b = O.trace_camera(far, met, ig, tm), O.trace_camera(exact, met, ig, tm)
        assert np.array_equal(a.cls, b.cls) and np.array_equal(a.x, b.x)
        ra = O.trace_render_camera(far, met, rtsim.render_integrator(), tm)
        rb = O.trace_render_camera(exact, met, rtsim.render_integrator(), tm)
        assert np.array_equal(ra.cls, rb.cls) and np.array_equal(ra.steps, rb.steps)
    # count 0: the whole frame, the offset ignored
    whole = rtsim.default_camera(30.0, 1.1, 0.2, 0.9, 12, 20)
    shifted = rtsim.default_camera(30.0, 1.1, 0.2, 0.9, 12, 20, row_offset=5)
    assert rtsim.n_rays_of(shifted) == 20 * 12
    dw, ds = rtsim.critical_distance(whole, met), rtsim.critical_distance(shifted, met)
    assert np.array_equal(dw, ds) and np.array_equal(dw[5 * 12:], rtsim.critical_distance(exact, met))
