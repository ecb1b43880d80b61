"""Edge cases on the GPU: frames smaller than one 8x8 tile and ragged sizes,
explicit-ray counts that do not fill a wave, a row window whose count runs
past the last row, requests the C ABI must refuse,
a one-request batch, and a full-size frame of the precision configuration
(Kerr a = 0.998, i = 85 deg, 4096^2) checked through properties that do not
depend on its size plus every 64th row against the oracle."""
import math

import numpy as np
import pytest

import oracle_lib as O
import rtsim
from test_gpu_parity import compare

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("w,h", [(1, 1), (1, 9), (9, 1), (7, 9), (13, 3)])
def test_tiny_and_ragged_frames(ctx, w, h):
    cam = rtsim.default_camera(30.0, math.radians(70.0), 0.4, math.radians(40.0), w, h)
    met = rtsim.kerr_metric(1.0, 0.7)
    ig, tm = rtsim.default_integrator(), rtsim.default_termination(escape_radius=100.0)
    compare(ctx.trace_camera(cam, met, ig, tm), O.trace_camera(cam, met, ig, tm), f"{w}x{h}")


@pytest.mark.parametrize("n", [1, 63, 65, 130])
def test_ragged_ray_counts(ctx, n):
    cam = rtsim.default_camera(40.0, math.radians(60.0), 0.0, math.radians(50.0), 16, 16)
    met = rtsim.kerr_metric(1.0, 0.9)
    ig, tm = rtsim.default_integrator(), rtsim.default_termination(escape_radius=200.0)
    x, p = np.zeros((4, n)), np.zeros((4, n))
    for i in range(n):
        ok, y, _ = O.camera_state(cam, met, tm, (7 * i) % 16, (3 * i) % 16)
        x[:, i], p[:, i] = y[:4], y[4:]
    compare(ctx.trace_rays(x, p, met, ig, tm), O.trace_rays(x, p, met, ig, tm), f"{n} rays")


def test_row_count_past_the_last_row(ctx):
    """row_offset + row_count beyond 2^32: the window ends at the last row (the
    reference's 64-bit row window), alone and in a batch; row_count 0 traces
    the whole frame whatever row_offset says."""
    met, ig = rtsim.kerr_metric(1.0, 0.7), rtsim.default_integrator()
    tm = rtsim.default_termination(escape_radius=100.0)
    far = rtsim.default_camera(30.0, math.radians(70.0), 0.4, math.radians(40.0), 24, 20, row_offset=7,
                               row_count=0xFFFFFFFF)
    exact = rtsim.default_camera(30.0, math.radians(70.0), 0.4, math.radians(40.0), 24, 20, row_offset=7,
                                 row_count=13)
    a, b = ctx.trace_camera(far, met, ig, tm), ctx.trace_camera(exact, met, ig, tm)
    assert a.n == b.n == 13 * 24
    for f in ["cls", "steps", "x", "p", "bl", "g"]:
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
    outs = ctx.trace_batch([rtsim.make_request(far, met, ig, tm)])
    assert np.array_equal(outs[0].cls, b.cls) and np.array_equal(outs[0].x, b.x)
    # count 0 is the whole frame whatever the offset (resolve_row_window)
    whole = rtsim.default_camera(30.0, math.radians(70.0), 0.4, math.radians(40.0), 24, 20)
    shifted = rtsim.default_camera(30.0, math.radians(70.0), 0.4, math.radians(40.0), 24, 20, row_offset=7)
    c, d = ctx.trace_camera(shifted, met, ig, tm), ctx.trace_camera(whole, met, ig, tm)
    assert c.n == d.n == 20 * 24 and np.array_equal(c.x, d.x) and np.array_equal(c.cls, d.cls)
    assert np.array_equal(d.x[:, 7 * 24:], b.x)


def test_invalid_requests_are_refused(ctx):
    met, ig, tm = rtsim.kerr_metric(1.0, 0.5), rtsim.default_integrator(), rtsim.default_termination()
    cam = rtsim.default_camera(100.0, 1.0, 0.0, 0.3, 8, 8)
    bad_cams = []
    c = rtsim.default_camera(100.0, 1.0, 0.0, 0.3, 0, 8)
    bad_cams.append(c)
    c = rtsim.default_camera(100.0, 1.0, 0.0, 0.0, 8, 8)  # zero field of view
    bad_cams.append(c)
    c = rtsim.default_camera(100.0, 1.0, 0.0, 0.3, 8, 8, row_offset=8, row_count=1)  # empty row window
    bad_cams.append(c)
    for c in bad_cams:
        with pytest.raises(rtsim.RtError):
            ctx.trace_camera(c, met, ig, tm)
    with pytest.raises(rtsim.RtError):
        ctx.trace_camera(cam, met, rtsim.default_integrator(rtol=0.0), tm)
    with pytest.raises(rtsim.RtError):
        ctx.trace_camera(cam, rtsim.kerr_metric(0.0, 0.0), ig, tm)
    too_many = rtsim.multi_metric([{"mass": 1.0}] + [{"mass": 0.1, "position": (5.0 * i, 0, 0)} for i in range(1, 16)])
    too_many.n_sources = 17
    with pytest.raises(rtsim.RtError):
        ctx.trace_camera(cam, too_many, ig, tm)
    with pytest.raises(rtsim.RtError):
        ctx.trace_rays(np.zeros((4, 0)), np.zeros((4, 0)), met, ig, tm)
    with pytest.raises(rtsim.RtError):
        ctx.trace_batch([])
    # the context still works after the refusals
    compare(ctx.trace_camera(cam, met, ig, tm), O.trace_camera(cam, met, ig, tm), "after errors")


def test_pixel_records_match_results(ctx):
    """rt_fetch_pixels: the reference's 32-byte GpuPixelOutput records packed
    on the device from the same resident results rt_fetch_results returns."""
    cam = rtsim.default_camera(1000.0, math.radians(60.0), 0.0, math.radians(1.6), 37, 21)
    met, ig, tm = rtsim.kerr_metric(1.0, 0.9), rtsim.default_integrator(), rtsim.default_termination()
    n = ctx.trace_camera_resident(cam, met, ig, tm)
    res, pix = ctx.fetch(n), ctx.fetch_pixels(n)
    assert {1, 2, 4} <= set(res.klass.tolist())
    assert np.array_equal(pix["status_flags"], res.cls & 0xFF)
    assert np.array_equal(pix["iterations_used"], res.steps)
    assert np.array_equal(pix["redshift"], res.g.astype(np.float32))
    assert np.array_equal(pix["affine_parameter"], res.lam.astype(np.float32))
    assert np.all(pix["a"] == 1.0) and not np.any(pix["r"]) and not np.any(pix["g"]) and not np.any(pix["b"])


def test_one_request_batch(ctx):
    cam = rtsim.default_camera(25.0, math.radians(80.0), 0.0, math.radians(30.0), 3, 5)
    req = rtsim.make_request(cam, rtsim.kerr_metric(1.0, 0.3), rtsim.default_integrator(),
                             rtsim.default_termination(escape_radius=100.0))
    out = ctx.trace_batch([req])[0]
    compare(out, O.trace_camera(req.camera, req.metric, req.integrator, req.termination), "one-request batch")


def test_full_size_precision_frame(ctx):
    """4096^2, Kerr a = 0.998, i = 85 deg (16.8 M rays): every ray classified,
    flags clean, redshift per class, step totals equal the kernel's counters,
    and every 64th row identical to the oracle's to the parity bar."""
    side = 4096
    cam = rtsim.default_camera(1000.0, math.radians(85.0), 0.0, math.radians(20.0), side, side)
    met, ig, tm = rtsim.kerr_metric(1.0, 0.998), rtsim.default_integrator(), rtsim.default_termination()
    n = ctx.trace_camera_resident(cam, met, ig, tm)
    st = ctx.stats()
    assert n == side * side
    res = ctx.fetch(n)
    print(f"[4096^2] kernel {st.kernel_ms:.1f} ms, classes {np.bincount(res.klass, minlength=5)[[1, 2, 4]]}", flush=True)
    assert set(np.unique(res.klass).tolist()) <= {1, 2, 4}
    assert not np.any(res.cls & (rtsim.RT_FLAG_NONFINITE | rtsim.RT_FLAG_INVALID))
    assert int(res.steps.sum()) == st.accepted_steps and int(res.rejected.sum()) == st.rejected_steps
    assert np.all(res.steps >= 1)
    esc, cap, disk = res.klass == 2, res.klass == 1, res.klass == 4
    assert np.all(res.g[cap] == 0.0) and np.allclose(res.g[esc], 1.0 / -res.p[0, esc])
    assert np.all((res.g[disk] >= 0.05) & (res.g[disk] <= 4.0))
    assert disk.sum() > 1000 and cap.sum() > 1000
    rows = np.arange(0, side, 64)
    ref = O.trace_camera_rows(cam, met, ig, tm, 0, 64, rows.size)
    idx = (rows[:, None] * side + np.arange(side)[None, :]).ravel()
    sub = rtsim.Results(idx.size)
    sub.cls[:] = res.cls[idx]
    sub.x[:], sub.p[:], sub.bl[:] = res.x[:, idx], res.p[:, idx], res.bl[:, idx]
    sub.lam[:], sub.g[:], sub.steps[:], sub.rejected[:] = res.lam[idx], res.g[idx], res.steps[idx], res.rejected[idx]
    compare(sub, ref, "4096^2 every 64th row")
