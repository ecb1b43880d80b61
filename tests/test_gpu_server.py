"""Serving loop (rt_server_*): requests admitted by HBM footprint run at once
next to those on the device.

Every request's result must be the same alone (single-request kernel) and in
the mix of concurrently served requests (class identical, coordinates within
1e-12 relative); a request too large for the budget is refused with
RT_ERR_BUSY; a cancelled request reports RT_ERR_CANCELLED and one past its
deadline RT_ERR_EXPIRED; only the outputs asked for come back; requests of
the renderer's RK4 method are served with identical records. The request
states, admission and deadlines themselves are checked on the CPU with an
injected clock (tests/test_serve_core.py)."""
import math
import time

import numpy as np
import pytest

import rtsim

pytestmark = pytest.mark.gpu


def req(side, a, inc, r_obs=1000.0, rtol=1e-10):
    cam = rtsim.default_camera(r_obs, math.radians(inc), 0.0, math.radians(20.0), side, side)
    return rtsim.make_request(cam, rtsim.kerr_metric(1.0, a), rtsim.default_integrator(rtol=rtol),
                              rtsim.default_termination(escape_radius=max(1000.0, r_obs)))


def test_alone_equals_mix(ctx):
    reqs = [req(48, 0.9, 60), req(33, 0.0, 20), req(64, 0.998, 85), req(40, 0.5, 45, 300.0, 1e-8),
            req(16, 0.3, 70)]
    alone = [ctx.trace_camera(r.camera, r.metric, r.integrator, r.termination) for r in reqs]
    with rtsim.Server(0, hbm_budget_bytes=64 << 20, max_inflight=8) as srv:
        ids = []
        for r in reqs:
            ids.append(srv.submit(r))
            time.sleep(0.002)  # arrivals while earlier requests run
        for i, rid in enumerate(ids):
            srv.wait(rid, 60000)
            got = srv.take(rid)
            assert np.array_equal(got.cls, alone[i].cls), f"request {i}: class differs"
            for name in ("x", "p", "bl", "lam", "g"):
                u, v = getattr(got, name), getattr(alone[i], name)
                d = (np.abs(u - v) / np.maximum(np.abs(v), 1.0)).max(initial=0)
                assert d < 1e-12, f"request {i}: {name} differs by {d:.3e}"


def test_admission_and_cancel():
    with rtsim.Server(0, hbm_budget_bytes=2 << 20, max_inflight=1) as srv:
        with pytest.raises(rtsim.RtError) as e:
            srv.submit(req(512, 0.5, 60))  # ~30 MB of results: never fits 2 MiB
        assert e.value.status == 7
        first = srv.submit(req(48, 0.9, 60))
        second = srv.submit(req(40, 0.9, 60))
        srv.cancel(second)
        srv.wait(first, 60000)
        srv.take(first)
        with pytest.raises(rtsim.RtError) as e2:
            srv.wait(second, 60000)
        assert e2.value.status == 5
        with pytest.raises(rtsim.RtError):
            srv.poll(123456)


def test_outputs_subset_and_take_errors(ctx):
    """Only the outputs asked for at submit come back (the pixel record's four
    here), equal to the full run's; asking take for another output is refused
    and leaves the request in place."""
    import ctypes as C
    r = req(64, 0.9, 70)
    with rtsim.Server(0, hbm_budget_bytes=64 << 20, max_inflight=8) as srv:
        full_id = srv.submit(r)
        pix_id = srv.submit(r, outputs=rtsim.RT_OUT_PIXEL)
        srv.wait(full_id, 60000)
        srv.wait(pix_id, 60000)
        full = srv.take(full_id)
        n = rtsim.n_rays_of(r.camera)
        bad = rtsim.Results(n)  # every field requested: x, p, bl and rejected were not asked for
        assert srv.lib.rt_server_take(srv.h, pix_id, C.byref(bad.struct)) == 1
        pix = srv.take(pix_id)
    for f in ("cls", "lam", "g", "steps"):
        assert np.array_equal(getattr(pix, f), getattr(full, f)), f
    assert not pix.x.any() and not pix.p.any() and not pix.rejected.any()


def test_deadline_cancel_and_shutdown_with_work_on_the_device(ctx):
    """A deadline shorter than the trace ends the request RT_ERR_EXPIRED; a
    cancelled large request ends RT_ERR_CANCELLED; requests queued after them
    still finish with the right results; destroying the server with work on
    the device waits for it."""
    big = req(1024, 0.998, 85)
    small = req(48, 0.5, 60)
    alone = ctx.trace_camera(small.camera, small.metric, small.integrator, small.termination)
    with rtsim.Server(0, hbm_budget_bytes=1 << 30, max_inflight=8) as srv:
        expired = srv.submit(big, deadline_ms=1)
        cancelled = srv.submit(big)
        srv.cancel(cancelled)
        after = srv.submit(small)
        for rid, status in ((expired, 9), (cancelled, 5)):
            with pytest.raises(rtsim.RtError) as e:
                srv.wait(rid, 60000)
            assert e.value.status == status
            with pytest.raises(rtsim.RtError) as e2:
                srv.take(rid)
            assert e2.value.status == status
        srv.wait(after, 60000)
        got = srv.take(after)
        assert np.array_equal(got.cls, alone.cls) and np.array_equal(got.x, alone.x)
        for _ in range(3):
            srv.submit(big)  # left on the device: the server's exit waits for them


def test_render_requests_served(ctx):
    """RT_METHOD_RENDER_RK4 requests served beside DP45 ones: records and end
    states identical to the same request traced alone (the DP45 ones within
    1e-12, as above: the serving loop traces them with its queue kernel)."""
    def render_req(side, a, inc, phi):
        cam = rtsim.default_camera(30.0, math.radians(inc), phi, math.radians(60.0), side, side + 7)
        tm = rtsim.default_termination(escape_radius=100.0, disk_outer=24.0)
        return rtsim.make_request(cam, rtsim.kerr_metric(1.0, a), rtsim.render_integrator(), tm)

    reqs = [render_req(40, 0.9, 65, 0.3), req(48, 0.9, 60), render_req(56, -0.6, 80, -1.0), req(24, 0.5, 30)]
    alone = [ctx.trace_camera(r.camera, r.metric, r.integrator, r.termination) for r in reqs]
    with rtsim.Server(0, hbm_budget_bytes=64 << 20, max_inflight=8) as srv:
        ids = [srv.submit(r) for r in reqs]
        for i, rid in enumerate(ids):
            srv.wait(rid, 60000)
            got = srv.take(rid)
            assert np.array_equal(got.cls, alone[i].cls) and np.array_equal(got.steps, alone[i].steps), i
            for name in ("x", "p", "bl", "lam", "g"):
                u, v = getattr(got, name), getattr(alone[i], name)
                if reqs[i].integrator.method == rtsim.RT_METHOD_RENDER_RK4:
                    assert np.array_equal(u, v), f"request {i}: {name} differs"  # the same kernel
                else:
                    d = (np.abs(u - v) / np.maximum(np.abs(v), 1.0)).max(initial=0)
                    assert d < 1e-12, f"request {i}: {name} differs by {d:.3e}"
