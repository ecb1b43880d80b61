"""Batched tracing: one persistent work-queue kernel over uneven requests
(whole 8x8 tiles per wave, or per-lane refill), captured into a hipGraph and
replayed.

Each request's results must equal the single-request kernel's results for the
same request (class, flags and step counts identical; coordinates within 1e-12
relative: both kernels evaluate the same formulas, which the compiler may
contract into FMAs differently, so the last bits can differ), on the first
(captured) call and on every replay, with eager work between the replays."""
import math

import numpy as np
import pytest

import oracle_lib as O
import rtsim

pytestmark = pytest.mark.gpu


def uneven_requests():
    reqs = []
    specs = [(37, 29, 0.0, 30.0, 300.0, 1e-10), (64, 64, 0.9, 60.0, 1000.0, 1e-10),
             (50, 71, 0.998, 85.0, 2000.0, 1e-8), (96, 40, 0.5, 12.0, 60.0, 1e-8)]
    for w, h, a, inc, r, rtol in specs:
        cam = rtsim.default_camera(r, math.radians(inc), 0.1, math.radians(20.0), w, h)
        reqs.append(rtsim.make_request(cam, rtsim.kerr_metric(1.0, a), rtsim.default_integrator(rtol=rtol),
                                       rtsim.default_termination(escape_radius=max(1000.0, r))))
    cam = rtsim.default_camera(150.0, math.radians(70.0), 0.0, math.radians(25.0), 24, 20)
    met = rtsim.multi_metric([{"mass": 1.0, "spin": 0.4}, {"mass": 0.4, "spin": 0.8, "position": (15, 3, -2),
                                                          "axis": (0, 1, 1)}])
    reqs.append(rtsim.make_request(cam, met, rtsim.default_integrator(), rtsim.default_termination(escape_radius=300.0)))
    return reqs


def same(a, b, label):
    assert np.array_equal(a.cls, b.cls), f"{label}: class/flags differ"
    for name in ("x", "p", "bl", "lam", "g"):
        u, v = getattr(a, name), getattr(b, name)
        r = (np.abs(u - v) / np.maximum(np.abs(v), 1.0)).max(initial=0)
        assert r < 1e-12, f"{label}: {name} differs by {r:.3e}"
    assert np.array_equal(a.steps, b.steps) and np.array_equal(a.rejected, b.rejected), f"{label}: step counts"


@pytest.mark.parametrize("kind", [0, 1, 2], ids=["tile_queue", "lane_refill_queue", "per_request_streams"])
def test_batch_matches_single_and_graph_replays(ctx, kind):
    plain, queue = ctx.launch_shapes()
    ctx.set_launch_shapes(queue=rtsim.rt_launch_shape(128, 2, 16, kind))
    try:
        check_batch_replays(ctx)
    finally:
        ctx.set_launch_shapes(queue=queue)


def check_batch_replays(ctx):
    reqs = uneven_requests()
    eager = [ctx.trace_camera(r.camera, r.metric, r.integrator, r.termination) for r in reqs]
    for rep in range(3):
        print(f"batch call {rep}", flush=True)
        outs = ctx.trace_batch(reqs)
        st = ctx.stats()
        assert st.rays == sum(e.n for e in eager)
        assert st.accepted_steps == sum(int(e.steps.sum()) for e in eager)
        for i, (o, e) in enumerate(zip(outs, eager)):
            same(o, e, f"call {rep} request {i}")
        # eager work between replays
        ctx.trace_camera(reqs[1].camera, reqs[1].metric, reqs[1].integrator, reqs[1].termination)


def test_autotune_then_batch(ctx):
    reqs = uneven_requests()[:3]
    best = ctx.autotune(reqs)
    print(f"autotuned block={best.block} waves={best.waves} refill={best.refill_min} queue_kind={best.queue_kind}",
          flush=True)
    assert best.block in (64, 128, 256) and best.waves in (1, 2) and best.queue_kind in (0, 1, 2)
    eager = [ctx.trace_camera(r.camera, r.metric, r.integrator, r.termination) for r in reqs]
    for o, e in zip(ctx.trace_batch(reqs), eager):
        same(o, e, "after autotune")


def test_launch_shape_roundtrip(ctx):
    plain, queue = ctx.launch_shapes()
    s = rtsim.rt_launch_shape(128, 2, 8, 1)
    ctx.set_launch_shapes(queue=s)
    _, q2 = ctx.launch_shapes()
    assert (q2.block, q2.waves, q2.refill_min, q2.queue_kind) == (128, 2, 8, 1)
    with pytest.raises(rtsim.RtError):
        ctx.set_launch_shapes(queue=rtsim.rt_launch_shape(100, 3, 0, 0))
    with pytest.raises(rtsim.RtError):
        ctx.set_launch_shapes(queue=rtsim.rt_launch_shape(64, 1, 16, 3))
    ctx.set_launch_shapes(plain=plain, queue=queue)


@pytest.mark.parametrize("kind", [0, 1], ids=["tile_queue", "lane_refill_queue"])
def test_queue_kernels_vs_oracle(ctx, kind):
    """The persistent queue kernels' results against the oracle directly, at
    the FP64 parity bar (not only against the single-request kernel)."""
    from test_gpu_multi import compare as compare_multi
    from test_gpu_parity import compare

    _, keep = ctx.launch_shapes()
    ctx.set_launch_shapes(queue=rtsim.rt_launch_shape(64, 2, 16, kind))
    try:
        reqs = uneven_requests()
        outs = ctx.trace_batch(reqs)
    finally:
        ctx.set_launch_shapes(queue=keep)
    for i, (r, o) in enumerate(zip(reqs, outs)):
        ref = O.trace_camera(r.camera, r.metric, r.integrator, r.termination)
        if r.metric.n_sources >= 2:
            compare_multi(o, ref, f"kind {kind} request {i} (superposition)")
        else:
            compare(o, ref, f"kind {kind} request {i}")
