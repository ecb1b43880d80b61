"""RT_METHOD_RENDER_RK4 (the reference renderer's tracer) and the camera
projections on the GPU.

- The pixel records of every fixture frame (rt_fetch_pixels: status_flags,
  iterations_used, redshift, affine_parameter) against the reference's own
  dispatch_fp64 records in tests/golden/render_frames.json: identical.
- End states (KS position and momentum, BL point) against the oracle's
  restatement: identical records, end states within 1e-9 (the FP64 parity
  bar; projections that go through atan2, asin, sin and cos start from
  directions that can differ from the host's in the last bit, which the
  trajectory carries to ~1e-10).
- Larger frames against the oracle: records identical.
- Render requests inside rt_trace_batch (beside DP45 requests) and inside
  rt_trace_tiles: the same results as traced alone.
- The DP45 kernels under every projection mode against the oracle at the
  FP64 parity bar."""
import json
import math
import os
import sys

import numpy as np
import pytest

import oracle_lib as O
import rtsim
from test_gpu_parity import compare

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from make_render_frames import frame_inputs, unpack  # noqa: E402

pytestmark = pytest.mark.gpu
FIXTURE = json.load(open(os.path.join(HERE, "golden", "render_frames.json")))
FIELDS = [("status_flags", "status", np.uint8), ("iterations_used", "iters", np.uint16),
          ("redshift", "redshift", np.float32), ("affine_parameter", "affine", np.float32)]


def rel(a, b):
    return np.abs(a - b) / np.maximum(np.abs(b), 1.0)


@pytest.mark.parametrize("f", FIXTURE["frames"], ids=lambda f: f["name"])
def test_render_records_match_reference(ctx, f):
    cam, met, ig, tm = frame_inputs(f["spec"])
    print(f"[render] {f['name']}: {cam.width}x{cam.height} projection {cam.projection}", flush=True)
    n = ctx.trace_camera_resident(cam, met, ig, tm)
    assert n == f["n"]
    pix = ctx.fetch_pixels(n)
    bad = {}
    for pf, key, dt in FIELDS:
        want = unpack(f[key], dt).astype(pix[pf].dtype)
        bad[pf] = int((pix[pf] != want).sum())
    print(f"[render] {f['name']}: mismatches {bad}", flush=True)
    assert all(v == 0 for v in bad.values()), bad
    assert np.all(pix["a"] == 1.0)
    res = ctx.fetch(n)
    ref = O.trace_render_camera(cam, met, ig, tm)
    assert np.array_equal(res.cls, ref.cls) and np.array_equal(res.steps, ref.steps)
    errs = {"x": rel(res.x, ref.x).max(), "p": rel(res.p, ref.p).max(), "bl": rel(res.bl, ref.bl).max(),
            "g": rel(res.g, ref.g).max()}
    print(f"[render] {f['name']}: end state vs oracle " + " ".join(f"{k}={v:.1e}" for k, v in errs.items()),
          flush=True)
    for k, v in errs.items():
        assert v < 1e-9, f"{k} differs by {v:.2e}"


@pytest.mark.parametrize("scene", ["r1000_i80_fov20", "r30_i70_fov60_disk24"])
def test_render_frames_vs_oracle(ctx, scene):
    if scene == "r1000_i80_fov20":
        cam = rtsim.default_camera(1000.0, math.radians(80.0), 0.0, math.radians(20.0), 256, 256)
        tm = rtsim.default_termination(escape_radius=2000.0, disk_outer=24.0)
        ig = rtsim.render_integrator(4096)
    else:
        cam = rtsim.default_camera(30.0, math.radians(70.0), 0.4, math.radians(60.0), 320, 200)
        tm = rtsim.default_termination(escape_radius=100.0, disk_outer=24.0)
        ig = rtsim.render_integrator(2048)
    met = rtsim.kerr_metric(1.0, 0.9)
    gpu = ctx.trace_camera(cam, met, ig, tm)
    ref = O.trace_render_camera(cam, met, ig, tm)
    same = {"status": (gpu.cls == ref.cls).mean(), "iters": (gpu.steps == ref.steps).mean(),
            "g": (gpu.g == ref.g).mean()}
    print(f"[render {scene}] {gpu.n} px, identical: " + " ".join(f"{k}={v:.6f}" for k, v in same.items()),
          f"status counts {np.bincount(ref.cls & 0xFF, minlength=8).tolist()}", flush=True)
    assert same["status"] == 1.0 and same["iters"] == 1.0
    assert rel(gpu.g, ref.g).max() < 1e-12


def test_render_in_batches_and_tiles(ctx):
    f = next(fr for fr in FIXTURE["frames"] if fr["name"] == "disk_pinhole")
    cam, met, ig, tm = frame_inputs(f["spec"])
    alone = ctx.trace_camera(cam, met, ig, tm)
    dp45 = rtsim.make_request(rtsim.default_camera(1000.0, 1.0, 0.0, 0.3, 40, 24), rtsim.kerr_metric(1.0, 0.5),
                              rtsim.default_integrator(), rtsim.default_termination())
    dp45_alone = ctx.trace_camera(dp45.camera, dp45.metric, dp45.integrator, dp45.termination)
    render = rtsim.make_request(cam, met, ig, tm)
    _, keep = ctx.launch_shapes()
    try:
        for kind in (0, 2):
            ctx.set_launch_shapes(queue=rtsim.rt_launch_shape(64, 2, 16, kind))
            for rep in range(2):
                outs = ctx.trace_batch([dp45, render, dp45])
                for name in ("cls", "steps", "x", "p", "bl", "lam", "g"):
                    assert np.array_equal(getattr(outs[1], name), getattr(alone, name)), (kind, rep, name)
                assert np.array_equal(outs[0].cls, dp45_alone.cls) and np.array_equal(outs[2].cls, dp45_alone.cls)
    finally:
        ctx.set_launch_shapes(queue=keep)
    # the frame as 32x32 tiles of render tasks
    tasks = []
    for i in range(((cam.width + 31) // 32) * ((cam.height + 31) // 32)):
        t = rtsim.rt_tile_task()
        t.tile_index, t.camera, t.metric, t.integrator, t.termination = i, cam, met, ig, tm
        t.tile = rtsim.tile_geometry(i, cam.width, cam.height)
        tasks.append(t)
    n = ctx.trace_tiles(tasks)
    res = ctx.fetch(n)
    off = 0
    for t in tasks:
        g = t.tile
        rows = np.arange(g.origin_y, g.origin_y + g.height)
        cols = np.arange(g.origin_x, g.origin_x + g.width)
        idx = (rows[:, None] * cam.width + cols[None, :]).ravel()
        k = idx.size
        assert np.array_equal(res.cls[off:off + k], alone.cls[idx]), t.tile_index
        assert np.array_equal(res.steps[off:off + k], alone.steps[idx]), t.tile_index
        assert np.array_equal(res.g[off:off + k], alone.g[idx]), t.tile_index
        off += k


@pytest.mark.parametrize("proj", range(8))
def test_dp45_projections_vs_oracle(ctx, proj):
    fov = math.radians(20.0) if proj in (0, 1, 6) else math.radians(40.0)
    cam = rtsim.default_camera(60.0, math.radians(75.0), 0.5, fov, 40, 30, projection=proj)
    met = rtsim.kerr_metric(1.0, 0.9)
    ig, tm = rtsim.default_integrator(), rtsim.default_termination(escape_radius=200.0)
    gpu = ctx.trace_camera(cam, met, ig, tm)
    ref = O.trace_camera(cam, met, ig, tm)
    compare(gpu, ref, f"DP45 projection {proj}")
