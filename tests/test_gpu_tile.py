"""32x32 tile tasks on the GPU: a decoded RenderTileTaskPayload traced with
rt_trace_tile gives, pixel for pixel, the whole frame's results (bit for bit:
the same kernel on the same pixels) and the oracle's to the parity bar; the
tile's pixel records; superposition windows; refused regions."""
import math

import numpy as np
import pytest

import oracle_lib as O
import rtsim
from test_gpu_multi import compare as compare_multi
from test_gpu_parity import compare
from test_tile_task import task_bytes, tasks

pytestmark = pytest.mark.gpu

FIELDS = ["cls", "lam", "g", "steps", "rejected"]


def window(res, W, t):
    """Results of the pixels of rectangle t from a row-major W-wide frame."""
    rows = np.arange(t.origin_y, t.origin_y + t.height)
    cols = np.arange(t.origin_x, t.origin_x + t.width)
    idx = (rows[:, None] * W + cols[None, :]).ravel()
    sub = rtsim.Results(idx.size)
    sub.cls[:], sub.lam[:], sub.g[:] = res.cls[idx], res.lam[idx], res.g[idx]
    sub.steps[:], sub.rejected[:] = res.steps[idx], res.rejected[idx]
    sub.x[:], sub.p[:], sub.bl[:] = res.x[:, idx], res.p[:, idx], res.bl[:, idx]
    return sub


def same_bits(a, b, label):
    for f in FIELDS:
        assert np.array_equal(getattr(a, f), getattr(b, f)), f"{label}: {f}"
    for f in ["x", "p", "bl"]:
        assert np.array_equal(getattr(a, f), getattr(b, f)), f"{label}: {f}"


def oracle_window(cam, met, ig, tm, t):
    rows = O.trace_camera_rows(cam, met, ig, tm, t.origin_y, 1, t.height)
    return window(rows, cam.width, rtsim.rt_tile(t.origin_x, 0, t.width, t.height))


@pytest.mark.parametrize("name", ["kerr_centre_tile", "ragged_corner_tile", "worker_test_task", "scaled_disk_tile"])
def test_decoded_task_matches_frame_and_oracle(ctx, name):
    task = rtsim.decode_tile_task(task_bytes(next(e for e in tasks(True) if e["name"] == name)))
    cam, met, ig, tm, t = task.camera, task.metric, task.integrator, task.termination, task.tile
    print(f"[tile] {name}: {cam.width}x{cam.height} tile {t.origin_x},{t.origin_y} {t.width}x{t.height}", flush=True)
    n = ctx.trace_tile(cam, met, ig, tm, t)
    assert n == t.width * t.height
    tile = ctx.fetch(n)
    frame = ctx.trace_camera(cam, met, ig, tm)
    same_bits(tile, window(frame, cam.width, t), name)
    ref = oracle_window(cam, met, ig, tm, t)
    compare(tile, ref, name)
    if name == "kerr_centre_tile":
        assert {1, 2, 4} <= set(tile.klass.tolist())


def test_tile_pixel_records(ctx):
    task = rtsim.decode_tile_task(task_bytes(next(e for e in tasks(True) if e["name"] == "kerr_centre_tile")))
    n = ctx.trace_tile(task.camera, task.metric, task.integrator, task.termination, task.tile)
    res, pix = ctx.fetch(n), ctx.fetch_pixels(n)
    assert np.array_equal(pix["status_flags"], res.cls & 0xFF)
    assert np.array_equal(pix["iterations_used"], res.steps)
    assert np.array_equal(pix["redshift"], res.g.astype(np.float32))


def test_superposition_window(ctx):
    """Any rectangle, here of a three-source superposition, equals the whole
    frame's pixels (the superposition kernel's window) and the oracle."""
    met = rtsim.multi_metric([{"mass": 1.0, "spin": 0.6},
                              {"mass": 0.4, "spin": 0.3, "position": (8.0, 3.0, 1.0), "axis": (0.2, 0.1, 1.0)},
                              {"mass": 0.3, "position": (-6.0, -7.0, 2.0)}])
    cam = rtsim.default_camera(60.0, math.radians(70.0), 0.2, math.radians(45.0), 96, 64)
    ig, tm = rtsim.default_integrator(), rtsim.default_termination(escape_radius=300.0)
    t = rtsim.rt_tile(33, 17, 20, 13)
    n = ctx.trace_tile(cam, met, ig, tm, t)
    tile = ctx.fetch(n)
    frame = ctx.trace_camera(cam, met, ig, tm)
    same_bits(tile, window(frame, cam.width, t), "superposition window")
    compare_multi(tile, oracle_window(cam, met, ig, tm, t), "superposition window")


def test_refused_regions(ctx):
    cam = rtsim.default_camera(100.0, 1.0, 0.0, 0.3, 40, 30)
    met, ig, tm = rtsim.kerr_metric(1.0, 0.5), rtsim.default_integrator(), rtsim.default_termination()
    for t in [rtsim.rt_tile(0, 0, 0, 5), rtsim.rt_tile(0, 0, 5, 0), rtsim.rt_tile(40, 0, 1, 1),
              rtsim.rt_tile(0, 30, 1, 1), rtsim.rt_tile(30, 0, 11, 1), rtsim.rt_tile(0, 20, 1, 11),
              rtsim.rt_tile(2 ** 32 - 1, 0, 2, 1)]:
        with pytest.raises(rtsim.RtError):
            ctx.trace_tile(cam, met, ig, tm, t)
    t = rtsim.rt_tile(8, 4, 32, 26)
    n = ctx.trace_tile(cam, met, ig, tm, t)
    compare(ctx.fetch(n), oracle_window(cam, met, ig, tm, t), "after refusals")
