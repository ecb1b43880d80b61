"""32x32 tile tasks on the GPU: a decoded RenderTileTaskPayload traced with
rt_trace_tile gives, pixel for pixel, the whole frame's results (bit for bit:
the same kernel on the same pixels) and the oracle's to the parity bar; the
tile's pixel records; superposition windows; refused regions."""
import math

import numpy as np
import pytest

import oracle_lib as O
import rtsim
from test_gpu_batch import same
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


def fixture_task(name):
    return rtsim.decode_tile_task(task_bytes(next(e for e in tasks(True) if e["name"] == name)))


def split(res, sizes):
    out, off = [], 0
    for n in sizes:
        idx = np.arange(off, off + n)
        sub = rtsim.Results(n)
        sub.cls[:], sub.lam[:], sub.g[:] = res.cls[idx], res.lam[idx], res.g[idx]
        sub.steps[:], sub.rejected[:] = res.steps[idx], res.rejected[idx]
        sub.x[:], sub.p[:], sub.bl[:] = res.x[:, idx], res.p[:, idx], res.bl[:, idx]
        out.append(sub)
        off += n
    return out


@pytest.fixture(params=[0, 1, 2], ids=["tile_queue", "lane_refill_queue", "per_request_kind"])
def batch_kind(ctx, request):
    """The context's batch shape for one test, restored afterwards. Kind 2
    (per-request kernels) makes rt_trace_tiles use the shared tile queue at
    two waves per SIMD; kind 1 runs the lane-refill kernel over the
    concatenated windows."""
    _, keep = ctx.launch_shapes()
    ctx.set_launch_shapes(queue=rtsim.rt_launch_shape(64, 2, 16, request.param))
    yield request.param
    ctx.set_launch_shapes(queue=keep)


def test_tile_batch_equals_single_tiles(ctx, batch_kind):
    """A batch of tasks of different frames, metrics (one a superposition),
    tolerances and tile shapes: each task's slice of the batch equals
    rt_trace_tile of it (class, flags and step counts identical, coordinates
    within 1e-12: the queue kernels run the same formulas, which the compiler
    contracts into FMAs differently), on two replays and after a smaller
    batch."""
    names = ["kerr_centre_tile", "ragged_corner_tile", "worker_test_task", "scaled_disk_tile", "disk_off_tile"]
    batch = [fixture_task(n) for n in names]
    multi = fixture_task("kerr_centre_tile")
    multi.metric = rtsim.multi_metric([{"mass": 1.0, "spin": 0.9},
                                       {"mass": 0.3, "spin": 0.5, "position": (12.0, -4.0, 2.0)}])
    batch.append(multi)
    loose = fixture_task("kerr_centre_tile")
    loose.integrator.rtol = 1e-8
    batch.append(loose)
    singles = []
    for t in batch:
        n = ctx.trace_tile(t.camera, t.metric, t.integrator, t.termination, t.tile)
        singles.append(ctx.fetch(n))
    sizes = [t.tile.width * t.tile.height for t in batch]
    for rep in range(2):
        print(f"[tile batch] kind {batch_kind}: {len(batch)} tasks, replay {rep}", flush=True)
        n = ctx.trace_tiles(batch)
        assert n == sum(sizes)
        for i, part in enumerate(split(ctx.fetch(n), sizes)):
            same(part, singles[i], f"task {i} replay {rep}")
    small = batch[1:3]
    n = ctx.trace_tiles(small)
    for i, part in enumerate(split(ctx.fetch(n), sizes[1:3])):
        same(part, singles[1 + i], f"smaller batch task {i}")
    pix = ctx.fetch_pixels(n)
    assert np.array_equal(pix["iterations_used"], np.concatenate([s.steps for s in singles[1:3]]))


def test_all_tiles_of_a_frame(ctx, batch_kind):
    """The 64 tiles of the 256^2 CPU-runnable configuration, as one batch,
    reassemble the frame rt_trace_camera traces (to the same 1e-12)."""
    base = fixture_task("kerr_centre_tile")
    cam, W, H = base.camera, base.camera.width, base.camera.height
    batch = []
    for i in range(((W + 31) // 32) * ((H + 31) // 32)):
        t = rtsim.rt_tile_task.from_buffer_copy(base)
        t.tile_index = i
        t.tile = rtsim.tile_geometry(i, W, H)
        batch.append(t)
    n = ctx.trace_tiles(batch)
    assert n == W * H
    res = ctx.fetch(n)
    frame = ctx.trace_camera(cam, base.metric, base.integrator, base.termination)
    parts = split(res, [t.tile.width * t.tile.height for t in batch])
    for t, part in zip(batch, parts):
        same(part, window(frame, W, t.tile), f"tile {t.tile_index}")


def test_render_tasks_reproduce_reference_tile_records(ctx):
    """Decoded tasks traced with RT_METHOD_RENDER_RK4, one by one
    (rt_trace_tile) and as one batch (rt_trace_tiles): the pixel records
    (rt_fetch_pixels) equal those of the reference's own tile path on the same
    wire bytes (tests/golden/tile_tasks.json)."""
    from test_tile_task import render_task, tile_records

    entries = [e for e in tasks(True) if "reference_tile_records" in e]
    assert len(entries) >= 6
    decoded = [render_task(rtsim.decode_tile_task(task_bytes(e))) for e in entries]
    want_all = []
    for e, t in zip(entries, decoded):
        n = ctx.trace_tile(t.camera, t.metric, t.integrator, t.termination, t.tile)
        pix = ctx.fetch_pixels(n)
        want = tile_records(e)
        want_all.append(want)
        bad = {"status": int((pix["status_flags"] != want["status"]).sum()),
               "iters": int((pix["iterations_used"] != want["iters"]).sum()),
               "redshift": int((pix["redshift"] != want["redshift"]).sum()),
               "affine": int((pix["affine_parameter"] != want["affine"]).sum())}
        print(f"[render tile] {e['name']}: projection {t.camera.projection} mismatches {bad}", flush=True)
        assert all(v == 0 for v in bad.values()), (e["name"], bad)
    n = ctx.trace_tiles(decoded)
    pix = ctx.fetch_pixels(n)
    assert np.array_equal(pix["status_flags"], np.concatenate([w["status"] for w in want_all]))
    assert np.array_equal(pix["iterations_used"], np.concatenate([w["iters"] for w in want_all]))
    assert np.array_equal(pix["redshift"], np.concatenate([w["redshift"] for w in want_all]))
