"""The 32x32 tile task boundary on the host (no GPU): tile geometry and the
decoding of RenderTileTaskPayload wire bytes, against the reference's own
task_payloads.cpp — through the committed fixtures (tests/golden/tile_tasks.json,
written by tests/golden/make_tile_tasks.py) and, where oracle/_ref is built,
directly on mutated tasks."""
import base64
import ctypes as C
import json
import math
import os
import struct
import zlib

import numpy as np
import pytest

import oracle_lib as O
import rtsim

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = json.load(open(os.path.join(HERE, "golden", "tile_tasks.json")))


def task_bytes(entry):
    return zlib.decompress(base64.b64decode(entry["bytes"]))


def tasks(accepted=None):
    return [t for t in FIXTURE["tasks"] if accepted is None or t["accepted"] == accepted]


def test_wire_sizes_match_the_reference_layout():
    lay = FIXTURE["layout"]
    assert lay["task_bytes"] == rtsim.RT_TILE_TASK_BYTES == 11080
    assert lay["header_bytes"] == 16 and lay["push_offset"] == 24 and lay["push_bytes"] == 11056
    assert rtsim.RT_TILE_EDGE == 32
    assert (rtsim.RT_TILE_MAX_SCREEN_WIDTH, rtsim.RT_TILE_MAX_SCREEN_HEIGHT) == (7680, 4320)


def test_tile_geometry_matches_reference():
    for case in FIXTURE["geometry"]:
        if case["ok"]:
            t = rtsim.tile_geometry(case["index"], case["width"], case["height"])
            assert [t.origin_x, t.origin_y, t.width, t.height] == case["tile"], case
        else:
            with pytest.raises(rtsim.RtError):
                rtsim.tile_geometry(case["index"], case["width"], case["height"])


def test_tile_geometry_known_answers():
    """The tiles the reference's tests use, and the ragged last row/column."""
    t = rtsim.tile_geometry(78, 384, 384)  # test_cluster_worker_execution.cpp:70
    assert (t.origin_x, t.origin_y, t.width, t.height) == (192, 192, 32, 32)
    t = rtsim.tile_geometry(42, 1920, 1080)  # test_cluster_serialization_codecs.cpp:194
    assert (t.origin_x, t.origin_y, t.width, t.height) == (1344, 0, 32, 32)
    t = rtsim.tile_geometry(60 * 34 - 1, 1920, 1080)
    assert (t.origin_x, t.origin_y, t.width, t.height) == (1888, 1056, 32, 24)
    # the tiles of a frame cover it exactly once
    W, H = 100, 70
    cover = np.zeros((H, W), int)
    for i in range(4 * 3):
        t = rtsim.tile_geometry(i, W, H)
        cover[t.origin_y:t.origin_y + t.height, t.origin_x:t.origin_x + t.width] += 1
    assert np.all(cover == 1)


def check_decoded(t, e):
    assert t.tile_index == e["tile_index"]
    assert [t.tile.origin_x, t.tile.origin_y, t.tile.width, t.tile.height] == e["tile"]
    assert (t.lod_reduced_steps, t.flags, t.compression_type) == (e["lod_reduced_steps"], e["flags"],
                                                                  e["compression_type"])
    assert t.dispatch_timestamp == e["dispatch_timestamp"]
    assert (t.metric_type, t.body_count) == (e["metric_type"], e["body_count"])
    c = t.camera
    assert list(c.observer) == e["observer"]
    assert list(c.forward) == e["forward"] and list(c.right) == e["right"] and list(c.up) == e["up"]
    assert c.fov_y == e["fov_y"] and (c.width, c.height) == (e["width"], e["height"])
    assert (c.row_offset, c.row_count) == (0, 0)
    assert t.metric.n_sources == 0 and t.metric.mass == e["mass"] and t.metric.spin == e["spin"]
    assert t.integrator.max_steps == e["max_steps"]
    d = rtsim.default_integrator()
    assert (t.integrator.rtol, t.integrator.atol, t.integrator.h0) == (d.rtol, d.atol, d.h0)
    tm = t.termination
    assert tm.escape_radius == e["escape_radius"] and tm.disk_enabled == e["disk_enabled"]
    assert tm.horizon_eps == 1e-6 and tm.flags == 0
    assert tm.disk_inner == pytest.approx(e["disk_inner"], rel=1e-14, abs=0.0)
    assert tm.disk_outer == pytest.approx(e["disk_outer"], rel=1e-14)
    assert c.projection == e["projection"] and tm.time == e["time"]
    ig = t.integrator
    assert (ig.min_step, ig.max_step, ig.step_size_factor, ig.far_field_step_scale) == (
        e["min_step"], e["max_step"], e["step_size_factor"], e["far_field_step_scale"])
    assert [getattr(tm.disk_profile, f) for f in rtsim.DISK_PROFILE_FIELDS] == e["disk_profile"]


@pytest.mark.parametrize("entry", tasks(), ids=lambda e: e["name"])
def test_decode_matches_reference_verdict_and_fields(entry):
    data = task_bytes(entry)
    if not entry["accepted"]:
        with pytest.raises(rtsim.RtError):
            rtsim.decode_tile_task(data)
        return
    check_decoded(rtsim.decode_tile_task(data), entry["expected"])


def test_decode_ignores_trailing_bytes():
    entry = next(t for t in tasks(True) if t["name"] == "kerr_centre_tile")
    t = rtsim.decode_tile_task(task_bytes(entry) + b"\x00" * 100)
    check_decoded(t, entry["expected"])


def test_metric_family_and_boost_refusals_are_the_only_extra_rules():
    extra = [t["name"] for t in FIXTURE["tasks"] if t["reference_accepts"] and not t["accepted"]]
    assert sorted(extra) == ["metric_flat", "metric_reissner_nordstrom", "observer_boost_task"]


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference tree absent)")
def test_random_mutations_match_reference_verdict():
    """Random byte changes over the header, the timestamp and the first 800
    bytes of the push constants: the decoder accepts exactly what the
    reference's deserialize_render_tile_task accepts, except a metric_type
    outside 1..3, a field of view the task's projection does not accept
    ((0, pi) for the pinhole-type ones, (0, 4 pi] for the others) and an
    escape radius that is not positive. (The observer boost lies beyond the
    mutated bytes.)"""
    ref = O.ref()
    ref.ref_accepts_tile_task.argtypes = [C.c_void_p, C.c_size_t]
    ref.ref_accepts_tile_task.restype = C.c_int
    base = task_bytes(next(t for t in tasks(True) if t["name"] == "kerr_centre_tile"))
    push = FIXTURE["layout"]["push_offset"]
    mt_off = push + FIXTURE["layout"]["metric_type"]
    rng = np.random.default_rng(7)
    n_refused = 0
    for _ in range(3000):
        data = bytearray(base)
        for _ in range(int(rng.integers(1, 4))):
            off = int(rng.integers(0, push + 800))
            data[off] = int(rng.integers(0, 256))
        data = bytes(data)
        want = bool(ref.ref_accepts_tile_task(data, len(data)))
        metric_type = int.from_bytes(data[mt_off:mt_off + 4], "little")
        fov = struct.unpack_from("<d", data, push + FIXTURE["layout"]["field_of_view_rad"])[0]
        escape = struct.unpack_from("<d", data, push + FIXTURE["layout"]["escape_radius"])[0]
        proj = struct.unpack_from("<I", data, push + FIXTURE["layout"]["projection_mode"])[0]
        proj = proj if proj <= 7 else 0
        fov_ok = (0.0 < fov < 3.14159) if proj in (0, 1, 6) else (0.0 < fov <= 4.0 * math.pi)
        want = want and 1 <= metric_type <= 3 and fov_ok and escape > 0.0
        try:
            rtsim.decode_tile_task(data)
            got = True
        except rtsim.RtError:
            got = False
        n_refused += not got
        assert got == want, data[:32].hex()
    assert n_refused > 100


def changed(name, field, fmt, value, extra=0):
    """Wire bytes of fixture task `name` with one push-constant field (at
    byte `extra` inside it) replaced."""
    data = bytearray(task_bytes(next(t for t in tasks(True) if t["name"] == name)))
    struct.pack_into(fmt, data, FIXTURE["layout"]["push_offset"] + FIXTURE["layout"][field] + extra, value)
    return bytes(data)


def test_changed_fields_decode_to_traceable_tasks():
    """Single fields of an accepted task changed to values the reference
    accepts: each decodes to a task the tracer accepts (the reference's clamp
    of the observer radius, the disk off where a NaN radius means it is never
    hit), or is refused with a message (field of view outside (0, pi),
    escape radius not positive)."""
    base = rtsim.decode_tile_task(changed("kerr_centre_tile", "escape_radius", "<d", 1000.0))
    m, a = base.metric.mass, base.metric.spin
    r_h = m + math.sqrt(m * m - a * a)
    for r in (0.0, -5.0, 0.5 * r_h):  # observer_position[1] is r
        t = rtsim.decode_tile_task(changed("kerr_centre_tile", "observer_position", "<d", r, extra=8))
        assert t.camera.observer[1] == pytest.approx(1.02 * r_h, rel=1e-15)
    for fov in (0.0, -0.1, math.pi, 4.0):
        with pytest.raises(rtsim.RtError, match="fov_y"):
            rtsim.decode_tile_task(changed("kerr_centre_tile", "field_of_view_rad", "<d", fov))
    for esc in (0.0, -10.0):
        with pytest.raises(rtsim.RtError, match="escape radius"):
            rtsim.decode_tile_task(changed("kerr_centre_tile", "escape_radius", "<d", esc))
    for field in ("disk_outer_radius_mass_units", "disk_inner_radius_scale"):
        t = rtsim.decode_tile_task(changed("kerr_centre_tile", field, "<f", float("nan")))
        assert t.termination.disk_enabled == 0
        t = rtsim.decode_tile_task(changed("kerr_centre_tile", field, "<f", 2.0))
        assert t.termination.disk_enabled == base.termination.disk_enabled
        assert math.isfinite(t.termination.disk_inner) and math.isfinite(t.termination.disk_outer)


def tile_records(entry):
    r = entry["reference_tile_records"]

    def un(key, dt):
        return np.frombuffer(zlib.decompress(base64.b64decode(r[key])), dtype=dt)

    return {"status": un("status", np.uint8).astype(np.uint32), "iters": un("iters", np.uint16).astype(np.uint32),
            "redshift": un("redshift", np.float32), "affine": un("affine", np.float32)}


def render_task(task):
    """The decoded task with the renderer's tracer selected."""
    task.integrator.method = rtsim.RT_METHOD_RENDER_RK4
    return task


@pytest.mark.parametrize("entry", [t for t in tasks(True) if "reference_tile_records" in t], ids=lambda e: e["name"])
def test_decoded_task_reproduces_reference_tile_records(entry):
    """A decoded task traced with RT_METHOD_RENDER_RK4 (here by the oracle,
    whose records equal the GPU kernel's) gives the pixel records of the
    reference's own tile path (deserialize_render_tile_task, the worker's
    push-constant adjustments, dispatch_tile): the decode mapping and the
    renderer's tracer together, bit for bit."""
    t = render_task(rtsim.decode_tile_task(task_bytes(entry)))
    g = t.tile
    cam = rtsim.rt_camera.from_buffer_copy(t.camera)
    cam.row_offset, cam.row_count = g.origin_y, g.height
    rows = O.trace_render_camera(cam, t.metric, t.integrator, t.termination)
    idx = (np.arange(g.height)[:, None] * cam.width + np.arange(g.origin_x, g.origin_x + g.width)[None, :]).ravel()
    got = {"status": rows.cls[idx] & 0xFF, "iters": rows.steps[idx], "redshift": rows.g[idx].astype(np.float32),
           "affine": rows.lam[idx].astype(np.float32)}
    want = tile_records(entry)
    bad = {k: int((got[k] != want[k]).sum()) for k in want}
    assert all(v == 0 for v in bad.values()), bad
