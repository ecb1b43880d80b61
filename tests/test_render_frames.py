"""The oracle's restatement of the reference renderer's tracer
(RT_METHOD_RENDER_RK4) against the reference's own pixel records.

tests/golden/render_frames.json holds SoftwareComputeEngine::dispatch_fp64's
status_flags, iterations_used, redshift and affine_parameter for every pixel
of twelve small Kerr frames (tests/golden/make_render_frames.py). The oracle
must reproduce all four bit for bit. The frames are asymmetric (spin, disk,
off-equator cameras), so a camera whose right or up vector is flipped, or
whose forward vector points the other way, changes the records: planting
such a mistake on the oracle's input must break the match. Where
oracle/_ref is built, the projections, the pixel record's layout and random
frames are also compared with the reference directly."""
import ctypes as C
import json
import math
import os
import sys

import numpy as np
import pytest

import oracle_lib as O
import rtsim

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from make_render_frames import frame_inputs, unpack  # noqa: E402

FIXTURE = json.load(open(os.path.join(HERE, "golden", "render_frames.json")))


def frames():
    return FIXTURE["frames"]


def records(f):
    return {"status": unpack(f["status"], np.uint8).astype(np.uint32),
            "iters": unpack(f["iters"], np.uint16).astype(np.uint32),
            "redshift": unpack(f["redshift"], np.float32), "affine": unpack(f["affine"], np.float32)}


def pixel_view(res):
    """The four record fields of a trace, as the pixel record holds them."""
    return {"status": res.cls & 0xFF, "iters": res.steps, "redshift": res.g.astype(np.float32),
            "affine": res.lam.astype(np.float32)}


def mismatches(got, want):
    return {k: int((got[k] != want[k]).sum()) for k in ("status", "iters", "redshift", "affine")}


@pytest.mark.parametrize("f", frames(), ids=lambda f: f["name"])
def test_oracle_matches_reference_records(f):
    cam, met, ig, tm = frame_inputs(f["spec"])
    res = O.trace_render_camera(cam, met, ig, tm)
    want = records(f)
    assert res.n == f["n"]
    bad = mismatches(pixel_view(res), want)
    assert all(v == 0 for v in bad.values()), f"{f['name']}: {bad}"


def test_fixtures_cover_every_rule():
    """The fixtures hold captured, escaped and disk records, a disk record that
    went on to escape, a budget that ran out near the hole (status 0), and
    crossings recorded on escaped rays (redshift != 1 with status 2)."""
    st = np.concatenate([records(f)["status"] for f in frames()])
    rz = np.concatenate([records(f)["redshift"] for f in frames()])
    assert {0, 1, 2, 4, 6} <= set(np.unique(st).tolist())
    assert ((st == 2) & (rz != 1.0)).sum() > 1000
    assert {f["spec"][8] for f in frames()} == set(range(8))


@pytest.mark.parametrize("mistake", ["mirror_right", "flip_up", "backward"])
def test_planted_camera_mistakes_break_the_match(mistake):
    """A mirrored, upside-down or reversed camera fails the reference records."""
    broken = 0
    for f in frames()[:4]:
        cam, met, ig, tm = frame_inputs(f["spec"])
        for i in range(3):
            if mistake == "mirror_right":
                cam.right[i] = -cam.right[i]
            elif mistake == "flip_up":
                cam.up[i] = -cam.up[i]
            else:
                cam.forward[i] = -cam.forward[i]
        bad = mismatches(pixel_view(O.trace_render_camera(cam, met, ig, tm)), records(f))
        broken += bad["status"] + bad["iters"]
    assert broken > 100, f"{mistake}: only {broken} records changed"


def test_render_records_are_reference_consistent():
    """Properties of every record: affine = iters x 0.05 as a float, iters
    within the halved budget, redshift 1 exactly when no crossing was
    recorded on a disk-free frame."""
    for f in frames():
        want = records(f)
        assert np.array_equal(want["affine"], (want["iters"] * 0.05).astype(np.float32))
        cam, met, ig, tm = frame_inputs(f["spec"])
        assert want["iters"].max() <= max(ig.max_steps // 2, 256)
        if not tm.disk_enabled:
            assert np.all(want["redshift"] == 1.0)


ref_only = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference tree absent)")


@ref_only
def test_screen_directions_bitwise():
    """Every projection of the oracle against CameraProjector::compute_ray_direction."""
    rng = np.random.default_rng(5)
    ref, orc = O.ref(), O.oracle()
    a, b = np.zeros(3), np.zeros(3)
    for mode in range(8):
        for _ in range(400):
            u, v = rng.uniform(-1.8, 1.8, size=2)
            fov = float(rng.uniform(0.02, 3.1 if mode in (0, 1, 6) else 5.0))
            ref.ref_screen_direction(mode, u, v, fov, a.ctypes.data_as(C.POINTER(C.c_double)))
            orc.rto_screen_direction(mode, u, v, fov, b.ctypes.data_as(C.POINTER(C.c_double)))
            assert np.array_equal(a, b), (mode, u, v, fov, a, b)
        for u, v in [(0.0, 0.0), (1e-16, 0.0)]:
            ref.ref_screen_direction(mode, u, v, 1.0, a.ctypes.data_as(C.POINTER(C.c_double)))
            orc.rto_screen_direction(mode, u, v, 1.0, b.ctypes.data_as(C.POINTER(C.c_double)))
            assert np.array_equal(a, b), (mode, u, v)


@ref_only
def test_pixel_record_layout_matches_reference():
    """rt_pixel has GpuPixelOutput's field offsets and size (render/gpu_types.hpp:1111-1121)."""
    out = (C.c_uint64 * 9)()
    assert O.ref().ref_pixel_layout(out) == 9
    fields = ["r", "g", "b", "a", "redshift", "affine_parameter", "status_flags", "iterations_used"]
    got = [getattr(rtsim.rt_pixel, f).offset for f in fields] + [C.sizeof(rtsim.rt_pixel)]
    assert got == list(out)
    assert FIXTURE["pixel_layout"] == list(out)


@ref_only
def test_random_frames_against_reference():
    """Random small frames (spin, inclination, azimuth, distance, field, size,
    projection, disk on or off, time, profile) through the oracle and the
    reference's dispatch_fp64: the records are identical."""
    rng = np.random.default_rng(11)
    total = 0
    for k in range(10):
        a = float(rng.uniform(-0.99, 0.99))
        proj = int(rng.integers(0, 8))
        fov = float(rng.uniform(0.2, 1.6 if proj in (0, 1, 6) else 3.0))
        r = float(np.exp(rng.uniform(np.log(15.0), np.log(300.0))))
        cam = rtsim.default_camera(r, float(rng.uniform(0.1, 3.0)), float(rng.uniform(-3, 3)), fov,
                                   int(rng.integers(8, 40)), int(rng.integers(8, 40)), projection=proj)
        met = rtsim.kerr_metric(1.0, a)
        ig = rtsim.render_integrator(int(rng.integers(300, 3000)))
        tm = rtsim.default_termination(escape_radius=max(60.0, 1.5 * r), disk_outer=float(rng.choice([12.0, 24.0])),
                                       disk_enabled=bool(rng.random() < 0.8))
        tm.time = float(rng.uniform(0, 10))
        tm.disk_profile.opacity_scale = float(np.float32(rng.uniform(0.2, 1.0)))
        want = O.ref_render_frame(cam, met, ig, tm)
        got = pixel_view(O.trace_render_camera(cam, met, ig, tm))
        bad = mismatches(got, want)
        assert all(v == 0 for v in bad.values()), (k, bad)
        total += got["status"].size
    assert total > 2000


@ref_only
def test_row_windows_against_reference():
    """Row windows through the oracle and the reference's dispatch_fp64
    (resolve_row_window): count 0 with and without an offset, an inner window,
    a count running past the last row, and offset + count beyond 2^32. The
    reference writes nothing when its buffer is smaller than its window, so a
    window of the wrong size shows up as mismatched records."""
    met, ig = rtsim.kerr_metric(1.0, 0.8), rtsim.render_integrator(600)
    tm = rtsim.default_termination(escape_radius=80.0, disk_outer=20.0)
    whole = None
    for off, cnt in [(0, 0), (5, 0), (3, 7), (14, 50), (10, 0xFFFFFFFF), (19, 1)]:
        cam = rtsim.default_camera(35.0, 1.2, 0.3, 0.8, 16, 20, row_offset=off, row_count=cnt)
        want = O.ref_render_frame(cam, met, ig, tm)
        got = pixel_view(O.trace_render_camera(cam, met, ig, tm))
        n_rows = 20 if cnt == 0 else min(off + cnt, 20) - off
        assert got["status"].size == want["status"].size == n_rows * 16, (off, cnt)
        bad = mismatches(got, want)
        assert all(v == 0 for v in bad.values()), (off, cnt, bad)
        if (off, cnt) == (0, 0):
            whole = want
        else:
            first = 0 if cnt == 0 else off
            assert np.array_equal(want["iters"], whole["iters"][first * 16:(first + n_rows) * 16]), (off, cnt)
