"""Writes tests/golden/tile_tasks.json: the reference's 32x32 tile geometry for
a set of (index, width, height) cases, the byte layout of its tile task, and
serialized RenderTileTaskPayload wire bytes with the reference's own verdict
on each (accepted or refused), all computed through oracle/_ref/libref.so,
which oracle/Makefile links from the reference's src/distributed/task_payloads.cpp.
Refused tasks are accepted ones with bytes changed at the given offsets.
Run after `make -C oracle`."""
import base64
import ctypes as C
import json
import math
import os
import struct
import sys
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402
import rtsim  # noqa: E402

U32 = C.POINTER(C.c_uint32)


def lib():
    ref = O.ref()
    ref.ref_tile_geometry.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, U32]
    ref.ref_tile_geometry.restype = C.c_int
    ref.ref_tile_layout.argtypes = [C.POINTER(C.c_uint64)]
    ref.ref_tile_layout.restype = C.c_int
    ref.ref_make_tile_task.argtypes = [C.POINTER(C.c_double), U32, C.POINTER(C.c_float), C.POINTER(C.c_double), U32,
                                       C.c_void_p, C.c_size_t]
    ref.ref_make_tile_task.restype = C.c_size_t
    ref.ref_accepts_tile_task.argtypes = [C.c_void_p, C.c_size_t]
    ref.ref_accepts_tile_task.restype = C.c_int
    ref.ref_render_tile_task.argtypes = [C.c_void_p, C.c_size_t, U32, U32, C.POINTER(C.c_float),
                                         C.POINTER(C.c_float)]
    ref.ref_render_tile_task.restype = C.c_size_t
    return ref


def tile_records(ref, data):
    """The reference's tile path (deserialize, the worker's adjustments,
    dispatch_tile) on the task's bytes: its pixel records, tile-local."""
    n = 32 * 32
    st, it = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
    rz, af = np.zeros(n, np.float32), np.zeros(n, np.float32)
    k = ref.ref_render_tile_task(data, len(data), st.ctypes.data_as(U32), it.ctypes.data_as(U32),
                                 rz.ctypes.data_as(C.POINTER(C.c_float)), af.ctypes.data_as(C.POINTER(C.c_float)))
    assert k > 0
    return {"status": pack(st[:k].astype(np.uint8).tobytes()), "iters": pack(it[:k].astype(np.uint16).tobytes()),
            "redshift": pack(rz[:k].tobytes()), "affine": pack(af[:k].tobytes())}


LAYOUT_NAMES = ["task_bytes", "header_bytes", "push_offset", "push_bytes", "observer_position", "tetrad_e1",
                "tetrad_e2", "tetrad_e3", "field_of_view_rad", "metric_mass", "metric_spin", "escape_radius",
                "screen_width", "screen_height", "metric_type", "max_integration_steps", "body_count",
                "disk_enabled", "disk_inner_radius_scale", "disk_outer_radius_mass_units", "projection_mode", "time",
                "step_size_factor", "min_step_size", "max_step_size", "far_field_step_scale", "render_flags",
                "lod_distance_threshold", "lod_reduced_steps", "observer_boost", "disk_profile_bytes",
                "use_lod_system_flag"]

# push-constant defaults (render/gpu_types.hpp:966-1109) the specs below leave alone
DEFAULTS = {"projection": 3, "steps": [0.004, 3.5, 0.5, 7.8], "time": 0.0, "lod_threshold": 0.0, "boost": 0.0,
            "render_flags": 0, "push_lod": 256}


def make_task(ref, spec):
    d = np.zeros(21)
    d[0:4] = spec["observer"]
    d[4:8] = spec["e1"]
    d[8:12] = spec["e2"]
    d[12:16] = spec["e3"]
    d[16:21] = [spec["fov"], spec["mass"], spec["spin"], spec["escape"], spec["timestamp"]]
    u = np.array([spec["tile_index"], spec["width"], spec["height"], spec["lod"], spec["flags"], spec["mode"],
                  spec["max_steps"], spec["metric_type"], spec["body_count"]], np.uint32)
    f = np.array(spec["disk"], np.float32)
    sp = dict(DEFAULTS, **spec)
    d2 = np.array(list(sp["steps"]) + [sp["time"], sp["lod_threshold"], sp["boost"]], np.float64)
    u2 = np.array([sp["projection"], sp["render_flags"], sp["push_lod"]], np.uint32)
    buf = (C.c_uint8 * 16384)()
    n = ref.ref_make_tile_task(d.ctypes.data_as(C.POINTER(C.c_double)), u.ctypes.data_as(U32),
                               f.ctypes.data_as(C.POINTER(C.c_float)), d2.ctypes.data_as(C.POINTER(C.c_double)),
                               u2.ctypes.data_as(U32), buf, len(buf))
    assert n > 0, spec["name"]
    return bytes(buf[:n])


def camera_spec(cam):
    return {"observer": list(cam.observer), "e1": [0.0] + list(cam.forward), "e2": [0.0] + list(cam.right),
            "e3": [0.0] + list(cam.up), "fov": cam.fov_y, "width": cam.width, "height": cam.height}


def expected_fields(ref, spec, tile):
    """The request a worker traces for an accepted task (field mapping of the
    tracer's C ABI, from the values put into the task)."""
    sp = dict(DEFAULTS, **spec)
    m = max(spec["mass"], 1e-4)
    a = 0.0 if spec["metric_type"] == 1 else min(max(spec["spin"], -0.999 * m), 0.999 * m)
    enabled, inner_scale, outer_units = [float(np.float32(v)) for v in spec["disk"]]
    inner_scale = max(inner_scale, 1.0)
    inner = ref.ref_isco(m, a) * inner_scale
    # the step budget under the push constants' LOD rule (dispatch_fp64_scalar,
    # render/software_compute_engine.hpp:3821-3825), its r_+ from the clamped spin
    a_pc = min(max(spec["spin"], -0.999 * m), 0.999 * m)
    rh = m + math.sqrt(max(m * m - a_pc * a_pc, 0.0)) if abs(a_pc) > 1e-12 else 2.0 * m
    r_frame = max(spec["observer"][1], rh * 1.02)
    lod_steps = spec["lod"] if spec["lod"] != 0 else sp["push_lod"]
    steps = spec["max_steps"]
    if (sp["render_flags"] & 512) and sp["lod_threshold"] > 0.0 and r_frame > sp["lod_threshold"]:
        steps = min(steps, lod_steps)
    profile = [float(np.float32(v)) for v in rtsim.DISK_PROFILE_DEFAULTS]
    profile[0:3] = [enabled, float(np.float32(spec["disk"][1])), outer_units]
    return {"tile_index": spec["tile_index"], "tile": tile, "lod_reduced_steps": spec["lod"], "flags": spec["flags"],
            "compression_type": spec["mode"], "dispatch_timestamp": spec["timestamp"],
            "metric_type": spec["metric_type"], "body_count": spec["body_count"],
            "observer": spec["observer"], "forward": spec["e1"][1:], "right": spec["e2"][1:], "up": spec["e3"][1:],
            "fov_y": spec["fov"], "width": spec["width"], "height": spec["height"], "mass": m, "spin": a,
            "max_steps": max(steps, 1), "escape_radius": spec["escape"], "disk_enabled": int(enabled > 0.5),
            "disk_inner": inner if inner_scale > 1.0 else 0.0, "disk_outer": max(outer_units * m, 1.05 * inner),
            "projection": sp["projection"] if sp["projection"] <= 7 else 0, "time": sp["time"],
            "min_step": sp["steps"][0], "max_step": sp["steps"][1], "step_size_factor": sp["steps"][2],
            "far_field_step_scale": sp["steps"][3], "disk_profile": profile}


def pack(data):
    return base64.b64encode(zlib.compress(data, 9)).decode()


def main():
    ref = lib()
    lay = (C.c_uint64 * 32)()
    n = ref.ref_tile_layout(lay)
    layout = dict(zip(LAYOUT_NAMES, [int(v) for v in lay[:n]]))
    out = {"source": "reference task_payloads.cpp via oracle/_ref/libref.so", "layout": layout, "geometry": [],
           "tasks": []}

    rng = np.random.default_rng(20261205)
    cases = [(78, 384, 384), (42, 1920, 1080), (0, 1, 1), (59, 1920, 1080), (60 * 34 - 1, 1920, 1080),
             (60 * 34, 1920, 1080), (0, 0, 10), (0, 10, 0), (0, 7680, 4320), (240 * 135 - 1, 7680, 4320),
             (0, 7681, 100), (0, 100, 4321), (11, 100, 70), (12, 100, 70), (2 ** 32 - 1, 7680, 4320)]
    for _ in range(200):
        w, h = int(rng.integers(1, 8000)), int(rng.integers(1, 4500))
        tiles = ((w + 31) // 32) * ((h + 31) // 32)
        cases.append((int(rng.integers(0, tiles + 3)), w, h))
    for idx, w, h in cases:
        g = (C.c_uint32 * 4)()
        ok = ref.ref_tile_geometry(idx, w, h, g)
        out["geometry"].append({"index": idx, "width": w, "height": h, "ok": bool(ok), "tile": list(g) if ok else None})

    specs = []
    # the reference's own serialization test (tests/test_cluster_serialization_codecs.cpp:181-210)
    specs.append({"name": "serialization_test_task", "observer": [0.0, 10.0, 1.57, 0.0], "e1": [0.0, 1.0, 0.0, 0.0],
                  "e2": [0.0, 0.0, 1.0, 0.0], "e3": [0.0, 0.0, 0.0, 1.0], "fov": 1.0, "width": 1920,
                  "height": 1080, "mass": 1.0, "spin": 0.0, "escape": 50.0, "timestamp": 123.456, "tile_index": 42,
                  "lod": 0, "flags": 0, "mode": 2, "max_steps": 1024, "metric_type": 1, "body_count": 0,
                  "disk": [1.0, 1.0, 24.0]})
    # the worker-execution test's camera (tests/test_cluster_worker_execution.cpp:30-57), its tile 78
    specs.append({"name": "worker_test_task", "observer": [0.0, 18.0, 1.45, 0.0], "e1": [0.0, -1.0, 0.0, 0.0],
                  "e2": [0.0, 0.0, 0.0, 1.0], "e3": [0.0, 0.0, 1.0, 0.0], "fov": math.radians(60.0),
                  "width": 384, "height": 384, "mass": 1.0, "spin": 0.94, "escape": 60.0, "timestamp": 10.0,
                  "tile_index": 78, "lod": 0, "flags": 0, "mode": 0, "max_steps": 1024, "metric_type": 2,
                  "body_count": 0, "disk": [1.0, 1.0, 18.0], "steps": [0.005, 2.0, 0.45, 7.8]})
    # the CPU-runnable Kerr configuration (a = 0.9, i = 80 deg, r = 1000 M, 256^2, 20 deg), the tile
    # at the image centre: captured, disk and escaped rays
    cam = rtsim.default_camera(1000.0, math.radians(80.0), 0.0, math.radians(20.0), 256, 256)
    specs.append(dict(camera_spec(cam), name="kerr_centre_tile", mass=1.0, spin=0.9, escape=1000.0,
                      timestamp=0.5, tile_index=36, lod=0, flags=0, mode=1, max_steps=20000, metric_type=2,
                      body_count=0, disk=[1.0, 1.0, 20.0], projection=0))
    # ragged last tile of a 100 x 70 frame (4 x 6 pixels)
    cam = rtsim.default_camera(50.0, math.radians(60.0), 0.3, math.radians(40.0), 100, 70)
    specs.append(dict(camera_spec(cam), name="ragged_corner_tile", mass=1.0, spin=0.5, escape=200.0,
                      timestamp=1.0, tile_index=11, lod=64, flags=1, mode=0, max_steps=5000, metric_type=3,
                      body_count=0, disk=[1.0, 1.0, 20.0], projection=0))
    # disk radii scaled, spin above the clamp, mass 2, disk off, bodies announced
    cam = rtsim.default_camera(200.0, math.radians(75.0), -0.7, math.radians(25.0), 96, 64)
    specs.append(dict(camera_spec(cam), name="scaled_disk_tile", mass=2.0, spin=2.5, escape=400.0,
                      timestamp=2.0, tile_index=4, lod=0, flags=6, mode=2, max_steps=30000, metric_type=3,
                      body_count=3, disk=[1.0, 1.5, 12.0], projection=0))
    specs.append(dict(specs[-1], name="disk_off_tile", disk=[0.25, 1.0, 24.0], tile_index=0, body_count=0))
    # all-sky projection, disk pattern time, render step formula, LOD rule active (observer beyond the
    # LOD distance: budget min(max_integration_steps, the header's lod_reduced_steps))
    cam = rtsim.default_camera(200.0, math.radians(70.0), 0.9, math.radians(50.0), 128, 96)
    specs.append(dict(camera_spec(cam), name="allsky_lod_tile", mass=1.0, spin=0.7, escape=500.0, timestamp=3.0,
                      tile_index=5, lod=300, flags=0, mode=0, max_steps=5000, metric_type=2, body_count=0,
                      disk=[1.0, 1.0, 24.0], projection=7, time=2.5, steps=[0.01, 3.0, 0.4, 6.0],
                      render_flags=512, lod_threshold=100.0))
    # LOD flag set but the observer inside the LOD distance: budget unchanged
    specs.append(dict(specs[-1], name="lod_inactive_tile", observer=[0.0, 60.0, specs[-1]["observer"][2], 0.9],
                      projection=2))
    # a projection mode the reference does not know traces as its switch's default, the pinhole
    specs.append(dict(specs[-1], name="unknown_projection_tile", projection=9, render_flags=0))

    push = layout["push_offset"]
    for spec in specs:
        data = make_task(ref, spec)
        W, H = spec["width"], spec["height"]
        g = (C.c_uint32 * 4)()
        assert ref.ref_tile_geometry(spec["tile_index"], W, H, g)
        assert ref.ref_accepts_tile_task(data, len(data)) == 1
        entry = {"name": spec["name"], "bytes": pack(data), "reference_accepts": True, "accepted": True,
                 "expected": expected_fields(ref, spec, list(g))}
        m = max(spec["mass"], 1e-4)
        if spec["metric_type"] in (2, 3) and abs(spec["spin"]) > 1e-9 * m:
            # the exact Kerr path: the reference renderer's records of this tile
            entry["reference_tile_records"] = tile_records(ref, data)
        out["tasks"].append(entry)

    # refused variants of kerr_centre_tile: (name, byte offset, new bytes)
    base = make_task(ref, specs[2])
    nan = struct.pack("<d", float("nan"))
    inf = struct.pack("<d", float("inf"))
    muts = [("truncated_100_bytes", None, None),
            ("zero_step_budget", push + layout["max_integration_steps"], struct.pack("<I", 0)),
            ("compression_mode_3", 15, b"\x03"),
            ("header_tile_width_mismatch", 8, struct.pack("<H", 31)),
            ("header_tile_x_mismatch", 4, struct.pack("<H", 5)),
            ("header_tile_index_past_end", 0, struct.pack("<I", 64)),
            ("nan_field_of_view", push + layout["field_of_view_rad"], nan),
            ("inf_escape_radius", push + layout["escape_radius"], inf),
            ("nan_spin", push + layout["metric_spin"], nan),
            ("nan_mass", push + layout["metric_mass"], nan),
            ("nan_observer_t", push + layout["observer_position"], nan),
            ("nan_tetrad_e3_z", push + layout["tetrad_e3"] + 24, nan),
            ("nan_timestamp", layout["header_bytes"], nan),
            ("zero_screen_height", push + layout["screen_height"], struct.pack("<I", 0)),
            ("screen_width_8000", push + layout["screen_width"], struct.pack("<I", 8000)),
            ("metric_reissner_nordstrom", push + layout["metric_type"], struct.pack("<I", 4)),
            ("metric_flat", push + layout["metric_type"], struct.pack("<I", 0)),
            ("nan_tetrad_e0_ignored", push + 32, nan)]
    for name, off, new in muts:
        if off is None:
            data = base[:100]
        else:
            data = base[:off] + new + base[off + len(new):]
        verdict = bool(ref.ref_accepts_tile_task(data, len(data)))
        ours = verdict and not name.startswith("metric_")
        entry = {"name": name, "bytes": pack(data), "reference_accepts": verdict, "accepted": ours}
        if ours:
            g = (C.c_uint32 * 4)()
            assert ref.ref_tile_geometry(36, 256, 256, g)
            entry["expected"] = expected_fields(ref, specs[2], list(g))
        out["tasks"].append(entry)

    # an observer boost (aberration) the reference applies and this tracer does not: refused
    data = make_task(ref, dict(specs[2], name="observer_boost_task", boost=0.3))
    out["tasks"].append({"name": "observer_boost_task", "bytes": pack(data),
                         "reference_accepts": bool(ref.ref_accepts_tile_task(data, len(data))), "accepted": False})

    path = os.path.join(HERE, "tile_tasks.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=0)
    print(f"wrote {path}: {len(out['geometry'])} geometry cases, {len(out['tasks'])} tasks")


if __name__ == "__main__":
    main()
