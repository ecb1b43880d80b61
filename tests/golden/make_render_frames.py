"""Writes tests/golden/render_frames.json: pixel records of the reference
renderer, SoftwareComputeEngine::dispatch_fp64 (render/software_compute_engine.hpp:5546-5587),
on small Kerr frames, through oracle/_ref/libref.so (ref_render_frame, which
fills the push constants from the C ABI's camera, metric, integrator and
termination). Each frame keeps its inputs and, per pixel, status_flags,
iterations_used, redshift and affine_parameter as the reference wrote them;
"pixel_layout" holds offsetof(GpuPixelOutput, r, g, b, a, redshift,
affine_parameter, status_flags, iterations_used) and its size.

The scenes are chosen to pin the pixel -> ray mapping and the renderer's
tracer (RT_METHOD_RENDER_RK4): non-square frames, cameras off the equator
and at several azimuths, every projection mode, prograde and retrograde
spins, the translucent disk with the default and a changed profile, a
rotated disk pattern (time), a scaled inner edge, a row window and a step
budget that runs out. Spin and disk break the mirror symmetries of the
image, so a flipped or mirrored camera changes the records.

Run after `make -C oracle` (needs the reference tree)."""
import base64
import json
import math
import os
import sys
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402
import rtsim  # noqa: E402

FRAMES = [
    # name, spin, incl deg, r_obs, phi_obs, fov deg, width, height, projection, disk, extra
    ("disk_off_pinhole", 0.9, 80.0, 30.0, 0.3, 60.0, 64, 48, 0, False, {}),
    ("disk_pinhole", 0.9, 65.0, 30.0, -0.8, 60.0, 64, 48, 0, True, {}),
    ("near_extremal_portrait_rows", 0.998, 60.0, 50.0, 1.1, 40.0, 48, 64, 0, True, {"row_offset": 9, "row_count": 37}),
    ("far_observer_narrow", 0.5, 85.0, 1000.0, 0.0, 3.0, 64, 64, 0, True, {"escape_radius": 2000.0}),
    ("equirectangular_time", 0.7, 70.0, 40.0, 2.0, 60.0, 64, 48, 3, True, {"time": 3.7}),
    ("stereographic_profile", 0.3, 45.0, 25.0, -2.5, 90.0, 56, 40, 2, True,
     {"profile": {"opacity_scale": 0.5, "ring_frequency": 5.0, "spiral_arm_count": 3.0, "noise_seed": 7.0,
                  "turbulence_octaves": 2.0, "edge_softness": 2.0}}),
    ("retrograde_panini", -0.6, 75.0, 35.0, 0.7, 50.0, 60, 44, 6, True, {}),
    ("hammer_scaled_inner", 0.95, 30.0, 20.0, -1.3, 100.0, 64, 40, 7, True, {"disk_inner_scale": 1.5,
                                                                             "disk_outer": 15.0}),
    ("equidistant", 0.6, 55.0, 28.0, 0.9, 120.0, 40, 40, 4, True, {}),
    ("orthographic", 0.8, 88.0, 26.0, -0.2, 150.0, 40, 32, 5, True, {}),
    ("auto_zoom", 0.4, 62.0, 45.0, 1.9, 35.0, 48, 36, 1, True, {}),
    ("small_budget_fine_steps", 0.9, 70.0, 60.0, 0.4, 30.0, 48, 32, 0, True,
     {"max_steps": 300, "escape_radius": 400.0, "step_size_factor": 0.05, "max_step": 0.4}),
]


def frame_inputs(spec):
    name, a, inc, r, phi, fov, w, h, proj, disk, extra = spec
    cam = rtsim.default_camera(r, math.radians(inc), phi, math.radians(fov), w, h,
                               row_offset=extra.get("row_offset", 0), row_count=extra.get("row_count", 0),
                               projection=proj)
    met = rtsim.kerr_metric(1.0, a)
    ig = rtsim.render_integrator(extra.get("max_steps", 2048))
    ig.step_size_factor = extra.get("step_size_factor", ig.step_size_factor)
    ig.max_step = extra.get("max_step", ig.max_step)
    tm = rtsim.default_termination(escape_radius=extra.get("escape_radius", max(100.0, 2.0 * r)),
                                   disk_outer=extra.get("disk_outer", 24.0), disk_enabled=disk)
    tm.time = extra.get("time", 0.0)
    if "disk_inner_scale" in extra:
        tm.disk_inner = extra["disk_inner_scale"] * O.oracle().rto_isco(1.0, a)
    for k, v in extra.get("profile", {}).items():
        setattr(tm.disk_profile, k, v)
    return cam, met, ig, tm


def pack(arr):
    return base64.b64encode(zlib.compress(np.ascontiguousarray(arr).tobytes(), 9)).decode()


def unpack(text, dtype):
    return np.frombuffer(zlib.decompress(base64.b64decode(text)), dtype=dtype)


def main():
    import ctypes as C

    layout = (C.c_uint64 * 9)()
    O.ref().ref_pixel_layout(layout)
    out = {"source": "SoftwareComputeEngine::dispatch_fp64 via oracle/_ref/libref.so (ref_render_frame)",
           "pixel_layout": list(layout), "frames": []}
    for spec in FRAMES:
        cam, met, ig, tm = frame_inputs(spec)
        ref = O.ref_render_frame(cam, met, ig, tm)
        out["frames"].append({"name": spec[0], "spec": [spec[0], spec[1], spec[2], spec[3], spec[4], spec[5],
                                                        spec[6], spec[7], spec[8], spec[9], spec[10]],
                              "n": int(ref["status"].size), "status": pack(ref["status"].astype(np.uint8)),
                              "iters": pack(ref["iters"].astype(np.uint16)), "redshift": pack(ref["redshift"]),
                              "affine": pack(ref["affine"])})
        counts = np.bincount(ref["status"], minlength=8)
        print(f"{spec[0]}: {ref['status'].size} px, status counts {counts.tolist()}, iters max {ref['iters'].max()}, "
              f"redshift != 1: {(ref['redshift'] != 1.0).sum()}")
    with open(os.path.join(HERE, "render_frames.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
