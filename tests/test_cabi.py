"""The C ABI without a GPU: librtsim.so loads and exports every function that
include/rt_tracer.h declares; the ctypes mirror has the C structs' layout;
defaults are the documented ones; argument errors come back as statuses."""
import ctypes as C
import json
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import rtsim

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "rt_tracer.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rt_[a-z_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = rtsim.load()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) <= set(rtsim.EXPORTS) | {"rt_get_launch_shape", "rt_set_launch_shape", "rt_autotune"}


STRUCTS = ["rt_source", "rt_metric", "rt_integrator", "rt_termination", "rt_camera", "rt_results", "rt_stats",
           "rt_request", "rt_server_config", "rt_launch_shape", "rt_pixel", "rt_tile", "rt_tile_task",
           "rt_disk_profile", "rt_submit_options"]


def test_struct_layout_matches_ctypes():
    src = "#include <stdio.h>\n#include <stddef.h>\n#include \"rt_tracer.h\"\nint main(void){\n"
    for s in STRUCTS:
        src += f'printf("{s} %zu\\n", sizeof({s}));\n'
    for s, f in [("rt_integrator", "max_steps"), ("rt_camera", "width"), ("rt_metric", "src"),
                 ("rt_request", "termination"), ("rt_stats", "block_size"), ("rt_results", "rejected"),
                 ("rt_pixel", "redshift"), ("rt_pixel", "iterations_used"), ("rt_tile_task", "dispatch_timestamp"),
                 ("rt_tile_task", "camera"), ("rt_tile_task", "metric"), ("rt_tile_task", "termination"),
                 ("rt_camera", "projection"), ("rt_integrator", "method"), ("rt_integrator", "far_field_step_scale"),
                 ("rt_termination", "time"), ("rt_termination", "disk_profile"), ("rt_disk_profile", "edge_softness"),
                 ("rt_pixel", "r"), ("rt_pixel", "g"), ("rt_pixel", "b"), ("rt_pixel", "a"),
                 ("rt_pixel", "affine_parameter"), ("rt_pixel", "status_flags"),
                 ("rt_server_config", "staging_mib"), ("rt_submit_options", "deadline_ms")]:
        src += f'printf("{s}.{f} %zu\\n", offsetof({s}, {f}));\n'
    src += "return 0;}\n"
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "probe.c")
        exe = os.path.join(d, "probe")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-std=c99", "-I", os.path.dirname(HEADER), "-o", exe, c], check=True)
        lines = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")
    got = dict(l.split() for l in lines if l)
    for s in STRUCTS:
        assert int(got[s]) == C.sizeof(getattr(rtsim, s)), s
    assert int(got["rt_integrator.max_steps"]) == rtsim.rt_integrator.max_steps.offset
    assert int(got["rt_camera.width"]) == rtsim.rt_camera.width.offset
    assert int(got["rt_metric.src"]) == rtsim.rt_metric.src.offset
    assert int(got["rt_request.termination"]) == rtsim.rt_request.termination.offset
    assert int(got["rt_stats.block_size"]) == rtsim.rt_stats.block_size.offset
    assert int(got["rt_results.rejected"]) == rtsim.rt_results.rejected.offset
    assert int(got["rt_server_config.staging_mib"]) == rtsim.rt_server_config.staging_mib.offset
    assert int(got["rt_submit_options.deadline_ms"]) == rtsim.rt_submit_options.deadline_ms.offset
    # rt_pixel against the reference's GpuPixelOutput (render/gpu_types.hpp:1111-1121): the offsets and
    # size that offsetof/sizeof give on the reference's struct (tests/golden/render_frames.json)
    ref_layout = json.load(open(os.path.join(REPO, "tests", "golden", "render_frames.json")))["pixel_layout"]
    fields = ["r", "g", "b", "a", "redshift", "affine_parameter", "status_flags", "iterations_used"]
    assert [int(got[f"rt_pixel.{f}"]) for f in fields] + [int(got["rt_pixel"])] == ref_layout
    assert [rtsim.PIXEL_DTYPE.fields[f][1] for f in fields] + [rtsim.PIXEL_DTYPE.itemsize] == ref_layout
    for s, f in [("rt_camera", "projection"), ("rt_integrator", "method"), ("rt_integrator", "far_field_step_scale"),
                 ("rt_termination", "time"), ("rt_termination", "disk_profile"), ("rt_disk_profile", "edge_softness")]:
        assert int(got[f"{s}.{f}"]) == getattr(getattr(rtsim, s), f).offset, (s, f)
    assert int(got["rt_disk_profile"]) == 128
    for f in ["dispatch_timestamp", "camera", "metric", "termination"]:
        assert int(got[f"rt_tile_task.{f}"]) == getattr(rtsim.rt_tile_task, f).offset, f


def test_defaults_match_python_mirror():
    lib = rtsim.load()
    ig = rtsim.rt_integrator()
    lib.rt_default_integrator(C.byref(ig))
    py = rtsim.default_integrator()
    for f, _ in rtsim.rt_integrator._fields_:
        assert getattr(ig, f) == getattr(py, f), f
    tm = rtsim.rt_termination()
    lib.rt_default_termination(C.byref(tm))
    pt = rtsim.default_termination()
    for f, _ in rtsim.rt_termination._fields_:
        if f == "disk_profile":
            assert bytes(tm.disk_profile) == bytes(pt.disk_profile), f
        else:
            assert getattr(tm, f) == getattr(pt, f), f
    cam = rtsim.rt_camera()
    lib.rt_default_camera(C.byref(cam), 1000.0, 1.0, 0.3, 0.35, 64, 32)
    pc = rtsim.default_camera(1000.0, 1.0, 0.3, 0.35, 64, 32)
    for f in ("observer", "forward", "right", "up"):
        assert np.allclose(list(getattr(cam, f)), list(getattr(pc, f)), rtol=0, atol=1e-15)
    assert lib.rt_abi_version() == 3
    assert b"librtsim" in lib.rt_version()


def test_null_arguments_are_statuses_not_crashes():
    lib = rtsim.load()
    assert lib.rt_trace_camera(None, None, None, None, None, None) == 1
    assert lib.rt_trace_rays(None, None, None, 0, None, None, None, None) == 1
    assert lib.rt_trace_batch(None, None, 0, None) == 1
    assert lib.rt_get_stats(None, None) == 1
    assert lib.rt_context_destroy(None) == 0
    assert lib.rt_server_destroy(None) == 0
    assert lib.rt_server_submit(None, None, None) == 1
    assert len(lib.rt_last_error()) > 0


def test_context_without_device_reports_no_device():
    lib = rtsim.load()
    h = C.c_void_p()
    st = lib.rt_context_create(0, C.byref(h))
    if st == 0:
        lib.rt_context_destroy(h)
        pytest.skip("a GPU is visible here")
    assert st == 6  # RT_ERR_NO_DEVICE
