"""Python binding of librtsim.so, the MI355X null-geodesic tracer.

A thin ctypes mirror of include/rt_tracer.h: the same structs, the same entry
points, results as numpy arrays. The C++/HIP library does all the work; this
module only marshals arguments. It refuses to run when the library is missing
(there is no fallback path).

Reference interfaces mirrored (paths relative to the Relativity-Simulator tree):
  Context.trace_camera   SoftwareComputeEngine::dispatch_fp64
                         (include/relativistic/render/software_compute_engine.hpp:5546)
  Context.trace_rays     GeodesicRayProbe::trace
                         (include/relativistic/optics/geodesic_ray_probe.hpp:452)
  Context.trace_batch    TileRenderer::render (src/distributed/tile_renderer.cpp:133)
  Server                 GeodesicComputePipeline windowed worker
                         (include/relativistic/render/geodesic_compute_pipeline.hpp:529-728)
"""
from __future__ import annotations

import ctypes as C
import math
import os
from typing import Dict, List, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "librtsim.so")

RT_MAX_SOURCES = 16
RT_CLS_CAPTURED = 1
RT_CLS_ESCAPED = 2
RT_CLS_DISK = 4
RT_CLS_MASK = 7
RT_FLAG_BUDGET = 1 << 8
RT_FLAG_MIN_STEP = 1 << 9
RT_FLAG_NONFINITE = 1 << 10
RT_FLAG_PROJECTED = 1 << 11
RT_FLAG_FP32_STEPS = 1 << 12
RT_FLAG_CARTER = 1 << 13
RT_FLAG_INVALID = 1 << 15
RT_PRECISION_FP64 = 0
RT_PRECISION_MIXED = 1
RT_TERM_NO_G_CLAMP = 1
RT_TERM_CARTER_CAPTURE = 2
RT_METHOD_DP45 = 0
RT_METHOD_RENDER_RK4 = 1
RT_PROJ_PINHOLE = 0
RT_PROJ_AUTO_ZOOM = 1
RT_PROJ_FISHEYE_STEREOGRAPHIC = 2
RT_PROJ_EQUIRECTANGULAR_360 = 3
RT_PROJ_FISHEYE_EQUIDISTANT = 4
RT_PROJ_FISHEYE_ORTHOGRAPHIC = 5
RT_PROJ_PANINI_CYLINDRICAL = 6
RT_PROJ_HAMMER_AITOFF = 7

RT_OUT_CLS = 1 << 0
RT_OUT_X = 1 << 1
RT_OUT_P = 1 << 2
RT_OUT_BL = 1 << 3
RT_OUT_LAMBDA = 1 << 4
RT_OUT_G = 1 << 5
RT_OUT_STEPS = 1 << 6
RT_OUT_REJECTED = 1 << 7
RT_OUT_ALL = 0xFF
RT_OUT_PIXEL = RT_OUT_CLS | RT_OUT_LAMBDA | RT_OUT_G | RT_OUT_STEPS
# result fields in rt_results order, with their RT_OUT_* bit
OUT_FIELDS = (("cls", RT_OUT_CLS), ("x", RT_OUT_X), ("p", RT_OUT_P), ("bl", RT_OUT_BL), ("lam", RT_OUT_LAMBDA),
              ("g", RT_OUT_G), ("steps", RT_OUT_STEPS), ("rejected", RT_OUT_REJECTED))

STATUS_NAMES = {
    0: "RT_OK", 1: "RT_ERR_INVALID_ARG", 2: "RT_ERR_OOM", 3: "RT_ERR_DEVICE", 4: "RT_ERR_TIMEOUT",
    5: "RT_ERR_CANCELLED", 6: "RT_ERR_NO_DEVICE", 7: "RT_ERR_BUSY", 8: "RT_ERR_NOT_FOUND", 9: "RT_ERR_EXPIRED",
}


class RtError(RuntimeError):
    def __init__(self, status: int, message: str = ""):
        self.status = status
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")


class rt_source(C.Structure):
    _fields_ = [("position", C.c_double * 3), ("mass", C.c_double), ("spin", C.c_double),
                ("axis", C.c_double * 3), ("disk_inner", C.c_double), ("disk_outer", C.c_double),
                ("disk_enabled", C.c_uint32), ("reserved", C.c_uint32)]


class rt_metric(C.Structure):
    _fields_ = [("mass", C.c_double), ("spin", C.c_double), ("n_sources", C.c_uint32),
                ("reserved", C.c_uint32), ("src", rt_source * RT_MAX_SOURCES)]


class rt_integrator(C.Structure):
    _fields_ = [("rtol", C.c_double), ("atol", C.c_double), ("h0", C.c_double), ("h_min", C.c_double),
                ("h_max", C.c_double), ("invariant_tol", C.c_double), ("mixed_switch_radius", C.c_double),
                ("max_steps", C.c_uint32), ("precision", C.c_uint32), ("method", C.c_uint32),
                ("reserved", C.c_uint32), ("step_size_factor", C.c_double), ("min_step", C.c_double),
                ("max_step", C.c_double), ("far_field_step_scale", C.c_double)]


DISK_PROFILE_FIELDS = [
    "enabled", "inner_radius_scale", "outer_radius_mass_units", "peak_temperature_k", "floor_temperature_k",
    "temperature_exponent", "zero_torque_strength", "extra_beaming_exponent", "brightness", "color_saturation",
    "tint_strength", "opacity_scale", "tint_r", "tint_g", "tint_b", "rotation_speed_scale", "ring_amplitude",
    "ring_frequency", "ring_sharpness", "spiral_arm_count", "spiral_pitch", "spiral_amplitude",
    "turbulence_amplitude", "turbulence_scale", "turbulence_octaves", "radial_stretch", "grain_amplitude",
    "grain_scale", "reference_luminance", "noise_seed", "temperature_normalization", "edge_softness"]
# GpuDiskProfile's defaults (render/gpu_types.hpp:42-77)
DISK_PROFILE_DEFAULTS = [1.0, 1.0, 24.0, 9500.0, 1200.0, 0.75, 1.0, 2.0, 0.45, 1.0, 0.0, 1.0, 1.0, 1.0, 1.0, 4.0,
                         0.35, 9.0, 3.0, 2.0, 1.2, 0.25, 0.45, 5.0, 4.0, 3.0, 0.2, 40.0, 1.0, 1.0, 1.0, 1.0]


class rt_disk_profile(C.Structure):
    _fields_ = [(f, C.c_float) for f in DISK_PROFILE_FIELDS]


class rt_termination(C.Structure):
    _fields_ = [("horizon_eps", C.c_double), ("escape_radius", C.c_double), ("disk_inner", C.c_double),
                ("disk_outer", C.c_double), ("disk_enabled", C.c_uint32), ("flags", C.c_uint32),
                ("time", C.c_double), ("disk_profile", rt_disk_profile)]


class rt_camera(C.Structure):
    _fields_ = [("observer", C.c_double * 4), ("forward", C.c_double * 3), ("right", C.c_double * 3),
                ("up", C.c_double * 3), ("fov_y", C.c_double), ("width", C.c_uint32), ("height", C.c_uint32),
                ("row_offset", C.c_uint32), ("row_count", C.c_uint32), ("projection", C.c_uint32),
                ("reserved", C.c_uint32)]


class rt_pixel(C.Structure):
    """The reference's 32-byte per-pixel record (GpuPixelOutput)."""
    _fields_ = [("r", C.c_float), ("g", C.c_float), ("b", C.c_float), ("a", C.c_float),
                ("redshift", C.c_float), ("affine_parameter", C.c_float), ("status_flags", C.c_uint32),
                ("iterations_used", C.c_uint32)]


PIXEL_DTYPE = np.dtype([("r", np.float32), ("g", np.float32), ("b", np.float32), ("a", np.float32),
                        ("redshift", np.float32), ("affine_parameter", np.float32), ("status_flags", np.uint32),
                        ("iterations_used", np.uint32)])


RT_TILE_EDGE = 32
RT_TILE_MAX_SCREEN_WIDTH = 7680
RT_TILE_MAX_SCREEN_HEIGHT = 4320
RT_TILE_TASK_BYTES = 11080
RT_NOALLOC_MAX_TASKS = 64


class rt_tile(C.Structure):
    """A rectangle of a frame (the reference's TileGeometry)."""
    _fields_ = [("origin_x", C.c_uint32), ("origin_y", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32)]


class rt_tile_task(C.Structure):
    """A decoded RenderTileTaskPayload (rt_decode_tile_task)."""
    _fields_ = [("tile_index", C.c_uint32), ("lod_reduced_steps", C.c_uint16), ("flags", C.c_uint8),
                ("compression_type", C.c_uint8), ("dispatch_timestamp", C.c_double), ("tile", rt_tile),
                ("metric_type", C.c_uint32), ("body_count", C.c_uint32), ("camera", rt_camera),
                ("metric", rt_metric), ("integrator", rt_integrator), ("termination", rt_termination)]


class rt_results(C.Structure):
    _fields_ = [("cls", C.POINTER(C.c_uint32)), ("x", C.POINTER(C.c_double)), ("p", C.POINTER(C.c_double)),
                ("bl", C.POINTER(C.c_double)), ("lam", C.POINTER(C.c_double)), ("g", C.POINTER(C.c_double)),
                ("steps", C.POINTER(C.c_uint32)), ("rejected", C.POINTER(C.c_uint32))]


class rt_stats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("total_ms", C.c_double), ("rays", C.c_uint64),
                ("accepted_steps", C.c_uint64), ("rejected_steps", C.c_uint64), ("rhs_evals", C.c_uint64),
                ("fp32_steps", C.c_uint64), ("block_size", C.c_uint32), ("grid_size", C.c_uint32)]


class rt_request(C.Structure):
    _fields_ = [("camera", rt_camera), ("metric", rt_metric), ("integrator", rt_integrator),
                ("termination", rt_termination)]


class rt_launch_shape(C.Structure):
    _fields_ = [("block", C.c_uint32), ("waves", C.c_uint32), ("refill_min", C.c_uint32), ("queue_kind", C.c_uint32)]


class rt_server_config(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_inflight", C.c_uint32), ("staging_mib", C.c_uint32)]


class rt_submit_options(C.Structure):
    _fields_ = [("outputs", C.c_uint32), ("reserved", C.c_uint32), ("deadline_ms", C.c_int64)]


# ------------------------------------------------------------ defaults
def default_integrator(rtol: float = 1e-10, atol: float = 1e-12, precision: int = RT_PRECISION_FP64,
                       max_steps: int = 20000, switch_radius: float = 0.0,
                       method: int = RT_METHOD_DP45) -> rt_integrator:
    """Same values as rt_default_integrator()."""
    ig = rt_integrator()
    ig.rtol, ig.atol, ig.h0, ig.h_min, ig.h_max = rtol, atol, 0.0, 1e-12, 1e13
    ig.invariant_tol = 1e-12
    ig.mixed_switch_radius = switch_radius
    ig.max_steps = max_steps
    ig.precision = precision
    ig.method = method
    ig.step_size_factor, ig.min_step, ig.max_step, ig.far_field_step_scale = 0.5, 0.004, 3.5, 7.8
    return ig


def render_integrator(max_integration_steps: int = 2048) -> rt_integrator:
    """The reference renderer's tracer (RT_METHOD_RENDER_RK4) with the push
    constants' default step formula; max_steps is max_integration_steps."""
    return default_integrator(max_steps=max_integration_steps, method=RT_METHOD_RENDER_RK4)


def default_disk_profile() -> rt_disk_profile:
    return rt_disk_profile(*DISK_PROFILE_DEFAULTS)


def default_termination(escape_radius: float = 1000.0, disk_outer: float = 20.0, horizon_eps: float = 1e-6,
                        disk_enabled: bool = True, flags: int = 0) -> rt_termination:
    """Same values as rt_default_termination()."""
    t = rt_termination()
    t.horizon_eps, t.escape_radius, t.disk_inner, t.disk_outer = horizon_eps, escape_radius, 0.0, disk_outer
    t.disk_enabled = 1 if disk_enabled else 0
    t.flags = flags
    t.time = 0.0
    t.disk_profile = default_disk_profile()
    return t


def default_camera(r_obs: float, theta_obs: float, phi_obs: float, fov_y: float, width: int, height: int,
                   row_offset: int = 0, row_count: int = 0, projection: int = RT_PROJ_PINHOLE) -> rt_camera:
    """Observer looking at the hole: forward = -e_r, right = e_phi, up = -e_theta
    (same as rt_default_camera())."""
    c = rt_camera()
    st, ct, sp, cp = math.sin(theta_obs), math.cos(theta_obs), math.sin(phi_obs), math.cos(phi_obs)
    c.observer[:] = [0.0, r_obs, theta_obs, phi_obs]
    c.forward[:] = [-st * cp, -st * sp, -ct]
    c.right[:] = [-sp, cp, 0.0]
    c.up[:] = [-ct * cp, -ct * sp, st]
    c.fov_y = fov_y
    c.width, c.height, c.row_offset, c.row_count = width, height, row_offset, row_count
    c.projection = projection
    return c


def kerr_metric(mass: float, spin: float) -> rt_metric:
    m = rt_metric()
    m.mass, m.spin, m.n_sources = mass, spin, 0
    return m


def multi_metric(sources: Sequence[dict]) -> rt_metric:
    """Superposition of 2..16 Kerr-Schild sources; sources[0] is the central
    hole. Each dict: position (3), mass, spin (a/m), axis (3), disk_enabled,
    disk_inner, disk_outer."""
    if not 2 <= len(sources) <= RT_MAX_SOURCES:
        raise ValueError("multi_metric takes 2..16 sources")
    m = rt_metric()
    m.mass = sources[0]["mass"]
    m.spin = sources[0].get("spin", 0.0) * sources[0]["mass"]
    m.n_sources = len(sources)
    for i, s in enumerate(sources):
        d = m.src[i]
        d.position[:] = list(s.get("position", (0.0, 0.0, 0.0)))
        d.mass = s["mass"]
        d.spin = s.get("spin", 0.0)
        d.axis[:] = list(s.get("axis", (0.0, 0.0, 1.0)))
        d.disk_inner = s.get("disk_inner", 0.0)
        d.disk_outer = s.get("disk_outer", 0.0)
        d.disk_enabled = 1 if s.get("disk_enabled", True) else 0
    return m


# ------------------------------------------------------------ results
class Results:
    """Host SoA result arrays for n rays, pointed to by an rt_results."""

    def __init__(self, n: int):
        self.n = n
        self.cls = np.zeros(n, np.uint32)
        self.x = np.zeros((4, n), np.float64)
        self.p = np.zeros((4, n), np.float64)
        self.bl = np.zeros((3, n), np.float64)
        self.lam = np.zeros(n, np.float64)
        self.g = np.zeros(n, np.float64)
        self.steps = np.zeros(n, np.uint32)
        self.rejected = np.zeros(n, np.uint32)
        self.struct = rt_results()
        self.struct.cls = self.cls.ctypes.data_as(C.POINTER(C.c_uint32))
        self.struct.x = self.x.ctypes.data_as(C.POINTER(C.c_double))
        self.struct.p = self.p.ctypes.data_as(C.POINTER(C.c_double))
        self.struct.bl = self.bl.ctypes.data_as(C.POINTER(C.c_double))
        self.struct.lam = self.lam.ctypes.data_as(C.POINTER(C.c_double))
        self.struct.g = self.g.ctypes.data_as(C.POINTER(C.c_double))
        self.struct.steps = self.steps.ctypes.data_as(C.POINTER(C.c_uint32))
        self.struct.rejected = self.rejected.ctypes.data_as(C.POINTER(C.c_uint32))

    @property
    def klass(self) -> np.ndarray:
        return self.cls & RT_CLS_MASK

    def as_dict(self) -> Dict[str, np.ndarray]:
        return {"cls": self.cls, "x": self.x, "p": self.p, "bl": self.bl, "lam": self.lam, "g": self.g,
                "steps": self.steps, "rejected": self.rejected}


def make_request(cam: rt_camera, metric: rt_metric, integ: rt_integrator, term: rt_termination) -> rt_request:
    r = rt_request()
    r.camera, r.metric, r.integrator, r.termination = cam, metric, integ, term
    return r


def n_rays_of(cam: rt_camera) -> int:
    """Rays of a camera's row window: all rows when row_count is 0 (the
    offset ignored), else rows [offset, offset + count) clipped at the last."""
    if not cam.row_count:
        return cam.height * cam.width
    first = min(cam.row_offset, cam.height)
    return (min(first + cam.row_count, cam.height) - first) * cam.width


# ------------------------------------------------------------ library
_LIB = None

EXPORTS = [
    "rt_default_integrator", "rt_default_termination", "rt_default_camera", "rt_kerr_metric",
    "rt_abi_version", "rt_version", "rt_last_error", "rt_context_create", "rt_context_destroy",
    "rt_get_stats", "rt_trace_camera", "rt_trace_rays", "rt_trace_camera_resident", "rt_fetch_results",
    "rt_fetch_pixels", "rt_tile_geometry", "rt_decode_tile_task", "rt_trace_tile", "rt_trace_tiles",
    "rt_trace_batch", "rt_get_launch_shape", "rt_set_launch_shape", "rt_autotune", "rt_server_create", "rt_server_destroy", "rt_server_submit", "rt_server_submit_ex", "rt_server_poll",
    "rt_server_wait", "rt_server_take", "rt_server_cancel", "rt_trace_tiles_noalloc", "rt_critical_distance",
    "rt_shadow_boundary", "rt_impact_distance",
]


def load(path: Optional[str] = None) -> C.CDLL:
    """Load librtsim.so (built by __graft_entry__.build()); raises if absent."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"librtsim.so not built at {p}; run __graft_entry__.build()")
    lib = C.CDLL(p)
    P = C.POINTER
    lib.rt_version.restype = C.c_char_p
    lib.rt_last_error.restype = C.c_char_p
    lib.rt_abi_version.restype = C.c_uint32
    lib.rt_context_create.argtypes = [C.c_int, P(C.c_void_p)]
    lib.rt_context_destroy.argtypes = [C.c_void_p]
    lib.rt_get_stats.argtypes = [C.c_void_p, P(rt_stats)]
    lib.rt_trace_camera.argtypes = [C.c_void_p, P(rt_camera), P(rt_metric), P(rt_integrator),
                                    P(rt_termination), P(rt_results)]
    lib.rt_trace_rays.argtypes = [C.c_void_p, P(C.c_double), P(C.c_double), C.c_size_t, P(rt_metric),
                                  P(rt_integrator), P(rt_termination), P(rt_results)]
    lib.rt_trace_camera_resident.argtypes = [C.c_void_p, P(rt_camera), P(rt_metric), P(rt_integrator),
                                             P(rt_termination), P(C.c_size_t)]
    lib.rt_fetch_results.argtypes = [C.c_void_p, P(rt_results)]
    lib.rt_fetch_pixels.argtypes = [C.c_void_p, C.c_void_p]
    lib.rt_tile_geometry.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, P(rt_tile)]
    lib.rt_decode_tile_task.argtypes = [C.c_void_p, C.c_size_t, P(rt_tile_task)]
    lib.rt_trace_tile.argtypes = [C.c_void_p, P(rt_camera), P(rt_metric), P(rt_integrator), P(rt_termination),
                                  P(rt_tile), P(C.c_size_t)]
    lib.rt_trace_tiles.argtypes = [C.c_void_p, P(rt_tile_task), C.c_size_t, P(C.c_size_t)]
    lib.rt_trace_tiles_noalloc.argtypes = [C.c_void_p, P(rt_tile_task), C.c_size_t, P(rt_results), C.c_size_t,
                                           P(C.c_size_t)]
    lib.rt_critical_distance.argtypes = [P(rt_camera), P(rt_metric), P(C.c_double)]
    lib.rt_shadow_boundary.argtypes = [C.c_double, C.c_double, C.c_double, C.c_size_t, P(C.c_double), P(C.c_double)]
    lib.rt_impact_distance.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.c_double]
    lib.rt_impact_distance.restype = C.c_double
    lib.rt_trace_batch.argtypes = [C.c_void_p, P(rt_request), C.c_size_t, P(rt_results)]
    lib.rt_get_launch_shape.argtypes = [C.c_void_p, P(rt_launch_shape), P(rt_launch_shape)]
    lib.rt_set_launch_shape.argtypes = [C.c_void_p, P(rt_launch_shape), P(rt_launch_shape)]
    lib.rt_autotune.argtypes = [C.c_void_p, P(rt_request), C.c_size_t, P(rt_launch_shape)]
    lib.rt_default_integrator.argtypes = [P(rt_integrator)]
    lib.rt_default_termination.argtypes = [P(rt_termination)]
    lib.rt_default_camera.argtypes = [P(rt_camera), C.c_double, C.c_double, C.c_double, C.c_double,
                                      C.c_uint32, C.c_uint32]
    lib.rt_kerr_metric.argtypes = [P(rt_metric), C.c_double, C.c_double]
    lib.rt_server_create.argtypes = [C.c_int, P(rt_server_config), P(C.c_void_p)]
    lib.rt_server_destroy.argtypes = [C.c_void_p]
    lib.rt_server_submit.argtypes = [C.c_void_p, P(rt_request), P(C.c_uint64)]
    lib.rt_server_submit_ex.argtypes = [C.c_void_p, P(rt_request), P(rt_submit_options), P(C.c_uint64)]
    lib.rt_server_poll.argtypes = [C.c_void_p, C.c_uint64, P(C.c_int)]
    lib.rt_server_wait.argtypes = [C.c_void_p, C.c_uint64, C.c_int64]
    lib.rt_server_take.argtypes = [C.c_void_p, C.c_uint64, P(rt_results)]
    lib.rt_server_cancel.argtypes = [C.c_void_p, C.c_uint64]
    if path is None:
        _LIB = lib
    return lib


def _check(lib, status: int) -> None:
    if status != 0:
        raise RtError(status, lib.rt_last_error().decode())


def tile_geometry(tile_index: int, screen_width: int, screen_height: int) -> rt_tile:
    """Tile tile_index of a frame (resolve_tile_geometry); RtError where the
    reference returns nullopt. Host-only: needs no device."""
    lib = load()
    t = rt_tile()
    _check(lib, lib.rt_tile_geometry(tile_index, screen_width, screen_height, C.byref(t)))
    return t


def decode_tile_task(data: bytes) -> rt_tile_task:
    """Decode and validate a RenderTileTaskPayload's wire bytes; RtError where
    the reference's deserialize_render_tile_task returns false (and for a
    metric outside the Kerr family). Host-only: needs no device."""
    lib = load()
    buf = C.create_string_buffer(bytes(data), len(data))
    t = rt_tile_task()
    _check(lib, lib.rt_decode_tile_task(buf, len(data), C.byref(t)))
    return t


def critical_distance(cam: rt_camera, metric: rt_metric) -> np.ndarray:
    """Each traced pixel's distance (units of M) to the Kerr critical curve in
    the impact-parameter plane (rt_critical_distance). Host-only."""
    lib = load()
    out = np.zeros(n_rays_of(cam), np.float64)
    _check(lib, lib.rt_critical_distance(C.byref(cam), C.byref(metric), out.ctypes.data_as(C.POINTER(C.c_double))))
    return out


def shadow_boundary(mass: float, spin: float, theta_obs: float, n: int = 2000):
    """n points (alpha, beta) of the critical curve (rt_shadow_boundary)."""
    lib = load()
    al, be = np.zeros(n), np.zeros(n)
    _check(lib, lib.rt_shadow_boundary(mass, spin, theta_obs, n, al.ctypes.data_as(C.POINTER(C.c_double)),
                                       be.ctypes.data_as(C.POINTER(C.c_double))))
    return al, be


def impact_distance(mass: float, spin: float, theta_obs: float, alpha: float, beta: float) -> float:
    return load().rt_impact_distance(mass, spin, theta_obs, alpha, beta)


class Context:
    """One device context (device buffers, stream, kernel statistics)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        self.h = C.c_void_p()
        _check(self.lib, self.lib.rt_context_create(device, C.byref(self.h)))

    def close(self) -> None:
        if self.h:
            self.lib.rt_context_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def stats(self) -> rt_stats:
        s = rt_stats()
        _check(self.lib, self.lib.rt_get_stats(self.h, C.byref(s)))
        return s

    def trace_camera(self, cam: rt_camera, metric: rt_metric, integ: rt_integrator,
                     term: rt_termination) -> Results:
        res = Results(n_rays_of(cam))
        _check(self.lib, self.lib.rt_trace_camera(self.h, C.byref(cam), C.byref(metric), C.byref(integ),
                                                  C.byref(term), C.byref(res.struct)))
        return res

    def trace_camera_resident(self, cam: rt_camera, metric: rt_metric, integ: rt_integrator,
                              term: rt_termination) -> int:
        n = C.c_size_t()
        _check(self.lib, self.lib.rt_trace_camera_resident(self.h, C.byref(cam), C.byref(metric),
                                                           C.byref(integ), C.byref(term), C.byref(n)))
        return n.value

    def trace_tile(self, cam: rt_camera, metric: rt_metric, integ: rt_integrator, term: rt_termination,
                   tile: rt_tile) -> int:
        """Trace the rectangle tile of cam's frame; results stay resident
        (tile-local row-major) for fetch / fetch_pixels. Returns the ray count."""
        n = C.c_size_t()
        _check(self.lib, self.lib.rt_trace_tile(self.h, C.byref(cam), C.byref(metric), C.byref(integ),
                                                C.byref(term), C.byref(tile), C.byref(n)))
        return n.value

    def trace_tiles(self, tasks) -> int:
        """Trace decoded tile tasks (a sequence, or a ctypes array of
        rt_tile_task) as one batch; results stay resident, task after task
        (each tile-local row-major). Returns the total ray count."""
        arr = tasks if isinstance(tasks, C.Array) else (rt_tile_task * len(tasks))(*tasks)
        n = C.c_size_t()
        _check(self.lib, self.lib.rt_trace_tiles(self.h, arr, len(tasks), C.byref(n)))
        return n.value

    def fetch(self, n: int) -> Results:
        res = Results(n)
        _check(self.lib, self.lib.rt_fetch_results(self.h, C.byref(res.struct)))
        return res

    def fetch_pixels(self, n: int) -> np.ndarray:
        """The last resident trace as n 32-byte pixel records (PIXEL_DTYPE)."""
        out = np.zeros(n, PIXEL_DTYPE)
        _check(self.lib, self.lib.rt_fetch_pixels(self.h, out.ctypes.data_as(C.c_void_p)))
        return out

    def trace_rays(self, x: np.ndarray, p: np.ndarray, metric: rt_metric, integ: rt_integrator,
                   term: rt_termination) -> Results:
        x = np.ascontiguousarray(x, np.float64)
        p = np.ascontiguousarray(p, np.float64)
        if x.ndim != 2 or x.shape[0] != 4 or p.shape != x.shape:
            raise ValueError("x and p must be (4, n) arrays")
        n = x.shape[1]
        res = Results(n)
        _check(self.lib, self.lib.rt_trace_rays(self.h, x.ctypes.data_as(C.POINTER(C.c_double)),
                                                p.ctypes.data_as(C.POINTER(C.c_double)), n, C.byref(metric),
                                                C.byref(integ), C.byref(term), C.byref(res.struct)))
        return res

    def trace_batch(self, reqs: Sequence[rt_request], fetch: bool = True,
                    only: Optional[Sequence[int]] = None) -> Optional[List[Optional[Results]]]:
        """All requests in one work-queue kernel (captured hipGraph, replayed
        while the layout is unchanged). fetch=False keeps results on the device;
        only: copy back just these requests (the others' entries are None)."""
        arr = (rt_request * len(reqs))(*reqs)
        if not fetch:
            _check(self.lib, self.lib.rt_trace_batch(self.h, arr, len(reqs), None))
            return None
        keep = set(range(len(reqs))) if only is None else set(only)
        outs = [Results(n_rays_of(r.camera)) if i in keep else None for i, r in enumerate(reqs)]
        oarr = (rt_results * len(reqs))(*[o.struct if o is not None else rt_results() for o in outs])
        _check(self.lib, self.lib.rt_trace_batch(self.h, arr, len(reqs), oarr))
        return outs

    def launch_shapes(self):
        plain, queue = rt_launch_shape(), rt_launch_shape()
        _check(self.lib, self.lib.rt_get_launch_shape(self.h, C.byref(plain), C.byref(queue)))
        return plain, queue

    def set_launch_shapes(self, plain: Optional[rt_launch_shape] = None, queue: Optional[rt_launch_shape] = None):
        _check(self.lib, self.lib.rt_set_launch_shape(self.h, C.byref(plain) if plain else None,
                                                      C.byref(queue) if queue else None))

    def autotune(self, sample: Sequence[rt_request]) -> rt_launch_shape:
        arr = (rt_request * len(sample))(*sample)
        best = rt_launch_shape()
        _check(self.lib, self.lib.rt_autotune(self.h, arr, len(sample), C.byref(best)))
        return best


class Server:
    """Request loop on one device (rt_server_*): requests run as soon as memory
    and a lane allow, next to the ones already on the device."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 8 << 30, max_inflight: int = 64,
                 staging_mib: int = 0):
        self.lib = load()
        cfg = rt_server_config(hbm_budget_bytes, max_inflight, staging_mib)
        self.h = C.c_void_p()
        _check(self.lib, self.lib.rt_server_create(device, C.byref(cfg), C.byref(self.h)))
        self._sizes: Dict[int, tuple] = {}

    def close(self) -> None:
        if self.h:
            self.lib.rt_server_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, req: rt_request, outputs: int = RT_OUT_ALL, deadline_ms: int = 0) -> int:
        """Queue a request; outputs: RT_OUT_* bits to copy back; deadline_ms > 0: expire after it."""
        rid = C.c_uint64()
        if outputs == RT_OUT_ALL and deadline_ms == 0:
            _check(self.lib, self.lib.rt_server_submit(self.h, C.byref(req), C.byref(rid)))
        else:
            opt = rt_submit_options(outputs, 0, deadline_ms)
            _check(self.lib, self.lib.rt_server_submit_ex(self.h, C.byref(req), C.byref(opt), C.byref(rid)))
        self._sizes[rid.value] = (n_rays_of(req.camera), outputs)
        return rid.value

    def poll(self, rid: int) -> bool:
        d = C.c_int()
        _check(self.lib, self.lib.rt_server_poll(self.h, rid, C.byref(d)))
        return bool(d.value)

    def wait(self, rid: int, timeout_ms: int = -1) -> None:
        _check(self.lib, self.lib.rt_server_wait(self.h, rid, timeout_ms))

    def take(self, rid: int) -> Results:
        """Results of a finished request; fields not requested at submit stay zero."""
        n, outputs = self._sizes.get(rid, (0, RT_OUT_ALL))
        res = Results(n)
        for name, bit in OUT_FIELDS:
            if not outputs & bit:
                setattr(res.struct, name, None)
        st = self.lib.rt_server_take(self.h, rid, C.byref(res.struct))
        if st != 7:  # RT_ERR_BUSY: still running, keep the bookkeeping
            self._sizes.pop(rid, None)
        _check(self.lib, st)
        return res

    def cancel(self, rid: int) -> None:
        _check(self.lib, self.lib.rt_server_cancel(self.h, rid))
