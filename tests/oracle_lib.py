"""ctypes wrappers of the CPU oracle (oracle/liboracle.so) and of the shims over
the reference's own headers (oracle/_ref/libref.so, present only where the
reference tree was available at build time). Test infrastructure."""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "relativity-simulator_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)

import rtsim  # noqa: E402
from rtsim import (rt_camera, rt_integrator, rt_metric, rt_results, rt_termination)  # noqa: E402

ORACLE_PATH = os.path.join(REPO, "oracle", "liboracle.so")
REF_PATH = os.path.join(REPO, "oracle", "_ref", "libref.so")

D = C.c_double
PD = C.POINTER(C.c_double)
P = C.POINTER


def _arr(a):
    a = np.ascontiguousarray(a, np.float64)
    return a, a.ctypes.data_as(PD)


_ORACLE = None
_REF = None


def oracle():
    global _ORACLE
    if _ORACLE is None:
        if not os.path.exists(ORACLE_PATH):
            raise RuntimeError("oracle/liboracle.so not built; run __graft_entry__.build()")
        lib = C.CDLL(ORACLE_PATH)
        lib.rto_trace_camera.argtypes = [P(rt_camera), P(rt_metric), P(rt_integrator), P(rt_termination),
                                         P(rt_results), C.c_int]
        lib.rto_trace_camera_rows.argtypes = [P(rt_camera), P(rt_metric), P(rt_integrator), P(rt_termination),
                                              C.c_uint32, C.c_uint32, C.c_uint32, P(rt_results), C.c_int]
        lib.rto_trace_rays.argtypes = [PD, PD, C.c_size_t, P(rt_metric), P(rt_integrator), P(rt_termination),
                                       P(rt_results), C.c_int]
        lib.rto_camera_state.argtypes = [P(rt_camera), P(rt_metric), P(rt_termination), C.c_uint32, C.c_uint32,
                                         PD, PD]
        lib.rto_ks_field.argtypes = [PD, D, D, PD]
        lib.rto_rhs.argtypes = [PD, P(rt_metric), PD]
        lib.rto_constraint.argtypes = [PD, P(rt_metric)]
        lib.rto_constraint.restype = D
        lib.rto_project.argtypes = [PD, P(rt_metric), D]
        lib.rto_initial_momentum.argtypes = [PD, D, D, PD, PD]
        lib.rto_ms_table.argtypes = [P(rt_metric), P(rt_termination), PD]
        lib.rto_ms_table.restype = C.c_int
        lib.rto_ms_initial_momentum.argtypes = [P(rt_metric), PD, PD, PD]
        lib.rto_ks_from_bl.argtypes = [D, D, D, D, PD]
        lib.rto_bl_from_ks.argtypes = [PD, D, PD]
        lib.rto_isco.argtypes = [D, D]
        lib.rto_isco.restype = D
        lib.rto_carter_verdict.argtypes = [D, D, D, D, D]
        lib.rto_carter_verdict.restype = C.c_int
        lib.rto_redshift.argtypes = [D, D, D, D, C.c_int]
        lib.rto_redshift.restype = D
        lib.rto_dp45_attempt.argtypes = [PD, D, P(rt_metric), D, D, PD, PD]
        lib.rto_dp45_attempt.restype = D
        lib.rto_test_ode.argtypes = [C.c_int, D, PD, D, D, C.c_int, D, PD, PD]
        lib.rto_test_ode.restype = D
        lib.rto_test_root.argtypes = [PD]
        lib.rto_test_root.restype = D
        lib.rto_trace_one.argtypes = [PD, P(rt_metric), P(rt_integrator), P(rt_termination), P(C.c_uint32), PD, PD,
                                      PD, PD, P(C.c_uint32), P(C.c_uint32)]
        lib.rto_trace_render_camera.argtypes = [P(rt_camera), P(rt_metric), P(rt_integrator), P(rt_termination),
                                                P(rt_results), C.c_int]
        lib.rto_screen_direction.argtypes = [C.c_uint32, D, D, D, PD]
        if hasattr(lib, "rto_flop_counts"):
            lib.rto_flop_counts.argtypes = [PD]
        _ORACLE = lib
    return _ORACLE


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    global _REF
    if _REF is None:
        lib = C.CDLL(REF_PATH)
        lib.ref_ks_field.argtypes = [PD, D, D, PD]
        lib.ref_ks_rhs.argtypes = [PD, PD, D, D, PD, PD]
        lib.ref_ks_rk4_step.argtypes = [PD, PD, D, D, D]
        lib.ref_initial_momentum.argtypes = [PD, D, D, PD, PD]
        lib.ref_ks_from_bl.argtypes = [D, D, D, D, PD]
        lib.ref_bl_from_ks.argtypes = [PD, D, D, PD]
        lib.ref_isco.argtypes = [D, D]
        lib.ref_isco.restype = D
        lib.ref_redshift.argtypes = [D, D, D, D]
        lib.ref_redshift.restype = D
        lib.ref_ks_metric.argtypes = [PD, D, D, PD, PD]
        lib.ref_rk45_integrate.argtypes = [PD, PD, D, D, D, D, D, D, PD, PD, P(C.c_uint64), P(C.c_uint64)]
        lib.ref_rk45_integrate.restype = C.c_int
        lib.ref_ms_table.argtypes = [P(rt_metric), PD]
        lib.ref_ms_table.restype = C.c_int
        lib.ref_ms_rhs.argtypes = [P(rt_metric), PD, PD, PD, PD]
        lib.ref_ms_initial_momentum.argtypes = [P(rt_metric), PD, PD, PD]
        lib.ref_screen_direction.argtypes = [C.c_uint32, D, D, D, PD]
        lib.ref_render_frame.argtypes = [P(rt_camera), P(rt_metric), P(rt_integrator), P(rt_termination),
                                         P(C.c_uint32), P(C.c_uint32), P(C.c_float), P(C.c_float), P(C.c_float)]
        lib.ref_pixel_layout.argtypes = [P(C.c_uint64)]
        lib.ref_pixel_layout.restype = C.c_int
        _REF = lib
    return _REF


# ------------------------------------------------------------------ helpers
def trace_camera(cam, metric, integ, term, threads: int = 0) -> rtsim.Results:
    res = rtsim.Results(rtsim.n_rays_of(cam))
    st = oracle().rto_trace_camera(C.byref(cam), C.byref(metric), C.byref(integ), C.byref(term),
                                   C.byref(res.struct), threads)
    assert st == 0
    return res


def trace_camera_rows(cam, metric, integ, term, first: int, stride: int, count: int,
                      threads: int = 0) -> rtsim.Results:
    res = rtsim.Results(count * cam.width)
    st = oracle().rto_trace_camera_rows(C.byref(cam), C.byref(metric), C.byref(integ), C.byref(term),
                                        first, stride, count, C.byref(res.struct), threads)
    assert st == 0
    return res


def trace_rays(x, p, metric, integ, term, threads: int = 0) -> rtsim.Results:
    x, px = _arr(x)
    p, pp = _arr(p)
    n = x.shape[1]
    res = rtsim.Results(n)
    st = oracle().rto_trace_rays(px, pp, n, C.byref(metric), C.byref(integ), C.byref(term),
                                 C.byref(res.struct), threads)
    assert st == 0
    return res


def camera_state(cam, metric, term, px: int, py: int):
    y = np.zeros(8)
    xi = D()
    ok = oracle().rto_camera_state(C.byref(cam), C.byref(metric), C.byref(term), px, py,
                                   y.ctypes.data_as(PD), C.byref(xi))
    return bool(ok), y, xi.value


def rhs(y, metric):
    y, py = _arr(y)
    d = np.zeros(8)
    oracle().rto_rhs(py, C.byref(metric), d.ctypes.data_as(PD))
    return d


def constraint(y, metric) -> float:
    y, py = _arr(y)
    return oracle().rto_constraint(py, C.byref(metric))


def trace_one(y0, metric, integ, term):
    y0, py = _arr(y0)
    cls = C.c_uint32()
    yend = np.zeros(8)
    bl = np.zeros(3)
    lam, g = D(), D()
    acc, rej = C.c_uint32(), C.c_uint32()
    oracle().rto_trace_one(py, C.byref(metric), C.byref(integ), C.byref(term), C.byref(cls),
                           yend.ctypes.data_as(PD), bl.ctypes.data_as(PD), C.byref(lam), C.byref(g),
                           C.byref(acc), C.byref(rej))
    return {"cls": cls.value, "y": yend, "bl": bl, "lam": lam.value, "g": g.value, "acc": acc.value,
            "rej": rej.value}


def trace_render_camera(cam, metric, integ, term, threads: int = 0) -> rtsim.Results:
    """RT_METHOD_RENDER_RK4 frame (the renderer's tracer) on the oracle."""
    res = rtsim.Results(rtsim.n_rays_of(cam))
    st = oracle().rto_trace_render_camera(C.byref(cam), C.byref(metric), C.byref(integ), C.byref(term),
                                          C.byref(res.struct), threads)
    assert st == 0
    return res


def ref_render_frame(cam, metric, integ, term):
    """The reference's dispatch_fp64 pixel records of a Kerr frame (libref):
    status_flags, iterations_used, redshift, affine_parameter, rgb."""
    n = rtsim.n_rays_of(cam)
    out = {"status": np.zeros(n, np.uint32), "iters": np.zeros(n, np.uint32), "redshift": np.zeros(n, np.float32),
           "affine": np.zeros(n, np.float32), "rgb": np.zeros((n, 3), np.float32)}
    ref().ref_render_frame(C.byref(cam), C.byref(metric), C.byref(integ), C.byref(term),
                           out["status"].ctypes.data_as(P(C.c_uint32)), out["iters"].ctypes.data_as(P(C.c_uint32)),
                           out["redshift"].ctypes.data_as(P(C.c_float)), out["affine"].ctypes.data_as(P(C.c_float)),
                           out["rgb"].ctypes.data_as(P(C.c_float)))
    return out
