"""Superposition (2..16 Kerr-Schild sources) on the GPU vs the CPU oracle.

Bar: class identical; final KS position, BL (r, theta, phi; phi as a wrapped
difference), affine parameter and g within 1e-9 relative (max(|ref|, 1) denominator) for rays that are not
captured; captured rays compare class only (they end on a horizon)."""
import math

import numpy as np
import pytest

import oracle_lib as O
import rtsim

pytestmark = pytest.mark.gpu
TOL = 1e-9


def rel(a, b):
    return np.abs(a - b) / np.maximum(np.abs(b), 1.0)


def phi_err(a, b):
    d = np.remainder(a - b + np.pi, 2.0 * np.pi) - np.pi
    return np.abs(d) / np.maximum(np.abs(b), 1.0)


def two_holes():
    return rtsim.multi_metric([
        {"mass": 1.0, "spin": 0.5, "position": (0, 0, 0), "axis": (0, 0, 1)},
        {"mass": 0.5, "spin": 0.6, "position": (18.0, -6.0, 4.0), "axis": (0.3, 0.1, 1.0)},
    ])


def many_holes(n, seed):
    rng = np.random.default_rng(seed)
    src = [{"mass": 1.0, "spin": 0.7, "position": (0, 0, 0), "axis": (0, 0, 1)}]
    for _ in range(n - 1):
        ax = rng.normal(size=3)
        src.append({"mass": float(rng.uniform(0.1, 1.0)), "spin": float(rng.uniform(0, 0.95)),
                    "position": tuple(float(v) for v in rng.uniform(-25, 25, size=3)), "axis": tuple(ax)})
    return rtsim.multi_metric(src)


def compare(gpu, ref, label):
    print(f"[{label}] n={ref.n} classes {np.bincount(ref.klass, minlength=5)[[1, 2, 4]]}", flush=True)
    bad = np.where(gpu.klass != ref.klass)[0]
    assert bad.size == 0, f"{label}: class mismatch at {bad[:10]}"
    nc = ref.klass != rtsim.RT_CLS_CAPTURED
    errs = {"x": rel(gpu.x[1:, nc], ref.x[1:, nc]).max(initial=0),
            "r": rel(gpu.bl[0, nc], ref.bl[0, nc]).max(initial=0),
            "theta": rel(gpu.bl[1, nc], ref.bl[1, nc]).max(initial=0),
            "phi": phi_err(gpu.bl[2, nc], ref.bl[2, nc]).max(initial=0),
            "lambda": rel(gpu.lam[nc], ref.lam[nc]).max(initial=0),
            "g": rel(gpu.g, ref.g).max(initial=0)}
    print(f"[{label}] max rel err " + " ".join(f"{k}={v:.2e}" for k, v in errs.items()), flush=True)
    for k, v in errs.items():
        assert v < TOL, f"{label}: {k} rel err {v:.3e}"


def test_two_holes_camera(ctx):
    cam = rtsim.default_camera(200.0, math.radians(70.0), 0.2, math.radians(24.0), 32, 32)
    met = two_holes()
    ig = rtsim.default_integrator()
    tm = rtsim.default_termination(escape_radius=400.0)
    gpu = ctx.trace_camera(cam, met, ig, tm)
    ref = O.trace_camera(cam, met, ig, tm)
    assert (ref.klass == 1).sum() > 0 and (ref.klass == 2).sum() > 0
    compare(gpu, ref, "two holes")


# Every lane-group shape of the kernel: 2 lanes (2 holes, above), 4 lanes with
# an idle lane (3) and full (4), 8 lanes with idle lanes (5) and full (8),
# 8 lanes with a second source on one lane (9) and on every lane (16).
@pytest.mark.parametrize("n", [3, 4, 5, 8, 9, 16])
def test_many_holes_camera(ctx, n):
    cam = rtsim.default_camera(300.0, math.radians(55.0), -0.4, math.radians(20.0), 24, 24)
    met = many_holes(n, 7 + n)
    ig = rtsim.default_integrator()
    tm = rtsim.default_termination(escape_radius=600.0)
    gpu = ctx.trace_camera(cam, met, ig, tm)
    ref = O.trace_camera(cam, met, ig, tm)
    compare(gpu, ref, f"{n} holes")


def test_multi_rays_mode(ctx):
    cam = rtsim.default_camera(200.0, math.radians(70.0), 0.2, math.radians(24.0), 12, 12)
    met = two_holes()
    ig = rtsim.default_integrator()
    tm = rtsim.default_termination(escape_radius=400.0)
    n = 144
    x = np.zeros((4, n))
    p = np.zeros((4, n))
    for i in range(n):
        ok, y, _ = O.camera_state(cam, met, tm, i % 12, i // 12)
        assert ok
        x[:, i], p[:, i] = y[:4], y[4:]
    gpu = ctx.trace_rays(x, p, met, ig, tm)
    ref = O.trace_rays(x, p, met, ig, tm)
    compare(gpu, ref, "multi rays")
