"""Seeded synthetic trace requests of the tracer's workloads.

Shared by bench.py and the GPU tests so that both trace the same rays.

  frames4      four 2048^2 Kerr frames, a in {0, 0.5, 0.9, 0.998}, inclination
               60 deg, observer r = 1000 M, vertical FOV 20 deg, rtol 1e-10
  precision    Kerr a = 0.998, inclination 85 deg, 4096^2 (mixed vs FP64)
  uneven512    512 camera requests: side U[64, 1024], a U[0, 0.998],
               inclination U[5, 89] deg, r_obs log-uniform in [50, 5000] M,
               rtol in {1e-8, 1e-10}
  serving      Poisson-arrival requests: lognormal ray counts (median 256^2,
               sigma 1, at most 2048^2), 70 % single Kerr, 30 % superpositions
               of 2..16 holes (masses U[0.1, 1], positions U[-25, 25]^3 M,
               spins U[0, 0.95] on random axes)
  multi        superposition frames of 2, 4, 8 and 16 holes drawn as in serving
  render       the renderer's own tracer (RT_METHOD_RENDER_RK4) on a near
               observer: a = 0.9, r_obs = 30 M, inclination 65 deg, FOV 60 deg,
               translucent disk to 24 M, escape at 100 M, budget 2048 steps
"""
from __future__ import annotations

import math
from typing import List, Tuple

import numpy as np

import rtsim

SEED = 0xC0FFEE1234
SPINS = (0.0, 0.5, 0.9, 0.998)
R_OBS = 1000.0
FOV_DEG = 20.0


def frame_requests(side: int = 2048, precision: int = rtsim.RT_PRECISION_FP64, incl_deg: float = 60.0):
    """The four Kerr frames (one per spin) as (camera, metric, integrator, termination)."""
    out = []
    for a in SPINS:
        cam = rtsim.default_camera(R_OBS, math.radians(incl_deg), 0.0, math.radians(FOV_DEG), side, side)
        out.append((cam, rtsim.kerr_metric(1.0, a), rtsim.default_integrator(precision=precision),
                    rtsim.default_termination()))
    return out


def precision_request(side: int = 4096, precision: int = rtsim.RT_PRECISION_FP64, switch_radius: float = 0.0):
    """Kerr a = 0.998 at inclination 85 deg; switch_radius 0 means 3 r_+."""
    cam = rtsim.default_camera(R_OBS, math.radians(85.0), 0.0, math.radians(FOV_DEG), side, side)
    return (cam, rtsim.kerr_metric(1.0, 0.998),
            rtsim.default_integrator(precision=precision, switch_radius=switch_radius), rtsim.default_termination())


def uneven_requests(n: int = 512, seed: int = SEED) -> List[rtsim.rt_request]:
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        side = int(rng.integers(64, 1025))
        r_obs = float(math.exp(rng.uniform(math.log(50.0), math.log(5000.0))))
        cam = rtsim.default_camera(r_obs, math.radians(rng.uniform(5, 89)), 0.0, math.radians(FOV_DEG), side, side)
        rtol = 1e-8 if rng.random() < 0.5 else 1e-10
        out.append(rtsim.make_request(cam, rtsim.kerr_metric(1.0, float(rng.uniform(0, 0.998))),
                                      rtsim.default_integrator(rtol=rtol),
                                      rtsim.default_termination(escape_radius=max(1000.0, r_obs))))
    return out


def random_holes(rng, k: int) -> rtsim.rt_metric:
    """Central hole (mass 1) plus k - 1 holes drawn as in the serving mix."""
    src = [{"mass": 1.0, "spin": float(rng.uniform(0, 0.95))}]
    for _ in range(k - 1):
        src.append({"mass": float(rng.uniform(0.1, 1.0)), "spin": float(rng.uniform(0, 0.95)),
                    "position": tuple(float(v) for v in rng.uniform(-25, 25, size=3)),
                    "axis": tuple(float(v) for v in rng.normal(size=3))})
    return rtsim.multi_metric(src)


def serving_requests(n: int, seed: int = SEED, max_rays: int = 2048 * 2048) -> List[rtsim.rt_request]:
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        rays = min(max_rays, int(round(65536 * math.exp(rng.normal()))))
        side = max(8, int(round(math.sqrt(rays))))
        cam = rtsim.default_camera(R_OBS, math.radians(rng.uniform(5, 89)), rng.uniform(-math.pi, math.pi),
                                   math.radians(FOV_DEG), side, side)
        if rng.random() < 0.3:
            met = random_holes(rng, int(rng.integers(2, 17)))
        else:
            met = rtsim.kerr_metric(1.0, float(rng.uniform(0, 0.998)))
        out.append(rtsim.make_request(cam, met, rtsim.default_integrator(), rtsim.default_termination()))
    return out


def poisson_arrivals(n: int, qps: float, seed: int = SEED + 1) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / qps, size=n))


def multi_requests(side: int = 512, counts: Tuple[int, ...] = (2, 4, 8, 16), seed: int = SEED,
                   incl_deg: float = 60.0) -> List[rtsim.rt_request]:
    rng = np.random.default_rng(seed)
    out = []
    for k in counts:
        cam = rtsim.default_camera(R_OBS, math.radians(incl_deg), 0.0, math.radians(FOV_DEG), side, side)
        out.append(rtsim.make_request(cam, random_holes(rng, k), rtsim.default_integrator(),
                                      rtsim.default_termination()))
    return out


def render_request(side: int = 2048):
    """The render workload as (camera, metric, integrator, termination): the
    scene of the reference-record fixture disk_pinhole at side x side."""
    cam = rtsim.default_camera(30.0, math.radians(65.0), 0.0, math.radians(60.0), side, side)
    tm = rtsim.default_termination(escape_radius=100.0, disk_outer=24.0, disk_enabled=True)
    return cam, rtsim.kerr_metric(1.0, 0.9), rtsim.render_integrator(2048), tm
