"""Throughput benchmark of the MI355X null-geodesic tracer.

Workload (one step): four 2048x2048 Kerr frames, spins a in {0, 0.5, 0.9,
0.998}, M = 1, observer at r = 1000 M, inclination 60 deg, vertical FOV 20 deg,
DP45 rtol 1e-10 / atol 1e-12, capture at r_+ + 1e-6, escape at r = 1000 M
outgoing, terminal disk z = 0 for r in [r_isco, 20 M]. Ray states are
generated on the device and results stay in HBM: the four frames are one
batch (rt_trace_batch, one persistent kernel, replayed from a hipGraph);
--path plain traces them with one launch per frame (rt_trace_camera_resident).
--workload uneven512 times the 512-request batched workload instead.
The "tiles" object reports the 32x32 tile-task path on the a = 0.9 frame;
"render" the renderer's own RK4 tracer (RT_METHOD_RENDER_RK4) on a 2048^2
near-observer disk scene, with the CPU restatement of that tracer and the
reference's dispatch_fp64 timed on row bands of the same frame.

Prints one JSON line. value = rays/s over the timed steps. roofline: the
algorithmic FP64 FLOP per launch (right-hand-side evaluations x FLOP_RHS +
attempted steps x FLOP_ATTEMPT, op-counted on the CPU oracle) over the
launch's HIP-event duration on its stream, against the MI355X FP64 vector
peak. A batch is one launch: one replay of its captured graph, whose per-frame
kernels run as concurrent branches.
cpu_baseline: the CPU oracle (a clean-room C++ port, FP64), built on the host
that runs the bench with -O3 -march=native, on every k-th row of the same
frames, on every CPU this process may run on (affinity mask and cgroup
quota), with nproc and the CPU model stated.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "relativity-simulator_amd"))

import numpy as np  # noqa: E402

import rtsim  # noqa: E402
import workloads  # noqa: E402

SPINS = (0.0, 0.5, 0.9, 0.998)
SIDE = 2048
INCL_DEG = 60.0
FOV_DEG = 20.0
R_OBS = 1000.0
# Algorithmic FLOP per right-hand-side evaluation and per attempted DP45 step
# (stage sums, error norm and controller over the 7 evolving components), from
# the operation-counting instantiation of the CPU oracle; tests check that the
# oracle still counts these values.
FLOP_RHS = 178
FLOP_ATTEMPT = 432
# One source's share of a superposition's right-hand side (offset and
# rotations into the source's frame, its KS terms, rotations back, sums),
# from the same operation count.
FLOP_SOURCE = 244
PEAK_FP64_TFLOPS = 78.6
# HBM bytes per ray of the tracing kernels on MI355X: rocprofv3 FETCH_SIZE +
# WRITE_SIZE per dispatch, each scaled by what the counter reads for a plain
# device-to-device copy of known size (x2.000 and x1.000), over the rays of
# the dispatch (4 x 2048^2 Kerr frames; for the superposition kernel the four
# 512^2 frames of 2/4/8/16 holes, 2077/1536/2450/2344 B, averaged). A camera
# ray reads nothing and writes 116 B of results; the superposition kernel's
# excess is its per-lane scratch reaching HBM.
HBM_BYTES_PER_RAY = {"k_trace_camera_kerr": 119.86, "k_trace_tiles": 116.55, "k_trace_queue": 116.55,
                     "k_trace_multi": 2101.8}
METRIC = "rays/s Kerr 2048\u00b2 mixed vs FP64 at \u22641e-9 rel err; p50/p99 latency @QPS"


def frame_requests(side: int, precision: int):
    return workloads.frame_requests(side, precision, INCL_DEG)


def host_cpus():
    """The CPUs this process may run on: the affinity mask, bounded by a
    cgroup CPU quota when one is set; with nproc and the CPU model."""
    info = {"nproc": os.cpu_count() or 1}
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except AttributeError:
        info["affinity"] = info["nproc"]
    usable = info["affinity"]
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if quota != "max":
            info["cgroup_quota_cpus"] = int(quota) / int(period)
            usable = min(usable, max(1, math.ceil(info["cgroup_quota_cpus"])))
    except (OSError, ValueError):
        pass
    info["threads"] = usable
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    info["model"] = model
    return info


_NATIVE = []


def native_oracle():
    """Build the oracle for this host (-O3 -march=native, as the reference's
    own build flags) once per process and return its path; None if the build
    fails."""
    import subprocess

    if not _NATIVE:
        r = subprocess.run(["make", "-B", "-s", "-C", os.path.join(REPO, "oracle"), "liboracle_native.so"],
                           capture_output=True, text=True)
        path = os.path.join(REPO, "oracle", "liboracle_native.so")
        _NATIVE.append(path if r.returncode == 0 and os.path.exists(path) else None)
    return _NATIVE[0]


def cpu_baseline(side: int, stride: int, threads: int):
    """Time the CPU oracle on rows 0, stride, 2*stride, ... of each frame."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_lib

    cpus = host_cpus()
    threads = threads or cpus["threads"]
    native = native_oracle()
    if native:
        oracle_lib.ORACLE_PATH = native
    flags = "-O3 -march=native" if native else "-O2 (portable; the native build failed)"
    rows = (side + stride - 1) // stride
    n = 0
    t0 = time.perf_counter()
    for cam, met, ig, tm in frame_requests(side, rtsim.RT_PRECISION_FP64):
        oracle_lib.trace_camera_rows(cam, met, ig, tm, 0, stride, rows, threads=threads)
        n += rows * side
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "rays/s", "cores": threads, "kind": "port",
            "sample": f"every {stride}th row of the four {side}x{side} frames ({n} rays, {dt:.1f} s), "
                      f"oracle/{os.path.basename(oracle_lib.ORACLE_PATH)} built {flags} "
                      f"-frounding-math -fno-math-errno -ffp-contract=off on this host",
            "host": cpus}


def serve_latency(qps: float, n: int, seed: int = 0xC0FFEE1234):
    """Poisson arrivals at a fixed QPS into the serving loop; latency is
    submit -> results on the host (every output). Reported overall and per
    request kind."""
    reqs = workloads.serving_requests(n, seed)
    arrivals = workloads.poisson_arrivals(n, qps, seed + 1)
    lat, kind = [], []
    rays = 0
    with rtsim.Server(0, hbm_budget_bytes=16 << 30, max_inflight=256, staging_mib=4096) as srv:
        pending = {}
        t0 = time.perf_counter()
        i = 0
        while i < n or pending:
            now = time.perf_counter() - t0
            while i < n and arrivals[i] <= now:
                pending[srv.submit(reqs[i])] = (time.perf_counter(), rtsim.n_rays_of(reqs[i].camera),
                                                reqs[i].metric.n_sources >= 2)
                i += 1
            for rid in [r for r in pending if srv.poll(r)]:
                ts, nr, multi = pending.pop(rid)
                srv.take(rid)
                lat.append(time.perf_counter() - ts)
                kind.append(multi)
                rays += nr
            time.sleep(0.0002)
        wall = time.perf_counter() - t0
    lat_ms, kind = np.array(lat) * 1e3, np.array(kind)

    def pct(v):
        return {"p50_ms": float(np.percentile(v, 50)), "p99_ms": float(np.percentile(v, 99)), "n": int(v.size)}

    return {"qps": qps, "requests": n, "p50_ms": float(np.percentile(lat_ms, 50)),
            "p99_ms": float(np.percentile(lat_ms, 99)), "mean_ms": float(lat_ms.mean()),
            "single_kerr": pct(lat_ms[~kind]), "superposition": pct(lat_ms[kind]) if kind.any() else None,
            "rays_per_s": rays / wall, "wall_s": wall, "multi_source_fraction": 0.3,
            "workload": "Poisson arrivals, lognormal rays (median 256^2, sigma 1, max 2048^2), "
                        "70% Kerr / 30% 2-16 hole superpositions"}


def precision_report(side: int, switch_radius: float):
    """Mixed precision against FP64 on the precision workload: Kerr a = 0.998,
    inclination 85 deg, side x side camera (4096: 16.8 M rays), FP32
    right-hand side beyond switch_radius (0: 3 r_+). The FP64 kernel agrees
    with the FP64 oracle to ~1e-10 (GPU tests), so the error of the mixed run
    is measured against the FP64 run on every ray. Reports kernel times and
    their ratio, per-output error (non-captured rays; max(|ref|, 1)
    denominator; phi as a wrapped angle), and class agreement overall and by
    distance to the critical curve (rt_critical_distance: rays within 1e-6 M
    of it are exempt from the class-match requirement)."""
    cam, met, ig64, tm = workloads.precision_request(side)
    igmx = workloads.precision_request(side, rtsim.RT_PRECISION_MIXED, switch_radius)[2]
    ctx = rtsim.Context(0)
    rep = {"workload": f"kerr_a0.998_i85_robs1000_fov20_{side}sq", "rays": side * side,
           "switch_radius": switch_radius if switch_radius > 0 else "3 r_+"}
    ref = ctx.trace_camera(cam, met, ig64, tm)
    rep["fp64_kernel_ms"] = ctx.stats().kernel_ms
    mx = ctx.trace_camera(cam, met, igmx, tm)
    st = ctx.stats()
    ctx.close()
    rep["mixed_kernel_ms"] = st.kernel_ms
    rep["mixed_over_fp64_speed"] = rep["fp64_kernel_ms"] / rep["mixed_kernel_ms"]
    rep["fp32_step_fraction"] = st.fp32_steps / max(1, st.accepted_steps + st.rejected_steps)
    same = mx.klass == ref.klass
    rep["class_agreement"] = float(same.mean())
    rep["class_mismatches"] = int((~same).sum())
    dist = rtsim.critical_distance(cam, met)
    bands = {}
    for lo, hi, name in [(0.0, 1e-6, "within_1e-6"), (1e-6, 1e-3, "1e-6_to_1e-3"), (1e-3, 1e-1, "1e-3_to_0.1"),
                         (1e-1, np.inf, "beyond_0.1")]:
        sel = (dist >= lo) & (dist < hi)
        bands[name] = {"rays": int(sel.sum()), "mismatches": int((~same & sel).sum())}
    rep["class_by_critical_distance_M"] = bands
    rep["class_agreement_outside_1e-6"] = float(same[dist >= 1e-6].mean())
    ok = same & (ref.klass != rtsim.RT_CLS_CAPTURED)
    den = lambda v: np.maximum(np.abs(v), 1.0)  # noqa: E731
    errs = {"ks_position": np.max(np.abs(mx.x[1:, ok] - ref.x[1:, ok]) / den(ref.x[1:, ok]), axis=0),
            "r": np.abs(mx.bl[0, ok] - ref.bl[0, ok]) / den(ref.bl[0, ok]),
            "theta": np.abs(mx.bl[1, ok] - ref.bl[1, ok]) / den(ref.bl[1, ok]),
            "phi": np.abs(np.remainder(mx.bl[2, ok] - ref.bl[2, ok] + np.pi, 2 * np.pi) - np.pi) / den(ref.bl[2, ok]),
            "g": np.abs(mx.g[ok] - ref.g[ok]) / den(ref.g[ok])}
    rep["rel_err"] = {k: {"p50": float(np.median(v)), "p99": float(np.percentile(v, 99)), "max": float(v.max()),
                          "frac_le_1e-9": float((v <= 1e-9).mean())} for k, v in errs.items()}
    return rep


def tile_report(side: int = SIDE, spin: float = 0.9, n_sample: int = 64, seed: int = 7):
    """The 32x32 tile path on one bench frame (a = 0.9): every tile of the
    frame as one rt_trace_tiles batch (device time and rays/s, beside the
    same frame as one rt_trace_camera launch), and single-tile rt_trace_tile
    latency (device and host time) over the tiles around the hole's image
    plus a random sample of the rest."""
    cam = rtsim.default_camera(R_OBS, math.radians(INCL_DEG), 0.0, math.radians(FOV_DEG), side, side)
    met, ig, tm = rtsim.kerr_metric(1.0, spin), rtsim.default_integrator(), rtsim.default_termination()
    n_tiles = ((side + 31) // 32) ** 2
    tasks = []
    for i in range(n_tiles):
        t = rtsim.rt_tile_task()
        t.tile_index, t.camera, t.metric, t.integrator, t.termination = i, cam, met, ig, tm
        t.tile = rtsim.tile_geometry(i, side, side)
        tasks.append(t)
    ctx = rtsim.Context(0)
    ctx.trace_camera_resident(cam, met, ig, tm)
    frame_ms = []
    for _ in range(3):
        ctx.trace_camera_resident(cam, met, ig, tm)
        frame_ms.append(ctx.stats().kernel_ms)
    arr = (rtsim.rt_tile_task * n_tiles)(*tasks)
    ctx.trace_tiles(arr)  # captures the batch graph
    batch_ms, wall = [], []
    for _ in range(3):
        t0 = time.perf_counter()
        n = ctx.trace_tiles(arr)
        wall.append(time.perf_counter() - t0)
        batch_ms.append(ctx.stats().kernel_ms)
    cols = (side + 31) // 32
    c = cols // 2
    centre = [(c + dy) * cols + (c + dx) for dy in range(-2, 2) for dx in range(-2, 2)]
    rng = np.random.default_rng(seed)
    rest = [int(i) for i in rng.choice(n_tiles, size=n_sample - len(centre), replace=False) if i not in centre]
    dev, host = [], []
    for i in centre + rest:
        t = tasks[i]
        t0 = time.perf_counter()
        ctx.trace_tile(t.camera, t.metric, t.integrator, t.termination, t.tile)
        host.append((time.perf_counter() - t0) * 1e3)
        dev.append(ctx.stats().kernel_ms)
    ctx.close()

    def pct(v):
        v = np.sort(np.asarray(v))
        return {"p50": float(np.percentile(v, 50)), "p99": float(np.percentile(v, 99)), "max": float(v[-1])}

    return {"frame": f"kerr_{side}sq_a{spin}_i60_robs1000_fov20", "tiles": n_tiles,
            "batch_kernel_ms": float(np.median(batch_ms)), "batch_rays_per_s": n / float(np.median(wall)),
            "frame_kernel_ms": float(np.median(frame_ms)), "single_tile_sampled": len(dev),
            "single_tile_centre": len(centre), "single_tile_kernel_ms": pct(dev), "single_tile_host_ms": pct(host)}


def launch_flops(st, n_src: int) -> float:
    """Algorithmic FP64 FLOP of one launch from its counters."""
    per_rhs = FLOP_RHS if n_src < 2 else n_src * FLOP_SOURCE
    return st.rhs_evals * per_rhs + (st.accepted_steps + st.rejected_steps) * FLOP_ATTEMPT


def superposition_report(side: int = 512, steps: int = 5):
    """Superposition frames of 2, 4, 8 and 16 holes (workloads.multi_requests:
    masses U[0.1, 1], positions U[-25, 25]^3 M, spins U[0, 0.95] on random
    axes; observer 1000 M, inclination 60 deg, FOV 20 deg), each traced by
    k_trace_multi (one source per lane of a ray's lane group): kernel time,
    rays/s and algorithmic FP64 rate against the vector peak."""
    ctx = rtsim.Context(0)
    rows = []
    for r in workloads.multi_requests(side):
        k = r.metric.n_sources
        n = ctx.trace_camera_resident(r.camera, r.metric, r.integrator, r.termination)  # warm-up
        ms, flops = [], []
        for _ in range(steps):
            ctx.trace_camera_resident(r.camera, r.metric, r.integrator, r.termination)
            st = ctx.stats()
            ms.append(st.kernel_ms)
            flops.append(launch_flops(st, k))
        t = float(np.median(ms))
        achieved = float(np.median(flops)) / (t * 1e-3) / 1e12
        rows.append({"sources": k, "kernel_ms": t, "rays_per_s": n / (t * 1e-3),
                     "mean_attempts": (st.accepted_steps + st.rejected_steps) / n, "achieved_tflops": achieved,
                     "frac": achieved / PEAK_FP64_TFLOPS})
    ctx.close()
    return {"frames": f"{side}^2, 2/4/8/16 holes", "kernel": "k_trace_multi", "per_frame": rows}


def render_report(side: int, steps: int = 5, warmup: int = 5, band: int = 16, band_stride: int = 128,
                  threads: int = 0, cpu: bool = True):
    """RT_METHOD_RENDER_RK4 frame on the GPU (workloads.render_request:
    a = 0.9, observer at 30 M, inclination 65 deg, FOV 60 deg, disk to 24 M,
    escape at 100 M; k_trace_render, one launch, median kernel time of
    `steps` after `warmup` launches: a ~3 ms launch keeps getting faster over
    its first few repeats), beside two CPU timings of bands of `band` rows every
    `band_stride` rows of the same frame: the oracle's restatement
    of the same tracer (records only, like the kernel), and, where
    oracle/_ref/libref.so was built, the reference's dispatch_fp64 itself
    (which also shades each pixel's colour)."""
    cam, met, ig, tm = workloads.render_request(side)
    ctx = rtsim.Context(0)
    for _ in range(warmup):
        n = ctx.trace_camera_resident(cam, met, ig, tm)
    ms, st = [], None
    for _ in range(steps):
        ctx.trace_camera_resident(cam, met, ig, tm)
        st = ctx.stats()
        ms.append(st.kernel_ms)
    ctx.close()
    t = float(np.median(ms))
    out = {"frame": f"render_rk4_kerr_{side}sq_a0.9_i65_robs30_fov60_disk", "kernel": "k_trace_render",
           "rays": n, "kernel_ms": t, "rays_per_s": n / (t * 1e-3), "mean_iterations": st.accepted_steps / n}
    if not cpu:
        return out
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_lib

    cpus = host_cpus()
    threads = threads or cpus["threads"]
    native = native_oracle()
    if native:
        oracle_lib.ORACLE_PATH = native
    firsts = list(range(0, side, band_stride))
    sample = f"rows [k*{band_stride}, k*{band_stride}+{band}) of the frame ({len(firsts) * band * side} rays)"

    def timed(fn):
        t0 = time.perf_counter()
        for f in firsts:
            cam.row_offset, cam.row_count = f, min(band, side - f)
            fn()
        cam.row_offset, cam.row_count = 0, 0
        return time.perf_counter() - t0

    nr = sum(min(band, side - f) for f in firsts) * side
    dt = timed(lambda: oracle_lib.trace_render_camera(cam, met, ig, tm, threads=threads))
    out["cpu_port"] = {"value": nr / dt, "unit": "rays/s", "cores": threads, "kind": "port",
                       "sample": f"{sample}, {dt:.2f} s, oracle/{os.path.basename(oracle_lib.ORACLE_PATH)}"}
    if oracle_lib.ref_available():
        dt = timed(lambda: oracle_lib.ref_render_frame(cam, met, ig, tm))
        out["cpu_reference"] = {"value": nr / dt, "unit": "rays/s", "cores": os.cpu_count(), "kind": "reference",
                                "sample": f"{sample}, {dt:.2f} s, SoftwareComputeEngine::dispatch_fp64 "
                                          "(one thread per online CPU, records and colour) from "
                                          "oracle/_ref/libref.so built -O2"}
    return out


def copy_probe(mib: int):
    """Three plain device-to-device copies of a known size: what FETCH_SIZE
    and WRITE_SIZE read for them calibrates the counters against bytes."""
    import torch

    src = torch.ones(mib << 20, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    return {"copy_bytes_read": mib << 20, "copy_bytes_written": mib << 20, "copies": 3}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--side", type=int, default=SIDE)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--path", choices=["batch", "plain"], default="batch",
                    help="batch: the four frames in one persistent work-queue kernel, slow tiles first (hipGraph "
                         "replay); plain: one kernel per frame, one lane per pixel")
    ap.add_argument("--cpu-stride", type=int, default=8)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--qps", type=float, default=50.0, help="the latency object's QPS")
    ap.add_argument("--qps-levels", type=str, default="25,100",
                    help="further QPS levels reported in latency_levels (comma-separated; empty: none)")
    ap.add_argument("--serve-requests", type=int, default=150)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-tiles", action="store_true", help="skip the 32x32 tile-task report")
    ap.add_argument("--no-render", action="store_true", help="skip the RT_METHOD_RENDER_RK4 report")
    ap.add_argument("--workload", choices=["frames4", "uneven512", "multi"], default="frames4",
                    help="frames4: the four 2048^2 frames; uneven512: 512 seeded camera requests in one batch; "
                         "multi: superposition frames of 2, 4, 8, 16 holes (512^2), one launch each")
    ap.add_argument("--precision-report", type=int, default=0, metavar="SIDE",
                    help="print a mixed-vs-FP64 error report at this camera side (e.g. 4096) and exit")
    ap.add_argument("--switch-radius", type=float, default=0.0,
                    help="mixed precision: FP64 right-hand side inside this radius (0: 3 r_+)")
    ap.add_argument("--precision-side", type=int, default=4096,
                    help="camera side of the precision object in the default run (0: leave it out)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="run three device-to-device copies of MIB MiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this tracer runs on one GPU; --gpus {args.gpus} is not supported "
              "(rays are independent: a frame shards by tiles across GPUs with no exchange)", file=sys.stderr)
        return 2
    if args.precision_report:
        print(json.dumps(precision_report(args.precision_report, args.switch_radius)), flush=True)
        return 0
    if args.copy_probe:
        print(json.dumps(copy_probe(args.copy_probe)), flush=True)
        return 0
    prec = rtsim.RT_PRECISION_MIXED if args.precision == "mixed" else rtsim.RT_PRECISION_FP64
    reqs = frame_requests(args.side, prec)
    if args.workload == "uneven512":
        reqs = [(r.camera, r.metric, r.integrator, r.termination) for r in workloads.uneven_requests()]
    if args.workload == "multi":
        reqs = [(r.camera, r.metric, r.integrator, r.termination) for r in workloads.multi_requests(512)]
        args.path = "plain"
    ctx = rtsim.Context(0)
    tuned = None
    if args.workload == "uneven512" and args.path == "batch":
        # autotune the work-queue launch shape on every 16th request
        best = ctx.autotune([rtsim.make_request(*r) for r in reqs[::16]])
        tuned = {"block": best.block, "waves": best.waves, "refill_min": best.refill_min,
                 "queue_kind": best.queue_kind}
        print(f"bench: autotuned {tuned}", file=sys.stderr, flush=True)
    print(f"bench: {len(reqs)} frames of {args.side}^2, precision {args.precision}, "
          f"warmup {args.warmup}, steps {args.steps}", file=sys.stderr, flush=True)
    batch = [rtsim.make_request(*r) for r in reqs]

    def one_step():
        """Trace the four frames; returns [(rays, stats)] per kernel launch."""
        if args.path == "batch":
            ctx.trace_batch(batch, fetch=False)
            st = ctx.stats()
            return [(st.rays, st, 0)]
        out = []
        for r in reqs:
            n = ctx.trace_camera_resident(*r)
            out.append((n, ctx.stats(), r[1].n_sources))
        return out

    for w in range(args.warmup):
        one_step()
        print(f"bench: warmup step {w + 1} done", file=sys.stderr, flush=True)
    kernel_ms = 0.0
    flops = 0.0
    rays = 0
    launches = 0
    acc = rej = rhs = 0
    shape = None
    t0 = time.perf_counter()
    for s in range(args.steps):
        for n, st, n_src in one_step():
            kernel_ms += st.kernel_ms
            flops += launch_flops(st, n_src)
            rays += n
            launches += 1
            acc += st.accepted_steps
            rej += st.rejected_steps
            rhs += st.rhs_evals
            shape = (st.block_size, st.grid_size)
        print(f"bench: step {s + 1} done", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    value = rays / wall
    # the tracing kernel of the timed launches: per-frame kernels for the plain
    # path and for batches of queue_kind 2 (one per request, concurrent
    # streams), else the persistent queue kernel of the batch
    _, qs = ctx.launch_shapes()
    kernel = ("k_trace_camera_kerr" if args.path == "plain" or qs.queue_kind == 2
              else ("k_trace_queue" if qs.queue_kind == 1 else "k_trace_tiles"))
    if args.workload == "multi":
        kernel = "k_trace_multi"
    avg_kernel_s = kernel_ms / launches / 1e3
    achieved = flops / launches / avg_kernel_s / 1e12
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == rtsim.RT_PRECISION_FP64 else "f64+f32",
        "data": "synthetic (camera rays generated on device)",
        "config": {"workload": {"frames4": f"kerr_{args.side}sq_4spins_i60_robs1000_fov20",
                                "uneven512": "uneven512_seed_c0ffee1234",
                                "multi": "superposition_512sq_2_4_8_16_holes"}[args.workload], "side": args.side,
                   "spins": list(SPINS), "precision": args.precision, "rays_per_step": rays // args.steps,
                   "mean_accepted_steps": acc / rays, "mean_rejected_steps": rej / rays,
                   "rhs_evals_per_ray": rhs / rays, "path": args.path, "kernel_ms_per_launch": kernel_ms / launches,
                   "launches_per_step": launches // args.steps, "block": shape[0], "grid": shape[1],
                   "autotuned": tuned, "kernel": kernel},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS,
                     "traffic": (HBM_BYTES_PER_RAY[kernel] * rays / launches
                                 if HBM_BYTES_PER_RAY[kernel] is not None else None)},
    }
    ctx.close()
    if not args.no_serve:
        print(f"bench: serving {args.serve_requests} requests at {args.qps} QPS", file=sys.stderr, flush=True)
        out["latency"] = serve_latency(args.qps, args.serve_requests)
        levels = [float(q) for q in args.qps_levels.split(",") if q.strip()]
        if levels:
            out["latency_levels"] = [out["latency"]] + [serve_latency(q, args.serve_requests) for q in levels]
    if not args.no_tiles:
        print("bench: 32x32 tile tasks", file=sys.stderr, flush=True)
        out["tiles"] = tile_report(args.side)
    if args.workload == "frames4":
        print("bench: superposition frames", file=sys.stderr, flush=True)
        out["superposition"] = superposition_report()
    if not args.no_render and args.workload == "frames4":
        print("bench: renderer's RK4 tracer", file=sys.stderr, flush=True)
        out["render"] = render_report(args.side, cpu=not args.no_cpu, threads=args.cpu_threads)
    if args.precision_side and args.workload == "frames4":
        print(f"bench: precision, mixed vs FP64 on {args.precision_side}^2", file=sys.stderr, flush=True)
        out["precision"] = precision_report(args.precision_side, args.switch_radius)
    if not args.no_cpu:
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(args.side, args.cpu_stride, args.cpu_threads)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
