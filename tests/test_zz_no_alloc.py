"""rt_trace_tiles_noalloc: the allocation-free tile batch against
rt_trace_tiles on the same tasks, through the C ABI.

Both run the same work-queue kernel instance (64-thread blocks, two waves
per SIMD), so every result is compared bit for bit. Two calls in a row, the
second with fewer tasks and rays, check that the scratch array's counters
and queue head start from zero on every call and that nothing of the first
call leaks into the second. The result arrays are device memory the test allocates with
hipMalloc, as a caller that cannot allocate would hold them."""
import ctypes as C
import math

import numpy as np
import pytest

import rtsim
from test_tile_task import task_bytes, tasks

pytestmark = pytest.mark.gpu


def fixture_tasks():
    out = []
    for name in ["kerr_centre_tile", "ragged_corner_tile", "worker_test_task", "scaled_disk_tile"]:
        out.append(rtsim.decode_tile_task(task_bytes(next(e for e in tasks(True) if e["name"] == name))))
    return out


def frame_tasks(side, spin, incl, indices):
    cam = rtsim.default_camera(1000.0, math.radians(incl), 0.0, math.radians(20.0), side, side)
    met, ig, tm = rtsim.kerr_metric(1.0, spin), rtsim.default_integrator(), rtsim.default_termination()
    out = []
    for i in indices:
        t = rtsim.rt_tile_task()
        t.tile_index, t.camera, t.metric, t.integrator, t.termination = i, cam, met, ig, tm
        t.tile = rtsim.tile_geometry(i, side, side)
        out.append(t)
    return out


def hip():
    """The HIP runtime librtsim.so runs on (already loaded with it)."""
    rtsim.load()
    lib = C.CDLL("libamdhip64.so.7")
    lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    lib.hipFree.argtypes = [C.c_void_p]
    lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    lib.hipDeviceSynchronize.argtypes = []
    return lib


H2D, D2H = 1, 2


class DeviceResults:
    """Result arrays in device memory (the caller's, from hipMalloc), viewed
    through an rt_results of device pointers with leading dimension cap. They
    start filled with a pattern (NaN / all ones) so that unwritten entries
    show."""

    LAYOUT = [("cls", np.uint32, 1), ("x", np.float64, 4), ("p", np.float64, 4), ("bl", np.float64, 3),
              ("lam", np.float64, 1), ("g", np.float64, 1), ("steps", np.uint32, 1), ("rejected", np.uint32, 1)]

    def __init__(self, cap):
        self.hip = hip()
        self.cap = cap
        self.ptr = {}
        self.struct = rtsim.rt_results()
        for f, dt, k in self.LAYOUT:
            p = C.c_void_p()
            nbytes = k * cap * np.dtype(dt).itemsize
            assert self.hip.hipMalloc(C.byref(p), nbytes) == 0
            fill = np.full(k * cap, np.nan if dt == np.float64 else 0xFFFFFFFF, dtype=dt)
            assert self.hip.hipMemcpy(p, fill.ctypes.data_as(C.c_void_p), nbytes, H2D) == 0
            self.ptr[f] = p
            ct = C.c_double if dt == np.float64 else C.c_uint32
            setattr(self.struct, f, C.cast(p, C.POINTER(ct)))

    def host(self, n):
        r = rtsim.Results(n)
        for f, dt, k in self.LAYOUT:
            buf = np.empty(k * self.cap, dtype=dt)
            assert self.hip.hipMemcpy(buf.ctypes.data_as(C.c_void_p), self.ptr[f], buf.nbytes, D2H) == 0
            dst = getattr(r, f)
            dst[...] = buf.reshape(k, self.cap)[:, :n] if k > 1 else buf[:n]
        return r

    def free(self):
        for p in self.ptr.values():
            self.hip.hipFree(p)
        self.ptr = {}


def noalloc(ctx, task_list, out, cap):
    arr = (rtsim.rt_tile_task * len(task_list))(*task_list)
    n = C.c_size_t()
    st = ctx.lib.rt_trace_tiles_noalloc(ctx.h, arr, len(task_list), C.byref(out.struct), cap, C.byref(n))
    if st != 0:
        raise rtsim.RtError(st, ctx.lib.rt_last_error().decode())
    return n.value


def same_bits(a, b, label):
    for f in ["cls", "steps", "rejected", "lam", "g", "x", "p", "bl"]:
        assert np.array_equal(getattr(a, f), getattr(b, f)), f"{label}: {f} differs"


def test_noalloc_matches_trace_tiles_two_calls():
    first = fixture_tasks() + frame_tasks(256, 0.9, 80.0, [0, 27, 28, 35, 36, 63])
    second = frame_tasks(160, 0.998, 60.0, [12, 7])
    n1 = sum(t.tile.width * t.tile.height for t in first)
    n2 = sum(t.tile.width * t.tile.height for t in second)
    assert n2 < n1
    cap = n1 + 100
    with rtsim.Context(0) as ctx:
        out = DeviceResults(cap)
        for label, tl, n_expect in [("first call", first, n1), ("second call, smaller", second, n2)]:
            print(f"[noalloc] {label}: {len(tl)} tasks, {n_expect} rays", flush=True)
            n = noalloc(ctx, tl, out, cap)
            assert n == n_expect
            st = ctx.stats()
            assert st.rays == n and st.accepted_steps > 0
            got = out.host(n)
            counts = (st.accepted_steps, st.rejected_steps, st.rhs_evals)
            n_ref = ctx.trace_tiles(tl)
            assert n_ref == n
            ref = ctx.fetch(n)
            st_ref = ctx.stats()
            same_bits(got, ref, label)
            assert counts == (st_ref.accepted_steps, st_ref.rejected_steps, st_ref.rhs_evals), label
            assert int(got.steps.sum()) == st.accepted_steps
            assert {1, 2} <= set(got.klass.tolist())
        out.free()


def test_noalloc_refusals():
    with rtsim.Context(0) as ctx:
        out = DeviceResults(1024)
        many = frame_tasks(64, 0.5, 60.0, [0] * (rtsim.RT_NOALLOC_MAX_TASKS + 1))
        with pytest.raises(rtsim.RtError) as e:
            noalloc(ctx, many, out, 1024)
        assert "exceed the scratch capacity" in str(e.value)
        with pytest.raises(rtsim.RtError) as e:
            noalloc(ctx, frame_tasks(64, 0.5, 60.0, [0]), out, 1023)
        assert "result arrays" in str(e.value)
        # the context still traces after the refusals
        assert noalloc(ctx, frame_tasks(64, 0.5, 60.0, [1]), out, 1024) == 32 * 32
        out.free()
