"""Writes tests/golden/ref_vectors.json: inputs and outputs of the reference's
own functions (KS field, Hamiltonian RHS, initial momentum, coordinate maps,
ISCO, disk redshift, end points of its Christoffel-based RK45 integrator, and
the renderer's superposition source table, right-hand side and initial
momentum),
computed through oracle/_ref/libref.so, which oracle/Makefile compiles from
the reference headers. Run after `make -C oracle`; the fixture lets the oracle
be checked where the reference is not available."""
import ctypes as C
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402
import rtsim  # noqa: E402

PD = C.POINTER(C.c_double)


def ptr(a):
    return a.ctypes.data_as(PD)


def main():
    ref = O.ref()
    rng = np.random.default_rng(20261203)
    out = {"source": "reference functions via oracle/_ref/libref.so", "field": [], "rhs": [], "momentum": [],
           "ks_from_bl": [], "bl_from_ks": [], "isco": [], "redshift": [], "rk45": []}
    for _ in range(64):
        a = float(rng.uniform(-0.999, 0.999))
        R = 10 ** rng.uniform(0.2, 3.3)
        x = rng.normal(size=3)
        x *= R / np.linalg.norm(x)
        p = rng.normal(size=3)
        f = np.zeros(5)
        ref.ref_ks_field(ptr(x), 1.0, a, ptr(f))
        out["field"].append({"x": x.tolist(), "M": 1.0, "a": a, "out": f.tolist()})
        dx, dp = np.zeros(3), np.zeros(3)
        ref.ref_ks_rhs(ptr(x), ptr(p), 1.0, a, ptr(dx), ptr(dp))
        out["rhs"].append({"x": x.tolist(), "p": p.tolist(), "M": 1.0, "a": a, "dx": dx.tolist(), "dp": dp.tolist()})
        o = rng.normal(size=3)
        o /= np.linalg.norm(o)
        q = np.zeros(3)
        ref.ref_initial_momentum(ptr(x), 1.0, a, ptr(o), ptr(q))
        out["momentum"].append({"x": x.tolist(), "outward": o.tolist(), "M": 1.0, "a": a, "p": q.tolist()})
        r, th, ph = float(rng.uniform(1.5, 800)), float(rng.uniform(0.01, 3.13)), float(rng.uniform(-3.1, 3.1))
        k = np.zeros(3)
        ref.ref_ks_from_bl(r, th, ph, a, ptr(k))
        out["ks_from_bl"].append({"r": r, "theta": th, "phi": ph, "a": a, "x": k.tolist()})
        b = np.zeros(3)
        ref.ref_bl_from_ks(ptr(k), a, 0.0, ptr(b))
        out["bl_from_ks"].append({"x": k.tolist(), "a": a, "bl": b.tolist()})
        m = float(rng.uniform(0.1, 3.0))
        am = float(rng.uniform(-0.999, 0.999)) * m
        out["isco"].append({"M": m, "a": am, "r": ref.ref_isco(m, am)})
        rr, xi = float(rng.uniform(1.0, 30.0)) * m, float(rng.uniform(-8, 8)) * m
        out["redshift"].append({"M": m, "a": am, "r": rr, "xi": xi, "g": ref.ref_redshift(m, am, rr, xi)})
    for a in (0.0, 0.5, 0.9, 0.998):
        cam = rtsim.default_camera(60.0, math.radians(63.0), 0.3, math.radians(40.0), 9, 9)
        for px, py in [(4, 4), (1, 6), (7, 2)]:
            ok, y, _ = O.camera_state(cam, rtsim.kerr_metric(1.0, a), rtsim.default_termination(), px, py)
            yf = y.copy()
            yf[4:] = -yf[4:]
            g, gi = np.zeros(16), np.zeros(16)
            ref.ref_ks_metric(ptr(yf[:4].copy()), 1.0, a, ptr(g), ptr(gi))
            u = gi.reshape(4, 4) @ yf[4:]
            xo, uo = np.zeros(4), np.zeros(4)
            acc, rej = C.c_uint64(), C.c_uint64()
            ref.ref_rk45_integrate(ptr(yf[:4].copy()), ptr(u), 1.0, a, 40.0, 1e-12, 1e-14, 0.5, ptr(xo), ptr(uo),
                                   C.byref(acc), C.byref(rej))
            out["rk45"].append({"a": a, "y0": yf.tolist(), "L": 40.0, "x_end": xo.tolist(), "u_end": uo.tolist()})
    # superpositions of 2..16 sources (separate stream: the vectors above stay as they were)
    rng = np.random.default_rng(20261204)
    out["ms_table"], out["ms_rhs"], out["ms_momentum"] = [], [], []
    for mi in range(16):
        k = int(rng.integers(2, 17))
        src = [{"mass": float(rng.uniform(0.5, 2.0)), "spin": float(rng.uniform(-0.99, 0.99)),
                "disk_enabled": bool(rng.random() < 0.7)}]
        for _ in range(k - 1):
            src.append({"mass": float(rng.uniform(0.1, 1.0)), "spin": float(rng.uniform(-0.95, 0.95)),
                        "position": [float(v) for v in rng.uniform(-25, 25, size=3)],
                        "axis": [float(v) for v in rng.normal(size=3)],
                        "disk_enabled": bool(rng.random() < 0.7)})
        met = rtsim.multi_metric(src)
        t = np.zeros(17 * 16)
        n = ref.ref_ms_table(C.byref(met), ptr(t))
        out["ms_table"].append({"sources": src, "table": t[:17 * n].tolist()})
        for _ in range(4):
            c = np.array(src[int(rng.integers(0, k))].get("position", [0.0, 0.0, 0.0]))
            x = c + rng.normal(size=3) * float(rng.choice([3.0, 10.0, 100.0]))
            p = rng.normal(size=3)
            dx, dp = np.zeros(3), np.zeros(3)
            ref.ref_ms_rhs(C.byref(met), ptr(x), ptr(p), ptr(dx), ptr(dp))
            out["ms_rhs"].append({"metric": mi, "x": x.tolist(), "p": p.tolist(), "dx": dx.tolist(),
                                  "dp": dp.tolist()})
        x = rng.normal(size=3)
        x *= float(rng.uniform(40.0, 1000.0)) / np.linalg.norm(x)
        o = rng.normal(size=3)
        o /= np.linalg.norm(o)
        q = np.zeros(3)
        ref.ref_ms_initial_momentum(C.byref(met), ptr(x), ptr(o), ptr(q))
        out["ms_momentum"].append({"metric": mi, "x": x.tolist(), "outward": o.tolist(), "p": q.tolist()})
    path = os.path.join(HERE, "ref_vectors.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=0)
    print("wrote", path)


if __name__ == "__main__":
    main()
