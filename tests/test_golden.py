"""The CPU oracle against committed vectors of the reference's own functions
(tests/golden/ref_vectors.json, written by tests/golden/make_golden.py from the
reference headers). Runs without the reference tree."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import oracle_lib as O
import rtsim

PD = C.POINTER(C.c_double)
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.json")))


def ptr(a):
    return np.ascontiguousarray(a, np.float64).ctypes.data_as(PD)


def test_field():
    for v in GOLD["field"]:
        x = np.array(v["x"])
        o = np.zeros(5)
        O.oracle().rto_ks_field(ptr(x), v["M"], v["a"], o.ctypes.data_as(PD))
        assert np.array_equal(o, np.array(v["out"]))


def test_rhs():
    for v in GOLD["rhs"]:
        d = O.rhs(np.array([0.0, *v["x"], -1.0, *v["p"]]), rtsim.kerr_metric(v["M"], v["a"]))
        assert np.array_equal(d[1:4], np.array(v["dx"])) and np.array_equal(d[5:8], np.array(v["dp"]))


def test_initial_momentum():
    for v in GOLD["momentum"]:
        x, o = np.array(v["x"]), np.array(v["outward"])
        p = np.zeros(3)
        O.oracle().rto_initial_momentum(ptr(x), v["M"], v["a"], ptr(o), p.ctypes.data_as(PD))
        assert np.array_equal(p, np.array(v["p"]))


def test_coordinates():
    for v in GOLD["ks_from_bl"]:
        k = np.zeros(3)
        O.oracle().rto_ks_from_bl(v["r"], v["theta"], v["phi"], v["a"], k.ctypes.data_as(PD))
        assert np.array_equal(k, np.array(v["x"]))
    for v in GOLD["bl_from_ks"]:
        b = np.zeros(3)
        O.oracle().rto_bl_from_ks(ptr(np.array(v["x"])), v["a"], b.ctypes.data_as(PD))
        assert np.array_equal(b, np.array(v["bl"]))


def test_isco_redshift():
    for v in GOLD["isco"]:
        assert O.oracle().rto_isco(v["M"], v["a"]) == v["r"]
    for v in GOLD["redshift"]:
        assert O.oracle().rto_redshift(v["M"], v["a"], v["r"], v["xi"], 1) == v["g"]


def test_rk45_end_points():
    lib = O.oracle()
    lib.rto_integrate_to.argtypes = [PD, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.c_double, PD]
    lib.rto_integrate_to.restype = C.c_int
    for v in GOLD["rk45"]:
        out = np.zeros(8)
        assert lib.rto_integrate_to(ptr(np.array(v["y0"])), 1.0, v["a"], v["L"], 1e-12, 1e-14, 0.5,
                                    out.ctypes.data_as(PD)) > 0
        xe = np.array(v["x_end"])
        assert np.abs(out[:4] - xe).max() < 1e-10 * max(1.0, np.abs(xe).max())


def _metric(i):
    return rtsim.multi_metric(GOLD["ms_table"][i]["sources"])


def test_ms_table():
    tm = rtsim.default_termination()
    for v in GOLD["ms_table"]:
        met = rtsim.multi_metric(v["sources"])
        t = np.zeros(17 * 16)
        n = O.oracle().rto_ms_table(C.byref(met), C.byref(tm), t.ctypes.data_as(PD))
        assert n == len(v["sources"]) and np.array_equal(t[:17 * n], np.array(v["table"]))


def test_ms_rhs():
    # the oracle's superposition RHS sums the same per-source terms in another
    # algebraic order than the renderer's ks_source_terms: equal to rounding
    for v in GOLD["ms_rhs"]:
        d = O.rhs(np.array([0.0, *v["x"], -1.0, *v["p"]]), _metric(v["metric"]))
        ref = np.array(v["dx"] + v["dp"])
        got = np.concatenate([d[1:4], d[5:8]])
        assert np.max(np.abs(got - ref)) <= 1e-14 * np.max(np.abs(ref))


def test_ms_initial_momentum():
    for v in GOLD["ms_momentum"]:
        p = np.zeros(3)
        O.oracle().rto_ms_initial_momentum(C.byref(_metric(v["metric"])), ptr(np.array(v["x"])),
                                           ptr(np.array(v["outward"])), p.ctypes.data_as(PD))
        assert np.array_equal(p, np.array(v["p"]))
