"""bench.py's algorithmic FLOP constants are the ones the oracle's
operation-counting instantiation reports (one RHS evaluation; one DP45
attempt's stage sums, error norm and controller over the 7 evolving
components, p_t having a zero derivative); and bench.py refuses more than
one GPU."""
import ctypes as C
import importlib.util
import os

import oracle_lib as O

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_flop_constants():
    spec = importlib.util.spec_from_file_location("bench", os.path.join(REPO, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    out = (C.c_double * 4)()
    O.oracle().rto_flop_counts(out)
    rhs, attempt8 = out[0], out[1]
    controller = 5  # pow, reciprocal, two products, clamp arithmetic
    per_component = (attempt8 - controller) / 8
    assert per_component == 61  # 46 stage-sum ops + 15 error-norm ops
    assert bench.FLOP_RHS == rhs
    assert bench.FLOP_ATTEMPT == 7 * per_component + controller
    assert bench.FLOP_SOURCE == out[3]


def test_more_than_one_gpu_is_refused_with_a_message():
    """The tracer runs on one GPU: --gpus 2 exits with status 2 and says so,
    before touching a device."""
    import subprocess
    import sys

    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and "--gpus 2" in r.stderr
    assert r.stdout == ""
