"""The serving loop's core (relativity-simulator_amd/csrc/rt_serve_core.cpp) on
the CPU: admission by footprint, refusal, overtaking and its age limit, cancel
and deadlines in every request state with an injected clock, lanes, the take
protocol, shutdown, and a multi-threaded run with a fake device. The checks
(tests/serve_core/serve_core_check.cpp) are built with g++ under
ThreadSanitizer and under AddressSanitizer + UndefinedBehaviorSanitizer; any
sanitizer report fails the test. No GPU: the core makes no device calls."""
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(REPO, "relativity-simulator_amd", "csrc")
CHECK = os.path.join(REPO, "tests", "serve_core", "serve_core_check.cpp")


@pytest.mark.parametrize("sanitize", ["thread", "address,undefined"])
def test_serve_core_under_sanitizer(tmp_path, sanitize):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "serve_core_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", f"-fsanitize={sanitize}",
                    "-fno-sanitize-recover=all", "-I", CSRC, "-o", exe, CHECK, os.path.join(CSRC, "rt_serve_core.cpp"),
                    "-lpthread"], check=True)
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66",
               ASAN_OPTIONS="halt_on_error=1 detect_leaks=1", UBSAN_OPTIONS="halt_on_error=1 print_stacktrace=1")
    run = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    out = run.stdout + run.stderr
    print(out)
    assert run.returncode == 0, out
    assert "all serve core checks passed" in run.stdout
    assert "WARNING: ThreadSanitizer" not in out and "ERROR: AddressSanitizer" not in out
    assert "runtime error" not in out
