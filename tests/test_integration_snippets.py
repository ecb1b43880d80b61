"""The C++ adapters shown in INTEGRATION.md compile against the reference's
own headers (where the reference tree is present) and include/rt_tracer.h:
the executor next to VulkanComputeExecutor and the tile-worker backend. The
check is -fsyntax-only; the reference-side colour step the executor ends with
(`shade`) is declared, not defined."""
import os
import re
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
REF_INCLUDE = "/root/reference/include"

pytestmark = pytest.mark.skipif(not os.path.isdir(os.path.join(REF_INCLUDE, "relativistic")) or not shutil.which("g++"),
                                reason="reference headers or g++ absent")


def snippets():
    text = open(os.path.join(REPO, "INTEGRATION.md")).read()
    return re.findall(r"```cpp\n(.*?)```", text, re.S)


def test_integration_snippets_compile(tmp_path):
    blocks = snippets()
    assert len(blocks) == 2
    executor, tile = blocks
    src = "\n".join([
        "#include <algorithm>", "#include <atomic>", "#include <span>", "#include <string>", "#include <vector>",
        '#include "relativistic/render/gpu_types.hpp"',
        '#include "relativistic/optics/disk_thermal_profile.hpp"',
        '#include "relativistic/core/engine_log.hpp"',
        "namespace Relativistic::Render {",
        "void shade(std::vector<GpuPixelOutput>& pixels);",
        executor.replace("#include <rt_tracer.h>", ""),
        "}  // namespace Relativistic::Render",
        "namespace Relativistic::Distributed {",
        tile,
        "}  // namespace Relativistic::Distributed",
    ])
    path = tmp_path / "integration_check.cpp"
    path.write_text('#include <rt_tracer.h>\n' + src)
    r = subprocess.run(["g++", "-std=c++2b", "-fsyntax-only", "-Wall", "-I", REF_INCLUDE,
                        "-I", os.path.join(REPO, "include"), str(path)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
