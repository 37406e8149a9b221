"""The reference-side binding shown in INTEGRATION.md section 1 compiles.

The `GpuTaskDispatcher` class a maintainer would add to the reference's
dispatcher factory is extracted from INTEGRATION.md and compiled (syntax and
types only, no object code, nothing linked or run) against the reference's own
headers where they lie, together with this repository's include/nexusmd.h.
This checks that the snippet binds to the reference's real interfaces
(TaskDispatcher, the XML parsers, PeriodicBoundaryCondition, Reporter) and to
the C-ABI as declared.  The reference tree exists only in the build container;
without it (or without g++) the test is skipped.
"""
import os
import re
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/NexusMD/core"


def _snippets():
    text = open(os.path.join(REPO, "INTEGRATION.md")).read()
    blocks = re.findall(r"```cpp\n(.*?)```", text, flags=re.S)
    cls = next(b for b in blocks if "class GpuTaskDispatcher" in b)
    factory = next(b for b in blocks if "make_shared<GpuTaskDispatcher>" in b)
    return cls, factory


@pytest.mark.skipif(not os.path.isdir(REF) or shutil.which("g++") is None,
                    reason="the reference's headers (build container only) or g++ are absent")
def test_gpu_task_dispatcher_compiles_against_the_reference_headers(tmp_path):
    cls, factory = _snippets()
    (tmp_path / "GpuTaskDispatcher.h").write_text(cls)
    # the factory lines inside a function of the shape of TaskDispatcher::CreateDispatcher
    (tmp_path / "factory.cpp").write_text(
        '#include "GpuTaskDispatcher.h"\n#include <memory>\n#include <string>\n'
        "std::shared_ptr<TaskDispatcher> make_gpu(const std::string& processorChoice) {\n"
        + factory + "    return nullptr;\n}\n")
    incs = [os.path.join(REF, d) for d in ("dispatcher/headers", "common/headers", "baseline/headers")]
    pugi = os.path.join(REF, "libs", "pugixml-1.14", "src")
    cmd = ["g++", "-std=c++17", "-fsyntax-only", "-I", str(tmp_path), "-I", os.path.join(REPO, "include"), "-I", pugi]
    for d in incs:
        cmd += ["-I", d]
    for d in incs:
        forces = os.path.join(d, "Forces")
        if os.path.isdir(forces):
            cmd += ["-I", forces]
    cmd.append(str(tmp_path / "factory.cpp"))
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
