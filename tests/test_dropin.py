"""The drop-in claim, compiled (CPU only, where the reference tree exists): the
pars::gpu binding INTEGRATION.md gives a maintainer
(tests/cpp/dropin/pars/backend/hip/pipeline.hpp, the same text) is compiled
against the reference's own headers with plans from the reference compiler
(compile::make_scan_plan<G>() for the six bench grammars), and called the way
the reference's bench drivers call gpu::scanner<S> and the free functions
(tests/cpp/dropin_check.cpp, following bench/bench_boundary.cu)."""
from __future__ import annotations

import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
BINDING = os.path.join(ROOT, "tests", "cpp", "dropin", "pars", "backend", "hip", "pipeline.hpp")


def test_integration_guide_shows_the_compiled_binding():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    start = text.index("```cpp\n// include/pars/backend/hip/pipeline.hpp")
    block = text[start + len("```cpp\n"):text.index("```", start + 6)]
    assert block == open(BINDING).read()


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "include", "pars")), reason="reference tree absent")
def test_binding_compiles_against_the_reference(tmp_path):
    cmd = ["g++", "-std=c++20", "-c", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           f"-I{REF}/include", f"-I{REF}/bench", "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(ROOT, "tests", "cpp", "dropin"), os.path.join(ROOT, "tests", "cpp", "dropin_check.cpp"),
           "-o", str(tmp_path / "dropin.o")]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0, r.stdout
    assert os.path.getsize(tmp_path / "dropin.o") > 0
