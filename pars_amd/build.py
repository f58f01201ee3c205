"""Build the native parts in-tree.

  pars_amd/libpars_hip.so   HIP kernels (gfx950) + host orchestration + C ABI
  pars_amd/libpars_gen.so   seeded workload generators (plain C++)
  oracle/liboracle.so       CPU restatement of the reference (test infrastructure)
  oracle/_ref/ref_driver    the reference's own CPU backend, only where its sources exist

`python -m pars_amd.build` or `__graft_entry__.build()`.  Always compiles.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_SOURCES = [
    "scan_kernels.hip", "scan_wide_kernels.hip", "batch_kernels.hip", "event_kernels.hip", "event_batch_kernels.hip",
    "pair_kernels.hip", "pair_stack.hip",
    "dom_kernels.hip", "kv_kernels.hip", "api.hip", "server.hip",
]
CXX_SOURCES = ["plan_dev.cpp"]
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 instructions are kept out of every code object
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed: " + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _compile_hip(src: str) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def _compile_cxx(src: str) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    _run([HIPCC, "-O3", "-std=c++17", "-fPIC", "-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def build_hip(jobs: int = 4) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_compile_hip, s) for s in HIP_SOURCES]
        futs += [ex.submit(_compile_cxx, s) for s in CXX_SOURCES]
        objs = [f.result() for f in futs]
    out = os.path.join(HERE, "libpars_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, out)
    return out


def build_gen() -> str:
    out = os.path.join(HERE, "libpars_gen.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", out, os.path.join(CSRC, "gen.cpp")])
    return out


def build_oracle() -> None:
    make = shutil.which("make") or "make"
    _run([make, "-s", "-C", os.path.join(ROOT, "oracle"), "-B", "liboracle.so"])
    if os.path.isdir("/root/reference/include/pars"):
        _run([make, "-s", "-C", os.path.join(ROOT, "oracle"), "ref"])


def build_facade_check() -> str:
    """tests/cpp/facade_check: the C++ facade (include/pars_hip.hpp) driven like
    the reference's gpu::scanner<S>; linked against the in-tree library."""
    src = os.path.join(ROOT, "tests", "cpp", "facade_check.cpp")
    out = os.path.join(ROOT, "tests", "cpp", "facade_check")
    _run(["g++", "-std=c++20", "-O2", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
          "-I" + os.path.join(ROOT, "include"), src, "-L" + HERE, "-lpars_hip", "-L/opt/rocm/lib", "-lamdhip64",
          "-Wl,-rpath,$ORIGIN/../../pars_amd", "-o", out])
    return out


def build_replay_check() -> str:
    """tests/cpp/replay_check: the facade's consumer side (replay,
    structural_cursor, trim), host only."""
    src = os.path.join(ROOT, "tests", "cpp", "replay_check.cpp")
    out = os.path.join(ROOT, "tests", "cpp", "replay_check")
    _run(["g++", "-std=c++20", "-O2", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
          "-I" + os.path.join(ROOT, "include"), src, "-o", out])
    return out


def build_gap_check() -> str:
    """tests/cpp/gap_check: the JSON tape's gap parser (device code) built for
    the host, checked against the oracle by tests/test_gap_parser.py."""
    src = os.path.join(ROOT, "tests", "cpp", "gap_check.cpp")
    out = os.path.join(ROOT, "tests", "cpp", "gap_check")
    _run([HIPCC, *HIP_FLAGS, src, "-o", out])
    return out


def build_ring_check() -> str:
    """tests/cpp/ring_check: the serving runtime's byte ring (csrc/ring.hpp),
    host-only, under the server's retirement order."""
    src = os.path.join(ROOT, "tests", "cpp", "ring_check.cpp")
    out = os.path.join(ROOT, "tests", "cpp", "ring_check")
    _run(["g++", "-std=c++17", "-O2", "-Wall", src, "-o", out])
    return out


def build_event_check() -> str:
    """tests/cpp/event_check: the event path's tables (csrc/plan_dev.cpp) and its
    stepping, against byte-by-byte DFA steps, host-only."""
    src = os.path.join(ROOT, "tests", "cpp", "event_check.cpp")
    out = os.path.join(ROOT, "tests", "cpp", "event_check")
    _run(["g++", "-std=c++17", "-O2", "-Wall", src, os.path.join(CSRC, "plan_dev.cpp"), "-o", out])
    return out


def build_all() -> None:
    build_gen()
    build_oracle()
    build_hip()
    build_facade_check()
    build_replay_check()
    build_gap_check()
    build_ring_check()
    build_event_check()


if __name__ == "__main__":
    build_all()
    print("built:", os.path.join(HERE, "libpars_hip.so"), file=sys.stderr)
