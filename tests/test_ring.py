"""Host check of the serving runtime's byte ring (pars_amd/csrc/ring.hpp):
allocations never overlap a region whose batch is not yet retired, when every
batch - including one that failed to enqueue - is retired in launch order
(tests/cpp/ring_check.cpp).  CPU only."""
from __future__ import annotations

import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_ring_retirement_order():
    from pars_amd import build

    exe = build.build_ring_check()
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    assert r.stdout.startswith("ok"), r.stdout
