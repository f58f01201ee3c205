"""Host check of the event path (EventEval in pars_amd/csrc/evaluators.hpp,
its tables built by pars_amd/csrc/plan_dev.cpp): on every shipped plan and
every probe plan of the grammar compiler that the path takes, random lanes
stepped at their event bytes only give each descriptor's map for every entry
state, and the masked bits for random entry states, exactly as stepping every
byte through the descriptors' DFAs (model/scan.hpp:33-72 semantics) does
(tests/cpp/event_check.cpp).  CPU only."""
from __future__ import annotations

import glob
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_event_path_equals_byte_steps(tmp_path):
    from pars_amd import build

    exe = build.build_event_check()
    plans = sorted(glob.glob(os.path.join(ROOT, "pars_amd", "plans", "*.plan")))
    plans += sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "plans", "*.plan")))
    r = subprocess.run([exe, *plans], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0, r.stdout
    # clike, clike6, toml and xml at least (two or more descriptors)
    assert r.stdout.startswith("ok") and int(r.stdout.split()[1]) >= 4, r.stdout
