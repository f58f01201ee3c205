"""The stage-2 restatement (oracle/json_dom.py and its C twin json_dom.c)
against the reference's own JSON DOM hook: tests/golden/dom_golden.json holds
what the reference's json_parser (bench/bench_parse.cu:34-354), compiled from
its source and fed by the reference's CPU scanner and model::replay
(tests/golden/make_dom_golden.py, oracle/_ref/ref_dom), gives on each case -
its verdict, node count, hash and the tape's kinds, begins and ends.  The
cases include malformed documents (flipped, deleted and inserted bytes, token
soups, the hand-worked ones), where Python's json module cannot serve as the
check.  CPU only."""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

from oracle import json_dom as J
from oracle import oracle as O
from pars_amd import Plan
from tests.golden.make_dom_golden import inputs

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dom_golden.json")
CASES = json.load(open(GOLDEN))["cases"]
PLAN = Plan.load("json").blob


def got_of(r: dict) -> dict:
    return {"ok": bool(r["ok"]), "nodes": int(r["nodes"]), "hash": int(r["hash"]), "tape_len": len(r["kind"]),
            "kind_fnv": O.fnv(np.asarray(r["kind"], np.uint8)), "begin_fnv": O.fnv(np.asarray(r["begin"], np.uint32)),
            "end_fnv": O.fnv(np.asarray(r["end"], np.uint32))}


def test_cases_cover_both_verdicts():
    oks = [c["ref"]["ok"] for c in CASES]
    assert sum(oks) >= 20 and len(oks) - sum(oks) >= 200


@pytest.mark.parametrize("case", range(len(CASES)), ids=[c["name"] for c in CASES])
def test_restatement_equals_the_reference_hook(case):
    c = CASES[case]
    d = inputs(c["input"])
    pos = O.scan(PLAN, np.frombuffer(d, np.uint8))["positions"]
    assert got_of(J.json_tape(d, pos)) == c["ref"], c["name"] + " (json_dom.py)"
    assert got_of(O.json_tape(np.frombuffer(d, np.uint8), pos)) == c["ref"], c["name"] + " (json_dom.c)"
