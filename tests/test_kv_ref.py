"""The INI and TOML stage-2 restatement (oracle/kv_hooks.py) against the
reference's own hooks: tests/golden/kv_golden.json holds what the reference's
ini_parse and toml_parse (bench/bench_parse_host.cu:30-248), compiled from its
source and fed by the reference's CPU scanner and model::replay
(tests/golden/make_kv_golden.py, oracle/_ref/ref_kv), build on each case - the
hook's hash and, per list it fills, the count and an FNV-1a 64 over the items.
Cases: hand-made documents for each branch of both hooks, the reference
benchmark's own records, and random documents over each grammar's bytes.
CPU only."""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

from oracle import kv_hooks as K
from oracle import oracle as O
from pars_amd import Plan
from tests.golden.make_kv_golden import inputs

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "kv_golden.json")
CASES = json.load(open(GOLDEN))["cases"]
PLANS = {f: Plan.load(f).blob for f in ("ini", "toml")}


def test_cases_reach_every_role():
    for fmt, lists in (("ini", ("sections", "keys", "vals")), ("toml", ("sections", "entries", "elements"))):
        cs = [c["ref"] for c in CASES if c["fmt"] == fmt]
        assert len(cs) >= 50
        for name in lists:
            assert sum(c[name]["n"] > 0 for c in cs) >= 10, (fmt, name)


@pytest.mark.parametrize("case", range(len(CASES)), ids=[c["name"] for c in CASES])
def test_restatement_equals_the_reference_hook(case):
    c = CASES[case]
    d = inputs(c["input"])
    pos = O.scan(PLANS[c["fmt"]], np.frombuffer(d, np.uint8))["positions"]
    assert K.digest(c["fmt"], d, K.tokens(c["fmt"], d, pos)) == c["ref"], c["name"]


def test_planted_mistake_is_caught(monkeypatch):
    """A restatement that splits quoted tokens at whitespace (a plausible
    slip: each_token without its quote branch) disagrees with the reference
    on the goldens."""
    def plain_tokens(base, b, e):
        i = b
        while i < e:
            while i < e and base[i] in K.WS:
                i += 1
            if i >= e:
                break
            s = i
            while i < e and base[i] not in K.WS:
                i += 1
            yield s, i

    monkeypatch.setattr(K, "each_token", plain_tokens)
    for name in ("ini_hand_5", "toml_hand_8"):
        c = next(c for c in CASES if c["name"] == name)
        d = inputs(c["input"])
        pos = O.scan(PLANS[c["fmt"]], np.frombuffer(d, np.uint8))["positions"]
        assert K.digest(c["fmt"], d, K.tokens(c["fmt"], d, pos)) != c["ref"], name
