"""GPU parity on plans made by the grammar compiler (pars_amd/compiler.py).

Each probe grammar of oracle/grammars/compiler_cases.hpp is compiled at test
time, checked byte-equal to the reference compiler's plan
(tests/golden/plans/), and its plan driven through the HIP scanner — single
documents across tile boundaries, several rows per tile, and a batch — bit-exact
against the oracle, which tests/test_compiler.py pins to the reference CPU
backend on these plans.  The plans take shapes the shipped grammars do not:
no lexical descriptor at all, eight bracket pairs (no one-bit bracket
ordinals), a NUL punctuation byte, a bracket byte that also opens a mode, a
mode without an escape, two descriptors, and the empty plan.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from oracle import oracle as O
from pars_amd import Plan, Scanner
from pars_amd import _native as N
from pars_amd import compiler as C
from pars_amd import workloads as W
from tests.gpu_util import gpu_full, say
from tests.test_gpu_batch import check_batch, run_batch

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "plans")
with open(os.path.join(ROOT, "oracle", "grammars", "compiler_cases.hpp")) as _f:
    PROBES = C.parse_grammars(_f.read())
NAMES = sorted(PROBES)
ALPHABETS = {
    "sql_case": b"''''  ''(()),;*=-\n-NULxy19 @\"\"",
    "esc_case": b"|||%%``^^af09#\n{}{}<>,:! x",
    "many_case": b"()[]{}<>ABCDEFGHIJ,;:.!?@$%&=+xyz ",
    "caps_case": b'~abcdefgh;;$aAbBcC,,@@.."""k\x00[]]+ ',
    "tiny_case": b"xxy;{}",
    "unsup_case": b"'''()!!-x ",
    "prop_case": b",,,xxy(()) ",
}
_plans: dict = {}


def plan_of(name: str) -> Plan:
    if name not in _plans:
        blob = C.make_scan_plan(PROBES[name]).to_bytes()
        with open(os.path.join(GOLD, name + ".plan"), "rb") as f:
            assert blob == f.read(), f"{name}: compiled plan differs from the reference compiler's"
        _plans[name] = Plan(blob, name)
    return _plans[name]


def compare(plan: bytes, data: np.ndarray, got: dict, tag: str):
    exp = O.full(plan, data)
    assert got["n_struct"] == exp["positions"].size, tag
    assert got["n_brackets"] == exp["bracket_pos"].size, tag
    assert np.array_equal(got["positions"], exp["positions"]), tag
    if "chars" in got:
        assert np.array_equal(got["chars"], exp["chars"]), tag
    assert np.array_equal(got["bracket_pos"], exp["bracket_pos"]), tag
    assert np.array_equal(got["bracket_char"], exp["bracket_char"]), tag
    for k in ("balanced", "well_formed", "max_depth", "unclosed_mode"):
        assert got[k] == exp[k], f"{tag} {k}: gpu {got[k]} cpu {exp[k]}"
    if "pair" in got:
        assert np.array_equal(got["pair"], exp["pair"]), tag
    if "pair_pos" in got:
        assert np.array_equal(got["pair_pos"],
                              O.pair_pos_device(exp["positions"], exp["bracket_pos"], exp["pair"])), tag


@pytest.mark.parametrize("name", NAMES)
def test_compiled_plan_documents(torch_cuda, name):
    P = plan_of(name)
    sc = Scanner(P)
    for seed, n in [(21, 1), (22, 4095), (23, 65536 + 7), (24, 3 << 20)]:
        d = W.alpha(seed, n, ALPHABETS[name])
        say(f"[gpu] compiled {name} n={n}")
        compare(P.blob, d, gpu_full(torch_cuda, sc, d), f"{name} n={n}")


@pytest.mark.parametrize("rows", [1, 2, 3])
@pytest.mark.parametrize("name", ["sql_case", "caps_case", "many_case"])
def test_compiled_plan_rows(torch_cuda, name, rows):
    P = plan_of(name)
    sc = Scanner(P)
    sc.set_option(N.PARS_OPT_ROWS_PER_TILE, rows)
    d = W.alpha(31 + rows, 1 << 20, ALPHABETS[name])
    say(f"[gpu] compiled {name} rows={rows}")
    compare(P.blob, d, gpu_full(torch_cuda, sc, d), f"{name} rows={rows}")


@pytest.mark.parametrize("name", NAMES)
def test_compiled_plan_batch(torch_cuda, name):
    P = plan_of(name)
    rng = np.random.default_rng(41)
    docs = [W.alpha(500 + i, int(n), ALPHABETS[name])
            for i, n in enumerate([1, 63, 64, 4097, 16385] + list(rng.integers(1, 60_000, 40)))]
    say(f"[gpu] compiled {name} batch of {len(docs)}")
    sc = Scanner(P)
    check_batch(name, docs, run_batch(torch_cuda, sc, docs), "compiled", plan=P.blob)
