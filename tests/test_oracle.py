"""CPU tests: pin the oracle (oracle/pars_oracle.c) before it judges the GPU.

Three independent pins:
  1. golden vectors written by the reference's own CPU backend, compiled from
     its sources (tests/golden/make_golden.py -> golden.json);
  2. the reference's known-answer tests (tests/test_model.cpp:152-257), whose
     naive scanners are re-run here against the oracle;
  3. closed forms that differ from a DFA walk: regex region scanners per
     descriptor kind (masked bytes of a string / comment region), the CSV parity
     formula (doubled quotes), a stack matcher for pairing balanced input, and
     exhaustive strings over each plan's class alphabet.
"""
from __future__ import annotations

import itertools
import re

import numpy as np
import pytest

from oracle import oracle as O
from pars_amd import Plan
from pars_amd import workloads as W

PLANS = {g: Plan.load(g) for g in ["json", "csv", "clike", "clike6", "toml", "ini", "xml"]}


def golden_input(case: dict) -> np.ndarray:
    if "input" in case:
        return np.frombuffer(case["input"].encode("latin-1"), np.uint8)
    from tests.golden.make_golden import make_input

    d = make_input(case["spec"])
    assert O.fnv(d) == case["input_hash"], "generator output changed: golden input no longer reproducible"
    return d


# ---------------------------------------------------------------------------
# 1. golden vectors from the reference
# ---------------------------------------------------------------------------
def test_golden_cases_present(golden):
    assert len(golden["cases"]) >= 30


@pytest.mark.parametrize("name", sorted(__import__("json").load(open(__import__("os").path.join(
    __import__("os").path.dirname(__file__), "golden", "golden.json")))["cases"]))
def test_oracle_matches_reference_golden(golden, name):
    case = golden["cases"][name]
    d = golden_input(case)
    r = O.full(PLANS[case["plan"]].blob, d)
    assert r["positions"].size == case["n_struct"]
    assert r["bracket_pos"].size == case["n_brackets"]
    for k in ("unclosed_mode", "balanced", "well_formed", "max_depth"):
        assert r[k] == case[k], k
    for k, h in case["hash"].items():
        assert O.fnv(r[k]) == h, k
    if "arrays" in case:
        for k, v in case["arrays"].items():
            assert r[k].tolist() == v, k


# ---------------------------------------------------------------------------
# 2. the reference's known-answer tests (test_model.cpp), naive scanners restated
# ---------------------------------------------------------------------------
def naive_json(s: bytes):  # test_model.cpp:157-172
    pos, ch, in_str = [], [], False
    for i, c in enumerate(s):
        if c == ord('"'):
            esc, j = False, i
            while j > 0 and s[j - 1] == ord("\\"):
                esc = not esc
                j -= 1
            if not esc:
                in_str = not in_str
        if not in_str and c in b"{}[]:,":
            pos.append(i)
            ch.append(c)
    return pos, ch


def naive_csv(s: bytes):  # test_model.cpp:181-198
    pos, ch, in_q, i = [], [], False, 0
    while i < len(s):
        c = s[i]
        if c == ord('"'):
            if in_q and i + 1 < len(s) and s[i + 1] == ord('"'):
                i += 2
                continue
            in_q = not in_q
            i += 1
            continue
        if not in_q and c in b",\n":
            pos.append(i)
            ch.append(c)
        i += 1
    return pos, ch


def naive_clike(s: bytes):  # test_model.cpp:232-253
    pos, ch, i, n = [], [], 0, len(s)
    while i < n:
        if s[i] == ord("/") and i + 1 < n and s[i + 1] == ord("*"):
            i += 2
            while i + 1 < n and not (s[i] == ord("*") and s[i + 1] == ord("/")):
                i += 1
            i = i + 2 if i + 2 <= n else n
            continue
        if s[i] == ord("/") and i + 1 < n and s[i + 1] == ord("/"):
            i += 2
            while i < n and s[i] != ord("\n"):
                i += 1
            continue
        if s[i] in b";{}":
            pos.append(i)
            ch.append(s[i])
        i += 1
    return pos, ch


KAT_JSON = b'{"name":"a\\"b","n":[1,2.5e-3,true,null],"x":{}}'
KAT_CSV = b'a,"b, c","d""e",f\n1,2,3,4\n'
KAT_CLIKE = [b"a;b/* { ; } */c;// x { } ;\nd;{e;f}", b"/**/;", b"/***/;", b"/* * */x;", b"/*/**/;",
             b"a/*a**b*/;c", b"// /* ; { */\n;"]


def test_kat_json_scan():
    r = O.scan(PLANS["json"].blob, KAT_JSON)
    pos, ch = naive_json(KAT_JSON)
    assert r["positions"].tolist() == pos and r["chars"].tolist() == ch


def test_kat_csv_scan():
    r = O.scan(PLANS["csv"].blob, KAT_CSV)
    pos, ch = naive_csv(KAT_CSV)
    assert r["positions"].tolist() == pos and r["chars"].tolist() == ch


@pytest.mark.parametrize("s", KAT_CLIKE)
def test_kat_clike_scan(s):
    r = O.scan(PLANS["clike"].blob, s)
    pos, ch = naive_clike(s)
    assert r["positions"].tolist() == pos and r["chars"].tolist() == ch


def test_kat_pair():  # test_model.cpp:203-217
    r = O.full(PLANS["json"].blob, b'{"a":[1,{"b":2}],"c":[]}')
    assert r["balanced"] and r["well_formed"] and r["max_depth"] == 3
    p = r["pair"]
    assert p.size == r["bracket_char"].size
    for i, j in enumerate(p):
        assert j >= 0 and p[j] == i


# ---------------------------------------------------------------------------
# 3. closed forms: regex regions per descriptor kind
# ---------------------------------------------------------------------------
def _regions(pattern: bytes, s: bytes, skip: int = 1) -> np.ndarray:
    m = np.zeros(len(s), bool)
    for mt in re.finditer(pattern, s, re.S):
        m[mt.start() + skip : mt.end()] = True
    return m


def rx_string(q: bytes, esc: bytes | None):
    qe = re.escape(q)
    if esc is None:
        pat = qe + rb"[^" + qe + rb"]*(?:" + qe + rb"|\Z)"
    else:
        ee = re.escape(esc)
        pat = qe + rb"(?:" + ee + rb".|[^" + qe + ee + rb"])*(?:" + qe + rb"|" + ee + rb"?\Z)"
    return lambda s: _regions(pat, s)


def rx_comment(c: bytes):
    return lambda s: _regions(re.escape(c) + rb"[^\n]*", s)


rx_line = lambda s: _regions(rb"//[^\n]*", s)  # noqa: E731
rx_block = lambda s: _regions(rb"/\*.*?(?:\*/|\Z)", s)  # noqa: E731


def csv_parity(s: bytes) -> np.ndarray:
    """Appendix-B closed form: masked(i) = odd_before(i) | (Q[i] & Q[i-1])."""
    q = np.frombuffer(s, np.uint8) == ord('"')
    odd_incl = np.cumsum(q) & 1
    odd_before = np.concatenate([[0], odd_incl[:-1]]).astype(bool) if len(s) else np.zeros(0, bool)
    prev = np.concatenate([[False], q[:-1]]) if len(s) else np.zeros(0, bool)
    return odd_before | (q & prev)


MASKERS = {
    "json": [rx_string(b'"', b"\\")],
    "csv": [csv_parity],
    "clike": [rx_line, rx_block],
    "clike6": [rx_string(b'"', b"\\"), rx_string(b"'", b"\\"), rx_line, rx_block, rx_comment(b"#")],
    "toml": [rx_comment(b"#"), rx_string(b'"', b"\\"), rx_string(b"'", None)],
    "ini": [rx_comment(b";")],
    "xml": [rx_string(b'"', None), rx_string(b"'", None)],
}


def expected_mask(g: str, s: bytes) -> np.ndarray:
    m = np.zeros(len(s), bool)
    for f in MASKERS[g]:
        m |= f(s)
    return m


def bits_to_bool(words: np.ndarray, n: int) -> np.ndarray:
    return np.unpackbits(words.view(np.uint8), bitorder="little")[:n].astype(bool)


def check_against_regions(g: str, s: bytes):
    plan = PLANS[g].blob
    r = O.scan(plan, s, bitmaps=True)
    m = bits_to_bool(r["masked_bitmap"], len(s))
    exp = expected_mask(g, s)
    assert np.array_equal(m, exp), (g, s[:80], np.nonzero(m != exp)[0][:10])
    cls = O.classifier(plan)
    a = np.frombuffer(s, np.uint8)
    structural = (cls[a] & 1).astype(bool) & ~exp
    assert r["positions"].tolist() == np.nonzero(structural)[0].tolist()
    oc = (cls[a] & 6).astype(bool) & ~exp
    assert r["bracket_pos"].tolist() == np.nonzero(oc)[0].tolist()


@pytest.mark.parametrize("g", sorted(MASKERS))
@pytest.mark.parametrize("seed", [12345, 999, 7])
def test_regions_random_alphabet(g, seed):
    s = bytes(W.alpha(seed, 16 << 10, W.ALPHABETS[g]))
    check_against_regions(g, s)


@pytest.mark.parametrize("g,gen", [("json", "json"), ("csv", "csv"), ("clike6", "clike"), ("clike", "clike")])
def test_regions_generated(g, gen):
    s = bytes({"json": lambda: W.json(11, 256 << 10), "csv": lambda: W.csv(11, 256 << 10),
               "clike": lambda: W.clike(11, 256 << 10)}[gen]())
    check_against_regions(g, s)


# exhaustive short strings over each plan's event bytes
EXHAUSTIVE = {
    "json": (b'a"\\,', 8),
    "csv": (b'a",', 9),
    "clike": (b"a/*\n;", 7),
    "clike6": (b"a\"'\\/*#\n", 5),
    "toml": (b"a#\"'\\\n", 6),
    "xml": (b"a\"'<", 7),
}


@pytest.mark.parametrize("g", sorted(EXHAUSTIVE))
def test_exhaustive_short_strings(g):
    alphabet, L = EXHAUSTIVE[g]
    syms = [bytes([c]) for c in alphabet]
    count = 0
    for n in range(L + 1):
        for t in itertools.product(syms, repeat=n):
            check_against_regions(g, b"".join(t))
            count += 1
    assert count > 1000


# ---------------------------------------------------------------------------
# pairing: stack matcher on balanced input; invariants everywhere
# ---------------------------------------------------------------------------
def stack_pairs(plan: Plan, bch: np.ndarray) -> np.ndarray:
    opens = {o: c for o, c in reversed(plan.brackets)}
    closes = {c for _, c in plan.brackets}
    out = np.full(bch.size, -1, np.int32)
    st = []
    for i, c in enumerate(bch.tolist()):
        if c in opens:
            st.append(i)
        elif c in closes:
            j = st.pop()
            if opens[bch[j]] == c:
                out[i], out[j] = j, i
    return out


@pytest.mark.parametrize("g,data", [("json", lambda: W.json(1, 1 << 20)),
                                    ("json", lambda: W.json(2, 512 << 10, 32, 0.30)),
                                    ("clike6", lambda: W.clike(3, 512 << 10))])
def test_pairing_equals_stack_matcher_on_balanced(g, data):
    d = data()
    r = O.full(PLANS[g].blob, d)
    assert r["balanced"] and r["well_formed"]
    assert np.array_equal(r["pair"], stack_pairs(PLANS[g], r["bracket_char"]))
    depth = np.cumsum(np.where(np.isin(r["bracket_char"], [o for o, _ in PLANS[g].brackets]), 1, -1))
    assert r["max_depth"] == max(0, int(depth.max(initial=0)))


@pytest.mark.parametrize("g", ["json", "clike", "clike6", "toml"])
def test_pairing_invariants_malformed(g):
    d = W.alpha(4242, 32 << 10, W.ALPHABETS[g])
    r = O.full(PLANS[g].blob, d)
    p, b = r["pair"], r["bracket_char"]
    opens = {o for o, _ in PLANS[g].brackets}
    for i, j in enumerate(p.tolist()):
        if j >= 0:
            assert p[j] == i
            a, c = (i, j) if i < j else (j, i)
            assert b[a] in opens and b[c] not in opens
    assert not r["well_formed"] or r["balanced"]


# ---------------------------------------------------------------------------
# workload generators
# ---------------------------------------------------------------------------
def test_generators_exact_size_and_deterministic():
    for f in (lambda: W.json(5, 100_003), lambda: W.csv(5, 77_777), lambda: W.clike(5, 55_555),
              lambda: W.json(5, 3_000_001, 32, 0.3)):
        a, b = f(), f()
        assert a.size == b.size and np.array_equal(a, b)
    assert W.json(5, 100_003).size == 100_003


def test_json_1mib_shape():
    import json

    d = W.json(1, 1 << 20)
    obj = json.loads(bytes(d).decode("utf-8"))
    assert 1500 <= len(obj) <= 2600 and all(isinstance(o, dict) and 4 <= len(o) <= 12 for o in obj)


def test_json_large_string_fraction_and_depth():
    d = W.json(2, 8 << 20, 32, 0.30)
    frac = W.json.last_string_bytes / d.size
    assert 0.28 <= frac <= 0.33
    r = O.full(PLANS["json"].blob, d)
    assert r["balanced"] and r["max_depth"] == 32


def test_sizes_lognormal():
    s = W.sizes(4, 100_000)
    assert s.min() >= 200 and s.max() <= 4 << 20
    med = float(np.median(s))
    assert 3500 < med < 4700
    assert 1.6e9 < float(s.sum()) < 2.8e9
