"""The stage-2 consumer restatement (oracle/json_dom.py) pinned before it is
used as the device tape's checker.

  * FNV-1a 64: the published test vectors.
  * replay / StructuralCursor: event streams and gaps worked out by hand
    (include/pars/model/visit.hpp:35-48, cursor.hpp:16-57).
  * JsonTape on valid documents: the tape (kinds in document order, decoded
    strings, numbers) against Python's own json parser, an independent
    implementation of the same grammar.
  * JsonTape on malformed documents: where the hook stops, how many nodes it
    pushed and what it hashed, worked out by hand from bench_parse.cu:86-332.
"""
from __future__ import annotations

import json
import random

import numpy as np
import pytest

from oracle import json_dom as J
from oracle import oracle as O
from pars_amd import Plan
from pars_amd import workloads as W

PLAN = Plan.load("json").blob


def positions(d: bytes):
    return O.scan(PLAN, np.frombuffer(d, np.uint8))["positions"]


def run(d: bytes) -> dict:
    return J.json_tape(d, positions(d))


def test_fnv_vectors():
    """The mixing step against the published FNV-1a 64 vectors; the hook's own
    seed is its 19-digit constant (bench_parse.cu:57)."""
    assert J.HOOK_SEED == 1469598103934665603 and J.HOOK_SEED != J.FNV_OFFSET
    assert J.fnv1a64(b"") == 0xCBF29CE484222325
    assert J.fnv1a64(b"a") == 0xAF63DC4C8601EC8C
    assert J.fnv1a64(b"foobar") == 0x85944171F73967E8


class Rec:
    def __init__(self):
        self.ev = []

    def scalar(self, b, e, ev):
        self.ev.append(("s", b, e, ev))

    def symbol(self, p, c, m, ev):
        self.ev.append(("y", p, chr(c), m, ev))


def test_replay_by_hand():
    d = b' {"a": [1, 2]} x'
    pos = positions(d)
    assert list(pos) == [1, 5, 7, 9, 12, 13]
    pair = np.array([5, -1, 4, -1, 2, 0])  # partner structural indices
    r = Rec()
    J.replay(d, pos, pair, r)
    assert r.ev == [("s", 0, 1, 0), ("y", 1, "{", 13, 1), ("s", 2, 5, 2), ("y", 5, ":", -1, 3),
                    ("s", 6, 7, 4), ("y", 7, "[", 12, 5), ("s", 8, 9, 6), ("y", 9, ",", -1, 7),
                    ("s", 10, 12, 8), ("y", 12, "]", 7, 9), ("y", 13, "}", 1, 11), ("s", 14, 16, 12)]
    r = Rec()
    J.replay(b"[]", [0, 1], None, r)  # no pairs, no gaps
    assert r.ev == [("y", 0, "[", -1, 1), ("y", 1, "]", -1, 3)]


def test_cursor_by_hand():
    d = b'{"k": 12 }'
    c = J.StructuralCursor(d, positions(d), np.array([2, -1, 0]))
    assert c.n == 3 and c.valid(2) and not c.valid(3)
    assert [c.off(i) for i in range(3)] == [0, 4, 9]
    assert chr(c.ch(1)) == ":" and c.match_index(0) == 2 and c.match_index(1) == -1
    assert c.gap(0) == b"" and c.gap(1) == b'"k"' and c.gap(2) == b" 12 "
    assert (c.gap_begin(2), c.gap_end(2)) == (5, 9) and J.trim(c.gap(2)) == b"12"
    assert J.StructuralCursor(d, positions(d)).match_index(0) == -1


# ---------------------------------------------------------------------------
def _expected_from_json(d: bytes):
    """Tape kinds, decoded strings and number values in document order, from
    Python's json parser (object pairs kept in order)."""
    kinds, strings, numbers = [], [], []

    def walk(v):
        if isinstance(v, _Obj):
            kinds.append(J.J_OBJ)
            for k, x in v:
                kinds.append(J.J_STR)
                strings.append(_units(k))
                walk(x)
        elif isinstance(v, list):
            kinds.append(J.J_ARR)
            for x in v:
                walk(x)
        elif isinstance(v, str):
            kinds.append(J.J_STR)
            strings.append(_units(v))
        elif v is True:
            kinds.append(J.J_TRUE)
        elif v is False:
            kinds.append(J.J_FALSE)
        elif v is None:
            kinds.append(J.J_NULL)
        else:
            kinds.append(J.J_NUM)
            numbers.append(v)

    walk(json.loads(d.decode("utf-8"), object_pairs_hook=_Obj))
    return kinds, strings, numbers


class _Obj(list):
    pass


def _units(s: str) -> bytes:
    """A string as UTF-16 code units: the hook encodes every \\u unit on its own
    (a surrogate pair becomes two 3-byte sequences) where Python joins pairs."""
    return s.encode("utf-16-le", "surrogatepass")


def _valid_docs():
    docs = [W.json(s, n, 8).tobytes() for s, n in [(1, 30_000), (5, 120_000)]]
    docs.append(W.json(9, 60_000, 24, 0.6).tobytes())
    rec = b'{"id":123,"name":"x \\"y\\"","a":[1,2,3],"n":{"z":null}}'  # bench_parse.cu:405-412
    docs.append(b"[" + b",".join([rec] * 300) + b"]")
    docs += [b'"\\u00e9\\u20ac\\/\\b\\f\\n\\r\\t"', b" -0.5e+7 ", b"[[], {}, [[[]]], {\"\": {}}]",
             b'{"a": [true, false, null, 0, -1, 1.25, 2E-3]}']
    return docs


@pytest.mark.parametrize("i", range(8))
def test_hook_matches_json_module(i):
    d = _valid_docs()[i]
    r = run(d)
    kinds, strings, numbers = _expected_from_json(d)
    assert r["ok"] and r["fail_event"] == -1
    assert list(r["kind"]) == kinds
    assert [_units(x.decode("utf-8", "surrogatepass")) for x in r["decoded"]] == strings
    nums = [json.loads(d[b:e]) for k, b, e in zip(r["kind"], r["begin"], r["end"]) if k == J.J_NUM]
    assert nums == numbers
    # every node's span is its source text; containers end at their close
    for k, b, e in zip(r["kind"], r["begin"], r["end"]):
        if k in (J.J_OBJ, J.J_ARR):
            assert d[b] in b"{[" and d[e] == {b"{"[0]: ord("}"), b"["[0]: ord("]")}[d[b]]
        else:
            json.loads(d[b:e])
    assert J.tape_hash(d, r["kind"], r["begin"], r["end"], r["decoded"]) == r["hash"]


def _h(*parts: bytes) -> int:
    return J.fnv1a64(b"".join(parts), J.HOOK_SEED)


# (document, ok, fail_event, nodes, hash), worked out by hand
KAT = [
    (b"", False, -1, 0, J.HOOK_SEED),
    (b"   ", False, -1, 0, J.HOOK_SEED),
    (b"[1,]", True, -1, 2, _h(b"N1")),                 # a trailing comma is accepted
    (b'{"a":1,}', True, -1, 3, _h(b"Sa\0", b"N1")),
    (b'{"a" 1}', False, 2, 2, _h(b"Sa\0")),            # key then a value in one gap
    (b"[1 2]", False, 2, 2, _h(b"N1")),
    (b"{} {}", False, 5, 1, J.HOOK_SEED),             # a second top-level value
    (b"1 2", False, 0, 1, _h(b"N1")),
    (b"[}", False, 3, 1, J.HOOK_SEED),
    (b'"abc', False, 0, 1, _h(b"Sabc")),               # unclosed: no terminating 0 mixed
    (b'{"a":1,"b"}', False, 7, 4, _h(b"Sa\0", b"N1", b"Sb\0")),
    (b"[1]]", False, 5, 2, _h(b"N1")),
    (b"[", False, -1, 1, J.HOOK_SEED),                # still open at the end
    (b'  "x"  ', True, -1, 1, _h(b"Sx\0")),
    (b"tru", False, 0, 0, J.HOOK_SEED),
    (b"[01]", False, 2, 2, _h(b"N0")),
    (b'{"a":"\\q"}', False, 4, 3, _h(b"Sa\0", b"S")),  # an unknown escape
    (b"{1:2}", False, 2, 1, J.HOOK_SEED),             # a key must be a string
    (b"[:]", False, 3, 1, J.HOOK_SEED),
    (b'{"a"}', False, 3, 2, _h(b"Sa\0")),
    (b'{"a":}', False, 5, 2, _h(b"Sa\0")),
    (b"[,1]", False, 3, 1, J.HOOK_SEED),
    (b'["\\u00e9"]', True, -1, 2, _h(b"S\xc3\xa9\0")),
    (b'"\\ud800"', True, -1, 1, _h(b"S\xed\xa0\x80\0")),  # each \u unit encoded alone
    (b'"a\x01"', False, 0, 1, _h(b"Sa")),              # a control byte
    (b"{}}", False, 5, 1, J.HOOK_SEED),
    (b"nul", False, 0, 0, J.HOOK_SEED),
    (b"[-]", False, 2, 1, J.HOOK_SEED),
    (b"[1.]", False, 2, 1, J.HOOK_SEED),
    (b"[1e+]", False, 2, 1, J.HOOK_SEED),
    (b"[true false]", False, 2, 2, _h(b"T")),
]


@pytest.mark.parametrize("case", range(len(KAT)))
def test_hook_known_answers(case):
    d, ok, fe, nodes, h = KAT[case]
    r = run(d)
    assert (r["ok"], r["fail_event"], r["nodes"], r["hash"]) == (ok, fe, nodes, h), d
    assert J.tape_hash(d, r["kind"], r["begin"], r["end"], r["decoded"]) == r["hash"]


def test_hook_partial_tape_by_hand():
    r = run(b'[{"a": [1, 2}]')  # '}' closes nothing open: the array is innermost
    assert r["fail_event"] == 2 * 5 + 1 and not r["ok"]
    assert [tuple(x) for x in zip(r["kind"], r["begin"], r["end"])] == [
        (J.J_ARR, 0, 0), (J.J_OBJ, 1, 0), (J.J_STR, 2, 5), (J.J_ARR, 7, 0), (J.J_NUM, 8, 9), (J.J_NUM, 11, 12)]
    r = run(b'{"a": [1]} ')
    assert [tuple(x) for x in zip(r["kind"], r["begin"], r["end"])] == [
        (J.J_OBJ, 0, 9), (J.J_STR, 1, 4), (J.J_ARR, 6, 8), (J.J_NUM, 7, 8)]


def test_hook_random_documents_hash_from_tape():
    rng = random.Random(3)
    parts = [b"{", b"}", b"[", b"]", b",", b":", b'"a"', b'"', b"1", b"-2.5e3", b"true", b"null", b" ",
             b"x", b'\\"', b"01", b'"\\u00e9"']
    for _ in range(300):
        d = b"".join(rng.choice(parts) for _ in range(rng.randint(0, 25)))
        r = run(d)
        assert J.tape_hash(d, r["kind"], r["begin"], r["end"], r["decoded"]) == r["hash"]
        assert r["nodes"] == len(r["kind"])


# ---------------------------------------------------------------------------
def _same_as_python(d: bytes):
    pos = positions(d)
    a, b = run(d), O.json_tape(d, pos)
    assert (a["ok"], a["nodes"], a["hash"], a["fail_event"]) == (b["ok"], b["nodes"], b["hash"], b["fail_event"]), d[:60]
    assert np.array_equal(a["kind"], b["kind"]) and np.array_equal(a["begin"], b["begin"])
    assert np.array_equal(a["end"], b["end"]) and b"".join(a["decoded"]) == b["dec"]


def test_c_twin_matches_restatement():
    """oracle/json_dom.c (the fast twin) against oracle/json_dom.py."""
    for d, *_ in KAT:
        _same_as_python(d)
    for d in _valid_docs():
        _same_as_python(d)
    rng = random.Random(4)
    parts = [b"{", b"}", b"[", b"]", b",", b":", b'"a"', b'"', b"1", b"-2.5e3", b"true", b"null", b" ", b"x",
             b'\\"', b"01", b'"\\u00e9"', b"fals", b"\\u12"]
    for _ in range(500):
        _same_as_python(b"".join(rng.choice(parts) for _ in range(rng.randint(0, 30))))
