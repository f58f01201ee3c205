"""The JSON tape's gap parser — the device automaton of
pars_amd/csrc/dom_kernels.hip, built for the host as tests/cpp/gap_check —
against the oracle's restatement of the hook's value parsers
(bench_parse.cu:86-130, 219-332): kind, acceptance, first byte, end and the
decoded bytes, on generated gaps (strings with every escape form, numbers in
every grammar position, literals, truncations, control bytes, whitespace)."""
from __future__ import annotations

import os
import random
import struct
import subprocess

import pytest

from oracle import json_dom as J

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "gap_check")
WS = b" \t\n\r"


def oracle_gap(d: bytes):
    e, i = len(d), 0
    while i < e and d[i] in WS:
        i += 1
    if i >= e:
        return (0, 0, 0, -1, b"")
    t = J.JsonTape(d)
    c = d[i]
    if c == ord('"'):
        j, kind, dec = t.parse_string(i, e), J.J_STR, bytes(t.decoded[0])
    elif c == ord("-") or 48 <= c <= 57:
        j = t.parse_number(i, e)
        kind, dec = (J.J_NUM if j is not None else 1), b""
    elif c in b"tfn":
        j = t.parse_lit(i, e)
        kind, dec = (t.tape[0][0] if j is not None else 1), b""
    else:
        j, kind, dec = None, 1, b""
    end = t.tape[0][2] if t.tape and t.tape[0][2] else -1
    ok = 0
    if j is not None:
        k = j
        while k < e and d[k] in WS:
            k += 1
        ok = int(k == e)
    return (kind, ok, i, end, dec)


def test_gap_parser_matches_oracle(tmp_path):
    if not os.access(EXE, os.X_OK):
        pytest.skip("tests/cpp/gap_check not built (run __graft_entry__.build())")
    rng = random.Random(7)
    parts = [b'"', b'"ab"', b"\\", b'\\"', b"\\u00e9", b"\\u12", b"\\uZZ", b"\\n", b"\\q", b" ", b"\t", b"\n",
             b"\x01", b"1", b"0", b"-", b".", b"e", b"E", b"+", b"12.5e-3", b"true", b"false", b"null", b"tru",
             b"fals", b"nul", b"x", b"a", b"/", b"\xc3\xa9", b"\\ud83d\\ude00", b"01", b"-0", b"1.", b"1e", b"abc",
             b"\\uABCD", b"\\u0800", b"\\/", b"\\b\\f\\r\\t"]
    gaps = [b"".join(rng.choice(parts) for _ in range(rng.randint(0, 8))) for _ in range(20000)]
    gaps += [b'"' + bytes(rng.randrange(32, 256) for _ in range(rng.randint(0, 80))).replace(b'"', b"x")
             .replace(b"\\", b"y") + b'"' for _ in range(1000)]
    # long strings whose plain runs (skipped a window at a time in the count
    # pass) end at every kind of byte that stops them, at every window offset
    stops = [b'\\n', b'\\"', b"\\u00e9", b"\x01", b"\t", b"\x1f", b'"', b"\\"]
    for _ in range(3000):
        body = b"".join(bytes(rng.randrange(32, 256) for _ in range(rng.randint(0, 70))).replace(b'"', b"x")
                        .replace(b"\\", b"y") + rng.choice(stops) for _ in range(rng.randint(1, 4)))
        gaps.append(b" " * rng.randint(0, 40) + b'"' + body + (b'"' if rng.random() < 0.8 else b"") + b" " * rng.randint(0, 3))
    path = tmp_path / "gaps.bin"
    with open(path, "wb") as f:
        for g in gaps:
            f.write(struct.pack("<I", len(g)))
            f.write(g)
    r = subprocess.run([EXE, str(path)], stdout=subprocess.PIPE, text=True, timeout=300)
    assert r.returncode == 0
    rows = r.stdout.split("\n")
    for g, row in zip(gaps, rows):
        f = row.split(" ")
        got = (int(f[0]), int(f[1]), int(f[2]), int(f[3]), bytes.fromhex(f[5]) if len(f) > 5 else b"")
        assert int(f[4]) == len(got[4])
        assert got == oracle_gap(g), g
