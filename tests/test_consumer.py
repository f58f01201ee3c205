"""The consumer side over a structural index, host only: pars::hip::replay,
structural_cursor and trim (include/pars_hip.hpp, driven by
tests/cpp/replay_check) and their Python twins (pars_amd.replay,
StructuralCursor, trim) against the reference's definitions
(include/pars/model/visit.hpp:35-48, cursor.hpp:16-66), restated here from a
byte mask: the scalar events are the maximal runs of non-structural bytes,
the symbols every structural byte in order with its partner's offset."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from pars_amd import Plan, StructuralCursor, replay, trim
from pars_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "replay_check")


def index_of(fmt: str, d: bytes):
    plan = Plan.load(fmt).blob
    r = O.full(plan, np.frombuffer(d, np.uint8))
    pos = r["positions"].astype(np.int32)
    pairs = O.pair_pos_device(r["positions"], r["bracket_pos"], r["pair"]) if r["bracket_pos"].size else None
    return pos, pairs


def expected_events(d: bytes, pos, pairs):
    """From the byte mask: runs of unmarked bytes, and marked bytes in order."""
    mark = np.zeros(len(d) + 1, bool)
    mark[np.asarray(pos, np.int64)] = True
    mark[len(d)] = True
    ev = []
    run = 0
    for i in range(len(d) + 1):
        if mark[i]:
            if i > run:
                ev.append(("S", run, i))
            run = i + 1
    sym = []
    for k, p in enumerate(pos):
        m = -1 if pairs is None or pairs[k] < 0 else int(pos[pairs[k]])
        sym.append(("Y", int(p), d[int(p)], m))
    # merge by offset: a scalar run precedes the structural at its end
    out, si = [], 0
    for s in ev:
        while si < len(sym) and sym[si][1] < s[1]:
            out.append(sym[si])
            si += 1
        out.append(s)
    return out + sym[si:]


DOCS = [
    ("json", b'{"a": [1, 2, {"b": null}], "c": "x,y"}'),
    ("json", b"[]"),
    ("json", b"  [1,2]  "),
    ("json", b""),
    ("json", b"x"),
    ("json", b",,,"),
    ("json", b'[1, "unterminated'),
    ("ini", W.format_doc("ini", 3000).tobytes()),
    ("toml", W.format_doc("toml", 3000).tobytes()),
    ("json", W.json(7, 20_000, 6).tobytes()),
    ("csv", W.alpha(3, 5000, W.ALPHABETS["csv"]).tobytes()),
]


class Recorder:
    def __init__(self):
        self.ev = []

    def scalar(self, b, e):
        self.ev.append(("S", b, e))

    def symbol(self, p, c, m):
        self.ev.append(("Y", p, c, m))


@pytest.mark.parametrize("case", range(len(DOCS)))
def test_replay_and_cursor(case, tmp_path):
    fmt, d = DOCS[case]
    pos, pairs = index_of(fmt, d)
    want = expected_events(d, pos, pairs)
    # Python
    rec = Recorder()
    replay(d, pos, rec, pairs)
    assert rec.ev == want
    # C++ header
    (tmp_path / "in").write_bytes(d)
    (tmp_path / "pos").write_bytes(np.asarray(pos, np.int32).tobytes())
    args = [EXE, str(tmp_path / "in"), str(tmp_path / "pos")]
    if pairs is not None:
        (tmp_path / "pair").write_bytes(np.asarray(pairs, np.int32).tobytes())
        args.append(str(tmp_path / "pair"))
    out = subprocess.run(args, stdout=subprocess.PIPE, text=True, check=True).stdout.splitlines()
    cpp_ev = []
    cur = []
    for line in out:
        f = line.split()
        if f[0] == "S":
            cpp_ev.append(("S", int(f[1]), int(f[2])))
        elif f[0] == "Y":
            cpp_ev.append(("Y", int(f[1]), int(f[2]), int(f[3])))
        else:
            cur.append(tuple(int(x) for x in f[1:]))
    assert cpp_ev == want
    # the cursor, both ways
    c = StructuralCursor(d, pos, pairs)
    assert len(cur) == c.n == len(pos)
    for i in range(c.n):
        b = 0 if i == 0 else int(pos[i - 1]) + 1
        g = d[b:int(pos[i])]
        t = g.strip(b" \t\n\r")
        tb = b + len(g) - len(g.lstrip(b" \t\n\r"))
        m = -1 if pairs is None else int(pairs[i])
        assert cur[i] == (i, int(pos[i]), d[int(pos[i])], m, b, int(pos[i]), tb, len(t))
        assert (c.off(i), c.ch(i), c.match_index(i), c.gap(i), trim(c.gap(i))) == (int(pos[i]), d[int(pos[i])], m, g, t)
    assert not c.valid(c.n)


def test_trim():
    assert trim(b" \t\r\n a b \n") == b"a b"
    assert trim(b"   ") == b""
    assert trim(b"\x0bx\x0b") == b"\x0bx\x0b"  # vertical tab is not whitespace here
