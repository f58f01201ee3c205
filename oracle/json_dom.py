"""CPU restatement of the reference's stage-2 consumer path for JSON.

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's cpu_baseline leg,
never by the product.

  replay            include/pars/model/visit.hpp:35-48 — the structural index
                    as a covering event stream: scalar(begin, end) for each
                    maximal run of non-structural bytes, symbol(pos, c, match)
                    for each structural byte
  StructuralCursor  include/pars/model/cursor.hpp:16-45 — random access to the
                    same index (offset, byte, partner, the gap before a token)
  JsonTape          bench/bench_parse.cu:41-333 — the JSON DOM hook the
                    reference drives through replay: a tape of
                    (kind, begin, end) nodes, every value decoded and folded
                    into an FNV-1a style hash (the hook's own seed, HOOK_SEED),
                    `ok` cleared at the first event
                    the grammar rejects (nothing is consumed after it)

Events are numbered as the device tape reports them: the run of bytes before
structural i is event 2i, structural i is event 2i + 1, the run after the
last structural is event 2n.  `fail_event` is the event at which the hook
stopped (-1 when every event was accepted; `ok` can still be false from the
final checks).
"""
from __future__ import annotations

import numpy as np

# node kinds, in the order of the reference's jkind (bench_parse.cu:33)
J_OBJ, J_ARR, J_STR, J_NUM, J_TRUE, J_FALSE, J_NULL = range(7)

FNV_OFFSET = 0xCBF29CE484222325  # the published FNV-1a 64 offset basis
FNV_PRIME = 0x100000001B3
# The hook seeds its hash with 1469598103934665603 (bench_parse.cu:57, 209), a
# 19-digit constant rather than the 20-digit offset basis above; the mixing
# step is FNV-1a's.
HOOK_SEED = 1469598103934665603
_M64 = (1 << 64) - 1
_WS = b" \t\n\r"
_ESC = {ord('"'): ord('"'), ord("\\"): ord("\\"), ord("/"): ord("/"), ord("b"): 8, ord("f"): 12,
        ord("n"): 10, ord("r"): 13, ord("t"): 9}


def fnv1a64(data: bytes, h: int = FNV_OFFSET) -> int:
    for b in data:
        h = ((h ^ b) * FNV_PRIME) & _M64
    return h


def replay(base: bytes, positions, pair_pos, visitor) -> None:
    """visit.hpp:35-48: scalar runs and structural symbols in byte order;
    `match` is the byte offset of the partner (pair_pos holds partner indices)."""
    prev = 0
    for i, p in enumerate(positions):
        p = int(p)
        if p > prev:
            visitor.scalar(prev, p, 2 * i)
        m = -1
        if pair_pos is not None and int(pair_pos[i]) >= 0:
            m = int(positions[int(pair_pos[i])])
        visitor.symbol(p, base[p], m, 2 * i + 1)
        prev = p + 1
    if len(base) > prev:
        visitor.scalar(prev, len(base), 2 * len(positions))


class StructuralCursor:
    """cursor.hpp:16-45."""

    def __init__(self, base: bytes, positions, pair_pos=None):
        self.base, self.pos, self.pair, self.n = base, positions, pair_pos, len(positions)

    def valid(self, i: int) -> bool:
        return 0 <= i < self.n

    def off(self, i: int) -> int:
        return int(self.pos[i])

    def ch(self, i: int) -> int:
        return self.base[int(self.pos[i])]

    def match_index(self, i: int) -> int:
        return int(self.pair[i]) if self.pair is not None else -1

    def gap_begin(self, i: int) -> int:
        return int(self.pos[i - 1]) + 1 if i else 0

    def gap_end(self, i: int) -> int:
        return int(self.pos[i])

    def gap(self, i: int) -> bytes:
        return self.base[self.gap_begin(i):self.gap_end(i)]


def trim(s: bytes) -> bytes:
    """cursor.hpp:48-57."""
    return s.strip(_WS)


class JsonTape:
    """The reference's replay-driven JSON DOM builder, restated
    (bench_parse.cu:41-333; member names follow it)."""

    def __init__(self, base: bytes):
        self.base = base
        self.tape: list = []          # [kind, begin, end]
        self.decoded: list = []       # per string node: the decoded bytes it mixed
        self.node_stack: list = []
        self.stack: list = []         # [open byte, phase]
        self.ok = True
        self.top_done = False
        self.hash = HOOK_SEED
        self.fail_event = -1
        self._ev = -1

    # -- plumbing ---------------------------------------------------------
    def _fail(self):
        if self.ok:
            self.ok = False
            self.fail_event = self._ev

    def mix(self, b: int):
        self.hash = ((self.hash ^ b) * FNV_PRIME) & _M64

    def in_value(self) -> bool:
        if not self.stack:
            return not self.top_done
        o, ph = self.stack[-1]
        return ph == 2 if o == ord("{") else ph == 0

    def value_done(self):
        if not self.stack:
            self.top_done = True
        elif self.stack[-1][0] == ord("{"):
            self.stack[-1][1] = 3
        else:
            self.stack[-1][1] = 1

    def push_node(self, k: int, b: int):
        self.tape.append([k, b, 0])

    # -- replay events (bench_parse.cu:86-194) -----------------------------
    def scalar(self, b: int, e: int, ev: int = -1):
        if not self.ok:
            return
        self._ev = ev
        base = self.base
        i = b
        while i < e and base[i] in _WS:
            i += 1
        if i >= e:
            return
        as_key = bool(self.stack) and self.stack[-1][0] == ord("{") and self.stack[-1][1] == 0
        if as_key:
            if base[i] != ord('"'):
                return self._fail()
            j = self.parse_string(i, e)
        else:
            if not self.in_value():
                return self._fail()
            j = self.parse_value(i, e)
        if j is None:
            return self._fail()
        while j < e and base[j] in _WS:
            j += 1
        if j != e:
            return self._fail()
        if as_key:
            self.stack[-1][1] = 1
        else:
            self.value_done()

    def symbol(self, p: int, c: int, match: int = -1, ev: int = -1):
        if not self.ok:
            return
        self._ev = ev
        st = self.stack
        if c in (ord("{"), ord("[")):
            if not self.in_value():
                return self._fail()
            self.push_node(J_OBJ if c == ord("{") else J_ARR, p)
            self.node_stack.append(len(self.tape) - 1)
            st.append([c, 0])
        elif c == ord("}"):
            if not st or st[-1][0] != ord("{") or st[-1][1] not in (0, 3):
                return self._fail()
            self.tape[self.node_stack.pop()][2] = p
            st.pop()
            self.value_done()
        elif c == ord("]"):
            if not st or st[-1][0] != ord("[") or st[-1][1] not in (0, 1):
                return self._fail()
            self.tape[self.node_stack.pop()][2] = p
            st.pop()
            self.value_done()
        elif c == ord(":"):
            if not st or st[-1][0] != ord("{") or st[-1][1] != 1:
                return self._fail()
            st[-1][1] = 2
        elif c == ord(","):
            if not st:
                return self._fail()
            if st[-1][0] == ord("{"):
                if st[-1][1] != 3:
                    return self._fail()
            elif st[-1][1] != 1:
                return self._fail()
            st[-1][1] = 0
        else:
            self._fail()

    def finish(self):
        if self.stack or not self.top_done:
            self.ok = False

    # -- values (bench_parse.cu:219-332) -----------------------------------
    def parse_value(self, i: int, e: int):
        c = self.base[i]
        if c == ord('"'):
            return self.parse_string(i, e)
        if c == ord("-") or ord("0") <= c <= ord("9"):
            return self.parse_number(i, e)
        if c in (ord("t"), ord("f"), ord("n")):
            return self.parse_lit(i, e)
        return None

    def parse_string(self, i: int, e: int):
        base = self.base
        self.push_node(J_STR, i)
        dec = bytearray()
        self.decoded.append(dec)
        self.mix(ord("S"))

        def put(b):
            dec.append(b)
            self.mix(b)

        j = i + 1
        while j < e:
            c = base[j]
            j += 1
            if c == ord('"'):
                self.tape[-1][2] = j
                self.mix(0)
                return j
            if c == ord("\\"):
                if j >= e:
                    return None
                esc = base[j]
                j += 1
                if esc == ord("u"):
                    if j + 4 > e:
                        return None
                    v = 0
                    for _ in range(4):
                        h = base[j]
                        j += 1
                        v <<= 4
                        if ord("0") <= h <= ord("9"):
                            v |= h - ord("0")
                        elif ord("a") <= h <= ord("f"):
                            v |= h - ord("a") + 10
                        elif ord("A") <= h <= ord("F"):
                            v |= h - ord("A") + 10
                        else:
                            return None
                    if v < 0x80:
                        put(v)
                    elif v < 0x800:
                        put(0xC0 | (v >> 6))
                        put(0x80 | (v & 0x3F))
                    else:
                        put(0xE0 | (v >> 12))
                        put(0x80 | ((v >> 6) & 0x3F))
                        put(0x80 | (v & 0x3F))
                    continue
                if esc not in _ESC:
                    return None
                put(_ESC[esc])
            elif c < 0x20:
                return None
            else:
                put(c)
        return None

    def parse_number(self, i: int, e: int):
        base = self.base
        dig = lambda k: k < e and ord("0") <= base[k] <= ord("9")  # noqa: E731
        j = i
        if base[j] == ord("-"):
            j += 1
        if j >= e:
            return None
        if base[j] == ord("0"):
            j += 1
        elif ord("1") <= base[j] <= ord("9"):
            while dig(j):
                j += 1
        else:
            return None
        if j < e and base[j] == ord("."):
            j += 1
            if not dig(j):
                return None
            while dig(j):
                j += 1
        if j < e and base[j] in (ord("e"), ord("E")):
            j += 1
            if j < e and base[j] in (ord("+"), ord("-")):
                j += 1
            if not dig(j):
                return None
            while dig(j):
                j += 1
        self.push_node(J_NUM, i)
        self.tape[-1][2] = j
        self.mix(ord("N"))
        for k in range(i, j):
            self.mix(base[k])
        return j

    def parse_lit(self, i: int, e: int):
        for word, kind, tag in ((b"true", J_TRUE, ord("T")), (b"false", J_FALSE, ord("F")),
                                (b"null", J_NULL, ord("0"))):
            if self.base[i:i + len(word)] == word and i + len(word) <= e:
                self.push_node(kind, i)
                self.tape[-1][2] = i + len(word)
                self.mix(tag)
                return i + len(word)
        return None


def json_tape(data, positions, pair_pos=None) -> dict:
    """The reference pipeline's consumer stage (bench_parse.cu:336-348): replay
    the structural index into the JSON hook, then its final checks."""
    base = bytes(np.asarray(data, dtype=np.uint8)) if not isinstance(data, (bytes, bytearray)) else bytes(data)
    t = JsonTape(base)
    replay(base, positions, pair_pos, t)
    t.finish()
    tape = np.asarray(t.tape, dtype=np.int64).reshape(-1, 3)
    return {"ok": t.ok, "nodes": len(t.tape), "hash": t.hash, "fail_event": t.fail_event,
            "kind": tape[:, 0].astype(np.uint8), "begin": tape[:, 1].astype(np.uint32),
            "end": tape[:, 2].astype(np.uint32), "decoded": [bytes(d) for d in t.decoded]}


def tape_hash(data, kind, begin, end, strings) -> int:
    """The hook's FNV-1a 64, recomputed from a tape: per node, in order, 'S' +
    decoded bytes (+ 0 once the string closed), 'N' + the number's bytes, 'T',
    'F' or '0' for the literals; containers mix nothing (bench_parse.cu:219-332).
    `strings` yields each string node's decoded bytes in node order."""
    base = bytes(np.asarray(data, dtype=np.uint8)) if not isinstance(data, (bytes, bytearray)) else bytes(data)
    h = HOOK_SEED
    it = iter(strings)
    for k, b, e in zip(kind, begin, end):
        k = int(k)
        if k == J_STR:
            h = fnv1a64(b"S" + next(it), h)
            if int(e):
                h = fnv1a64(b"\0", h)
        elif k == J_NUM:
            h = fnv1a64(b"N" + base[int(b):int(e)], h)
        elif k == J_TRUE:
            h = fnv1a64(b"T", h)
        elif k == J_FALSE:
            h = fnv1a64(b"F", h)
        elif k == J_NULL:
            h = fnv1a64(b"0", h)
    return h
