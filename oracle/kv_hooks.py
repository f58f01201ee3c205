"""CPU restatement of the reference's INI and TOML stage-2 hooks.

TEST INFRASTRUCTURE ONLY: the checker of the device key/value consumer
(pars_hip_kv_parse); only tests/ and bench.py's cpu_baseline leg use it.

The reference's full-parse benchmark (bench/bench_parse_host.cu) replays a
structural index (model/visit.hpp:35-48) into a user hook per format:
  * each_token (bench_parse_host.cu:34-56): a gap's whitespace-separated
    tokens, a token that starts with a quote running to the same quote
    (a backslash skips the next byte) or the gap's end;
  * ini_parse (:69-134): '\\n' ends a line and a comment; ';' or '#' starts a
    comment; '[' makes the next token a section, '=' the next token a value;
    every other token is a key; every token is hashed (FNV-1a 64 seeded with
    1469598103934665603, bench_parse_host.cu:58-65);
  * toml_parse (:202-248): '\\n' ends a line and a comment, '#' starts a
    comment, '[' at a value position opens an array (each later token of the
    line is an element), elsewhere a section; '=' makes the next token the
    value of the last key token (keys persist across lines); sections, key and
    value of each entry, and elements are hashed.

Each token the hook meets is reported with its role, in document order:
SECTION, KEY (INI: a key; TOML: a key token, which is hashed only through the
entry it names), VALUE (TOML: with the index of its key token) and ELEMENT.
"""
from __future__ import annotations

import numpy as np

SECTION, KEY, VALUE, ELEMENT = 0, 1, 2, 3
HOOK_SEED = 1469598103934665603
FNV_PRIME = 1099511628211
MASK = (1 << 64) - 1
WS = b" \t\n\r"


def fnv(h: int, s: bytes) -> int:
    for c in s:
        h = ((h ^ c) * FNV_PRIME) & MASK
    return h


def each_token(base: bytes, b: int, e: int):
    """(begin, end) of each token of base[b:e] (bench_parse_host.cu:34-56)."""
    i = b
    while i < e:
        while i < e and base[i] in WS:
            i += 1
        if i >= e:
            break
        s = i
        if base[i] in b'"\'':
            q = base[i]
            i += 1
            while i < e and base[i] != q:
                if base[i] == 0x5C and i + 1 < e:
                    i += 1
                i += 1
            if i < e:
                i += 1
        else:
            while i < e and base[i] not in WS:
                i += 1
        yield s, i


def _replay(base: bytes, positions, scalar, symbol) -> None:
    prev = 0
    for p in positions:
        p = int(p)
        if p > prev:
            scalar(prev, p)
        symbol(base[p])
        prev = p + 1
    if len(base) > prev:
        scalar(prev, len(base))


def ini_tokens(base: bytes, positions) -> list[tuple[int, int, int, int]]:
    """(role, begin, end, key) of every token ini_parse meets (key = -1)."""
    out = []
    st = {"expect": 0, "comment": False}

    def scalar(b, e):
        if st["comment"]:
            return
        for s, t in each_token(base, b, e):
            if st["expect"] == 1:
                out.append((SECTION, s, t, -1))
                st["expect"] = 0
            elif st["expect"] == 2:
                out.append((VALUE, s, t, -1))
                st["expect"] = 0
            else:
                out.append((KEY, s, t, -1))

    def symbol(c):
        if c == 0x0A:
            st["comment"], st["expect"] = False, 0
        elif c in b";#":
            st["comment"] = True
        elif st["comment"]:
            return
        elif c == 0x5B:
            st["expect"] = 1
        elif c == 0x5D:
            st["expect"] = 0
        elif c == 0x3D:
            st["expect"] = 2

    _replay(base, positions, scalar, symbol)
    return out


def toml_tokens(base: bytes, positions) -> list[tuple[int, int, int, int]]:
    """(role, begin, end, key) of every token toml_parse meets; a VALUE's key
    is the index of the last KEY token before it (-1: none yet)."""
    out = []
    st = {"mode": 0, "comment": False, "key": -1}

    def scalar(b, e):
        if st["comment"]:
            return
        for s, t in each_token(base, b, e):
            m = st["mode"]
            if m == 1:
                out.append((SECTION, s, t, -1))
                st["mode"] = 0
            elif m == 2:
                out.append((VALUE, s, t, st["key"]))
                st["mode"] = 0
            elif m == 3:
                out.append((ELEMENT, s, t, -1))
            else:
                st["key"] = len(out)
                out.append((KEY, s, t, -1))

    def symbol(c):
        if c == 0x0A:
            st["comment"], st["mode"] = False, 0
        elif c == 0x23:
            st["comment"] = True
        elif st["comment"]:
            return
        elif c == 0x5B:
            st["mode"] = 3 if st["mode"] == 2 else 1
        elif c == 0x5D:
            st["mode"] = 0
        elif c == 0x3D:
            st["mode"] = 2

    _replay(base, positions, scalar, symbol)
    return out


def digest(fmt: str, base: bytes, tokens) -> dict:
    """What the reference hook builds, from a token list: the hash it folds,
    and each list it fills (count, FNV-1a 64 over u32 length + bytes of each
    item in order).  INI: sections, keys, vals; TOML: sections, entries
    (key bytes, then value bytes), and the number of elements (the reference
    hook counts and hashes them without keeping them)."""
    def lst(items):
        h = HOOK_SEED
        for s in items:
            h = fnv(h, len(s).to_bytes(4, "little") + s)
        return {"n": len(items), "fnv": h}

    h = HOOK_SEED
    tok = [(r, base[b:e], k) for r, b, e, k in tokens]
    if fmt == "ini":
        for r, s, _ in tok:
            h = fnv(h, s)
        return {"hash": h, "sections": lst([s for r, s, _ in tok if r == SECTION]),
                "keys": lst([s for r, s, _ in tok if r == KEY]), "vals": lst([s for r, s, _ in tok if r == VALUE])}
    entries, elems, sections = [], [], []
    for r, s, k in tok:
        if r == SECTION:
            h = fnv(h, s)
            sections.append(s)
        elif r == VALUE:
            key = tok[k][1] if k >= 0 else b""
            h = fnv(fnv(h, key), s)
            entries.append(key)
            entries.append(s)
        elif r == ELEMENT:
            h = fnv(h, s)
            elems.append(s)
    return {"hash": h, "sections": lst(sections), "entries": lst(entries), "elements": {"n": len(elems)}}


def tokens(fmt: str, data, positions) -> list[tuple[int, int, int, int]]:
    base = bytes(np.asarray(data, dtype=np.uint8)) if not isinstance(data, (bytes, bytearray)) else bytes(data)
    return (ini_tokens if fmt == "ini" else toml_tokens)(base, positions)
