"""CPU model of the key-value table kernels (pars_amd/csrc/kv_kernels.hip),
step for step: the gap tokenizer's eight-state maps per 64-byte lane
(structural bytes reset every state), their scan over lanes and tiles, the
walk from each lane's entry state with its masks (token starts, first starts
of a gap, ends before a byte or with a closing quote), the hook's maps over
first starts and structural bytes, the token offsets, the roles and ends as
the table writer derives them (the k-th end of a lane closing its k-th start
or the token open at its start), and TOML keys as a max-scan.  The transition
tables are the kernels' constants, packed as they pack them.  Checked against
the restatement of the reference hooks (oracle/kv_hooks.py, itself pinned by
tests/test_kv_ref.py) on the golden documents and on random ones whose
tokens, quotes and gaps cross lane and tile boundaries (small tiles here, so
that every level carries work)."""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

from oracle import kv_hooks as K
from oracle import oracle as O
from pars_amd import Plan
from tests.golden.make_kv_golden import inputs

# the kernels' constants: per byte class (lo, hi) v_perm maps and the nibble table
WS = (0x01010100, 0x06060404, 0x66441110)
DQ = (0x03040404, 0x06060402, 0x66423444)
SQ = (0x03060606, 0x06020404, 0x62443666)
BS = (0x03030303, 0x06070405, 0x67453333)
OT = (0x03030303, 0x06060404, 0x66443333)
HOOK_FIRST = [0, 0, 0, 3, 4, 5, 6, 7]
ID = list(range(8))


def unbytes(lo: int, hi: int) -> list[int]:
    return [(lo >> 8 * i) & 0xFF for i in range(4)] + [(hi >> 8 * i) & 0xFF for i in range(4)]


def tok_class(c: int):
    if c in b" \t\n\r":
        return WS
    return {0x22: DQ, 0x27: SQ, 0x5C: BS}.get(c, OT)


def hook_symbol(c: int, toml: bool) -> list[int]:
    idh = [4, 5, 6, 7]
    if c == 0x0A:
        return [0] * 8
    if c == 0x23 or (not toml and c == 0x3B):
        return [4, 5, 6, 7] * 2
    if c == 0x5B:
        return ([1, 1, 3, 1] if toml else [1, 1, 1, 1]) + idh
    if c == 0x5D:
        return [0, 0, 0, 0] + idh
    if c == 0x3D:
        return [2, 2, 2, 2] + idh
    return ID


def hook_role(h: int, toml: bool) -> int:
    tab = 0x3201 if toml else 0x1201
    return 4 if h & 4 else (tab >> (4 * h)) & 0xF


def comp(outer, inner):
    return [outer[inner[s]] for s in range(8)]


def model(fmt: str, d: bytes, pos, lanes_per_tile: int):
    toml = fmt == "toml"
    n = len(d)
    sb = np.zeros(n, bool)
    sb[np.asarray(pos, np.int64)] = True
    nl = (n + 63) // 64
    ntile = (nl + lanes_per_tile - 1) // lanes_per_tile

    # pass 1: lane maps, their in-tile prefixes and tile maps
    lmap = []
    for L in range(nl):
        m = ID
        for j in range(L * 64, min(n, L * 64 + 64)):
            t = [0] * 8 if sb[j] else unbytes(*tok_class(d[j])[:2])
            m = comp(t, m)
        lmap.append(m)
    tile_map, lane_pre = [], []
    for t in range(ntile):
        pre = ID
        for L in range(t * lanes_per_tile, min(nl, (t + 1) * lanes_per_tile)):
            lane_pre.append(pre)
            pre = comp(lmap[L], pre)
        tile_map.append(pre)
    # chain: tile entry states
    tin, s = [], 0
    for t in range(ntile):
        tin.append(s)
        s = tile_map[t][s]

    # pass 3: walk, masks, counts, lane hook maps
    lanes, open_end = [], False
    for L in range(nl):
        base, valid = L * 64, min(64, n - L * 64)
        s_in = lane_pre[L][tin[L // lanes_per_tile]]
        s, nib = s_in, []
        for j in range(valid):
            w = 0 if sb[base + j] else tok_class(d[base + j])[2]
            s = (w >> 4 * s) & 0xF
            nib.append(s)
        vm = (1 << valid) - 1
        in_tok = sum((1 << j) for j, x in enumerate(nib) if x >= 3)
        gap0 = sum((1 << j) for j, x in enumerate(nib) if x == 0)
        cq = sum((1 << j) for j, x in enumerate(nib) if x == 2)
        ine = 1 if s_in >= 3 else 0
        prev_i = ((in_tok << 1) | ine) & ((1 << 64) - 1)
        prev_z = ((gap0 << 1) | (1 if s_in == 0 else 0)) & ((1 << 64) - 1)
        start = in_tok & ~prev_i & vm
        first = start & prev_z
        end = ((prev_i & ~in_tok & ~cq) | cq) & vm
        if valid and base + valid == n and s >= 3:
            open_end = True
        sbm = sum(1 << j for j in range(valid) if sb[base + j])
        h = ID
        for j in range(64):
            if (sbm | first) >> j & 1:
                h = comp(hook_symbol(d[base + j], toml) if sbm >> j & 1 else HOOK_FIRST, h)
        lanes.append(dict(start=start, first=first, end=end, endq=cq & vm, ine=ine, h=h, sbm=sbm,
                          cnt=bin(start).count("1")))
    # block scans and chains: lane hook prefixes, tile hook states, token offsets
    hin, lane_h, lane_off, s, off = [], [], [], 0, 0
    for t in range(ntile):
        hin.append(s)
        pre, cnt = ID, 0
        for L in range(t * lanes_per_tile, min(nl, (t + 1) * lanes_per_tile)):
            lane_h.append(pre[s])
            lane_off.append(off + cnt)
            pre = comp(lanes[L]["h"], pre)
            cnt += lanes[L]["cnt"]
        s, off = pre[s], off + cnt
    ntok = off

    # pass 4: the table
    role, beg, end = [None] * ntok, [None] * ntok, [None] * ntok
    for L, ln in enumerate(lanes):
        base, h, k = L * 64, lane_h[L], lane_off[L]
        for j in range(64):
            if ln["sbm"] >> j & 1:
                h = hook_symbol(d[base + j], toml)[h]
            elif ln["start"] >> j & 1:
                role[k], beg[k] = hook_role(h, toml), base + j
                k += 1
                if ln["first"] >> j & 1:
                    h = HOOK_FIRST[h]
        ke = lane_off[L] - ln["ine"]
        for j in range(64):
            if ln["end"] >> j & 1:
                end[ke] = base + j + (ln["endq"] >> j & 1)
                ke += 1
    if open_end:
        end[ntok - 1] = n
    assert None not in role and None not in end
    key, last = [-1] * ntok, -1
    for i in range(ntok):  # TOML: the max-scan of KEY indices
        if toml and role[i] == 2:
            key[i] = last
        if role[i] == 1:
            last = i
    return [(role[i], beg[i], end[i], key[i]) for i in range(ntok)]


def check(fmt: str, d: bytes, lanes_per_tile: int):
    pos = O.scan(Plan.load(fmt).blob, np.frombuffer(d, np.uint8))["positions"]
    got = [t for t in model(fmt, d, pos, lanes_per_tile) if t[0] != 4]
    exp = K.tokens(fmt, d, pos)
    if fmt == "ini":
        exp = [(r, b, e, -1) for r, b, e, _ in exp]
    assert got == exp


CASES = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "kv_golden.json")))["cases"]


@pytest.mark.parametrize("case", [c["name"] for c in CASES if "format_65536" not in c["name"]])
def test_model_on_the_goldens(case):
    c = next(c for c in CASES if c["name"] == case)
    check(c["fmt"], inputs(c["input"]), 2)


@pytest.mark.parametrize("fmt", ["ini", "toml"])
@pytest.mark.parametrize("seed", range(4))
def test_model_on_tokens_across_lanes_and_tiles(fmt, seed):
    rng = np.random.default_rng(seed)
    parts = [b"k = \"" + b"a b \\\" c" * int(rng.integers(5, 40)) + b"\"\n", b"[s t]\n", b"x = 'q w'z\n",
             b"   " * int(rng.integers(1, 50)), b"a = [1, 2, 'x y']\n", b"plain" * int(rng.integers(1, 30)),
             b"= v w\n", b"\"open " * 3, b"#c = d\n", b";c = d\n", b"\\\\ '\\' \n"]
    d = b"".join(parts[int(i)] for i in rng.integers(0, len(parts), 60))
    check(fmt, d, 1 + seed)


def test_nibble_table_equals_the_maps():
    for t in (WS, DQ, SQ, BS, OT):
        assert unbytes(t[0], t[1]) == [(t[2] >> 4 * s) & 0xF for s in range(8)]
