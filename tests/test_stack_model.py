"""CPU model of the stack pairing path (pars_amd/csrc/pair_stack.hip), index
for index: threads of T brackets with a bit-mask stack, chunks of W threads
whose unmatched closes find the last earlier thread with a small enough
minimum depth by walking each thread's previous-smaller chain (found once per
thread by binary lifting over a sparse table), chunk residues laid out
as the kernels lay them out (closes from the entry level down, then opens from
the minimum up) and resolved across chunks by walking back to the nearest
chunk whose minimum is below the highest open level, skipping groups of G and
G*G chunks by their minima (G = 64 in the kernels; smaller here, so that
every level of the walk runs).  Checked against the oracle's pair_brackets (nest.hpp:45-92) on
balanced documents, mismatched bracket bytes included, with small T and W so
that every level - thread, chunk, document - carries work; every residue index
is asserted in range as the kernels' writes require."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from pars_amd import Plan

JSON = Plan.load("json").blob
OPEN = {ord("["): ord("]"), ord("{"): ord("}")}


def sparse(m: list[int], levels: int) -> list[list[int]]:
    """`levels` levels, as the kernels build them: log2(threads) in a chunk
    (10 for 1024 threads), stack_levels(chunks) across chunks."""
    st = [list(m)]
    for j in range(1, levels):
        h, prev = 1 << (j - 1), st[-1]
        st.append([min(prev[c], prev[c - h]) if c >= h else prev[c] for c in range(len(m))])
    return st


def stack_levels(chunks: int) -> int:
    """pair_stack.hip stack_levels: the smallest L with 2^(L-1) >= chunks."""
    L = 1
    while (1 << (L - 1)) < chunks:
        L += 1
    return L


def last_at_most(st, t: int, k: int) -> int:
    pos = t - 1
    for j in range(len(st) - 1, -1, -1):
        if pos - (1 << j) + 1 >= 0 and st[j][pos] > k:
            pos -= 1 << j
    return pos


def range_min(st, lo: int, hi: int) -> int:
    j = min((hi - lo + 1).bit_length() - 1, len(st) - 1)
    assert (1 << j) * 2 >= hi - lo + 1, "two blocks must cover the range"
    return min(st[j][lo + (1 << j) - 1], st[j][hi])


def stack_pairs(chars: np.ndarray, T: int, W: int, G: int = 64):
    nb = chars.size
    pair = np.full(nb, -2, np.int64)  # -2: never written
    mismatched = False
    CH = T * W
    nch = (nb + CH - 1) // CH
    res = np.full(nb, -1, np.int64)
    crec = []
    for c in range(nch):
        base = c * CH
        th = []
        for t in range(W):
            i0 = base + t * T
            cnt = max(0, min(nb - i0, T))
            mask, stack, rcl, rol = 0, [], [], []
            for q in range(cnt):
                ch = int(chars[i0 + q])
                if ch in OPEN:
                    mask |= 1 << q
                    stack.append((q, OPEN[ch]))
                elif mask:
                    p, exp = stack.pop()
                    assert p == mask.bit_length() - 1  # the highest set bit is the top of the stack
                    mask ^= 1 << p
                    ok = exp == ch
                    mismatched |= not ok
                    pair[i0 + p] = i0 + q if ok else -1
                    pair[i0 + q] = i0 + p if ok else -1
                else:
                    rcl.append(q)
            rol = [q for q in range(T) if (mask >> q) & 1]
            th.append((i0, cnt, rcl, rol))
        deltas = [len(ro) - len(rc) for _, _, rc, ro in th]
        e = np.concatenate([[0], np.cumsum(deltas)[:-1]]).astype(int)
        xc = int(sum(deltas))
        NOMIN = 1 << 30
        m = [int(e[t]) - len(th[t][2]) if th[t][1] else NOMIN for t in range(W)]
        st = sparse(m, max(1, (W - 1).bit_length()))
        mc = range_min(st, 0, W - 1)
        assert mc <= 0
        first = int(chars[base])
        entry = int(depth_before[base])
        # each thread's previous smaller minimum (one lifting per thread); a
        # thread's closes descend in level and walk that chain back
        psm = [last_at_most(st, t, m[t] - 1) for t in range(W)]
        for t in range(W):
            i0, cnt, rcl, rol = th[t]
            u = t - 1
            for j, q in enumerate(rcl):
                k = int(e[t]) - 1 - j
                while u >= 0 and m[u] > k:
                    u = psm[u]
                assert u == last_at_most(st, t, k)
                gc = i0 + q
                if u >= 0:
                    jo = k - m[u]
                    assert 0 <= jo < len(th[u][3])
                    go = base + u * T + th[u][3][jo]
                    ok = OPEN[int(chars[go])] == int(chars[gc])
                    mismatched |= not ok
                    pair[gc] = go if ok else -1
                    pair[go] = gc if ok else -1
                else:
                    idx = -1 - k
                    assert 0 <= idx < -mc and base + idx < nb
                    res[base + idx] = gc
            suf = range_min(st, t + 1, W - 1) if t + 1 < W else NOMIN
            for j, q in enumerate(rol):
                k = m[t] + j
                if k < suf:
                    idx = -mc + (k - mc)
                    assert -mc <= idx < -mc + (xc - mc) and base + idx < nb
                    res[base + idx] = i0 + q
        crec.append((entry, entry + mc, entry + xc))
        del first
    M0 = [r[1] for r in crec]
    M1 = [min(M0[i:i + G]) for i in range(0, nch, G)]
    M2 = [min(M0[i:i + G * G]) for i in range(0, nch, G * G)]

    def scan(arr, hi, thr):  # highest index <= hi in hi's G-aligned window with arr < thr
        lo = hi - hi % G
        for i in range(hi, lo - 1, -1):
            if arr[i] < thr:
                return i
        return -1

    def prev_below(cur, thr):
        steps = 0
        while cur >= 0:
            steps += 1
            p = scan(M0, cur, thr)
            if p >= 0:
                return p
            g = cur // G - 1
            if g < 0:
                return -1
            Gi = scan(M1, g, thr)
            if Gi < 0:
                h, H = g // G - 1, -1
                while h >= 0:
                    H = scan(M2, h, thr)
                    if H >= 0:
                        break
                    h = h - h % G - 1
                if H < 0:
                    return -1
                Gi = scan(M1, H * G + G - 1, thr)
                assert Gi >= 0
            p = scan(M0, Gi * G + G - 1, thr)
            assert p >= 0
            return p
        return -1

    for c in range(nch):
        E, M, X = crec[c]
        rm, cur = E, c - 1
        while rm > M:
            p = prev_below(cur, rm)
            assert p >= 0, "a residue close without an open on a balanced document"
            Ep, Mp, Xp = crec[p]
            lo = max(Mp, M)
            for k in range(lo, rm):
                assert k < Xp
                gc = int(res[c * CH + (E - 1 - k)])
                go = int(res[p * CH + (Ep - Mp) + (k - Mp)])
                ok = OPEN[int(chars[go])] == int(chars[gc])
                mismatched |= not ok
                pair[gc] = go if ok else -1
                pair[go] = gc if ok else -1
            rm, cur = lo, p - 1
    return pair, mismatched


def balanced_brackets(rng, n_pairs: int, max_depth: int, p_swap: float) -> np.ndarray:
    out, stack = [], []
    opened = 0
    while opened < n_pairs or stack:
        if opened < n_pairs and len(stack) < max_depth and (not stack or rng.random() < 0.5):
            o = rng.choice([ord("["), ord("{")])
            out.append(o)
            stack.append(o)
            opened += 1
        else:
            o = stack.pop()
            cl = OPEN[o]
            if rng.random() < p_swap:  # the other closing byte: a mismatched pair
                cl = ord("}") if cl == ord("]") else ord("]")
            out.append(cl)
    return np.array(out, np.uint8)


depth_before: np.ndarray = np.zeros(0, np.int64)


@pytest.mark.parametrize("T,W,G", [(32, 64, 64), (32, 256, 64), (4, 8, 2), (2, 4, 3), (16, 2, 2), (1, 256, 4), (32, 4, 64)])
@pytest.mark.parametrize("seed", range(6))
def test_stack_model_equals_depth_sort(T, W, G, seed):
    global depth_before
    rng = np.random.default_rng(seed)
    for trial in range(4):
        n_pairs = int(rng.integers(1, 3000))
        depth = int(rng.integers(1, 80)) if trial % 2 else int(rng.integers(1, 6))
        chars = balanced_brackets(rng, n_pairs, depth, 0.02 if trial == 3 else 0.0)
        step = np.where(np.isin(chars, list(OPEN)), 1, -1)
        depth_before = np.concatenate([[0], np.cumsum(step)[:-1]])
        exp = O.pair(JSON, chars)
        assert exp["balanced"]
        got, mism = stack_pairs(chars, T, W, G)
        assert (got != -2).all(), "some bracket was never written"
        assert np.array_equal(got, exp["pair"]), (T, W, seed, trial)
        assert (not mism) == exp["well_formed"]
