"""GPU parity of R6's stack path (pair_stack.hip) against the oracle's
pair_brackets (nest.hpp:45-92), and its hand-over to the sorting path.

Balanced documents take the stack: bracket soups with mismatched closing
bytes, depths from 1 to 20,000, sizes around the thread (16 brackets) and chunk
(16,384 brackets) boundaries, a 2^20-deep nest whose every chunk is all
residue, text between the brackets, and pair_pos.  Unbalanced documents (a
close below depth 0, an open never closed) must come out of the sorting path
identically.  The chunked and streamed boundaries pair through the same path,
and forcing the sort (a radix digit cap) gives the same outputs.
"""
from __future__ import annotations

import numpy as np
import pytest

from pars_amd import Plan, Scanner
from pars_amd import _native as N
from tests.gpu_util import gpu_full, say
from tests.test_gpu_parity import compare

pytestmark = pytest.mark.gpu

OPEN = {ord("["): ord("]"), ord("{"): ord("}")}


def soup(seed: int, n_pairs: int, max_depth: int, p_swap: float, p_text: float = 0.0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    out, stack, opened = [], [], 0
    while opened < n_pairs or stack:
        if p_text and rng.random() < p_text:
            out.append(int(rng.choice(list(b'1 ,:"x\\'))))
            if out[-1] == ord('"'):
                out.append(ord("a"))
                out.append(ord('"'))
            continue
        if opened < n_pairs and len(stack) < max_depth and (not stack or rng.random() < 0.5):
            o = int(rng.choice([ord("["), ord("{")]))
            out.append(o)
            stack.append(o)
            opened += 1
        else:
            cl = OPEN[stack.pop()]
            if rng.random() < p_swap:
                cl = ord("}") if cl == ord("]") else ord("]")
            out.append(cl)
    return np.array(out, np.uint8)


def scanner() -> Scanner:
    sc = Scanner(Plan.load("json"))
    sc.set_option(N.PARS_OPT_WANT_CHARS, 1)
    return sc


@pytest.mark.parametrize("n_pairs", [1, 8, 9, 8192, 8193, 16384, 40_000, 300_000])
@pytest.mark.parametrize("depth", [1, 7, 300, 20_000])
def test_balanced_soups(torch_cuda, n_pairs, depth):
    sc = scanner()
    for p_swap in (0.0, 0.001):
        d = soup(n_pairs * 31 + depth, n_pairs, depth, p_swap)
        say(f"[gpu] stack soup pairs={n_pairs} depth={depth} swap={p_swap} ({d.size} B)")
        compare("json", d, gpu_full(torch_cuda, sc, d, want_pairs=True), f"soup {n_pairs}/{depth}/{p_swap}")
    sc.close()


def test_balanced_with_text_and_offsets(torch_cuda):
    sc = scanner()
    for seed in range(4):
        d = soup(900 + seed, 120_000, 64, 0.0005, p_text=0.4)
        say(f"[gpu] stack soup with text seed={seed} ({d.size} B)")
        compare("json", d, gpu_full(torch_cuda, sc, d, offset=seed * 5, want_pairs=True), f"text {seed}")
    sc.close()


def test_deep_nest_all_residue(torch_cuda):
    sc = scanner()
    depth = 1 << 20
    d = np.frombuffer(b"[" * depth + b"{}" * 3 + b"]" * depth, np.uint8).copy()
    say(f"[gpu] stack deep nest {d.size} B")
    got = gpu_full(torch_cuda, sc, d, want_pairs=True)
    compare("json", d, got, "deep nest")
    assert got["max_depth"] == depth + 1 and got["well_formed"]
    assert sc.get_option(N.PARS_OPT_LAST_PAIRING) == N.PARS_PAIRING_STACK
    sc.close()


@pytest.mark.parametrize("kind", ["extra_close", "unclosed", "close_first", "swap_at_zero"])
def test_unbalanced_documents_take_the_sort(torch_cuda, kind):
    sc = scanner()
    base = soup(77, 50_000, 40, 0.0)
    if kind == "extra_close":
        d = np.concatenate([base[:40_000], np.frombuffer(b"]]", np.uint8), base[40_000:]])
    elif kind == "unclosed":
        d = np.concatenate([base, np.frombuffer(b"[[{", np.uint8)])
    elif kind == "close_first":
        d = np.concatenate([np.frombuffer(b"}", np.uint8), base, np.frombuffer(b"[", np.uint8)])
    else:
        d = np.concatenate([base, np.frombuffer(b"][", np.uint8), base])
    say(f"[gpu] stack fallback {kind} ({d.size} B)")
    got = gpu_full(torch_cuda, sc, d, want_pairs=True)
    assert not got["balanced"]
    compare("json", d, got, kind)
    assert sc.get_option(N.PARS_OPT_LAST_PAIRING) == N.PARS_PAIRING_SORT
    sc.close()


def test_sorting_path_gives_the_same_pairs(torch_cuda):
    d = soup(5, 200_000, 500, 0.0002)
    a, b = scanner(), scanner()
    b.set_option(N.PARS_OPT_RADIX_BITS, 8)  # a digit cap always takes the sort
    ga = gpu_full(torch_cuda, a, d, want_pairs=True)
    gb = gpu_full(torch_cuda, b, d, want_pairs=True)
    for k in ("pair", "pair_pos"):
        assert np.array_equal(ga[k], gb[k]), k
    assert (ga["well_formed"], ga["max_depth"]) == (gb["well_formed"], gb["max_depth"])
    compare("json", d, ga, "stack vs sort")
    a.close()
    b.close()


def test_stream_and_chunks_pair_through_the_stack(torch_cuda):
    from tests.gpu_util import d2h

    torch = torch_cuda
    d = soup(11, 400_000, 90, 0.0003, p_text=0.2)
    sc = scanner()
    st = N.StageTimes()
    r = sc.scan_stream(d, chunk_bytes=64 << 10, copy_threads=2, want_pairs=True, stage_times=st)
    got = {"n_struct": r.n_struct, "n_brackets": r.n_brackets,
           "positions": d2h(r.structural, r.n_struct, np.uint32),
           "bracket_pos": d2h(r.bracket_pos, r.n_brackets, np.uint32),
           "bracket_char": d2h(r.bracket_char, r.n_brackets, np.uint8),
           "balanced": r.balanced, "well_formed": r.well_formed, "max_depth": r.max_depth,
           "unclosed_mode": r.unclosed_mode, "pair": d2h(r.bracket_pair, r.n_brackets, np.int32),
           "pair_pos": d2h(r.pair_pos, r.n_struct, np.int32)}
    compare("json", d, got, "stream")
    # chunks of one document, then R6 over all of them
    t = torch.from_numpy(d).to("cuda")
    torch.cuda.synchronize()
    sc2 = Scanner(Plan.load("json"))
    sc2.reserve_struct(d.size)
    carry = N.ChunkCarry()
    cut = [0, 12_345, 400_001, d.size]
    for a0, b0 in zip(cut[:-1], cut[1:]):
        carry = sc2.scan_chunk(t.data_ptr() + a0, b0 - a0, carry, byte_base=a0)
    rr = sc2.pair_chunks(carry, want_pairs=True)
    got2 = dict(got)
    got2.update({"pair": d2h(rr.bracket_pair, rr.n_brackets, np.int32),
                 "pair_pos": d2h(rr.pair_pos, rr.n_struct, np.int32),
                 "balanced": rr.balanced, "well_formed": rr.well_formed, "max_depth": rr.max_depth})
    compare("json", d, got2, "chunks")
    sc.close()
    sc2.close()


def test_sort_after_sparse_keys(torch_cuda):
    """A scan meant for the stack path leaves keys only at chunk starts; every
    sort of such a document (an unbalanced one falling back, chunks of one,
    a digit cap set between the chunks' scans and their pairing) rebuilds the
    other keys first and gives the oracle's pairs."""
    from tests.gpu_util import d2h

    torch = torch_cuda
    base = soup(123, 90_000, 300, 0.0005, p_text=0.1)
    cases = {
        "extra_close": np.concatenate([base[:70_001], np.frombuffer(b"]}]", np.uint8), base[70_001:]]),
        "close_first": np.concatenate([np.frombuffer(b"]", np.uint8), base]),
        "unclosed": np.concatenate([base, np.frombuffer(b"{[[", np.uint8)]),
    }
    for kind, d in cases.items():
        sc = scanner()
        say(f"[gpu] sparse keys, whole document: {kind} ({d.size} B)")
        got = gpu_full(torch, sc, d, want_pairs=True)
        assert not got["balanced"]
        compare("json", d, got, f"sparse {kind}")
        sc.close()
        for radix_late in (False, True):
            t = torch.from_numpy(d).to("cuda")
            torch.cuda.synchronize()
            sc2 = Scanner(Plan.load("json"))
            sc2.reserve_struct(d.size)
            carry = N.ChunkCarry()
            cut = [0, 33_333, 150_000, d.size]
            for a0, b0 in zip(cut[:-1], cut[1:]):
                carry = sc2.scan_chunk(t.data_ptr() + a0, b0 - a0, carry, byte_base=a0)
            if radix_late:
                sc2.set_option(N.PARS_OPT_RADIX_BITS, 6)
            say(f"[gpu] sparse keys, chunks: {kind} radix_late={radix_late}")
            rr = sc2.pair_chunks(carry, want_pairs=True)
            exp = got  # the whole-document path, checked against the oracle above
            assert np.array_equal(d2h(rr.bracket_pair, rr.n_brackets, np.int32), exp["pair"]), kind
            assert np.array_equal(d2h(rr.pair_pos, rr.n_struct, np.int32), exp["pair_pos"]), kind
            assert (rr.balanced, rr.well_formed, rr.max_depth) == (exp["balanced"], exp["well_formed"],
                                                                   exp["max_depth"])
            sc2.close()
