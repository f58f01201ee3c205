"""GPU parity of the INI / TOML key-value tables (pars_hip_kv_parse) with the
reference's stage-2 hooks (bench/bench_parse_host.cu: ini_parse :69-134,
toml_parse :202-248) as restated in oracle/kv_hooks.py, itself pinned against
the reference's own hooks (tests/test_kv_ref.py):
  * every golden case: the table's digest (hash and lists) equals the
    reference hook's, and every token (role, begin, end, key) equals the
    restatement's;
  * generated documents across the kernels' 16 KiB tiles and 64-byte lanes,
    misaligned inputs, gaps and quoted tokens that span many tiles, a token
    open at the document's end, empty and one-byte documents;
  * a plan whose structural bytes include ';', so the INI hook enters its
    comment state (tokens it does not see are SKIPPED in the table);
  * at 64 MiB, the reference benchmark's own documents: the table repeats the
    record's tokens exactly, record after record.
"""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

from oracle import kv_hooks as K
from oracle import oracle as O
from pars_amd import Plan, Scanner
from pars_amd import workloads as W
from tests.golden.make_kv_golden import SOUP, inputs
from tests.gpu_util import d2h, say, to_device

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "kv_golden.json")
_sc: dict = {}


def scanner(plan: str) -> Scanner:
    if plan not in _sc:
        _sc[plan] = Scanner(plan)
    return _sc[plan]


def gpu_table(torch, fmt: str, d: bytes, plan: str | None = None, offset: int = 0) -> dict:
    arr = np.frombuffer(d, np.uint8)
    t, ptr = to_device(torch, arr, offset)
    r = scanner(plan or fmt).kv_parse(ptr, arr.size, fmt, torch.cuda.current_stream().cuda_stream)
    n = r.n_tokens
    out = {"n": n, "n_role": r.n_role, "role": d2h(r.role, n, np.uint8), "begin": d2h(r.begin, n, np.uint32),
           "end": d2h(r.end, n, np.uint32), "n_struct": r.n_struct}
    out["key"] = d2h(r.key, n, np.int32) if fmt == "toml" else None
    if fmt == "ini":
        assert not r.key
    del t
    return out


def rows(fmt: str, g: dict) -> list[tuple[int, int, int, int]]:
    """The table as the restatement's token list (SKIPPED rows dropped)."""
    keep = [i for i in range(g["n"]) if g["role"][i] != 4]
    at = {i: k for k, i in enumerate(keep)}
    out = []
    for i in keep:
        k = -1
        if fmt == "toml" and g["key"][i] >= 0:
            k = at[int(g["key"][i])]
        out.append((int(g["role"][i]), int(g["begin"][i]), int(g["end"][i]), k))
    return out


def check(torch, fmt: str, d: bytes, tag: str = "", offset: int = 0, plan: str | None = None):
    pos = O.scan(Plan.load(plan or fmt).blob, np.frombuffer(d, np.uint8))["positions"]
    exp = K.tokens(fmt, d, pos)
    g = gpu_table(torch, fmt, d, plan, offset)
    tag = tag or repr(d[:60])
    assert g["n_struct"] == len(pos), tag
    assert g["n_role"] == [int((g["role"] == r).sum()) for r in range(5)], tag
    got = rows(fmt, g)
    if fmt == "ini":
        exp = [(r, b, e, -1) for r, b, e, _ in exp]
    assert len(got) == len(exp), tag
    assert got == exp, tag
    if fmt == "toml":  # non-values carry no key
        assert all(g["key"][i] == -1 for i in range(g["n"]) if g["role"][i] != 2), tag
    return g


CASES = json.load(open(GOLDEN))["cases"]


def test_golden_cases_equal_the_reference_hooks(torch_cuda):
    for c in CASES:
        d = inputs(c["input"])
        say(f"[gpu] kv golden {c['name']} n={len(d)}")
        g = check(torch_cuda, c["fmt"], d, c["name"])
        toks = [(int(g["role"][i]), int(g["begin"][i]), int(g["end"][i]),
                 int(g["key"][i]) if c["fmt"] == "toml" else -1) for i in range(g["n"])]
        assert K.digest(c["fmt"], d, toks) == c["ref"], c["name"]


@pytest.mark.parametrize("fmt", ["ini", "toml"])
@pytest.mark.parametrize("n", [1, 63, 64, 65, 16383, 16384, 16385, 3 * 16384 + 7, 1 << 20, (3 << 20) + 11])
def test_reference_documents_across_tiles(torch_cuda, fmt, n):
    d = W.format_doc(fmt, n).tobytes()[:n]
    say(f"[gpu] kv {fmt} format doc n={n}")
    check(torch_cuda, fmt, d)


@pytest.mark.parametrize("fmt", ["ini", "toml"])
@pytest.mark.parametrize("seed", range(6))
def test_byte_soups(torch_cuda, fmt, seed):
    alphabet = SOUP[fmt] if seed % 2 else W.ALPHABETS[fmt]
    d = W.alpha(900 + seed, 40_000 + 9_001 * seed, alphabet).tobytes()
    say(f"[gpu] kv {fmt} soup seed={seed} n={len(d)}")
    check(torch_cuda, fmt, d, offset=seed % 4)


@pytest.mark.parametrize("fmt", ["ini", "toml"])
def test_misaligned_inputs(torch_cuda, fmt):
    d = W.format_doc(fmt, 200_000).tobytes()
    for off in (1, 3, 7, 15):
        say(f"[gpu] kv {fmt} offset={off}")
        check(torch_cuda, fmt, d, offset=off)


def test_gaps_and_tokens_spanning_many_tiles(torch_cuda):
    big = b"x" * 70_000
    docs = {
        # a TOML string masked by the scanner: one gap of ~200 KB, one quoted token
        "toml": [b"k = \"" + b"a b \\\" " * 30_000 + b"\"\nz = 1\n",
                 b"a = [" + b" 1," * 60_000 + b"]\n[t]\nk = v",
                 b"k = 'open quote to the end " + big],
        # INI: a plain token, and a quoted one, longer than a tile; no newline at the end
        "ini": [b"[s]\nk = " + big + b"\nx = y", b"k = \"" + big + b" " + big, b"  " + big + b"   ", big],
    }
    for fmt, ds in docs.items():
        for d in ds:
            say(f"[gpu] kv {fmt} long gap n={len(d)}")
            check(torch_cuda, fmt, d)
            check(torch_cuda, fmt, d, offset=5)


def test_empty_and_tiny_documents(torch_cuda):
    for fmt in ("ini", "toml"):
        sc = scanner(fmt)
        r = sc.kv_parse(0, 0, fmt)
        assert r.n_tokens == 0 and r.n_role == [0] * 5
        for d in (b" ", b"\n", b"a", b"\"", b"=", b"[", b"a\n", b"\"\\"):
            say(f"[gpu] kv {fmt} tiny {d!r}")
            check(torch_cuda, fmt, d)


def test_comment_symbols_skip_tokens(torch_cuda):
    """With the clike plan ';' is a structural byte, so the INI hook enters its
    comment state at it: those tokens are SKIPPED, the rest equal the hook's."""
    d = b"[s] a = b ; c d = e\nk = v; x\n" * 500 + W.alpha(5, 30_000, b"[]=;\n abc").tobytes()
    say("[gpu] kv ini under the clike plan")
    g = check(torch_cuda, "ini", d, plan="clike")
    assert g["n_role"][4] > 500


def test_reference_documents_at_64_mib_repeat_the_record(torch_cuda):
    """The reference benchmark's documents (bench_parse_host.cu:305-312) are one
    record repeated: the table is the record's table, shifted per record (the
    last record is whole, as repeat_rec appends whole records)."""
    for fmt in ("ini", "toml"):
        rec = W.FORMAT_RECORDS[fmt]
        d = W.format_doc(fmt, 64 << 20)
        assert d.size % len(rec) == 0
        reps = d.size // len(rec)
        one = K.tokens(fmt, rec, O.scan(Plan.load(fmt).blob, np.frombuffer(rec, np.uint8))["positions"])
        say(f"[gpu] kv {fmt} 64 MiB reps={reps} tokens/record={len(one)}")
        g = gpu_table(torch_cuda, fmt, d.tobytes())
        assert g["n"] == reps * len(one)
        shift = (np.arange(reps, dtype=np.int64) * len(rec))[:, None]
        role = np.array([t[0] for t in one], np.uint8)
        beg = np.array([t[1] for t in one], np.int64)
        end = np.array([t[2] for t in one], np.int64)
        assert np.array_equal(g["role"].reshape(reps, -1), np.broadcast_to(role, (reps, len(one))))
        assert np.array_equal(g["begin"].reshape(reps, -1).astype(np.int64), beg + shift)
        assert np.array_equal(g["end"].reshape(reps, -1).astype(np.int64), end + shift)
        if fmt == "toml":
            key = np.array([t[3] for t in one], np.int64)
            want = np.where(key >= 0, key + np.arange(reps, dtype=np.int64)[:, None] * len(one), -1)
            assert np.array_equal(g["key"].reshape(reps, -1).astype(np.int64), want)


@pytest.mark.parametrize("fmt", ["ini", "toml"])
def test_dense_and_sparse_tiles_mixed(torch_cuda, fmt):
    """Tiles with few token starts write their rows through LDS, dense ones
    (more than 3072 starts in 16 KiB) from each lane: a document alternating
    the two, with tokens running across the tiles between them."""
    dense = (b"a=b c\n" if fmt == "ini" else b"a = [b, c]\n") * 6000
    sparse = W.format_doc(fmt, 40_000).tobytes()
    d = sparse + b"k = " + dense + b" long" * 5000 + sparse[:33_333] + dense + b"tail"
    say(f"[gpu] kv {fmt} dense/sparse n={len(d)}")
    check(torch_cuda, fmt, d)
    check(torch_cuda, fmt, d, offset=9)
