"""The largest documents the boundary accepts (2^32-1 bytes: positions are
uint32 byte offsets, so every offset from 2^31 up is in use):
  * a JSON document of exactly 2^32-1 bytes through scan_device with pairing,
    every position and every pair equal to the oracle's (model::scanner and
    model::pair_brackets restated in C);
  * the reference benchmark's INI and TOML records repeated to just under
    2^32 bytes through the key-value table, the table equal to the record's
    table shifted record by record.
Both print a line before every step, so a slow step names itself."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import kv_hooks as K
from oracle import oracle as O
from pars_amd import Plan, Scanner
from pars_amd import workloads as W
from tests.gpu_util import d2h, say

pytestmark = pytest.mark.gpu

MAX_N = (1 << 32) - 1


def test_json_at_the_maximum_size(torch_cuda):
    torch = torch_cuda
    say(f"[gpu] max-size JSON: generating {MAX_N} bytes")
    doc = W.json(11, MAX_N, 16, 0.30)
    assert doc.size == MAX_N
    d = torch.from_numpy(doc).to("cuda")
    torch.cuda.synchronize()
    sc = Scanner(Plan.load("json"))
    say("[gpu] max-size JSON: scan_device with pairs")
    r = sc.scan_device(d.data_ptr(), MAX_N, torch.cuda.current_stream().cuda_stream, want_pairs=False,
                       run_pairs=True)
    k, kb = r.n_struct, r.n_brackets
    pos = d2h(r.structural, k, np.uint32)
    bpos = d2h(r.bracket_pos, kb, np.uint32)
    pair = d2h(r.bracket_pair, kb, np.int32)
    del d
    sc.close()
    torch.cuda.empty_cache()
    say(f"[gpu] max-size JSON: oracle scan (k={k}, k_b={kb})")
    ref = O.scan(Plan.load("json").blob, doc, cap=k + 1)
    assert ref["positions"].size == k and pos[-1] >= (1 << 31)
    assert np.array_equal(pos, ref["positions"])
    assert np.array_equal(bpos, ref["bracket_pos"])
    say("[gpu] max-size JSON: oracle pairs")
    rp = O.pair(Plan.load("json").blob, ref["bracket_char"])
    assert np.array_equal(pair, rp["pair"])
    assert bool(r.balanced) == rp["balanced"] and r.max_depth == rp["max_depth"]


@pytest.mark.parametrize("fmt", ["ini", "toml"])
def test_key_value_table_at_the_maximum_size(torch_cuda, fmt):
    torch = torch_cuda
    rec = W.FORMAT_RECORDS[fmt]
    reps = MAX_N // len(rec)
    n = reps * len(rec)
    say(f"[gpu] max-size {fmt}: {reps} records, {n} bytes")
    doc = np.frombuffer(rec, np.uint8)
    d = torch.from_numpy(doc.copy()).to("cuda").repeat(reps)
    torch.cuda.synchronize()
    assert d.numel() == n and n > (1 << 32) - len(rec)
    one = K.tokens(fmt, rec, O.scan(Plan.load(fmt).blob, doc)["positions"])
    sc = Scanner(Plan.load(fmt))
    say(f"[gpu] max-size {fmt}: kv_parse")
    r = sc.kv_parse(d.data_ptr(), n, fmt, torch.cuda.current_stream().cuda_stream)
    assert r.n_tokens == reps * len(one)
    role = d2h(r.role, r.n_tokens, np.uint8).reshape(reps, -1)
    assert np.array_equal(role, np.broadcast_to(np.array([t[0] for t in one], np.uint8), role.shape))
    del role
    shift = np.arange(reps, dtype=np.int64)[:, None] * len(rec)
    for name, ptr, col in (("begin", r.begin, 1), ("end", r.end, 2)):
        say(f"[gpu] max-size {fmt}: {name}")
        got = d2h(ptr, r.n_tokens, np.uint32).reshape(reps, -1)
        assert np.array_equal(got, (np.array([t[col] for t in one], np.int64) + shift).astype(np.uint32))
        assert int(got[-1, -1]) >= (1 << 31)
        del got
    if fmt == "toml":
        key = d2h(r.key, r.n_tokens, np.int32).reshape(reps, -1).astype(np.int64)
        k1 = np.array([t[3] for t in one], np.int64)
        want = np.where(k1 >= 0, k1 + np.arange(reps, dtype=np.int64)[:, None] * len(one), -1)
        assert np.array_equal(key, want)
    sc.close()
    del d
    torch.cuda.empty_cache()
