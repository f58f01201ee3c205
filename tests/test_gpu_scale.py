"""GPU parity at the configured batch and serving workloads, and a hostile
batchmate.

- The 100k-document batch the benchmark times (seed 4: 70 % JSON / 30 % CSV,
  log-normal sizes, median 4 KiB, sigma 1.85, clipped to [200 B, 4 MiB]),
  run as the benchmark runs it (one captured and replayed launch chain per
  plan on its own stream), every document checked against the CPU oracle run
  on that document alone.
- The serving mix (60 % JSON / 25 % CSV / 15 % C-like, the same size law)
  through a server with the configured 4 GiB device ring, 1 GiB staging and
  64 GiB HBM budget, submitted under sustained pressure so the staging ring
  wraps; every result checked against the oracle; and a server with a small
  device ring so that ring wraps too.
- A 4 MiB document nesting to depth 2^21 in a batch of 2,048 documents and in
  a server batch: every batchmate and the deep document itself bit-exact (the
  batch pairing keys are per-document ranges, so one deep document cannot
  overflow the sort key of the others).
"""
from __future__ import annotations

import time

import numpy as np
import pytest

from oracle import oracle as O
from pars_amd import Plan, Scanner, Server, pack_batch
from pars_amd import _native as N
from pars_amd import workloads as W
from tests.gpu_util import d2h, say

pytestmark = pytest.mark.gpu

PLANS = ["json", "csv", "clike"]


def fetch_batch(r, nd, pairs=True):
    out = {
        "positions": d2h(r.structural, r.n_struct, np.uint32),
        "struct_start": d2h(r.struct_start, nd + 1, np.uint32),
        "bracket_pos": d2h(r.bracket_pos, r.n_brackets, np.uint32),
        "bracket_char": d2h(r.bracket_char, r.n_brackets, np.uint8),
        "bracket_start": d2h(r.bracket_start, nd + 1, np.uint32),
        "flags": d2h(r.doc_flags, nd, np.uint8),
        "max_depth": d2h(r.max_depth, nd, np.int32),
    }
    if pairs:
        out["pair"] = d2h(r.bracket_pair, r.n_brackets, np.int32)
    return out


def compare_batch(got, exp, tag):
    for k in ("struct_start", "bracket_start"):
        assert np.array_equal(got[k].astype(np.uint64), exp[k]), f"{tag}: {k}"
    for k in ("positions", "bracket_pos", "bracket_char", "pair", "flags", "max_depth"):
        if not np.array_equal(got[k], exp[k]):
            bad = int(np.nonzero(got[k] != exp[k])[0][0]) if got[k].size == exp[k].size else -1
            raise AssertionError(f"{tag}: {k} differs (first index {bad}, sizes {got[k].size} / {exp[k].size})")


def test_c4_batch_100k_documents(torch_cuda):
    torch = torch_cuda
    t0 = time.perf_counter()
    b = W.batch(4, 100_000, {"json": 70, "csv": 30}, median=4096.0, sigma=1.85)
    say(f"c4: {int(b['sizes'].sum())} bytes generated in {time.perf_counter() - t0:.1f} s")
    stream = torch.cuda.Stream()
    for kind in ("json", "csv"):
        buf, spans, idx = b[kind]
        d = torch.from_numpy(buf).to("cuda")
        torch.cuda.synchronize()
        sc = Scanner(kind)
        sc.set_option(N.PARS_OPT_GRAPH, 1)
        first = None
        for call in range(3):  # eager, captured, replayed
            say(f"c4 {kind}: {len(idx)} documents, call {call}")
            r = sc.scan_batch(d.data_ptr(), d.numel(), spans, stream.cuda_stream, want_pairs=False, run_pairs=True)
            got = fetch_batch(r, len(idx))
            if first is None:
                t0 = time.perf_counter()
                exp = O.batch_full(Plan.load(kind).blob, buf, spans)
                say(f"c4 {kind}: oracle in {time.perf_counter() - t0:.1f} s, {exp['positions'].size} structural")
                first = got
            compare_batch(got, exp, f"c4 {kind} call {call}")
        sc.close()
        del d
        torch.cuda.empty_cache()


def deep_doc(depth: int) -> np.ndarray:
    return np.frombuffer(b"[" * depth + b"]" * depth, dtype=np.uint8).copy()


def test_deep_document_among_2048_batchmates(torch_cuda):
    torch = torch_cuda
    rng = np.random.default_rng(17)
    sz = W.sizes(17, 2047, median=4096.0, sigma=1.85)
    docs = [W.make_doc("json", W.doc_seed(17, i), int(sz[i])) for i in range(2047)]
    deep = deep_doc(1 << 21)  # 4 MiB, depth 2^21
    at = int(rng.integers(0, 2048))
    docs.insert(at, deep)
    buf, spans = pack_batch(docs)
    say(f"deep: {len(docs)} documents, the deep one at {at}, {buf.size} bytes")
    d = torch.from_numpy(buf).to("cuda")
    torch.cuda.synchronize()
    sc = Scanner("json")
    r = sc.scan_batch(d.data_ptr(), d.numel(), spans, torch.cuda.current_stream().cuda_stream, want_pairs=False,
                      run_pairs=True)
    got = fetch_batch(r, len(docs))
    exp = O.batch_full(Plan.load("json").blob, buf, spans)
    compare_batch(got, exp, "deep batch")
    assert int(got["max_depth"][at]) == 1 << 21 and got["flags"][at] & 2
    sc.close()


def request_stream(seed: int, count: int):
    kinds = W.doc_kinds(seed, count, {"json": 60, "csv": 25, "clike": 15})
    sz = W.sizes(seed, count, median=4096.0, sigma=1.85)
    return [(PLANS.index(k), W.make_doc(k, W.doc_seed(seed, i), int(sz[i]))) for i, k in enumerate(kinds)]


def fetch_request(r):
    return (d2h(r.structural, r.n_struct, np.uint32), d2h(r.bracket_pos, r.n_brackets, np.uint32),
            d2h(r.bracket_char, r.n_brackets, np.uint8), d2h(r.bracket_pair, r.n_brackets, np.int32),
            int(r.flags), int(r.max_depth))


def expected_request(plan: str, data):
    e = O.full(Plan.load(plan).blob, data)
    f = (1 if e["balanced"] else 0) | (2 if e["well_formed"] else 0) | (4 if e["unclosed_mode"] else 0)
    return (e["positions"], e["bracket_pos"], e["bracket_char"], e["pair"], f, e["max_depth"])


def same(a, b) -> bool:
    return all(np.array_equal(x, y) if isinstance(x, np.ndarray) else x == y for x, y in zip(a, b))


def pressure_serve(sv: Server, docs, check, tag: str, budget_s: float = 90.0):
    """Submit until the staging ring refuses, fetch and release what completed,
    repeat: the rings stay occupied the whole time.  check(i, result) runs on
    every completed request before it is released."""
    pending, i, done, t0, last = {}, 0, 0, time.monotonic(), time.monotonic()
    multi = 0
    while done < len(docs):
        assert time.monotonic() - t0 < budget_s, f"{tag}: {done}/{len(docs)} done after {budget_s} s"
        assert time.monotonic() - last < 30, f"{tag}: no progress for 30 s"
        while i < len(docs):
            rid = sv.submit(docs[i][0], docs[i][1])
            if rid is None:
                break
            pending[rid] = i
            i += 1
        for r in sv.poll(cap=1024, wait_us=1000):
            assert r.status == 0, f"{tag}: request {pending.get(r.id)} failed with status {r.status}"
            k = pending.pop(r.id)
            check(k, r)
            multi += r.batch_docs > 1
            sv.release(r.id)
            done += 1
            last = time.monotonic()
    return multi


def test_c5_serving_configured_rings(torch_cuda):
    """4 GiB device ring, 1 GiB staging, 64 GiB budget, 64 MiB batches: 110k
    requests of the serving mix (~2.5 GB; the JSON share passes its 1.43 GB ring
    region) under sustained pressure, every result exact, and each request's
    input readable on the device beside its results."""
    docs = request_stream(5, 110_000)
    t0 = time.perf_counter()
    exp = [expected_request(PLANS[p], d) for p, d in docs]
    say(f"c5: {len(docs)} requests, {sum(d.size for _, d in docs)} bytes, oracle in {time.perf_counter() - t0:.1f} s")
    sv = Server(PLANS)  # the configured sizes are the defaults

    def check(k, r):
        assert same(fetch_request(r), exp[k]), f"c5 request {k} ({PLANS[docs[k][0]]}, {docs[k][1].size} B)"
        if k % 997 == 0:
            assert np.array_equal(d2h(r.input, r.n_bytes, np.uint8), docs[k][1]), f"c5 request {k} input"

    multi = pressure_serve(sv, docs, check, "c5", budget_s=240.0)
    st = sv.stats()
    say(f"c5: {st.batches} batches, {multi} requests shared a batch, ring wraps {st.ring_wraps}, "
        f"staging wraps {st.staging_wraps}, busy rejects {st.busy_rejects}")
    assert st.completed == st.released == len(docs) and st.inflight_reserved == 0
    assert multi > len(docs) // 2, "continuous batching was not exercised"
    assert st.staging_wraps > 0, "the staging ring never wrapped"
    assert st.ring_wraps > 0, "the device ring never wrapped"
    # the same requests alone: identical outputs
    for k in range(0, len(docs), 6000):
        got = {}
        pressure_serve(sv, [docs[k]], lambda _, r: got.setdefault(0, fetch_request(r)), "c5 alone")
        assert same(got[0], exp[k]), f"c5 request {k} alone"
    sv.close()


def test_serving_small_ring_wraps(torch_cuda):
    """A device ring of two largest batches per plan (requests up to 4 MiB): the
    ring wraps many times, and batches wait for ring space their held
    predecessors free on release."""
    docs = request_stream(8, 12_000)
    exp = [expected_request(PLANS[p], d) for p, d in docs]
    sv = Server(PLANS, ring_bytes=3 * (8 << 20), staging_bytes=3 * (16 << 20), max_batch_bytes=4 << 20,
                max_batch_docs=1024)

    def check(k, r):
        assert same(fetch_request(r), exp[k]), f"small ring request {k}"

    pressure_serve(sv, docs, check, "small ring")
    st = sv.stats()
    say(f"small ring: {st.batches} batches, ring wraps {st.ring_wraps}, staging wraps {st.staging_wraps}")
    assert st.ring_wraps > 0 and st.staging_wraps > 0
    sv.close()


def test_deep_request_in_a_server_batch(torch_cuda):
    """The 4 MiB depth-2^21 request submitted in a burst with 4,096 others:
    it shares a batch, and it and every batchmate come back exact."""
    sz = W.sizes(19, 4096, median=4096.0, sigma=1.85)
    docs = [(0, W.make_doc("json", W.doc_seed(19, i), int(sz[i]))) for i in range(4096)]
    docs.insert(2048, (0, deep_doc(1 << 21)))
    exp = [expected_request("json", d) for _, d in docs]
    sv = Server(["json"])
    shared = {}

    def check(k, r):
        assert same(fetch_request(r), exp[k]), f"deep burst request {k} ({docs[k][1].size} B)"
        if k == 2048:
            shared["docs"] = int(r.batch_docs)

    pressure_serve(sv, docs, check, "deep burst")
    say(f"deep burst: the deep request shared a batch of {shared.get('docs')} documents")
    assert shared.get("docs", 0) > 1, "the deep request was served alone"
    sv.close()
