"""GPU tests of the serving runtime (pars_hip_server_*): requests of three
plans submitted while others are in flight come back bit-exact against the
CPU oracle run on each request alone, and equal to the same request served
alone; the staging ring refuses work when full and recovers; the load driver
replays Poisson arrivals to completion."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from pars_amd import Plan, Server
from pars_amd import workloads as W
from tests.gpu_util import d2h, say

pytestmark = pytest.mark.gpu

PLANS = ["json", "csv", "clike"]


def request_docs(seed: int, count: int, max_size: int = 200_000):
    rng = np.random.default_rng(seed)
    kinds = W.doc_kinds(seed, count, {"json": 60, "csv": 25, "clike": 15})
    docs = []
    for i, k in enumerate(kinds):
        n = int(min(max(rng.lognormal(np.log(4096), 1.3), 1), max_size))
        if i % 7 == 6:  # adversarial
            docs.append((PLANS.index(k), W.alpha(seed * 131 + i, n, W.ALPHABETS[k])))
        else:
            docs.append((PLANS.index(k), W.make_doc(k, W.doc_seed(seed, i), n)))
    return docs


def fetch(r):
    out = {
        "positions": d2h(r.structural, r.n_struct, np.uint32),
        "bracket_pos": d2h(r.bracket_pos, r.n_brackets, np.uint32),
        "bracket_char": d2h(r.bracket_char, r.n_brackets, np.uint8),
        "flags": r.flags,
        "max_depth": r.max_depth,
    }
    if r.bracket_pair:
        out["pair"] = d2h(r.bracket_pair, r.n_brackets, np.int32)
    return out


def check(plan: str, data, got, tag):
    exp = O.full(Plan.load(plan).blob, data)
    assert np.array_equal(got["positions"], exp["positions"]), tag
    assert np.array_equal(got["bracket_pos"], exp["bracket_pos"]), tag
    assert np.array_equal(got["bracket_char"], exp["bracket_char"]), tag
    assert bool(got["flags"] & 1) == exp["balanced"], tag
    assert bool(got["flags"] & 2) == exp["well_formed"], tag
    assert bool(got["flags"] & 4) == exp["unclosed_mode"], tag
    assert got["max_depth"] == exp["max_depth"], tag
    assert np.array_equal(got["pair"], exp["pair"]), tag


def serve_all(sv: Server, docs, max_inflight: int | None = None):
    """Submit everything (throttled to max_inflight), fetch each result as it completes."""
    import time

    results, pending, i = {}, {}, 0
    last = time.monotonic()
    while len(results) < len(docs):
        assert time.monotonic() - last < 30, "server made no progress for 30 s"
        while i < len(docs) and (max_inflight is None or len(pending) < max_inflight):
            rid = sv.submit(docs[i][0], docs[i][1])
            if rid is None:
                break
            pending[rid] = i
            i += 1
        for r in sv.poll(wait_us=2000):
            assert r.status == 0
            results[pending.pop(r.id)] = (fetch(r), r.batch_docs)
            sv.release(r.id)
            last = time.monotonic()
    return results


def test_server_results_equal_alone_and_in_the_mix(torch_cuda):
    docs = request_docs(5, 240)
    sv = Server(PLANS, ring_bytes=64 << 20, staging_bytes=32 << 20, max_batch_bytes=2 << 20,
                max_batch_docs=512)
    mixed = serve_all(sv, docs)
    batched = sum(1 for v in mixed.values() if v[1] > 1)
    say(f"server: {len(docs)} requests, {batched} served in multi-request batches, stats {sv.stats().batches} batches")
    assert batched > 0, "no request shared a batch: continuous batching was not exercised"
    for i, (p, d) in enumerate(docs):
        check(PLANS[p], d, mixed[i][0], f"request {i} in the mix")
    # a sample served alone: identical outputs
    for i in range(0, len(docs), 24):
        alone = serve_all(sv, [docs[i]])[0][0]
        for k in ("positions", "bracket_pos", "bracket_char", "pair"):
            assert np.array_equal(alone[k], mixed[i][0][k]), f"request {i} alone vs mix: {k}"
        assert alone["flags"] == mixed[i][0]["flags"] and alone["max_depth"] == mixed[i][0]["max_depth"]
    st = sv.stats()
    assert st.completed == st.released and st.inflight_reserved == 0
    sv.close()


def test_server_staging_full_then_recovers(torch_cuda):
    sv = Server(["json"], ring_bytes=8 << 20, staging_bytes=4 << 20, max_batch_bytes=2 << 20, max_batch_docs=64)
    doc = W.json(9, 1 << 20, 6)
    ids, busy = [], 0
    for _ in range(16):
        rid = sv.submit(0, doc)
        if rid is None:
            busy += 1
        else:
            ids.append(rid)
    assert busy > 0, "a 4 MiB staging ring cannot hold sixteen 1 MiB requests"
    import time

    got, t0 = 0, time.monotonic()
    while got < len(ids):
        assert time.monotonic() - t0 < 60, "staged requests did not complete"
        for r in sv.poll(wait_us=5000):
            assert r.status == 0
            check("json", doc, fetch(r), "staged request")
            sv.release(r.id)
            got += 1
    assert sv.submit(0, doc) is not None
    sv.close()


def test_server_replay_poisson(torch_cuda):
    docs = request_docs(6, 600, max_size=64_000)
    rng = np.random.default_rng(6)
    gaps = rng.exponential(20_000.0, len(docs))  # mean 20 us between arrivals
    arrival = np.cumsum(gaps).astype(np.uint64)
    arrival -= arrival[0]
    sv = Server(PLANS, ring_bytes=64 << 20, staging_bytes=32 << 20, max_batch_bytes=2 << 20, max_batch_docs=512)
    lat, elapsed = sv.replay([p for p, _ in docs], [d for _, d in docs], arrival)
    say(f"replay: {len(docs)} requests in {elapsed / 1e6:.2f} ms, p50 {np.percentile(lat, 50) / 1e3:.1f} us, "
        f"p99 {np.percentile(lat, 99) / 1e3:.1f} us")
    assert (lat > 0).all() and elapsed > 0
    st = sv.stats()
    assert st.completed == len(docs) and st.released == len(docs)
    sv.close()


def test_concurrent_submitters(torch_cuda):
    """Four threads submit at once (the copies into staging run outside the
    server lock, the request becomes batchable when its copy is done) while
    this thread polls and releases: every result bit-exact, ids unique, and
    the staging and device rings, budget and request pool back to empty."""
    import threading
    import time

    docs = request_docs(17, 600)
    sv = Server(PLANS, ring_bytes=96 << 20, staging_bytes=48 << 20, max_batch_bytes=2 << 20, max_batch_docs=256)
    ids: dict[int, int] = {}
    lock = threading.Lock()
    errors: list[str] = []

    def submitter(k: int):
        try:
            for i in range(k, len(docs), 4):
                t0 = time.monotonic()
                while (rid := sv.submit(docs[i][0], docs[i][1])) is None:
                    if time.monotonic() - t0 > 30:
                        raise RuntimeError("staging stayed full for 30 s")
                    time.sleep(0.0005)
                with lock:
                    assert rid not in ids, "an id was handed out twice"
                    ids[rid] = i
        except Exception as e:  # reported by the main thread
            errors.append(repr(e))

    th = [threading.Thread(target=submitter, args=(k,)) for k in range(4)]
    for t in th:
        t.start()
    results, last = {}, time.monotonic()
    while len(results) < len(docs) and not errors:
        assert time.monotonic() - last < 30, "server made no progress for 30 s"
        for r in sv.poll(wait_us=2000):
            assert r.status == 0
            while True:  # the submitter records the id right after submit returns
                with lock:
                    i = ids.get(r.id)
                if i is not None:
                    break
                time.sleep(0.0001)
            results[i] = fetch(r)
            sv.release(r.id)
            last = time.monotonic()
    for t in th:
        t.join()
    assert not errors, errors
    say(f"server: {len(docs)} requests from 4 threads in {sv.stats().batches} batches")
    for i, (p, d) in enumerate(docs):
        check(PLANS[p], d, results[i], f"request {i} (concurrent submit)")
    st = sv.stats()
    assert st.submitted == st.completed == st.released == len(docs) and st.inflight_reserved == 0
    sv.close()
