"""GPU parity of the batched scan (pars_hip_scan_batch): every document of a
batch against the CPU oracle run on that document alone — structural
positions and bytes, bracket positions/bytes, bracket pairs (document-local,
nest.hpp convention), pair_pos, and the balanced / well_formed / max_depth /
unclosed_mode flags.  So each result is the same alone and in the mix.

Batches mix tiny documents (one byte, lane and tile boundaries) with
documents spanning many tiles, adversarial random-alphabet documents
(unterminated strings, unbalanced and mismatched brackets), and run on the
fast evaluators, the general path, one- and two-descriptor plans, each tile
height, and through the captured HIP graph on repeated calls.  Batches whose
documents are all balanced pair by stack matching (mismatched closing bytes
included); one unbalanced document sends the whole batch through the sort.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from pars_amd import Plan, Scanner, pack_batch
from pars_amd import _native as N
from pars_amd import workloads as W
from tests.gpu_util import d2h, say, to_device

pytestmark = pytest.mark.gpu


def mixed_docs(g: str, seed: int, count: int, big: int = 0) -> list[np.ndarray]:
    """Documents of assorted sizes: boundary sizes, log-normal sizes, adversarial
    alphabets, and `big` documents of several tiles."""
    rng = np.random.default_rng(seed)
    kind = {"json": "json", "csv": "csv", "clike": "clike"}.get(g)
    docs = []
    boundary = [1, 2, 63, 64, 65, 127, 128, 200, 4095, 4096, 4097, 16383, 16384, 16385]
    for i in range(count):
        r = i % 4
        if i < len(boundary):
            n = boundary[i]
        else:
            n = int(min(max(rng.lognormal(np.log(3000), 1.2), 1), 200_000))
        if r == 3 or kind is None:
            docs.append(W.alpha(seed * 1000 + i, n, W.ALPHABETS[g]))
        else:
            docs.append(W.make_doc(kind, W.doc_seed(seed, i), n))
    for j in range(big):
        n = [70_000, 150_001, 300_000][j % 3]
        docs.append(W.make_doc(kind, W.doc_seed(seed + 7, j), n) if kind else
                    W.alpha(seed * 77 + j, n, W.ALPHABETS[g]))
    order = rng.permutation(len(docs))
    return [docs[k] for k in order]


def run_batch(torch, sc: Scanner, docs, want_pairs=True, run_pairs=True, stream=None):
    buf, spans = pack_batch(docs)
    t, ptr = to_device(torch, buf)
    s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    r = sc.scan_batch(ptr, buf.size, spans, s, want_pairs=want_pairs, run_pairs=run_pairs)
    nd = len(docs)
    out = {
        "r": r,
        "structural": d2h(r.structural, r.n_struct, np.uint32),
        "struct_start": d2h(r.struct_start, nd + 1, np.uint32),
        "bracket_pos": d2h(r.bracket_pos, r.n_brackets, np.uint32),
        "bracket_char": d2h(r.bracket_char, r.n_brackets, np.uint8),
        "bracket_start": d2h(r.bracket_start, nd + 1, np.uint32),
        "flags": d2h(r.doc_flags, nd, np.uint8),
        "max_depth": d2h(r.max_depth, nd, np.int32),
    }
    if run_pairs:
        out["pair"] = d2h(r.bracket_pair, r.n_brackets, np.int32)
    if want_pairs:
        out["pair_pos"] = d2h(r.pair_pos, r.n_struct, np.int32)
    del t
    return out


def check_batch(g: str, docs, got, label="", plan: bytes | None = None):
    plan = plan if plan is not None else Plan.load(g).blob
    ss, bs = got["struct_start"], got["bracket_start"]
    assert ss[0] == 0 and bs[0] == 0, label
    assert ss[-1] == got["r"].n_struct and bs[-1] == got["r"].n_brackets, label
    for d, data in enumerate(docs):
        exp = O.full(plan, data)
        tag = f"{g}:{label} doc {d} ({data.size} B)"
        pos = got["structural"][ss[d]:ss[d + 1]]
        assert np.array_equal(pos, exp["positions"]), tag
        bpos = got["bracket_pos"][bs[d]:bs[d + 1]]
        assert np.array_equal(bpos, exp["bracket_pos"]), tag
        assert np.array_equal(got["bracket_char"][bs[d]:bs[d + 1]], exp["bracket_char"]), tag
        f = int(got["flags"][d])
        assert bool(f & 1) == exp["balanced"], tag + " balanced"
        assert bool(f & 4) == exp["unclosed_mode"], tag + " unclosed_mode"
        assert int(got["max_depth"][d]) == exp["max_depth"], tag + " max_depth"
        if "pair" in got:
            assert bool(f & 2) == exp["well_formed"], tag + " well_formed"
            assert np.array_equal(got["pair"][bs[d]:bs[d + 1]], exp["pair"]), tag + " pair"
        if "pair_pos" in got:
            pp = got["pair_pos"][ss[d]:ss[d + 1]]
            assert np.array_equal(pp, O.pair_pos_device(exp["positions"], exp["bracket_pos"], exp["pair"])), tag


@pytest.mark.parametrize("g", ["json", "csv", "clike", "clike6", "toml", "xml"])
def test_batch_mixed_sizes(torch_cuda, g):
    docs = mixed_docs(g, 11, 120, big=3)
    say(f"batch {g}: {len(docs)} docs, {sum(d.size for d in docs)} bytes")
    sc = Scanner(g)
    got = run_batch(torch_cuda, sc, docs)
    check_batch(g, docs, got, "default")


@pytest.mark.parametrize("rows", [1, 2, 3, 4])
@pytest.mark.parametrize("path", [0, 1])
def test_batch_rows_and_paths(torch_cuda, rows, path):
    docs = mixed_docs("json", 23, 80, big=2)
    sc = Scanner("json")
    sc.set_option(N.PARS_OPT_ROWS_PER_TILE, rows)
    sc.set_option(N.PARS_OPT_PATH, path)
    say(f"batch json rows={rows} path={path}")
    check_batch("json", docs, run_batch(torch_cuda, sc, docs), f"rows{rows} path{path}")


def test_batch_single_document_equals_plain(torch_cuda):
    data = W.json(5, 300_000, 12)
    sc = Scanner("json")
    got = run_batch(torch_cuda, sc, [data])
    check_batch("json", [data], got, "single")


def test_batch_without_pairs(torch_cuda):
    docs = mixed_docs("csv", 31, 60)
    sc = Scanner("csv")
    got = run_batch(torch_cuda, sc, docs, want_pairs=False, run_pairs=False)
    check_batch("csv", docs, got, "no pairs")


def test_batch_graph_replays(torch_cuda):
    """The captured graph: the first call of a shape runs eagerly, the second
    captures, later ones replay; every call must match, including after the
    document table changes under the same shape."""
    torch = torch_cuda
    sc = Scanner("json")
    sc.set_option(N.PARS_OPT_GRAPH, 1)
    stream = torch.cuda.Stream()
    docs_a = mixed_docs("json", 41, 64, big=1)
    docs_b = mixed_docs("json", 42, 64, big=1)
    buf_a, spans_a = pack_batch(docs_a)
    buf_b, spans_b = pack_batch(docs_b)
    n = max(buf_a.size, buf_b.size)
    t = torch.empty(n + 64, dtype=torch.uint8, device="cuda")
    ptr = t.data_ptr()
    for it, (docs, buf, spans) in enumerate([(docs_a, buf_a, spans_a)] * 3 + [(docs_b, buf_b, spans_b)] * 2):
        t[: buf.size].copy_(torch.from_numpy(buf))
        torch.cuda.synchronize()
        say(f"graph call {it}")
        r = sc.scan_batch(ptr, n, spans, stream.cuda_stream, want_pairs=True, run_pairs=True)
        nd = len(docs)
        got = {
            "r": r,
            "structural": d2h(r.structural, r.n_struct, np.uint32),
            "struct_start": d2h(r.struct_start, nd + 1, np.uint32),
            "bracket_pos": d2h(r.bracket_pos, r.n_brackets, np.uint32),
            "bracket_char": d2h(r.bracket_char, r.n_brackets, np.uint8),
            "bracket_start": d2h(r.bracket_start, nd + 1, np.uint32),
            "flags": d2h(r.doc_flags, nd, np.uint8),
            "max_depth": d2h(r.max_depth, nd, np.int32),
            "pair": d2h(r.bracket_pair, r.n_brackets, np.int32),
            "pair_pos": d2h(r.pair_pos, r.n_struct, np.int32),
        }
        check_batch("json", docs, got, f"graph call {it}")


def test_batch_rejects_bad_spans(torch_cuda):
    torch = torch_cuda
    sc = Scanner("json")
    buf, spans = pack_batch([b"[1]", b"{}"])
    t, ptr = to_device(torch, buf)
    bad = (N.DocSpan * 2)()
    bad[0].offset, bad[0].size = 0, 3
    bad[1].offset, bad[1].size = 32, 2  # not on a 64-byte boundary
    with pytest.raises(Exception):
        sc.scan_batch(ptr, buf.size, bad, torch.cuda.current_stream().cuda_stream)
    bad[1].offset, bad[1].size = 64, 0  # empty
    with pytest.raises(Exception):
        sc.scan_batch(ptr, buf.size, bad, torch.cuda.current_stream().cuda_stream)
    got = sc.scan_batch(ptr, buf.size, spans, torch.cuda.current_stream().cuda_stream)
    assert got.n_docs == 2 and got.n_struct == 4


def balanced_docs(seed: int, count: int) -> list[np.ndarray]:
    """Balanced JSON-grammar documents over 1 MiB in all (so the batch pairs on
    the device-wide paths): valid JSON of log-normal sizes, bracket soups whose
    closing bytes are sometimes swapped (balanced but not well formed), and two
    soups deep and long enough that their residues cross stack chunks."""
    from tests.test_gpu_stack import soup

    rng = np.random.default_rng(seed)
    docs = []
    for i in range(count):
        if i % 5 == 4:
            docs.append(soup(seed * 1000 + i, int(rng.integers(1, 4000)), int(rng.integers(1, 300)), 0.01,
                             p_text=0.3))
        else:
            n = int(min(max(rng.lognormal(np.log(4096), 1.0), 200), 300_000))
            docs.append(W.make_doc("json", W.doc_seed(seed, i), n))
    docs.append(soup(seed + 1, 40_000, 20_000, 0.0005))
    docs.append(soup(seed + 2, 30_000, 9_000, 0.0))
    order = rng.permutation(len(docs))
    return [docs[k] for k in order]


@pytest.mark.parametrize("graph", [0, 1])
def test_batch_of_balanced_documents_pairs_by_stack(torch_cuda, graph):
    docs = balanced_docs(61, 300)
    total = sum(d.size for d in docs)
    assert total > (1 << 20)
    sc = Scanner("json")
    sc.set_option(N.PARS_OPT_GRAPH, graph)
    s = torch_cuda.cuda.Stream() if graph else None
    for call in range(3 if graph else 1):  # graph: eager, capture, replay
        say(f"batch stack path: {len(docs)} docs, {total} B, graph={graph} call {call}")
        got = run_batch(torch_cuda, sc, docs, stream=s.cuda_stream if s else None)
        assert sc.get_option(N.PARS_OPT_LAST_PAIRING) == N.PARS_PAIRING_STACK
        check_batch("json", docs, got, f"stack graph={graph} call {call}")
    assert any(not (f & 2) for f in got["flags"]), "the swapped closing bytes leave some document ill-formed"


@pytest.mark.parametrize("kind", ["extra_close", "unclosed", "below_zero"])
def test_batch_with_one_unbalanced_document_sorts(torch_cuda, kind):
    docs = balanced_docs(62, 200)
    bad = {"extra_close": b"[1,{}]]", "unclosed": b"[[{}", "below_zero": b"}{"}[kind]
    docs.insert(97, np.frombuffer(bad, np.uint8).copy())
    sc = Scanner("json")
    say(f"batch stack fallback: {kind}")
    got = run_batch(torch_cuda, sc, docs)
    assert sc.get_option(N.PARS_OPT_LAST_PAIRING) == N.PARS_PAIRING_SORT
    check_batch("json", docs, got, f"fallback {kind}")
    assert not (got["flags"][97] & 1)
