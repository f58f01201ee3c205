"""GPU parity of the JSON tape (pars_hip_json_tape) with the reference's
replay-driven JSON DOM hook as restated in oracle/json_dom.py: the verdict,
the event the hook stopped at, every node (kind, begin, end), every decoded
string and the hook's FNV-1a hash, bit-exact, on valid documents (generated,
the reference's own bench payload, deep nesting, sizes across the kernels'
4096-token blocks), on malformed ones (hand-made cases, random byte soup,
corrupted documents) and on edge cases (empty, whitespace, one value).
"""
from __future__ import annotations

import random

import numpy as np
import pytest

from oracle import json_dom as J
from oracle import oracle as O
from pars_amd import Plan, Scanner
from pars_amd import workloads as W
from tests.gpu_util import d2h, say, to_device
from tests.test_json_dom import KAT

pytestmark = pytest.mark.gpu

PLAN = Plan.load("json").blob
_sc: dict = {}


def scanner() -> Scanner:
    if "json" not in _sc:
        _sc["json"] = Scanner("json")
    return _sc["json"]


def gpu_tape(torch, d: bytes, offset: int = 0) -> dict:
    arr = np.frombuffer(d, np.uint8)
    t, ptr = to_device(torch, arr, offset)
    r = scanner().json_tape(ptr, arr.size, torch.cuda.current_stream().cuda_stream)
    n = r.n_nodes
    out = {"ok": r.ok, "fail_event": r.fail_event, "nodes": n,
           "kind": d2h(r.kind, n, np.uint8), "begin": d2h(r.begin, n, np.uint32), "end": d2h(r.end, n, np.uint32),
           "str_off": d2h(r.str_off, n, np.uint32), "arena": d2h(r.arena, r.arena_bytes, np.uint8).tobytes(),
           "arena_bytes": r.arena_bytes}
    del t
    return out


def check(torch, d: bytes, tag: str = "", offset: int = 0):
    exp = J.json_tape(d, O.scan(PLAN, np.frombuffer(d, np.uint8))["positions"])
    got = gpu_tape(torch, d, offset)
    tag = tag or repr(d[:60])
    assert (got["ok"], got["fail_event"], got["nodes"]) == (exp["ok"], exp["fail_event"], exp["nodes"]), tag
    assert np.array_equal(got["kind"], exp["kind"]), tag
    assert np.array_equal(got["begin"], exp["begin"]), tag
    assert np.array_equal(got["end"], exp["end"]), tag
    so = got["str_off"]
    bounds = list(so) + [got["arena_bytes"]]
    strings = [got["arena"][bounds[k]:bounds[k + 1]] for k in range(got["nodes"]) if got["kind"][k] == J.J_STR]
    assert strings == exp["decoded"], tag
    for k in range(got["nodes"]):  # non-strings decode to nothing
        if got["kind"][k] != J.J_STR:
            assert bounds[k + 1] == bounds[k], tag
    assert J.tape_hash(d, got["kind"], got["begin"], got["end"], strings) == exp["hash"], tag
    return got


def test_known_answers(torch_cuda):
    for d, ok, fe, nodes, h in KAT:
        say(f"[gpu] tape KAT {d[:40]!r}")
        got = check(torch_cuda, d)
        assert (got["ok"], got["fail_event"], got["nodes"]) == (ok, fe, nodes)
    check(torch_cuda, b'[{"a": [1, 2}]')


@pytest.mark.parametrize("seed,n,depth,sf", [(1, 4096, 8, None), (2, 65536 + 13, 8, None),
                                             (3, 1 << 20, 12, None), (4, 3 << 20, 32, 0.5)])
def test_generated_documents(torch_cuda, seed, n, depth, sf):
    d = W.json(seed, n, depth, sf).tobytes()
    say(f"[gpu] tape json seed={seed} n={len(d)}")
    got = check(torch_cuda, d)
    assert got["ok"]


def test_reference_bench_payload(torch_cuda):
    rec = b'{"id":123,"name":"x \\"y\\"","a":[1,2,3],"n":{"z":null}}'  # bench_parse.cu:405-412
    d = b"[" + b",".join([rec] * 40000) + b"]"
    say(f"[gpu] tape bench payload n={len(d)}")
    got = check(torch_cuda, d)
    assert got["ok"] and got["nodes"] == 1 + 40000 * 14


@pytest.mark.parametrize("depth", [1, 4095, 4097, 30000])
def test_deep_nesting(torch_cuda, depth):
    for d in (b"[" * depth + b"]" * depth, b"[" * depth + b"]" * (depth - 1),
              b'{"k":' * depth + b"1" + b"}" * depth, b"[" * depth + b"}" + b"]" * depth):
        say(f"[gpu] tape nesting depth={depth} n={len(d)}")
        check(torch_cuda, d, f"depth {depth} {d[:8]!r}")


def test_edge_cases(torch_cuda):
    for d in (b"", b" ", b"1", b'"', b"[]", b"{}", b"  true  ", b"[" + b" " * 10000 + b"]",
              b'"' + b"a" * 100000 + b'"', b"[" + b"1," * 20000 + b"1]"):
        say(f"[gpu] tape edge n={len(d)}")
        check(torch_cuda, d)


def test_random_byte_soup(torch_cuda):
    rng = random.Random(11)
    parts = [b"{", b"}", b"[", b"]", b",", b":", b'"a"', b'"', b"1", b"-2.5e3", b"true", b"null", b" ", b"x",
             b'\\"', b"01", b'"\\u00e9"', b"fals", b"\n"]
    say("[gpu] tape random soup x400")
    for _ in range(400):
        check(torch_cuda, b"".join(rng.choice(parts) for _ in range(rng.randint(0, 40))))


def test_random_token_soup(torch_cuda):
    """Longer soups of JSON-ish tokens: escapes of every kind (surrogate pairs
    included), long strings, nested fragments, control bytes."""
    rng = random.Random(15)
    parts = [b"{", b"}", b"[", b"]", b",", b":", b'"a"', b'"', b"1", b"-2.5e3", b"true", b"null", b" ", b"x",
             b'\\"', b"01", b'"\\u00e9"', b"fals", b"\n", b'"k":', b'{"a":1}', b"[1,2]", b'"\\ud83d\\ude00"',
             b'"' + b"y" * 40 + b'"', b'"\\\\"', b"-0.5e-7", b"\t", b'"\x01"', b"[]", b"{}"]
    say("[gpu] tape random token soup x1000")
    for _ in range(1000):
        check(torch_cuda, b"".join(rng.choice(parts) for _ in range(rng.randint(0, 120))))


def test_corrupted_documents(torch_cuda):
    rng = random.Random(12)
    base = W.json(21, 200_000, 10).tobytes()
    say("[gpu] tape corrupted x24")
    for _ in range(24):
        i = rng.randrange(len(base))
        d = base[:i] + bytes([rng.choice(b'{}[],:"x 1\\')]) + base[i + 1:]
        check(torch_cuda, d, f"corrupt at {i}")


def test_string_runs_staged_and_direct(torch_cuda):
    """Runs of short strings (a block's decoded bytes fit the LDS staging
    buffer) alternate with runs of long ones (they do not: written directly),
    so blocks of both kinds meet and share 16-byte chunks of the arena; escapes
    make some strings decode again rather than copy."""
    rng = random.Random(14)
    pieces = ["a", "b", "z", " ", "0", "\\n", '\\"', "\\\\", "\\u00e9", "\\u4e2d", "\\/"]
    items = []
    for seg in range(8):
        for _ in range(2500):
            n = rng.randint(30, 300) if seg % 2 else rng.randint(0, 10)
            items.append('"' + "".join(rng.choice(pieces) if rng.random() < 0.1 else "q" for _ in range(n)) + '"')
    d = ("[" + ",".join(items) + "]").encode()
    say(f"[gpu] tape string runs n={len(d)}")
    got = check(torch_cuda, d)
    assert got["ok"] and got["nodes"] == 1 + len(items)
    check(torch_cuda, d[: len(d) * 2 // 3], "cut inside a run")  # the arena ends inside a block


def test_unaligned_inputs(torch_cuda):
    """Input pointers off 16-byte alignment: strings at the buffer's first and
    last bytes are read through the byte path, the rest in aligned chunks."""
    docs = (b'"ab\\ncd efgh ijkl mnop"', b'"' + b"x" * 40 + b'"', b'["s", "tu", "' + b"v" * 30 + b'"]',
            W.json(41, 50_000, 8).tobytes())
    for off in (1, 3, 7, 13):
        for d in docs:
            say(f"[gpu] tape unaligned offset={off} n={len(d)}")
            check(torch_cuda, d, f"off={off} {d[:20]!r}", offset=off)


def test_repeated_calls_and_refusal(torch_cuda):
    a, b = W.json(31, 300_000, 8).tobytes(), b'{"x": [1, 2, {"y": "z"}]}'
    check(torch_cuda, a)
    check(torch_cuda, b)  # a smaller document after a larger one
    check(torch_cuda, a)
    with pytest.raises(Exception):
        t, ptr = to_device(torch_cuda, np.frombuffer(b, np.uint8))
        Scanner("csv").json_tape(ptr, len(b))


def test_cpp_facade_tape(torch_cuda, tmp_path):
    """runtime_scanner::json_tape_device (include/pars_hip.hpp) from a C++
    program equals the oracle."""
    import os
    import subprocess

    root = os.path.dirname(os.path.dirname(__file__))
    d = W.json(72, 80_000, 12).tobytes() + b' {"tail": [1, }'  # valid prefix, then a rejected event
    inp = tmp_path / "in.json"
    inp.write_bytes(d)
    say("[gpu] C++ facade tape")
    r = subprocess.run([os.path.join(root, "tests", "cpp", "facade_check"), os.path.join(root, "pars_amd", "plans",
                        "json.plan"), str(inp), "tape"], stdout=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0
    rows = r.stdout.split("\n")
    n, ok, fe, _ = (int(v) for v in rows[0].split())
    exp = O.json_tape(d, O.scan(PLAN, np.frombuffer(d, np.uint8))["positions"])
    assert (n, bool(ok), fe) == (exp["nodes"], exp["ok"], exp["fail_event"])
    got = np.array([[int(v) for v in row.split()] for row in rows[1:1 + n]], dtype=np.int64).reshape(-1, 3)
    assert np.array_equal(got[:, 0], exp["kind"]) and np.array_equal(got[:, 1], exp["begin"])
    assert np.array_equal(got[:, 2], exp["end"])


def test_large_document_against_c_twin(torch_cuda):
    """256 MiB of the benchmark's JSON: every node and every decoded byte
    against the hook restated in C (oracle/json_dom.c)."""
    d = W.json(2, 256 << 20, 32)
    say(f"[gpu] tape 256 MiB n={d.size}")
    t, ptr = to_device(torch_cuda, d)
    r = scanner().json_tape(ptr, d.size, torch_cuda.cuda.current_stream().cuda_stream)
    exp = O.json_tape(d, O.scan(PLAN, d)["positions"])
    assert (r.ok, r.n_nodes, r.fail_event) == (exp["ok"], exp["nodes"], exp["fail_event"])
    assert np.array_equal(d2h(r.kind, r.n_nodes, np.uint8), exp["kind"])
    assert np.array_equal(d2h(r.begin, r.n_nodes, np.uint32), exp["begin"])
    assert np.array_equal(d2h(r.end, r.n_nodes, np.uint32), exp["end"])
    assert d2h(r.arena, r.arena_bytes, np.uint8).tobytes() == exp["dec"]
    del t
