"""Write tests/golden/dom_golden.json from the reference's own JSON DOM hook.

Runs only where the reference tree exists: it drives oracle/_ref/ref_dom, the
reference's json_parser (bench/bench_parse.cu:34-354) compiled from its source
file by oracle/Makefile and fed as the reference's pars_parse feeds it (its CPU
scanner, model::replay, finish()).  For every case it records how the input is
made (`inputs(spec)` below rebuilds it) and the hook's outputs: its verdict,
node count, FNV-1a hash and FNV-1a 64 hashes of the tape's kinds, begins and
ends.  Cases: the hand-worked malformed documents of tests/test_json_dom.py,
valid generated documents, the same documents with bytes flipped, deleted or
inserted, and token soups (valid tokens in random order, with stray bytes).

    python tests/golden/make_dom_golden.py
"""
from __future__ import annotations

import base64
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from pars_amd import workloads as W  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dom_golden.json")
REF_DOM = os.path.join(ROOT, "oracle", "_ref", "ref_dom")

TOKENS = [b"{", b"}", b"[", b"]", b":", b",", b'"k"', b'"a\\"b"', b'"\\u00e9x"', b'"\\n\\t"', b'"\\q"', b"1",
          b"-0.5e+3", b"012", b"1.", b"true", b"false", b"null", b"nul", b" ", b"\n", b'"unterminated', b"\\",
          b"\x01", b"x"]


def inputs(spec: dict) -> bytes:
    """The input bytes of a case (literal bytes, or a seeded generator)."""
    if "b64" in spec:
        return base64.b64decode(spec["b64"])
    rng = np.random.default_rng(spec["seed"])
    if spec["gen"] == "json":
        return W.json(spec["seed"], spec["n"], spec.get("max_depth", 8)).tobytes()
    if spec["gen"] == "corrupt":
        d = bytearray(W.json(spec["seed"], spec["n"], spec.get("max_depth", 8)).tobytes())
        for _ in range(spec["edits"]):
            i = int(rng.integers(0, len(d)))
            op = int(rng.integers(0, 3))
            b = int(rng.choice(list(b'{}[]:,"\\ 1tfn\x01')))
            if op == 0:
                d[i] = b
            elif op == 1:
                del d[i]
            else:
                d.insert(i, b)
        return bytes(d)
    if spec["gen"] == "soup":
        out = []
        for _ in range(spec["tokens"]):
            out.append(TOKENS[int(rng.integers(0, len(TOKENS)))])
        return b"".join(out)
    raise ValueError(spec)


def specs() -> list[tuple[str, dict]]:
    from tests.test_json_dom import KAT

    cases = [(f"kat_{i}", {"b64": base64.b64encode(d).decode()}) for i, (d, *_rest) in enumerate(KAT)]
    for s in range(6):
        cases.append((f"json_{s}", {"gen": "json", "seed": 100 + s, "n": 4096 << s, "max_depth": 3 + 2 * s}))
    for s in range(60):
        cases.append((f"corrupt_{s}", {"gen": "corrupt", "seed": 200 + s, "n": 2000 + 300 * s, "max_depth": 6,
                                        "edits": 1 + s % 5}))
    for s in range(200):
        cases.append((f"soup_{s}", {"gen": "soup", "seed": 400 + s, "tokens": 3 + (s * 7) % 90}))
    return cases


def run_ref(data: bytes) -> dict:
    with tempfile.TemporaryDirectory() as td:
        src, out = os.path.join(td, "in"), os.path.join(td, "out")
        with open(src, "wb") as f:
            f.write(data)
        subprocess.run([REF_DOM, src, out], check=True)
        b = open(out, "rb").read()
    assert b[:8] == b"PRSDOM01"
    ok = b[8]
    nodes, h, n = np.frombuffer(b[9:33], np.uint64)
    rec = np.frombuffer(b[33:], dtype=np.dtype([("k", "u1"), ("b", "<u4"), ("e", "<u4")]), count=int(n))
    return {"ok": bool(ok), "nodes": int(nodes), "hash": int(h), "tape_len": int(n),
            "kind_fnv": O.fnv(rec["k"].copy()), "begin_fnv": O.fnv(rec["b"].copy()),
            "end_fnv": O.fnv(rec["e"].copy())}


def main() -> None:
    if not os.path.exists(REF_DOM):
        sys.exit("oracle/_ref/ref_dom is missing: build it with `make -C oracle ref` (needs the reference tree)")
    out = []
    for name, spec in specs():
        out.append({"name": name, "input": spec, "ref": run_ref(inputs(spec))})
    with open(OUT, "w") as f:
        json.dump({"source": "the reference's json_parser (bench/bench_parse.cu:34-354) via oracle/_ref/ref_dom",
                   "cases": out}, f, indent=0)
    print(f"wrote {len(out)} cases to {OUT}")


if __name__ == "__main__":
    main()
