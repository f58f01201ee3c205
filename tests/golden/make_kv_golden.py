"""Write tests/golden/kv_golden.json from the reference's own INI and TOML hooks.

Runs only where the reference tree exists: it drives oracle/_ref/ref_kv, the
reference's ini_parse and toml_parse (bench/bench_parse_host.cu:30-248)
compiled from its source file by oracle/Makefile and fed as the reference's
full-parse benchmark feeds them (its CPU scanner, model::replay).  Per case:
how the input is made (`inputs(spec)` rebuilds it) and what the hook built -
its hash and, per list it fills, the count and an FNV-1a 64 over the items.
Cases: hand-made documents for the branches of both hooks, the reference
benchmark's own records repeated, and random documents over each grammar's
bytes (comment markers, brackets, '=', quotes, backslashes, newlines).

    python tests/golden/make_kv_golden.py
"""
from __future__ import annotations

import base64
import json
import os
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from pars_amd import workloads as W  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kv_golden.json")
REF_KV = os.path.join(ROOT, "oracle", "_ref", "ref_kv")

# bytes each grammar treats specially, plus text
SOUP = {"ini": b"[]=;#\n \t\"'\\ab1.", "toml": b"[]=#\n \t\"'\\,ab1."}

HAND = {
    "ini": [
        b"[sec]\nk = v\n", b"k = v", b"[a b]\nx = 1 2 3\n", b"; c [x] = y\nk=v\n", b"# c\n[s]\n",
        b"a = \"q w\" e\n", b"a = 'q\\' w' e\n", b"a = \"unterminated\n", b"]x = [\n=\n", b"k = v ; tail\nz = 9",
        b"", b"\n\n\n", b"[", b"=", b"   spaced   key   =   spaced value  \n", b"[s]k=v[t]x=y\n",
        b"a=\"\\\"\" b\n", b"x = 'a'b'c'\n", b"\tk\t=\tv\r\n",
    ],
    "toml": [
        b"[server]\nhost = \"localhost\"\nport = 8080\n", b"arr = [1, 2, 3]\n", b"k = [a b]\nx = 1\n",
        b"# c [x] = y\nk = v\n", b"= v\n", b"k =\nv\n", b"a = [\n1, 2\n]\n", b"[[t]]\nx = [[1], [2]]\n",
        b"k = 'a # b'\n", b"k = \"a \\\" b\"\n", b"a b c = d e\n", b"[s] k = v\n", b"", b"[", b"k = [ ] z\n",
        b"x = 1 # tail\ny = 2", b"k = \"multi\nline\"\n",
    ],
}


def inputs(spec: dict) -> bytes:
    """The input bytes of a case (literal bytes, or a seeded generator)."""
    if "b64" in spec:
        return base64.b64decode(spec["b64"])
    if spec["gen"] == "format":
        return W.format_doc(spec["fmt"], spec["n"]).tobytes()
    if spec["gen"] == "soup":
        return W.alpha(spec["seed"], spec["n"], SOUP[spec["fmt"]]).tobytes()
    raise ValueError(spec)


def specs() -> list[tuple[str, str, dict]]:
    cases = []
    for fmt in ("ini", "toml"):
        for i, d in enumerate(HAND[fmt]):
            cases.append((f"{fmt}_hand_{i}", fmt, {"b64": base64.b64encode(d).decode()}))
        for n in (100, 4096, 65536):
            cases.append((f"{fmt}_format_{n}", fmt, {"gen": "format", "fmt": fmt, "n": n}))
        for s in range(40):
            cases.append((f"{fmt}_soup_{s}", fmt, {"gen": "soup", "fmt": fmt, "seed": 500 + s, "n": 50 + 37 * s}))
    return cases


def run_ref(fmt: str, data: bytes) -> dict:
    with tempfile.TemporaryDirectory() as td:
        src = os.path.join(td, "in")
        with open(src, "wb") as f:
            f.write(data)
        r = subprocess.run([REF_KV, fmt, src], check=True, stdout=subprocess.PIPE, text=True)
    return json.loads(r.stdout)


def main() -> None:
    if not os.path.exists(REF_KV):
        sys.exit("oracle/_ref/ref_kv is missing: build it with `make -C oracle ref` (needs the reference tree)")
    out = [{"name": name, "fmt": fmt, "input": spec, "ref": run_ref(fmt, inputs(spec))}
           for name, fmt, spec in specs()]
    with open(OUT, "w") as f:
        json.dump({"source": "the reference's ini_parse / toml_parse (bench/bench_parse_host.cu:30-248) via "
                             "oracle/_ref/ref_kv", "cases": out}, f, indent=0)
    print(f"wrote {len(out)} cases to {OUT}")


if __name__ == "__main__":
    main()
