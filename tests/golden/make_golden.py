"""Write tests/golden/golden.json from the reference's own CPU backend.

Runs only where the reference tree exists (it drives oracle/_ref/ref_driver,
the reference's model::scanner + model::pair_brackets compiled from its
sources by oracle/Makefile).  For every case it records the input (literal
cases: the bytes; generated cases: generator spec + input hash) and the
reference's outputs: counts, flags and FNV-1a 64 hashes of positions, chars,
bracket_pos, bracket_char and pair; literal cases also keep the full arrays.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from pars_amd import workloads as W  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.json")

# tests/test_model.cpp:154, 179, 205, 221-227 (the reference's known-answer inputs)
KAT = [
    ("kat_json", "json", b'{"name":"a\\"b","n":[1,2.5e-3,true,null],"x":{}}'),
    ("kat_csv", "csv", b'a,"b, c","d""e",f\n1,2,3,4\n'),
    ("kat_pair", "json", b'{"a":[1,{"b":2}],"c":[]}'),
    ("kat_clike_0", "clike", b"a;b/* { ; } */c;// x { } ;\nd;{e;f}"),
    ("kat_clike_1", "clike", b"/**/;"),
    ("kat_clike_2", "clike", b"/***/;"),
    ("kat_clike_3", "clike", b"/* * */x;"),
    ("kat_clike_4", "clike", b"/*/**/;"),
    ("kat_clike_5", "clike", b"a/*a**b*/;c"),
    ("kat_clike_6", "clike", b"// /* ; { */\n;"),
    # traps the backend must reproduce (escape outside a string opens one;
    # descriptors are independent; malformed nesting; unclosed modes)
    ("trap_json_escape_outside", "json", b'[\\"a", 1] {"k": "v\\\\"} \\\\"x"'),
    ("trap_json_unclosed", "json", b'{"a": "abc'),
    ("trap_json_closed_at_end", "json", b'"abc"'),
    ("trap_json_malformed", "json", b"]]{[}(]{{]"),
    ("trap_clike6_independent", "clike6", b'x = "http://x"; { y; } /* // */ ; z;'),
    ("trap_clike6_char_quote", "clike6", b"c = '\"'; d = \"'\"; {a;}"),
    ("trap_csv_doubled", "csv", b'"a""b",c\n"""",d\n"x\n'),
    ("edge_empty", "json", b""),
    ("edge_one_open", "json", b"["),
    ("edge_one_close", "json", b"]"),
]


def generated_cases():
    cases = [
        ("json_1mib", "json", {"gen": "json", "seed": 1, "n": 1 << 20, "max_depth": 8}),
        ("json_deep_256k", "json", {"gen": "json", "seed": 7, "n": 256 << 10, "max_depth": 32,
                                    "string_fraction": 0.30}),
        ("csv_1mib", "csv", {"gen": "csv", "seed": 4, "n": 1 << 20}),
        ("clike6_1mib", "clike6", {"gen": "clike", "seed": 3, "n": 1 << 20, "max_depth": 24}),
        ("clike_1mib", "clike", {"gen": "clike", "seed": 3, "n": 1 << 20, "max_depth": 24}),
    ]
    for g in ["json", "csv", "clike", "clike6", "toml", "ini", "xml"]:
        for seed in (12345, 999):
            cases.append((f"rand_{g}_{seed}", g, {"gen": "alpha", "seed": seed, "n": 64 << 10, "alphabet": g}))
    return cases


def make_input(spec: dict) -> np.ndarray:
    g = spec["gen"]
    if g == "json":
        return W.json(spec["seed"], spec["n"], spec.get("max_depth", 8), spec.get("string_fraction"))
    if g == "csv":
        return W.csv(spec["seed"], spec["n"])
    if g == "clike":
        return W.clike(spec["seed"], spec["n"], spec.get("max_depth", 24))
    if g == "alpha":
        return W.alpha(spec["seed"], spec["n"], W.ALPHABETS[spec["alphabet"]])
    raise ValueError(g)


def record(r: dict, keep_arrays: bool) -> dict:
    rec = {
        "n_struct": int(r["positions"].size),
        "n_brackets": int(r["bracket_pos"].size),
        "unclosed_mode": bool(r["unclosed_mode"]),
        "balanced": bool(r["balanced"]),
        "well_formed": bool(r["well_formed"]),
        "max_depth": int(r["max_depth"]),
        "hash": {k: O.fnv(r[k]) for k in ("positions", "chars", "bracket_pos", "bracket_char", "pair")},
    }
    if keep_arrays:
        rec["arrays"] = {k: r[k].tolist() for k in ("positions", "chars", "bracket_pos", "bracket_char", "pair")}
    return rec


def main() -> None:
    if not O.ref_available():
        sys.exit("oracle/_ref/ref_driver is missing: build it with `make -C oracle ref` (needs the reference tree)")
    out = {"format": 1, "hash": "fnv1a64 of the little-endian array bytes", "cases": {}}
    for name, g, data in KAT:
        r = O.ref_full(g, data)
        out["cases"][name] = {"plan": g, "input": data.decode("latin-1"), **record(r, True)}
    for name, g, spec in generated_cases():
        d = make_input(spec)
        r = O.ref_full(g, d)
        out["cases"][name] = {"plan": g, "spec": spec, "input_hash": O.fnv(d), **record(r, False)}
        print(name, out["cases"][name]["n_struct"], out["cases"][name]["n_brackets"], file=sys.stderr)
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", OUT, file=sys.stderr)


if __name__ == "__main__":
    main()
