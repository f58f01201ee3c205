"""Write tests/golden/plans/ from the reference compiler and CPU backend.

Runs only where the reference tree exists.  Builds oracle/_ref/ref_plans and
oracle/_ref/ref_driver (oracle/Makefile compiles them from the reference's
sources) and writes:

  <name>.plan     the reference's compile::make_scan_plan bytes for each probe
                  grammar of oracle/grammars/compiler_cases.hpp
  manifest.json   its support verdict (compile::is_supported) and the plan hash
  scans.json      model::scanner + model::pair_brackets outputs over each probe
                  plan for seeded inputs drawn from the grammar's alphabet

    python tests/golden/make_plans.py
"""
from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "plans")
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from pars_amd import workloads as W  # noqa: E402
from tests.golden.make_golden import record  # noqa: E402

# Bytes each probe grammar gives meaning to, plus filler (repeats weight them).
ALPHABETS = {
    "sql_case": b"''''  ''(()),;*=-\n-NULxy19 @\"\"",
    "esc_case": b"|||%%``^^af09#\n{}{}<>,:! x",
    "many_case": b"()[]{}<>ABCDEFGHIJ,;:.!?@$%&=+xyz ",
    "caps_case": b'~abcdefgh;;$aAbBcC,,@@.."""k\x00[]]+ ',
    "tiny_case": b"xxy;{}",
    "unsup_case": b"'''()!!-x ",
    "prop_case": b",,,xxy(()) ",
}
SCANS = [(7, 4096), (8, 65536), (9, 200003)]


def main() -> int:
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_plans")
    if not os.path.exists(exe) or not O.ref_available():
        print("oracle/_ref is not built (reference tree absent)", file=sys.stderr)
        return 1
    os.makedirs(OUT, exist_ok=True)
    res = subprocess.run([exe, OUT], check=True, capture_output=True, text=True)
    manifest = {}
    for line in res.stdout.split("\n"):
        if not line.strip():
            continue
        name, sup = line.split()
        with open(os.path.join(OUT, name + ".plan"), "rb") as f:
            blob = f.read()
        manifest[name] = {"supported": sup == "1", "sha256": hashlib.sha256(blob).hexdigest()}
    with open(os.path.join(OUT, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
        f.write("\n")
    scans = {"format": 1, "hash": "fnv1a64 of the little-endian array bytes", "cases": {}}
    for name in manifest:
        for seed, n in SCANS:
            d = W.alpha(seed, n, ALPHABETS[name])
            rec = record(O.ref_full(name, d), n <= 4096)
            scans["cases"][f"{name}_{seed}_{n}"] = {
                "plan": name, "seed": seed, "n": n, "alphabet": ALPHABETS[name].decode("latin-1"),
                "input_hash": O.fnv(d), **rec}
            print(name, seed, n, rec["n_struct"], rec["n_brackets"], rec["max_depth"], file=sys.stderr)
    with open(os.path.join(OUT, "scans.json"), "w") as f:
        json.dump(scans, f, indent=1, sort_keys=True)
        f.write("\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
