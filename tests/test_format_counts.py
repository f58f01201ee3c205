"""The reference's per-format benchmark documents (bench/bench_boundary.cu:
312-337, restated in pars_amd/workloads.py format_doc) and its published
structural counts for them (the `tokens` column of paper/plots/data/pars.csv,
N1a rows): the oracle's scan of each document yields exactly the reference's
count at 8 MiB and 64 MiB, for all six grammars.  The same documents at 512
MiB go through the HIP path in tests/test_gpu_formats.py.  CPU only."""
from __future__ import annotations

import pytest

from oracle import oracle as O
from pars_amd import Plan
from pars_amd import workloads as W

# paper/plots/data/pars.csv: (format, scale_mb) -> tokens, N1a rows
TOKENS = {
    ("json", 8): 2592858, ("csv", 8): 1070892, ("toml", 8): 814430,
    ("clike", 8): 453440, ("ini", 8): 1613200, ("xml", 8): 1398103,
    ("json", 64): 20742755, ("csv", 64): 8567094, ("toml", 64): 6515430,
    ("clike", 64): 3627508, ("ini", 64): 12905560, ("xml", 64): 11184817,
    ("json", 512): 165941931, ("csv", 512): 68536716, ("toml", 512): 52123390,
    ("clike", 512): 29020050, ("ini", 512): 103244410, ("xml", 512): 89478487,
}


def test_json_document_shape():
    """format_doc("json") is the reference's loop (bench_boundary.cu:312-320)."""
    rec = W.FORMAT_RECORDS["json"]
    for target in [0, 1, 2, len(rec), len(rec) + 1, len(rec) + 2, 2 * len(rec) + 3, 5000]:
        js = bytearray(b"[")
        while len(js) < target:
            if len(js) > 1:
                js += b","
            js += rec
        js += b"]"
        assert W.format_doc("json", target).tobytes() == bytes(js), target


@pytest.mark.parametrize("scale", [8, 64])
@pytest.mark.parametrize("fmt", ["json", "csv", "toml", "clike", "ini", "xml"])
def test_oracle_equals_published_counts(fmt, scale):
    d = W.format_doc(fmt, scale << 20)
    r = O.full(Plan.load(fmt).blob, d)
    assert len(r["positions"]) == TOKENS[(fmt, scale)]
    assert r["balanced"] and r["well_formed"]
