"""Seeded synthetic documents (bindings for libpars_gen.so, csrc/gen.cpp).

Seeds per workload: the 1 MiB JSON document uses seed 1, the 2 GiB JSON
document seed 2, the 1 GiB C-like source seed 3, the 100k-document batch seed
4, the serving stream seed 5.  Per-document seeds are splitmix(seed ^ index).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N

# Adversarial alphabets of the reference's random validation
# (bench/bench_boundary.cu:349-354), one per grammar family.
ALPHABETS = {
    "json": b'{}[]:,"\\0123456789tfn.-eE \n',
    "csv": b',.\n\r"abc0123 ',
    "clike": b"{};/*\n abc123",
    "clike6": b"{}()[];,/*\n\"'\\# abc123",
    "toml": b"[]{}.,=\"'#\nabc0123 \n",
    "ini": b"[]=;\n abc0123",
    "xml": b"<>/=?!\"' ab0123",
}


# The reference's per-format benchmark documents (bench/bench_boundary.cu:
# 312-337): a record repeated until the document reaches the target size
# (repeat_rec, :35-40); JSON is an array of one record, comma-separated
# (:312-320).  Kept byte for byte: the reference's published token counts at
# each size (paper/plots/data/pars.csv) then pin the scan of these documents.
FORMAT_RECORDS = {
    "json": b'{"id":123,"name":"x \\"y\\"","a":[1,2,3],"n":{"z":null}}',
    "csv": b'alpha,beta,"quoted, field","has ""q""",42,true\n',
    "toml": b"[server]\nhost = \"localhost\"\nport = 8080\nname = 'a # b'\n"
            b"# comment { } , = with \"quotes\"\narr = [1, 2, 3]\n",
    "clike": b"int x; /* c { ; } */ y = 1; // { ; }\n",
    "ini": b"[section]\nkey = value\nnum = 123\n; comment\nother = x\n",
    "xml": b'<item id="1" name="x &amp; y">alpha</item>',
}


def format_doc(fmt: str, target: int) -> np.ndarray:
    """The reference's benchmark document of `fmt` for a target of `target`
    bytes (the reference's scale_mb * 2^20)."""
    rec = FORMAT_RECORDS[fmt]
    if fmt == "json":  # '[' rec (',' rec)* ']': records appended while shorter than the target
        reps = 0 if target <= 1 else 1 + max(0, -(-(target - 1 - len(rec)) // (1 + len(rec))))
        out = b"[" + b",".join([rec] * reps) + b"]"
    else:
        out = rec * max(1, -(-target // len(rec)))
    return np.frombuffer(out, dtype=np.uint8).copy()


def _buf(n: int) -> np.ndarray:
    return np.empty(n, dtype=np.uint8)


# Documents of n bytes are generated as min(SEGMENTS, n / 1 MiB) independently
# seeded regions (filled in parallel; the bytes do not depend on thread count).
SEGMENTS = 64


def json(seed: int, n: int, max_depth: int = 8, string_fraction: float | None = None) -> np.ndarray:
    """JSON array of objects, exactly n bytes.  string_fraction targets the share of
    bytes inside string literals (quotes included)."""
    out = _buf(n)
    sb = C.c_uint64(0)
    target = -1 if string_fraction is None else int(round(string_fraction * 1000))
    N.gen().pars_gen_json(seed, n, max_depth, target, SEGMENTS, out.ctypes.data, C.byref(sb))
    json.last_string_bytes = sb.value
    return out


json.last_string_bytes = 0


def csv(seed: int, n: int) -> np.ndarray:
    out = _buf(n)
    N.gen().pars_gen_csv(seed, n, SEGMENTS, out.ctypes.data)
    return out


def clike(seed: int, n: int, max_depth: int = 24) -> np.ndarray:
    out = _buf(n)
    N.gen().pars_gen_clike(seed, n, max_depth, SEGMENTS, out.ctypes.data)
    return out


def alpha(seed: int, n: int, alphabet: bytes) -> np.ndarray:
    out = _buf(n)
    N.gen().pars_gen_alpha(seed, n, alphabet, len(alphabet), out.ctypes.data)
    return out


def sizes(seed: int, count: int, median: float = 4096.0, sigma: float = 1.85, lo: int = 200,
          hi: int = 4 << 20) -> np.ndarray:
    """Log-normal document sizes, clipped to [lo, hi]."""
    out = np.empty(count, dtype=np.uint64)
    N.gen().pars_gen_sizes(seed, count, median, sigma, lo, hi, out.ctypes.data)
    return out


def doc_seed(seed: int, index: int) -> int:
    return N.gen().pars_gen_doc_seed(seed, index)


def doc_kinds(seed: int, count: int, mix: dict[str, int]) -> list[str]:
    """Kind of each document, drawn per document from splitmix(seed ^ index) by the
    integer weights in `mix` (e.g. {"json": 70, "csv": 30})."""
    names = list(mix)
    cuts = np.cumsum([mix[k] for k in names])
    total = int(cuts[-1])
    return [names[int(np.searchsorted(cuts, doc_seed(seed, i) % total, side="right"))] for i in range(count)]


def make_doc(kind: str, seed: int, n: int) -> np.ndarray:
    if kind == "json":
        return json(seed, n, 8)
    if kind == "csv":
        return csv(seed, n)
    if kind == "clike":
        return clike(seed, n, 24)
    raise ValueError(kind)


KIND_CODES = {"json": 0, "csv": 1, "clike": 2}


def pack_generated(kinds: list[str], seeds: np.ndarray, sz: np.ndarray, align: int = 64):
    """Generate documents straight into one batch buffer (each on a 64-byte
    boundary, spaces between), in parallel; returns (buffer, DocSpan array)."""
    count = len(kinds)
    sz = np.asarray(sz, dtype=np.uint64)
    offs = np.zeros(count, dtype=np.uint64)
    if count > 1:
        offs[1:] = np.cumsum((sz + (align - 1)) // align * align)[:-1]
    total = int(offs[-1] + sz[-1]) if count else 0
    buf = np.full(max((total + align - 1) // align * align, align), 0x20, dtype=np.uint8)
    codes = np.array([KIND_CODES[k] for k in kinds], dtype=np.uint8)
    seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
    rc = N.gen().pars_gen_docs(count, codes.ctypes.data, seeds.ctypes.data, sz.ctypes.data, offs.ctypes.data,
                               buf.ctypes.data)
    if rc:
        raise RuntimeError(f"pars_gen_docs failed ({rc})")
    spans = (N.DocSpan * count)()
    for i in range(count):
        spans[i].offset = int(offs[i])
        spans[i].size = int(sz[i])
    return buf, spans


def batch(seed: int, count: int, mix: dict[str, int] | None = None, **size_kw) -> dict:
    """The batched workload: `count` documents with log-normal sizes, kinds by
    `mix`; grouped by kind (one batch per plan), each group generated into its
    packed buffer.  Returns {kind: (buffer, spans, doc indices)} plus "sizes"
    and "kinds"."""
    mix = mix or {"json": 70, "csv": 30}
    sz = sizes(seed, count, **size_kw)
    kinds = doc_kinds(seed, count, mix)
    seeds = np.array([doc_seed(seed, i) for i in range(count)], dtype=np.uint64)
    out: dict = {"sizes": sz, "kinds": kinds}
    karr = np.array(kinds)
    for k in mix:
        idx = np.nonzero(karr == k)[0]
        buf, spans = pack_generated([k] * idx.size, seeds[idx], sz[idx])
        out[k] = (buf, spans, idx.tolist())
    return out
