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


def batch(seed: int, count: int, mix: dict[str, int] | None = None, total_bytes: int | None = None,
          **size_kw) -> dict:
    """The batched workload: `count` documents with log-normal sizes, kinds by
    `mix`; grouped by kind (one batch per plan), each group packed for
    pars_hip_scan_batch.  Returns {kind: (buffer, spans, doc indices)} plus
    "sizes" and "kinds".  total_bytes, when given, rescales the median so the
    sizes sum close to it."""
    from . import pack_batch

    mix = mix or {"json": 70, "csv": 30}
    sz = sizes(seed, count, **size_kw)
    if total_bytes is not None:
        med = size_kw.get("median", 4096.0)
        for _ in range(3):
            scale = total_bytes / max(float(sz.sum()), 1.0)
            med *= scale
            sz = sizes(seed, count, **{**size_kw, "median": med})
    kinds = doc_kinds(seed, count, mix)
    out: dict = {"sizes": sz, "kinds": kinds}
    for k in mix:
        idx = [i for i in range(count) if kinds[i] == k]
        docs = [make_doc(k, doc_seed(seed, i), int(sz[i])) for i in idx]
        buf, spans = pack_batch(docs)
        out[k] = (buf, spans, idx)
    return out
