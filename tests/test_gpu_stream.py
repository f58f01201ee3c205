"""GPU parity of the streaming host boundary (pars_hip_scan_stream): host bytes
staged chunk by chunk through pinned buffers, uploaded and scanned with the
carry chained on the device, R6 over the whole document - every output equal
to the CPU oracle on the whole document (so equal to scan_device), for chunks
far smaller than the document (strings, comments and brackets open across
chunk boundaries), chunk sizes that do not divide the document, one chunk,
every evaluator path, and adversarial random-alphabet inputs."""
from __future__ import annotations

import numpy as np
import pytest

from pars_amd import Plan, Scanner
from pars_amd import _native as N
from pars_amd import workloads as W
from tests.gpu_util import d2h, say
from tests.test_gpu_parity import compare

pytestmark = pytest.mark.gpu


def stream_full(sc: Scanner, data: np.ndarray, chunk: int, threads: int = 4, want_pairs: bool = True) -> dict:
    r = sc.scan_stream(data, chunk_bytes=chunk, copy_threads=threads, want_pairs=want_pairs)
    out = {
        "n_struct": r.n_struct,
        "n_brackets": r.n_brackets,
        "positions": d2h(r.structural, r.n_struct, np.uint32),
        "bracket_pos": d2h(r.bracket_pos, r.n_brackets, np.uint32),
        "bracket_char": d2h(r.bracket_char, r.n_brackets, np.uint8),
        "balanced": r.balanced,
        "well_formed": r.well_formed,
        "max_depth": r.max_depth,
        "unclosed_mode": r.unclosed_mode,
        "pair": d2h(r.bracket_pair, r.n_brackets, np.int32),
    }
    if r.chars:
        out["chars"] = d2h(r.chars, r.n_struct, np.uint8)
    if r.struct_bitmap:
        out["struct_bitmap"] = d2h(r.struct_bitmap, (data.size + 63) // 64, np.uint64)
    if want_pairs:
        out["pair_pos"] = d2h(r.pair_pos, r.n_struct, np.int32)
    return out


GEN = {
    "json": lambda n: W.json(31, n, 16, 0.3),
    "csv": lambda n: W.csv(32, n),
    "clike6": lambda n: W.clike(33, n, 12),
    "clike": lambda n: W.clike(34, n, 12),
}


@pytest.mark.parametrize("g", ["json", "csv", "clike6", "clike"])
@pytest.mark.parametrize("chunk", [64 << 10, 192 << 10, 1 << 20])
def test_stream_generated(torch_cuda, g, chunk):
    d = GEN[g]((3 << 20) + 12345)
    sc = Scanner(g)
    sc.set_option(N.PARS_OPT_WANT_CHARS, 1)
    say(f"[gpu] stream {g} n={d.size} chunk={chunk}")
    compare(g, d, stream_full(sc, d, chunk), f"stream chunk={chunk}")


@pytest.mark.parametrize("g", ["json", "csv", "clike6", "toml", "xml"])
def test_stream_adversarial(torch_cuda, g):
    d = W.alpha(900 + len(g), (1 << 20) + 77, W.ALPHABETS[g])
    sc = Scanner(g)
    say(f"[gpu] stream adversarial {g}")
    compare(g, d, stream_full(sc, d, 64 << 10, threads=2), "stream adversarial")


def test_stream_small_and_single_chunk(torch_cuda):
    sc = Scanner("json")
    for n in (1, 63, 64, 4097, 70000):
        d = W.json(35, n, 8, 0.3)
        say(f"[gpu] stream small n={n}")
        compare("json", d, stream_full(sc, d, 0, threads=1), f"stream n={n}")


def test_stream_equals_device_scan_and_repeats(torch_cuda):
    torch = torch_cuda
    d = GEN["json"](5 << 20)
    sc = Scanner("json")
    a = stream_full(sc, d, 256 << 10)
    b = stream_full(sc, d, 128 << 10, threads=8)  # again on the same scanner, other chunking
    for k in ("positions", "bracket_pos", "pair", "pair_pos", "struct_bitmap"):
        assert np.array_equal(a[k], b[k]), k
    compare("json", d, a, "stream vs oracle")
