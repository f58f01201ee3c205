"""GPU tests of the launch knobs and the autotuner (pars_hip_autotune).

Every knob value the tuner may choose - look-back window, resident blocks per
CU, rows per tile, radix digit cap - must leave every output bit-exact against
the CPU oracle, for single documents and for batches; a digit cap too narrow
for a key range is refused with an error (not a wrong sort); and a tuned
scanner gives the same outputs as an untuned one.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from pars_amd import Plan, Scanner, pack_batch
from pars_amd import _native as N
from pars_amd import workloads as W
from tests.gpu_util import gpu_full, say, to_device
from tests.test_gpu_batch import check_batch, mixed_docs, run_batch
from tests.test_gpu_parity import compare

pytestmark = pytest.mark.gpu


def _json_doc(n: int = 3 << 20):
    return W.json(77, n, 20)


@pytest.mark.parametrize("window", [8, 16, 48, 64])
@pytest.mark.parametrize("g", ["json", "csv", "clike6"])
def test_lookback_window(torch_cuda, g, window):
    gen = {"json": lambda: _json_doc(), "csv": lambda: W.csv(5, 2 << 20),
           "clike6": lambda: W.clike(6, 2 << 20, 16)}[g]
    d = gen()
    sc = Scanner(g)
    sc.set_option(N.PARS_OPT_LOOKBACK_WINDOW, window)
    assert sc.get_option(N.PARS_OPT_LOOKBACK_WINDOW) == window
    say(f"[gpu] window {window} {g} n={d.size}")
    compare(g, d, gpu_full(torch_cuda, sc, d), f"window {window}")


@pytest.mark.parametrize("bpc", [1, 2])
@pytest.mark.parametrize("rows", [1, 4])
def test_blocks_per_cu(torch_cuda, bpc, rows):
    d = _json_doc()
    sc = Scanner("json")
    sc.set_option(N.PARS_OPT_BLOCKS_PER_CU, bpc)
    sc.set_option(N.PARS_OPT_ROWS_PER_TILE, rows)
    say(f"[gpu] blocks/CU {bpc} rows {rows}")
    compare("json", d, gpu_full(torch_cuda, sc, d), f"bpc {bpc} rows {rows}")


@pytest.mark.parametrize("bits", [2, 3, 5, 8])
def test_radix_digit_cap_document(torch_cuda, bits):
    d = _json_doc(1 << 20)  # depth up to 20: keys need 5 bits
    sc = Scanner("json")
    sc.set_option(N.PARS_OPT_RADIX_BITS, bits)
    say(f"[gpu] radix bits {bits}")
    compare("json", d, gpu_full(torch_cuda, sc, d), f"radix {bits}")


def test_radix_digit_cap_batch_and_refusal(torch_cuda):
    docs = mixed_docs("json", 61, 90, big=1)
    sc = Scanner("json")
    sc.set_option(N.PARS_OPT_RADIX_BITS, 4)  # (document, key) needs more than 8 bits here
    check_batch("json", docs, run_batch(torch_cuda, sc, docs), "radix 4")
    sc.set_option(N.PARS_OPT_RADIX_BITS, 1)  # more than four passes: refused
    with pytest.raises(Exception):
        run_batch(torch_cuda, sc, docs)
    sc.set_option(N.PARS_OPT_RADIX_BITS, 0)  # and the scanner recovers
    check_batch("json", docs, run_batch(torch_cuda, sc, docs), "radix default after refusal")


def test_option_bounds(torch_cuda):
    sc = Scanner("json")
    for opt, bad in [(N.PARS_OPT_LOOKBACK_WINDOW, 7), (N.PARS_OPT_LOOKBACK_WINDOW, 65),
                     (N.PARS_OPT_BLOCKS_PER_CU, -1), (N.PARS_OPT_RADIX_BITS, 9),
                     (N.PARS_OPT_WAVES_PER_BLOCK, 5), (N.PARS_OPT_WAVES_PER_BLOCK, 16),
                     (N.PARS_OPT_LAST_PAIRING, 1)]:
        with pytest.raises(Exception):
            sc.set_option(opt, bad)


def test_autotune_document_keeps_parity(torch_cuda):
    torch = torch_cuda
    big = W.json(78, 48 << 20, 24)
    sc = Scanner("json")
    t, ptr = to_device(torch, big)
    res = sc.autotune(ptr, big.size, stream=torch.cuda.current_stream().cuda_stream, reps=2)
    say(f"[gpu] autotune json 48 MiB: {res}")
    assert res["rows"] in (0, 1, 2, 3, 4) and res["lb_window"] in (0, 16, 32, 48, 64)
    assert 0 <= res["blocks_per_cu"] < max(res["max_blocks_per_cu"], 1) or res["blocks_per_cu"] == 0
    assert res["radix_bits"] in (0, 4, 5, 6, 7, 8)
    assert res["waves_per_block"] in (0, 4, 6, 8)
    assert sc.get_option(N.PARS_OPT_WAVES_PER_BLOCK) == res["waves_per_block"]
    assert res["trials"] > 10 and res["scan_ms"] > 0
    assert sc.get_option(N.PARS_OPT_LOOKBACK_WINDOW) == res["lb_window"]
    del t
    d = _json_doc()
    compare("json", d, gpu_full(torch, sc, d), "after autotune")


def test_autotune_batch_keeps_parity(torch_cuda):
    torch = torch_cuda
    docs = mixed_docs("json", 71, 300, big=2)
    buf, spans = pack_batch(docs)
    t, ptr = to_device(torch, buf)
    sc = Scanner("json")
    res = sc.autotune(ptr, buf.size, spans, stream=torch.cuda.current_stream().cuda_stream, reps=2)
    say(f"[gpu] autotune batch: {res}")
    assert res["trials"] > 10
    del t
    check_batch("json", docs, run_batch(torch, sc, docs), "after autotune")
