"""GPU tests of where scanners and servers get device memory
(pars_hip_scanner_create_ex, pars_hip_budget_*, pars_server_config.budget):
the reference's DeviceCache (pipeline.cuh:562-714) exits the process when an
allocation fails (:520-555); here a refused or failed allocation returns
PARS_ERR_MEMORY, the scanner stays usable, and a budget shared by scanners
and a server bounds what they hold together.  Results through a caller's
allocator or under a budget equal the oracle's."""
from __future__ import annotations

import gc

import numpy as np
import pytest

from oracle import oracle as O
from pars_amd import Budget, ParsError, Plan, Scanner, Server, TorchAllocator
from pars_amd import _native as N
from pars_amd import workloads as W
from tests.gpu_util import gpu_full, say
from tests.test_gpu_parity import compare

pytestmark = pytest.mark.gpu


def test_budget_refuses_growth_and_the_scanner_recovers(torch_cuda):
    doc = W.json(7, 4 << 20, 12)
    probe_budget = Budget(1 << 40)
    probe = Scanner("json", budget=probe_budget)
    gpu_full(torch_cuda, probe, doc)
    need = probe_budget.usage()["in_use"]  # what one scanner holds for this document
    probe.close()
    assert probe_budget.usage()["in_use"] == 0
    probe_budget.close()
    say(f"one scanner holds {need} bytes for a {doc.size}-byte document")

    b = Budget(need + need // 2)  # room for one scanner's buffers, not two
    a, c = Scanner("json", budget=b), Scanner("json", budget=b)
    compare("json", doc, gpu_full(torch_cuda, a, doc), "first scanner under the budget")
    with pytest.raises(ParsError) as e:
        gpu_full(torch_cuda, c, doc)
    assert e.value.status == 3 and "budget" in str(e.value)  # PARS_ERR_MEMORY
    u = b.usage()
    assert u["in_use"] <= u["limit"] and u["peak"] <= u["limit"]
    a.close()  # its buffers go back to the budget
    compare("json", doc, gpu_full(torch_cuda, c, doc), "second scanner once the first is gone")
    with pytest.raises(ParsError):
        b.close()  # still used by c
    c.close()
    assert b.usage()["in_use"] == 0
    b.close()


def test_buffers_from_the_torch_caching_allocator(torch_cuda):
    torch = torch_cuda
    doc = W.json(8, 3 << 20, 10)
    ta = TorchAllocator()
    gc.collect()  # tensors earlier tests left for the collector go now, not part way through
    torch.cuda.synchronize()
    before = torch.cuda.memory_allocated()
    sc = Scanner("json", allocator=ta)
    sc.set_option(N.PARS_OPT_WANT_CHARS, 1)
    got = gpu_full(torch, sc, doc)
    assert ta.allocs > 10 and torch.cuda.memory_allocated() > before + doc.size
    compare("json", doc, got, "torch allocator")
    sc.close()
    assert ta.releases == ta.allocs and not ta.live
    assert torch.cuda.memory_allocated() == before


def test_a_failing_allocator_returns_an_error_not_an_exit(torch_cuda):
    doc = W.json(9, 2 << 20, 10)
    for fail_after in (0, 3, 12):  # at creation, at the first buffers, part way through
        ta = TorchAllocator(fail_after=fail_after)
        say(f"allocator failing after {fail_after} allocations")
        try:
            sc = Scanner("json", allocator=ta)
        except ParsError as e:
            assert e.status == 3 and "allocator" in str(e)
            continue
        with pytest.raises(ParsError) as e:
            gpu_full(torch_cuda, sc, doc)
        assert e.value.status == 3
        ta.fail_after = None  # memory is back: the same scanner works
        compare("json", doc, gpu_full(torch_cuda, sc, doc), f"after a failure at {fail_after}")
        sc.close()
        assert ta.releases == ta.allocs and not ta.live


def test_server_and_scanner_share_a_budget(torch_cuda):
    from tests.test_gpu_server import check, request_docs, serve_all

    docs = request_docs(21, 400, max_size=100_000)
    ring = 64 << 20
    b = Budget(3 << 30)
    sv = Server(["json", "csv", "clike"], slots_per_plan=1, ring_bytes=ring, staging_bytes=ring,
                max_batch_bytes=4 << 20, budget=b)
    held = b.usage()["in_use"]
    assert held >= ring  # the rings and the slots' scanners
    sc = Scanner("json", budget=b)
    doc = W.json(10, 1 << 20, 8)
    compare("json", doc, gpu_full(torch_cuda, sc, doc), "scanner beside the server")
    res = serve_all(sv, docs, max_inflight=64)
    for i, (p, data) in enumerate(docs):
        check(["json", "csv", "clike"][p], data, res[i][0], f"request {i}")
    u = b.usage()
    assert u["peak"] <= u["limit"] and u["in_use"] >= held  # every request released: its reserve is back
    sv.close()
    sc.close()
    assert b.usage()["in_use"] == 0
    # rings larger than the budget: refused at creation
    small = Budget(32 << 20)
    with pytest.raises(ParsError) as e:
        Server(["json"], slots_per_plan=1, ring_bytes=ring, staging_bytes=ring, max_batch_bytes=4 << 20,
               budget=small)
    assert e.value.status == 5  # PARS_ERR_BUDGET
    assert small.usage()["in_use"] == 0
    small.close()
    b.close()
