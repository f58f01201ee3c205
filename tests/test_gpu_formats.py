"""The reference's per-format benchmark documents (bench/bench_boundary.cu:
312-337, pars_amd/workloads.py format_doc) through the HIP path: at 8 MiB
every output bit-exact against the oracle for all six grammars (single
document, R1-R6 and R1-R5 only); at 512 MiB the structural count equals the
reference's published count for the same document (paper/plots/data/pars.csv
N1a rows), and R6 reports the document balanced and well formed."""
from __future__ import annotations

import numpy as np
import pytest

from pars_amd import Plan, Scanner
from pars_amd import workloads as W
from tests.gpu_util import gpu_full, say
from tests.test_format_counts import TOKENS
from tests.test_gpu_parity import compare

pytestmark = pytest.mark.gpu

FORMATS = ["json", "csv", "toml", "clike", "ini", "xml"]


@pytest.mark.parametrize("fmt", FORMATS)
def test_format_document_parity(torch_cuda, fmt):
    d = W.format_doc(fmt, 8 << 20)
    sc = Scanner(Plan.load(fmt))
    say(f"[gpu] format {fmt} 8 MiB ({d.size} B)")
    got = gpu_full(torch_cuda, sc, d, want_pairs=True)
    compare(fmt, d, got, f"format {fmt}")
    assert got["n_struct"] == TOKENS[(fmt, 8)]
    n1a = gpu_full(torch_cuda, sc, d, want_pairs=False, run_pairs=False)
    assert np.array_equal(n1a["positions"], got["positions"])
    sc.close()


@pytest.mark.parametrize("fmt", FORMATS)
def test_format_512mib_counts(torch_cuda, fmt):
    torch = torch_cuda
    d = W.format_doc(fmt, 512 << 20)
    t = torch.from_numpy(d).to("cuda")
    torch.cuda.synchronize()
    sc = Scanner(Plan.load(fmt))
    say(f"[gpu] format {fmt} 512 MiB ({d.size} B)")
    s = torch.cuda.current_stream().cuda_stream
    r = sc.scan_device(t.data_ptr(), d.size, s, want_pairs=False, run_pairs=False)
    assert r.n_struct == TOKENS[(fmt, 512)]
    r = sc.scan_device(t.data_ptr(), d.size, s, want_pairs=False, run_pairs=True)
    assert r.n_struct == TOKENS[(fmt, 512)] and r.balanced and r.well_formed
    sc.close()
    del t
