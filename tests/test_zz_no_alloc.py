"""The allocation-free entry point (pars_hip_scan_device_noalloc) through the
C ABI against pars_hip_scan_device on the same inputs, bit for bit: structural
positions and bitmap, bracket positions, bytes and pairs, counts and flags.
Calls run back to back on one GPU, each after a larger one (the static
scratch keeps the previous call's contents), covering the escape and parity
evaluators, the general path, the stack R6 and its sorting fallback on an
unbalanced document; outputs too small for a document are refused."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from pars_amd import Plan, Scanner
from pars_amd import _native as N
from pars_amd import workloads as W
from tests.gpu_util import d2h, gpu_full, say

pytestmark = pytest.mark.gpu


def noalloc(torch, sc: Scanner, data: np.ndarray, cap_struct: int | None = None, cap_br: int | None = None):
    n = data.size
    t = torch.from_numpy(np.ascontiguousarray(data)).to("cuda")
    cap_struct = n if cap_struct is None else cap_struct
    cap_br = min(n, 1 << 20) if cap_br is None else cap_br
    pos = torch.empty(max(cap_struct, 1), dtype=torch.int32, device="cuda")
    bm = torch.empty(max((n + 63) // 64, 1), dtype=torch.int64, device="cuda")
    bpos = torch.empty(max(cap_br, 1), dtype=torch.int32, device="cuda")
    bch = torch.empty(max(cap_br, 1), dtype=torch.uint8, device="cuda")
    pr = torch.empty(max(cap_br, 1), dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    r = N.DeviceResult()
    rc = N.lib().pars_hip_scan_device_noalloc(
        sc._h, C.c_void_p(t.data_ptr()), n, C.c_void_p(torch.cuda.current_stream().cuda_stream), 1,
        C.c_void_p(pos.data_ptr()), cap_struct, C.c_void_p(bm.data_ptr()), C.c_void_p(bpos.data_ptr()),
        C.c_void_p(bch.data_ptr()), C.c_void_p(pr.data_ptr()), cap_br, C.byref(r))
    if rc:
        return rc, N.lib().pars_hip_last_error().decode()
    out = {
        "n_struct": r.n_struct, "n_brackets": r.n_brackets, "balanced": bool(r.balanced),
        "well_formed": bool(r.well_formed), "max_depth": r.max_depth, "unclosed_mode": bool(r.unclosed_mode),
        "positions": d2h(r.structural, r.n_struct, np.uint32),
        "struct_bitmap": d2h(r.struct_bitmap, (n + 63) // 64, np.uint64),
        "bracket_pos": d2h(r.bracket_pos, r.n_brackets, np.uint32),
        "bracket_char": d2h(r.bracket_char, r.n_brackets, np.uint8),
        "pair": d2h(r.bracket_pair, r.n_brackets, np.int32),
    }
    return 0, out


def check(torch, plan: str, data: np.ndarray, tag: str):
    a = Scanner(Plan.load(plan))
    b = Scanner(Plan.load(plan))
    rc, got = noalloc(torch, b, data)
    assert rc == 0, f"{tag}: {got}"
    ref = gpu_full(torch, a, data, want_pairs=False)
    for k in ("n_struct", "n_brackets", "balanced", "well_formed", "max_depth", "unclosed_mode"):
        assert got[k] == ref[k], f"{tag}: {k} {got[k]} vs {ref[k]}"
    for k in ("positions", "struct_bitmap", "bracket_pos", "bracket_char", "pair"):
        assert np.array_equal(got[k], ref[k]), f"{tag}: {k}"
    exp = O.full(Plan.load(plan).blob, data)
    assert np.array_equal(got["pair"], exp["pair"]) and got["well_formed"] == exp["well_formed"], tag
    a.close()
    b.close()


def test_no_alloc_equals_scan_device(torch_cuda):
    torch = torch_cuda
    big = W.json(21, 12 << 20, 32, 0.30)  # ~450k brackets: stack path across chunks
    say(f"[gpu] noalloc json {big.size} B")
    check(torch, "json", big, "json 12 MiB")
    small = W.json(22, 300_000, 8)
    say(f"[gpu] noalloc json {small.size} B (second, smaller)")
    check(torch, "json", small, "json 300 KB")
    rng = np.random.default_rng(3)
    unb = np.frombuffer(bytes(rng.choice(list(b"[]{}a ,"), 200_000)), np.uint8).copy()
    say(f"[gpu] noalloc unbalanced soup {unb.size} B")
    check(torch, "json", unb, "unbalanced")
    csv = W.csv(23, 4 << 20)
    say(f"[gpu] noalloc csv {csv.size} B")
    check(torch, "csv", csv, "csv 4 MiB")
    src = W.clike(24, 2 << 20, 24)
    say(f"[gpu] noalloc clike6 {src.size} B")
    check(torch, "clike6", src, "clike6 2 MiB")
    tiny = np.frombuffer(b'{"a":[1,2]}', np.uint8).copy()
    check(torch, "json", tiny, "tiny")


def test_no_alloc_refuses_small_outputs(torch_cuda):
    torch = torch_cuda
    sc = Scanner("json")
    d = W.json(25, 1 << 20, 8)
    rc, msg = noalloc(torch, sc, d, cap_struct=1000)
    assert rc == 1 and "outputs too small" in msg
    rc, got = noalloc(torch, sc, d)  # the next call is whole again
    assert rc == 0 and got["n_struct"] == O.scan(Plan.load("json").blob, d)["positions"].size
    sc.close()
