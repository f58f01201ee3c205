"""CPU tests of the C-ABI boundary: the library loads and exports every entry
point include/pars_hip.h declares; plans are the reference's 3177-byte layout."""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np
import pytest

from pars_amd import Plan, _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols() -> list[str]:
    names = set()
    for h in ("pars_hip.h",):
        text = open(os.path.join(ROOT, "include", h)).read()
        names |= set(re.findall(r"\b(pars_hip_\w+)\s*\(", text))
    return sorted(names)


def test_header_declares_the_boundary():
    names = declared_symbols()
    for required in ("pars_hip_scanner_create", "pars_hip_scan_device", "pars_hip_scan_host",
                     "pars_hip_reserve_struct", "pars_hip_scan_chunk", "pars_hip_struct_ptr"):
        assert required in names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(N.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binds_every_declared_symbol():
    assert set(declared_symbols()) <= set(N.PROTOTYPES)


def test_version_string():
    assert b"gfx950" in N.lib().pars_hip_version()


@pytest.mark.parametrize("g", ["json", "csv", "clike", "clike6", "toml", "ini", "xml"])
def test_plan_blobs_have_scan_plan_layout(g):
    p = Plan.load(g)
    assert len(p.blob) == 3177
    assert 1 <= p.n_descs <= 8
    for d in range(p.n_descs):
        desc = p.desc(d)
        assert 1 <= desc["n_states"] <= 8 and 1 <= desc["n_classes"] <= 8


def test_plan_shapes_match_reference_static_asserts():
    # tests/test_model.cpp:32-40, 52-53, 74-75
    j, c, l = Plan.load("json"), Plan.load("csv"), Plan.load("clike")
    assert len(j.brackets) == 2 and len(j.punctuation) == 2 and j.desc(0)["kind"] == 0
    assert len(c.brackets) == 0 and len(c.punctuation) == 2 and c.desc(0)["kind"] == 2
    assert len(l.brackets) == 1 and len(l.punctuation) == 1 and any(l.desc(d)["kind"] == 3 for d in range(l.n_descs))
    six = Plan.load("clike6")
    assert six.n_descs == 5 and sum(six.desc(d)["n_states"] for d in range(5)) == 18  # > the reference's 16


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(N, "_lib", None)
    monkeypatch.setattr(N, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        N.lib()


def test_cpp_facade_builds_and_links():
    """include/pars_hip.hpp compiles against the C ABI (scanner<S> with a
    compile-time plan, runtime_scanner with a loaded one) and links to the
    in-tree library."""
    exe = os.path.join(ROOT, "tests", "cpp", "facade_check")
    assert os.access(exe, os.X_OK)
    import subprocess

    out = subprocess.run(["ldd", exe], stdout=subprocess.PIPE, text=True).stdout
    assert "libpars_hip.so" in out and "not found" not in out


def test_pack_batch_layout():
    """Documents start on 64-byte boundaries, in order, with spaces between."""
    from pars_amd import pack_batch

    docs = [b"x" * n for n in (1, 63, 64, 65, 200)]
    buf, spans = pack_batch(docs)
    assert buf.size % 64 == 0
    prev = 0
    for d, s in zip(docs, spans):
        assert s.offset % 64 == 0 and s.offset >= prev and s.size == len(d)
        assert bytes(buf[s.offset:s.offset + s.size]) == d
        prev = s.offset + s.size
    assert set(bytes(buf[1:64])) == {0x20}


def test_batch_workload_is_deterministic():
    from pars_amd import workloads as W

    a = W.batch(4, 50)
    b = W.batch(4, 50)
    assert a["kinds"] == b["kinds"]
    assert np.array_equal(a["sizes"], b["sizes"])
    for k in ("json", "csv"):
        assert np.array_equal(a[k][0], b[k][0])
        assert len(a[k][1]) == len(a[k][2])


def test_server_errors_reach_last_error():
    """A server call that fails before touching the device records its message
    where pars_hip_last_error() reads it (one error state for the library)."""
    lib = N.lib()
    cfg = N.ServerConfig(None, 0, 0, 0, 0, 0, 0, 0, 0)
    h = C.c_void_p()
    assert lib.pars_hip_server_create(C.byref(cfg), C.byref(h)) == 1  # PARS_ERR_ARGUMENT
    assert b"at least one plan" in lib.pars_hip_last_error()
    assert lib.pars_hip_server_submit(None, 0, None, 0, C.byref(C.c_uint64())) == 1
    assert b"submit" in lib.pars_hip_last_error()


def test_no_alloc_limits_refused_before_the_device():
    """pars_hip_scan_device_noalloc refuses sizes beyond its scratch with a
    clear error, before any device call (so this runs without a GPU)."""
    lib = N.lib()
    r = N.DeviceResult()
    dummy = C.c_void_p(1)
    rc = lib.pars_hip_scan_device_noalloc(dummy, dummy, (32 << 20) + 1, None, 1, dummy, 1, None, dummy, dummy,
                                          dummy, 1, C.byref(r))
    assert rc == 1 and b"at most" in lib.pars_hip_last_error()
    rc = lib.pars_hip_scan_device_noalloc(dummy, dummy, 1024, None, 1, dummy, 1024, None, dummy, dummy, dummy,
                                          (1 << 20) + 1, C.byref(r))
    assert rc == 1 and b"brackets" in lib.pars_hip_last_error()


def test_budget_accounting_without_a_device():
    """pars_hip_budget_*: host-only objects (no HIP call)."""
    lib = N.lib()
    h = C.c_void_p()
    assert lib.pars_hip_budget_create(1 << 20, C.byref(h)) == N.PARS_OK and h.value
    u, p, lim = C.c_uint64(7), C.c_uint64(7), C.c_uint64(0)
    assert lib.pars_hip_budget_usage(h, C.byref(u), C.byref(p), C.byref(lim)) == N.PARS_OK
    assert (u.value, p.value, lim.value) == (0, 0, 1 << 20)
    assert lib.pars_hip_budget_usage(h, None, None, None) == N.PARS_OK  # every output optional
    assert lib.pars_hip_budget_usage(None, C.byref(u), None, None) == 1  # PARS_ERR_ARGUMENT
    assert lib.pars_hip_budget_create(5, None) == 1
    assert lib.pars_hip_budget_destroy(h) == N.PARS_OK
    assert lib.pars_hip_budget_destroy(None) == N.PARS_OK


def test_allocator_needs_both_callbacks():
    lib = N.lib()
    a = N.Allocator(N.ALLOC_FN(lambda ctx, n: None), N.RELEASE_FN(), None)  # release missing
    out = C.c_void_p()
    assert lib.pars_hip_scanner_create_ex(Plan.load("json").blob, C.byref(a), None, C.byref(out)) == 1
    assert not out.value
    assert b"allocator" in lib.pars_hip_last_error()
