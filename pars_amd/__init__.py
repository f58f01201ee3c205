"""pars_amd — MI355X (gfx950) backend for pars scan plans.

Python mirror of the reference backend's boundary
(include/pars/backend/cuda/pipeline.cuh:1164-1281): a `Scanner` per plan with
`scan_device`, `scan_host`, `reserve_struct`, `scan_chunk` and `struct_ptr`,
and the free functions `scan_device` / `scan_host` / `scan_chunk` that keep one
scanner per plan per thread (pipeline.cuh:1260-1281).  Everything calls the C
ABI in include/pars_hip.h (libpars_hip.so); there is no CPU fallback: when
the library is missing the first call raises.

Device memory is passed as raw device pointers (ints); any allocator works
(the tests and bench use torch tensors' data_ptr()).
"""
from __future__ import annotations

import ctypes as C
import os
import sys
import threading

from . import _native as N

__all__ = [
    "ParsError",
    "Plan",
    "Scanner",
    "DeviceScan",
    "DeviceTape",
    "DeviceKvTable",
    "replay",
    "trim",
    "StructuralCursor",
    "HostScan",
    "ChunkCarry",
    "scan_device",
    "scan_host",
    "scan_chunk",
    "PLAN_DIR",
]

PLAN_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "plans")
ChunkCarry = N.ChunkCarry


class ParsError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{N.STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


def _check(status: int) -> None:
    if status != N.PARS_OK:
        raise ParsError(status, N.lib().pars_hip_last_error().decode(errors="replace"))


class Plan:
    """A compiled scan plan: the 3177 bytes of pars::scan_plan (include/pars/plan.hpp:39-79)."""

    def __init__(self, blob: bytes, name: str = "plan"):
        if len(blob) != N.PLAN_BYTES:
            raise ValueError(f"a plan is {N.PLAN_BYTES} bytes, got {len(blob)}")
        self.blob = bytes(blob)
        self.name = name

    @classmethod
    def load(cls, name_or_path: str) -> "Plan":
        path = name_or_path
        if not os.path.exists(path):
            path = os.path.join(PLAN_DIR, name_or_path + ".plan")
        with open(path, "rb") as f:
            return cls(f.read(), os.path.splitext(os.path.basename(path))[0])

    # plan.hpp field views
    @property
    def punctuation(self) -> bytes:
        return self.blob[0:8][: self.blob[44]]

    @property
    def brackets(self) -> list[tuple[int, int]]:
        nb = self.blob[45]
        return [(self.blob[8 + 2 * i], self.blob[9 + 2 * i]) for i in range(nb)]

    @property
    def n_descs(self) -> int:
        return self.blob[48 + 8 * 391]

    def desc(self, d: int) -> dict:
        o = 48 + 391 * d
        b = self.blob
        return {
            "kind": b[o],
            "n_states": b[o + 1],
            "n_classes": b[o + 2],
            "start": b[o + 3],
            "ok": b[o + 4],
            "delim": b[o + 5],
            "escape": b[o + 6],
            "rep_of": b[o + 7 : o + 263],
            "trans": [list(b[o + 263 + 8 * s : o + 271 + 8 * s]) for s in range(8)],
            "masked": [list(b[o + 327 + 8 * s : o + 335 + 8 * s]) for s in range(8)],
        }

    def __repr__(self) -> str:
        return f"Plan({self.name!r}, descs={self.n_descs}, punct={self.punctuation!r}, brackets={self.brackets})"


class DeviceScan:
    """pipeline.cuh:722-728 device_result; pointers alias the scanner's arena."""

    def __init__(self, r: N.DeviceResult, n: int):
        self.n = n
        self.structural = r.structural or 0
        self.pair_pos = r.pair_pos or 0
        self.n_struct = r.n_struct
        self.balanced = bool(r.balanced)
        self.well_formed = bool(r.well_formed)
        self.max_depth = r.max_depth
        self.chars = r.chars or 0
        self.struct_bitmap = r.struct_bitmap or 0
        self.bracket_pos = r.bracket_pos or 0
        self.bracket_char = r.bracket_char or 0
        self.bracket_pair = r.bracket_pair or 0
        self.n_brackets = r.n_brackets
        self.unclosed_mode = bool(r.unclosed_mode)

    def __repr__(self) -> str:
        return (
            f"DeviceScan(n_struct={self.n_struct}, n_brackets={self.n_brackets}, balanced={self.balanced}, "
            f"well_formed={self.well_formed}, max_depth={self.max_depth}, unclosed={self.unclosed_mode})"
        )


class DeviceTape:
    """The JSON tape (pars_hip_json_tape): device pointers into the scanner's
    arena, valid until its next call; `fail_event` is -1 when the hook
    accepted every event."""

    def __init__(self, r: N.JsonTape, n: int):
        self.n = n
        self.kind = r.kind or 0
        self.begin = r.begin or 0
        self.end = r.end or 0
        self.str_off = r.str_off or 0
        self.arena = r.arena or 0
        self.structural = r.structural or 0
        self.n_nodes = r.n_nodes
        self.arena_bytes = r.arena_bytes
        self.n_struct = r.n_struct
        self.fail_event = r.fail_event
        self.ok = bool(r.ok)

    def __repr__(self) -> str:
        return (f"DeviceTape(n_nodes={self.n_nodes}, arena_bytes={self.arena_bytes}, ok={self.ok}, "
                f"fail_event={self.fail_event})")


class DeviceKvTable:
    """An INI / TOML key-value table (pars_hip_kv_parse): device pointers into
    the scanner's arena, valid until its next call.  Per token: role (uint8,
    N.PARS_KV_*), begin and end (uint32 byte offsets), and for TOML key (int32:
    a VALUE's key token, -1 otherwise)."""

    def __init__(self, r: N.KvTable, fmt: str):
        self.format = fmt
        self.role = r.role or 0
        self.begin = r.begin or 0
        self.end = r.end or 0
        self.key = r.key or 0
        self.structural = r.structural or 0
        self.n_tokens = r.n_tokens
        self.n_struct = r.n_struct
        self.n_role = list(r.n_role)

    def __repr__(self) -> str:
        return f"DeviceKvTable({self.format}, n_tokens={self.n_tokens}, n_role={self.n_role})"


# ---------------------------------------------------------------------------
# Consumer side (host): the structural index handed to user code as the
# reference's model::replay (include/pars/model/visit.hpp:35-48) and
# model::structural_cursor (include/pars/model/cursor.hpp:16-66) hand it; the
# same functions as pars::hip::replay / structural_cursor in pars_hip.hpp.
# ---------------------------------------------------------------------------
_WS = b" \t\n\r"


def replay(base, positions, visitor, pairs=None) -> None:
    """One pass over the input as a covering, non-overlapping event stream:
    visitor.scalar(begin, end) for every non-empty run of bytes between
    structural bytes (and before the first / after the last), and
    visitor.symbol(offset, byte, partner offset or -1) for every structural
    byte.  `pairs`: per structural, its partner's structural index or -1."""
    mv = memoryview(base).cast("B")
    nxt = 0
    for i, at in enumerate(positions):
        at = int(at)
        if at > nxt:
            visitor.scalar(nxt, at)
        partner = -1
        if pairs is not None and int(pairs[i]) >= 0:
            partner = int(positions[int(pairs[i])])
        visitor.symbol(at, mv[at], partner)
        nxt = at + 1
    if len(mv) > nxt:
        visitor.scalar(nxt, len(mv))


def trim(s: bytes) -> bytes:
    """ASCII whitespace (space, tab, newline, carriage return) removed from both ends."""
    return bytes(s).strip(_WS)


class StructuralCursor:
    """Random access to a structural index: offset, byte and partner of
    structural i, and the run of bytes just before it (its gap)."""

    def __init__(self, base, positions, pairs=None):
        self.base = memoryview(base).cast("B")
        self.pos = positions
        self.pair = pairs
        self.n = len(positions)

    def valid(self, i: int) -> bool:
        return 0 <= i < self.n

    def off(self, i: int) -> int:
        return int(self.pos[i])

    def ch(self, i: int) -> int:
        return self.base[int(self.pos[i])]

    def match_index(self, i: int) -> int:
        return -1 if self.pair is None else int(self.pair[i])

    def gap_begin(self, i: int) -> int:
        return 0 if i == 0 else int(self.pos[i - 1]) + 1

    def gap_end(self, i: int) -> int:
        return int(self.pos[i])

    def gap(self, i: int) -> bytes:
        return bytes(self.base[self.gap_begin(i):self.gap_end(i)])


class HostScan:
    """pipeline.cuh:730-735 host_result (arrays copied out of the pinned buffers)."""

    def __init__(self, r: N.HostResult):
        import numpy as np

        n = r.n_struct
        self.structural = np.ctypeslib.as_array(r.structural, (n,)).copy() if n else np.zeros(0, np.uint32)
        if r.pair_pos:
            self.pair_pos = np.ctypeslib.as_array(r.pair_pos, (n,)).copy() if n else np.zeros(0, np.int32)
        else:
            self.pair_pos = None
        self.balanced = bool(r.balanced)
        self.well_formed = bool(r.well_formed)
        self.max_depth = r.max_depth
        self.unclosed_mode = bool(r.unclosed_mode)


class BatchScan:
    """pars_batch_result: device pointers into the scanner's arena, valid until its next call."""

    def __init__(self, r: N.BatchResult):
        self.n_docs = r.n_docs
        self.n_struct = r.n_struct
        self.n_brackets = r.n_brackets
        self.structural = r.structural or 0
        self.struct_start = r.struct_start or 0
        self.bracket_pos = r.bracket_pos or 0
        self.bracket_char = r.bracket_char or 0
        self.bracket_start = r.bracket_start or 0
        self.bracket_pair = r.bracket_pair or 0
        self.pair_pos = r.pair_pos or 0
        self.doc_flags = r.doc_flags or 0
        self.max_depth = r.max_depth or 0
        self.struct_bitmap = r.struct_bitmap or 0

    def __repr__(self) -> str:
        return f"BatchScan(n_docs={self.n_docs}, n_struct={self.n_struct}, n_brackets={self.n_brackets})"


def pack_batch(docs, align: int = 64):
    """Lay documents out for pars_hip_scan_batch: each starts on a 64-byte
    boundary; the gaps are spaces.  Returns (buffer uint8 array, DocSpan array)."""
    import numpy as np

    sizes = np.array([len(d) for d in docs], dtype=np.uint64)
    offs = np.zeros(len(docs), dtype=np.uint64)
    if len(docs) > 1:
        padded = (sizes + (align - 1)) // align * align
        offs[1:] = np.cumsum(padded)[:-1]
    total = int(offs[-1] + sizes[-1]) if len(docs) else 0
    total = (total + align - 1) // align * align
    buf = np.full(max(total, align), 0x20, dtype=np.uint8)
    for o, d in zip(offs, docs):
        buf[int(o):int(o) + len(d)] = np.frombuffer(bytes(d), dtype=np.uint8) if not isinstance(d, np.ndarray) else d
    spans = (N.DocSpan * len(docs))()
    for i in range(len(docs)):
        spans[i].offset = int(offs[i])
        spans[i].size = int(sizes[i])
    return buf, spans


class Budget:
    """A limit on the device memory its scanners and servers hold at once
    (pars_hip_budget_*): growth past it fails with PARS_ERR_MEMORY, a server's
    batches wait for it."""

    _teardown_order = 2  # after the scanners and servers that draw on it

    def __init__(self, limit_bytes: int):
        self._lib = N.lib()
        h = C.c_void_p()
        _check(self._lib.pars_hip_budget_create(limit_bytes, C.byref(h)))
        self._h = h
        N.track(self)

    def usage(self) -> dict:
        u, p, lim = C.c_uint64(), C.c_uint64(), C.c_uint64()
        _check(self._lib.pars_hip_budget_usage(self._h, C.byref(u), C.byref(p), C.byref(lim)))
        return {"in_use": u.value, "peak": p.value, "limit": lim.value}

    def close(self) -> None:
        if getattr(self, "_h", None):
            _check(self._lib.pars_hip_budget_destroy(self._h))
            self._h = None

    def __del__(self):
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


class TorchAllocator:
    """A pars_allocator over PyTorch's caching allocator, so a scanner's
    buffers come from (and return to) the pool the rest of the program uses.
    `fail_after` (for tests) makes every allocation past that many return NULL."""

    def __init__(self, fail_after: int | None = None):
        import torch

        self._torch = torch
        self.fail_after = fail_after
        self.allocs = self.releases = 0
        self.live: dict[int, int] = {}
        self._alloc = N.ALLOC_FN(self._on_alloc)
        self._release = N.RELEASE_FN(self._on_release)
        self.struct = N.Allocator(self._alloc, self._release, None)

    def _on_alloc(self, ctx, n):
        if self.fail_after is not None and self.allocs >= self.fail_after:
            return None
        p = self._torch.cuda.caching_allocator_alloc(int(n))
        self.allocs += 1
        self.live[p] = int(n)
        return p

    def _on_release(self, ctx, p, n):
        self.releases += 1
        assert self.live.pop(p) == int(n), "released with another size"
        self._torch.cuda.caching_allocator_delete(p)


class Scanner:
    """gpu::scanner<S> (pipeline.cuh:1164-1258) for a runtime plan; device
    memory from `allocator` (default hipMalloc), charged to `budget`."""

    def __init__(self, plan: Plan | str | bytes, allocator: "TorchAllocator | None" = None,
                 budget: Budget | None = None):
        if isinstance(plan, str):
            plan = Plan.load(plan)
        elif isinstance(plan, (bytes, bytearray)):
            plan = Plan(bytes(plan))
        self.plan = plan
        self._lib = N.lib()
        self._allocator, self._budget = allocator, budget  # kept alive as long as the scanner
        h = C.c_void_p()
        if allocator is None and budget is None:
            _check(self._lib.pars_hip_scanner_create(plan.blob, C.byref(h)))
        else:
            _check(self._lib.pars_hip_scanner_create_ex(
                plan.blob, C.byref(allocator.struct) if allocator is not None else None,
                budget._h if budget is not None else None, C.byref(h)))
        self._h = h
        N.track(self)

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.pars_hip_scanner_destroy(self._h)
            self._h = None

    def __del__(self):
        # at interpreter shutdown the exit handler has closed every handle
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_option(self, option: int, value: int) -> None:
        _check(self._lib.pars_hip_set_option(self._h, option, value))

    def scan_device(
        self, d_data: int, n: int, stream: int | None = None, want_pairs: bool = False,
        run_pairs: bool = True, stage_times: N.StageTimes | None = None,
    ) -> DeviceScan:
        r = N.DeviceResult()
        _check(
            self._lib.pars_hip_scan_device(
                self._h, C.c_void_p(d_data), n, C.c_void_p(stream or 0), int(want_pairs), int(run_pairs),
                C.byref(stage_times) if stage_times is not None else None, C.byref(r),
            )
        )
        return DeviceScan(r, n)

    def scan_host(self, data: bytes, stream: int | None = None, want_pairs: bool = False) -> HostScan:
        r = N.HostResult()
        _check(self._lib.pars_hip_scan_host(self._h, bytes(data), len(data), C.c_void_p(stream or 0),
                                            int(want_pairs), C.byref(r)))
        return HostScan(r)

    def scan_stream(self, data, chunk_bytes: int = 0, copy_threads: int = 0, stream: int | None = None,
                    want_pairs: bool = False, stage_times: N.StageTimes | None = None) -> DeviceScan:
        """Host bytes streamed through pinned staging in chunks, scanned as they
        arrive (pars_hip_scan_stream); results on the device as scan_device."""
        import numpy as np

        b = data if isinstance(data, np.ndarray) else np.frombuffer(bytes(data), dtype=np.uint8)
        b = np.ascontiguousarray(b, dtype=np.uint8)
        r = N.DeviceResult()
        _check(self._lib.pars_hip_scan_stream(
            self._h, C.c_void_p(b.ctypes.data), b.size, chunk_bytes, copy_threads, C.c_void_p(stream or 0),
            int(want_pairs), C.byref(stage_times) if stage_times is not None else None, C.byref(r)))
        return DeviceScan(r, b.size)

    def json_tape(self, d_data: int, n: int, stream: int | None = None,
                  stage_times: N.StageTimes | None = None) -> DeviceTape:
        """R1-R6 and the JSON tape of one device-resident document
        (pars_hip_json_tape): the reference's replay-driven JSON DOM hook
        (bench/bench_parse.cu:336-348) built on the device."""
        r = N.JsonTape()
        _check(self._lib.pars_hip_json_tape(
            self._h, C.c_void_p(d_data), n, C.c_void_p(stream or 0),
            C.byref(stage_times) if stage_times is not None else None, C.byref(r)))
        return DeviceTape(r, n)

    def kv_parse(self, d_data: int, n: int, fmt: str, stream: int | None = None,
                 stage_times: N.StageTimes | None = None) -> DeviceKvTable:
        """R1-R5 and the key-value table of one device-resident INI or TOML
        document (pars_hip_kv_parse): the reference's stage-2 hooks for these
        formats (bench/bench_parse_host.cu:69-134, 202-248) built on the device."""
        code = {"ini": N.PARS_KV_INI, "toml": N.PARS_KV_TOML}[fmt]
        r = N.KvTable()
        _check(self._lib.pars_hip_kv_parse(
            self._h, C.c_void_p(d_data), n, code, C.c_void_p(stream or 0),
            C.byref(stage_times) if stage_times is not None else None, C.byref(r)))
        return DeviceKvTable(r, fmt)

    def reserve_struct(self, cap: int) -> None:
        _check(self._lib.pars_hip_reserve_struct(self._h, cap))

    def scan_chunk(self, d_chunk: int, n: int, carry: N.ChunkCarry, stream: int | None = None,
                   byte_base: int = 0) -> N.ChunkCarry:
        out = N.ChunkCarry()
        _check(self._lib.pars_hip_scan_chunk(self._h, C.c_void_p(d_chunk), n, C.byref(carry),
                                             C.c_void_p(stream or 0), byte_base, C.byref(out)))
        return out

    def pair_chunks(self, last: N.ChunkCarry, stream: int | None = None, want_pairs: bool = False) -> DeviceScan:
        r = N.DeviceResult()
        _check(self._lib.pars_hip_pair_chunks(self._h, C.byref(last), C.c_void_p(stream or 0),
                                              int(want_pairs), C.byref(r)))
        return DeviceScan(r, 0)

    def scan_batch(self, d_data: int, n: int, spans, stream: int | None = None, want_pairs: bool = False,
                   run_pairs: bool = True, stage_times: N.StageTimes | None = None) -> BatchScan:
        """Many documents in one launch (pars_hip_scan_batch); spans from pack_batch."""
        r = N.BatchResult()
        _check(self._lib.pars_hip_scan_batch(
            self._h, C.c_void_p(d_data), n, spans, len(spans), C.c_void_p(stream or 0), int(want_pairs),
            int(run_pairs), C.byref(stage_times) if stage_times is not None else None, C.byref(r)))
        return BatchScan(r)

    def get_option(self, option: int) -> int:
        v = C.c_int64()
        _check(self._lib.pars_hip_get_option(self._h, option, C.byref(v)))
        return v.value

    def autotune(self, d_data: int, n: int, spans=None, stream: int | None = None, reps: int = 3) -> dict:
        """Time the launch knobs on this input (a document, or a batch with
        spans) and keep the fastest setting (pars_hip_autotune)."""
        r = N.TuneResult()
        _check(self._lib.pars_hip_autotune(
            self._h, C.c_void_p(d_data), n, spans, len(spans) if spans is not None else 0,
            C.c_void_p(stream or 0), reps, C.byref(r)))
        return r.as_dict()

    def struct_ptr(self) -> int:
        return self._lib.pars_hip_struct_ptr(self._h) or 0


class Server:
    """The serving runtime (pars_hip_server_*): submit requests as they arrive,
    poll completed ones, release them when their outputs are no longer needed."""

    def __init__(self, plans, slots_per_plan: int = 2, ring_bytes: int = 4 << 30, staging_bytes: int = 1 << 30,
                 hbm_budget: int = 64 << 30, max_batch_bytes: int = 64 << 20, max_batch_docs: int = 16384,
                 run_pairs: bool = True, submit_threads: int = 4, budget: Budget | None = None):
        import numpy as np

        self.plans = [Plan.load(p) if isinstance(p, str) else p for p in plans]
        blob = b"".join(p.blob for p in self.plans)
        self._plan_buf = np.frombuffer(blob, dtype=np.uint8).copy()
        cfg = N.ServerConfig(self._plan_buf.ctypes.data, len(self.plans), slots_per_plan, ring_bytes,
                             staging_bytes, hbm_budget, max_batch_bytes, max_batch_docs, int(run_pairs),
                             submit_threads, 0, budget._h if budget is not None else None)
        self._budget = budget
        self._lib = N.lib()
        h = C.c_void_p()
        _check(self._lib.pars_hip_server_create(C.byref(cfg), C.byref(h)))
        self._h = h
        N.track(self)

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.pars_hip_server_destroy(self._h)
            self._h = None

    def __del__(self):
        # at interpreter shutdown the exit handler has closed every handle
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def submit(self, plan: int, data) -> int | None:
        """Request id, or None when the staging ring is full (poll, release, retry)."""
        import numpy as np

        b = np.ascontiguousarray(np.frombuffer(bytes(data), dtype=np.uint8) if not isinstance(data, np.ndarray)
                                 else data)
        rid = C.c_uint64()
        rc = self._lib.pars_hip_server_submit(self._h, plan, b.ctypes.data, b.size, C.byref(rid))
        if rc == 6:  # PARS_ERR_BUSY
            return None
        _check(rc)
        return rid.value

    def poll(self, cap: int = 256, wait_us: int = 0) -> list[N.RequestResult]:
        out = (N.RequestResult * cap)()
        k = C.c_size_t()
        _check(self._lib.pars_hip_server_poll(self._h, out, cap, C.byref(k), wait_us))
        return [out[i] for i in range(k.value)]

    def release(self, rid: int) -> None:
        _check(self._lib.pars_hip_server_release(self._h, rid))

    def stats(self) -> N.ServerStats:
        s = N.ServerStats()
        _check(self._lib.pars_hip_server_stats(self._h, C.byref(s)))
        return s

    def replay(self, plan_idx, docs, arrival_ns):
        """Submit docs[i] (numpy uint8) at arrival_ns[i] after the start; returns
        (latency_ns per request, elapsed_ns)."""
        import numpy as np

        n = len(docs)
        plans = np.ascontiguousarray(plan_idx, dtype=np.uint32)
        ptrs = np.array([d.ctypes.data for d in docs], dtype=np.uint64)
        sizes = np.array([d.size for d in docs], dtype=np.uint64)
        arr = np.ascontiguousarray(arrival_ns, dtype=np.uint64)
        lat = np.zeros(n, dtype=np.uint64)
        el = C.c_uint64()
        _check(self._lib.pars_hip_server_replay(self._h, n, plans.ctypes.data, ptrs.ctypes.data, sizes.ctypes.data,
                                                arr.ctypes.data, lat.ctypes.data, C.byref(el)))
        return lat, el.value


_tls = threading.local()


def _thread_scanner(plan: Plan) -> Scanner:
    cache = getattr(_tls, "scanners", None)
    if cache is None:
        cache = _tls.scanners = {}
    sc = cache.get(plan.blob)
    if sc is None:
        sc = cache[plan.blob] = Scanner(plan)
    return sc


def scan_device(plan: Plan, d_data: int, n: int, stream: int | None = None, want_pairs: bool = False) -> DeviceScan:
    """pipeline.cuh:1260-1266 (one scanner per plan per thread)."""
    return _thread_scanner(plan).scan_device(d_data, n, stream, want_pairs)


def scan_host(plan: Plan, data: bytes, want_pairs: bool = False) -> HostScan:
    """pipeline.cuh:1268-1272."""
    return _thread_scanner(plan).scan_host(data, None, want_pairs)


def scan_chunk(plan: Plan, d_chunk: int, n: int, carry: N.ChunkCarry, stream: int | None = None,
               byte_base: int = 0) -> N.ChunkCarry:
    """pipeline.cuh:1274-1281."""
    return _thread_scanner(plan).scan_chunk(d_chunk, n, carry, stream, byte_base)
