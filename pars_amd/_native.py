"""ctypes view of include/pars_hip.h (structures and prototypes)."""
from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpars_hip.so")
GEN_PATH = os.path.join(HERE, "libpars_gen.so")

PLAN_BYTES = 3177

PARS_OK = 0
STATUS_NAMES = {
    0: "PARS_OK",
    1: "PARS_ERR_ARGUMENT",
    2: "PARS_ERR_PLAN",
    3: "PARS_ERR_MEMORY",
    4: "PARS_ERR_DEVICE",
    5: "PARS_ERR_BUDGET",
    6: "PARS_ERR_BUSY",
}

PARS_OPT_WANT_CHARS = 1
PARS_OPT_WANT_BITMAP = 2
PARS_OPT_PATH = 3
PARS_OPT_ROWS_PER_TILE = 4
PARS_OPT_GRAPH = 5
PARS_OPT_LOOKBACK_WINDOW = 6
PARS_OPT_BLOCKS_PER_CU = 7
PARS_OPT_RADIX_BITS = 8
PARS_OPT_CHASER = 9
PARS_OPT_LAST_PAIRING = 10  # read-only
PARS_OPT_WAVES_PER_BLOCK = 11
PARS_PAIRING_NONE, PARS_PAIRING_ONE_BLOCK, PARS_PAIRING_STACK, PARS_PAIRING_SORT = 0, 1, 2, 3


class StageTimes(C.Structure):
    _fields_ = [("p1", C.c_double), ("p2", C.c_double), ("p3", C.c_double), ("p4", C.c_double)]


class DeviceResult(C.Structure):
    _fields_ = [
        ("structural", C.c_void_p),
        ("pair_pos", C.c_void_p),
        ("n_struct", C.c_uint32),
        ("balanced", C.c_int32),
        ("well_formed", C.c_int32),
        ("max_depth", C.c_int32),
        ("chars", C.c_void_p),
        ("struct_bitmap", C.c_void_p),
        ("bracket_pos", C.c_void_p),
        ("bracket_char", C.c_void_p),
        ("bracket_pair", C.c_void_p),
        ("n_brackets", C.c_uint32),
        ("unclosed_mode", C.c_int32),
    ]


class JsonTape(C.Structure):
    _fields_ = [
        ("kind", C.c_void_p),
        ("begin", C.c_void_p),
        ("end", C.c_void_p),
        ("str_off", C.c_void_p),
        ("arena", C.c_void_p),
        ("structural", C.c_void_p),
        ("n_nodes", C.c_uint64),
        ("arena_bytes", C.c_uint64),
        ("n_struct", C.c_uint64),
        ("fail_event", C.c_int64),
        ("ok", C.c_int32),
        ("pad", C.c_int32),
    ]


class KvTable(C.Structure):
    _fields_ = [
        ("role", C.c_void_p),
        ("begin", C.c_void_p),
        ("end", C.c_void_p),
        ("key", C.c_void_p),
        ("structural", C.c_void_p),
        ("n_tokens", C.c_uint64),
        ("n_struct", C.c_uint64),
        ("n_role", C.c_uint64 * 5),
    ]


PARS_KV_INI, PARS_KV_TOML = 0, 1
PARS_KV_SECTION, PARS_KV_KEY, PARS_KV_VALUE, PARS_KV_ELEMENT, PARS_KV_SKIPPED = 0, 1, 2, 3, 4


class HostResult(C.Structure):
    _fields_ = [
        ("structural", C.POINTER(C.c_uint32)),
        ("pair_pos", C.POINTER(C.c_int32)),
        ("n_struct", C.c_uint32),
        ("balanced", C.c_int32),
        ("well_formed", C.c_int32),
        ("max_depth", C.c_int32),
        ("unclosed_mode", C.c_int32),
    ]


class ChunkCarry(C.Structure):
    _fields_ = [
        ("scan", C.c_uint64),
        ("out_base", C.c_uint32),
        ("bracket_base", C.c_uint32),
        ("depth", C.c_int32),
        ("min_depth", C.c_int32),
        ("max_depth", C.c_int32),
        ("reserved", C.c_uint32),
    ]


class DocSpan(C.Structure):
    _fields_ = [("offset", C.c_uint64), ("size", C.c_uint64)]


class BatchResult(C.Structure):
    _fields_ = [
        ("n_docs", C.c_uint64),
        ("n_struct", C.c_uint64),
        ("n_brackets", C.c_uint64),
        ("structural", C.c_void_p),
        ("struct_start", C.c_void_p),
        ("bracket_pos", C.c_void_p),
        ("bracket_char", C.c_void_p),
        ("bracket_start", C.c_void_p),
        ("bracket_pair", C.c_void_p),
        ("pair_pos", C.c_void_p),
        ("doc_flags", C.c_void_p),
        ("max_depth", C.c_void_p),
        ("struct_bitmap", C.c_void_p),
    ]


class ServerConfig(C.Structure):
    _fields_ = [
        ("plans", C.c_void_p),
        ("n_plans", C.c_uint32),
        ("slots_per_plan", C.c_uint32),
        ("ring_bytes", C.c_uint64),
        ("staging_bytes", C.c_uint64),
        ("hbm_budget", C.c_uint64),
        ("max_batch_bytes", C.c_uint64),
        ("max_batch_docs", C.c_uint32),
        ("run_pairs", C.c_uint32),
        ("submit_threads", C.c_uint32),
        ("reserved", C.c_uint32),
        ("budget", C.c_void_p),
    ]


class RequestResult(C.Structure):
    _fields_ = [
        ("id", C.c_uint64),
        ("status", C.c_int32),
        ("plan", C.c_uint32),
        ("n_bytes", C.c_uint64),
        ("n_struct", C.c_uint64),
        ("n_brackets", C.c_uint64),
        ("structural", C.c_void_p),
        ("bracket_pos", C.c_void_p),
        ("bracket_char", C.c_void_p),
        ("bracket_pair", C.c_void_p),
        ("flags", C.c_uint32),
        ("max_depth", C.c_int32),
        ("t_submit_ns", C.c_uint64),
        ("t_launch_ns", C.c_uint64),
        ("t_done_ns", C.c_uint64),
        ("batch_docs", C.c_uint32),
        ("reserved", C.c_uint32),
        ("input", C.c_void_p),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_uint64),
        ("completed", C.c_uint64),
        ("released", C.c_uint64),
        ("batches", C.c_uint64),
        ("bytes_completed", C.c_uint64),
        ("busy_rejects", C.c_uint64),
        ("budget_waits", C.c_uint64),
        ("device_bytes", C.c_uint64),
        ("inflight_reserved", C.c_uint64),
        ("ring_wraps", C.c_uint64),
        ("staging_wraps", C.c_uint64),
        ("worker_launch_ns", C.c_uint64),
        ("worker_retire_ns", C.c_uint64),
        ("submit_ns", C.c_uint64),
        ("submit_lock_ns", C.c_uint64),
    ]


class TuneResult(C.Structure):
    _fields_ = [
        ("rows", C.c_int32),
        ("lb_window", C.c_int32),
        ("blocks_per_cu", C.c_int32),
        ("radix_bits", C.c_int32),
        ("scan_ms", C.c_float),
        ("pair_ms", C.c_float),
        ("base_scan_ms", C.c_float),
        ("base_pair_ms", C.c_float),
        ("trials", C.c_uint32),
        ("max_blocks_per_cu", C.c_uint32),
        ("waves_per_block", C.c_int32),
    ]

    def as_dict(self) -> dict:
        return {k: (round(getattr(self, k), 4) if t is C.c_float else getattr(self, k)) for k, t in self._fields_}


# pars_allocator: device memory from the caller (alloc returns NULL when none is left)
ALLOC_FN = C.CFUNCTYPE(C.c_void_p, C.c_void_p, C.c_size_t)
RELEASE_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.c_size_t)


class Allocator(C.Structure):
    _fields_ = [("alloc", ALLOC_FN), ("release", RELEASE_FN), ("ctx", C.c_void_p)]


# name -> (restype, argtypes); every symbol include/pars_hip.h declares
PROTOTYPES = {
    "pars_hip_budget_create": (C.c_int, [C.c_uint64, C.POINTER(C.c_void_p)]),
    "pars_hip_budget_destroy": (C.c_int, [C.c_void_p]),
    "pars_hip_budget_usage": (
        C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    ),
    "pars_hip_scanner_create_ex": (
        C.c_int, [C.c_char_p, C.POINTER(Allocator), C.c_void_p, C.POINTER(C.c_void_p)]
    ),
    "pars_hip_scanner_create": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "pars_hip_scanner_destroy": (C.c_int, [C.c_void_p]),
    "pars_hip_set_option": (C.c_int, [C.c_void_p, C.c_int, C.c_int64]),
    "pars_hip_scan_device": (
        C.c_int,
        [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_int, C.c_int, C.POINTER(StageTimes),
         C.POINTER(DeviceResult)],
    ),
    "pars_hip_scan_device_noalloc": (
        C.c_int,
        [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_int, C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p,
         C.c_void_p, C.c_void_p, C.c_size_t, C.POINTER(DeviceResult)],
    ),
    "pars_hip_scan_host": (
        C.c_int, [C.c_void_p, C.c_char_p, C.c_size_t, C.c_void_p, C.c_int, C.POINTER(HostResult)]
    ),
    "pars_hip_scan_stream": (
        C.c_int,
        [C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t, C.c_int, C.c_void_p, C.c_int, C.POINTER(StageTimes),
         C.POINTER(DeviceResult)],
    ),
    "pars_hip_reserve_struct": (C.c_int, [C.c_void_p, C.c_size_t]),
    "pars_hip_scan_chunk": (
        C.c_int,
        [C.c_void_p, C.c_void_p, C.c_size_t, C.POINTER(ChunkCarry), C.c_void_p, C.c_size_t,
         C.POINTER(ChunkCarry)],
    ),
    "pars_hip_pair_chunks": (
        C.c_int, [C.c_void_p, C.POINTER(ChunkCarry), C.c_void_p, C.c_int, C.POINTER(DeviceResult)]
    ),
    "pars_hip_scan_batch": (
        C.c_int,
        [C.c_void_p, C.c_void_p, C.c_size_t, C.POINTER(DocSpan), C.c_size_t, C.c_void_p, C.c_int, C.c_int,
         C.POINTER(StageTimes), C.POINTER(BatchResult)],
    ),
    "pars_hip_server_create": (C.c_int, [C.POINTER(ServerConfig), C.POINTER(C.c_void_p)]),
    "pars_hip_server_destroy": (C.c_int, [C.c_void_p]),
    "pars_hip_server_submit": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_size_t, C.POINTER(C.c_uint64)]),
    "pars_hip_server_poll": (
        C.c_int, [C.c_void_p, C.POINTER(RequestResult), C.c_size_t, C.POINTER(C.c_size_t), C.c_int64]
    ),
    "pars_hip_server_release": (C.c_int, [C.c_void_p, C.c_uint64]),
    "pars_hip_server_stats": (C.c_int, [C.c_void_p, C.POINTER(ServerStats)]),
    "pars_hip_server_replay": (
        C.c_int,
        [C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
         C.POINTER(C.c_uint64)],
    ),
    "pars_hip_get_option": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int64)]),
    "pars_hip_autotune": (
        C.c_int,
        [C.c_void_p, C.c_void_p, C.c_size_t, C.POINTER(DocSpan), C.c_size_t, C.c_void_p, C.c_int,
         C.POINTER(TuneResult)],
    ),
    "pars_hip_struct_ptr": (C.c_void_p, [C.c_void_p]),
    "pars_hip_kv_parse": (
        C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p, C.POINTER(StageTimes), C.POINTER(KvTable)]
    ),
    "pars_hip_json_tape": (
        C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.POINTER(StageTimes), C.POINTER(JsonTape)]
    ),
    "pars_hip_last_error": (C.c_char_p, []),
    "pars_hip_version": (C.c_char_p, []),
}

GEN_PROTOTYPES = {
    "pars_gen_json": (
        C.c_int, [C.c_uint64, C.c_uint64, C.c_int, C.c_int, C.c_uint64, C.c_void_p, C.POINTER(C.c_uint64)]
    ),
    "pars_gen_csv": (C.c_int, [C.c_uint64, C.c_uint64, C.c_uint64, C.c_void_p]),
    "pars_gen_clike": (C.c_int, [C.c_uint64, C.c_uint64, C.c_int, C.c_uint64, C.c_void_p]),
    "pars_gen_alpha": (C.c_int, [C.c_uint64, C.c_uint64, C.c_char_p, C.c_uint64, C.c_void_p]),
    "pars_gen_sizes": (
        C.c_int, [C.c_uint64, C.c_uint64, C.c_double, C.c_double, C.c_uint64, C.c_uint64, C.c_void_p]
    ),
    "pars_gen_doc_seed": (C.c_uint64, [C.c_uint64, C.c_uint64]),
    "pars_gen_docs": (C.c_int, [C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
}


def _bind(lib, protos):
    for name, (res, args) in protos.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_lib = None
_gen = None
# Every object holding a native handle (Scanner, Server): shutdown() closes
# whatever is still open before the interpreter tears down.
_live: "weakref.WeakSet" = weakref.WeakSet()


def track(obj) -> None:
    """Register an object with a close() that releases a native handle."""
    _live.add(obj)


def shutdown() -> None:
    """Deterministic teardown, run at exit (before interpreter finalisation):
    servers first (their worker threads and streams), then scanners, then the
    library itself is unloaded while the HIP runtime is still fully alive, so
    no HIP call is left for the process's static destructors."""
    global _lib
    objs = list(_live)
    # servers, then scanners, then the budgets they drew on
    objs.sort(key=lambda o: 0 if hasattr(o, "submit") else getattr(o, "_teardown_order", 1))
    for o in objs:
        try:
            o.close()
        except Exception:
            pass
    _live.clear()
    if _lib is not None:
        handle, _lib = _lib._handle, None
        import _ctypes

        _ctypes.dlclose(handle)


def lib():
    """The HIP backend.  Raises if it was not built: there is no fallback."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m pars_amd.build` "
                "(the scan path has no CPU fallback)"
            )
        _lib = _bind(C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL), PROTOTYPES)
        atexit.unregister(shutdown)
        atexit.register(shutdown)
    return _lib


def gen():
    global _gen
    if _gen is None:
        if not os.path.exists(GEN_PATH):
            raise RuntimeError(f"{GEN_PATH} is missing: build it with `python -m pars_amd.build`")
        _gen = _bind(C.CDLL(GEN_PATH), GEN_PROTOTYPES)
    return _gen
