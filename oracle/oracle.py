"""Python binding of the CPU oracle (oracle/liboracle.so, oracle/pars_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product path never touches
it.  Also drives the reference's own CPU backend (oracle/_ref/ref_driver)
where it was built, to write golden vectors and to time the reference.
"""
from __future__ import annotations

import ctypes as C
import json as _json
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF = os.path.join(HERE, "_ref", "ref_driver")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: run `make -C oracle liboracle.so`")
        L = C.CDLL(LIB)
        L.oracle_classifier.argtypes = [C.c_char_p, C.c_void_p]
        L.oracle_scan.restype = C.c_int
        L.oracle_scan.argtypes = [C.c_char_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_pair.restype = C.c_int
        L.oracle_pair.argtypes = [C.c_char_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p]
        L.oracle_final_states.argtypes = [C.c_char_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p]
        L.oracle_fnv1a.restype = C.c_uint64
        L.oracle_fnv1a.argtypes = [C.c_uint64, C.c_void_p, C.c_size_t]
        L.oracle_json_tape.restype = C.c_int
        L.oracle_json_tape.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.POINTER(TapeInfo)]
        _lib = L
    return _lib


FNV0 = 0xCBF29CE484222325


class TapeInfo(C.Structure):
    _fields_ = [("n_nodes", C.c_uint64), ("dec_bytes", C.c_uint64), ("hash", C.c_uint64),
                ("fail_event", C.c_int64), ("ok", C.c_int32), ("pad", C.c_int32)]


def json_tape(data, positions) -> dict:
    """json_dom.c: the JSON hook (bench_parse.cu:41-348) replayed over a
    structural index; the same dict as oracle/json_dom.py json_tape, with the
    decoded strings as one buffer (`dec`) instead of a list."""
    d = _u8(data)
    pos = np.ascontiguousarray(positions, dtype=np.uint32)
    cap = 2 * pos.size + 1
    kind = np.zeros(cap, np.uint8)
    beg = np.zeros(cap, np.uint32)
    end = np.zeros(cap, np.uint32)
    dec = np.zeros(max(d.size, 1), np.uint8)
    info = TapeInfo()
    rc = lib().oracle_json_tape(d.ctypes.data, d.size, pos.ctypes.data, pos.size, kind.ctypes.data, beg.ctypes.data,
                                end.ctypes.data, cap, dec.ctypes.data, dec.size, C.byref(info))
    if rc:
        raise MemoryError("oracle_json_tape")
    k = int(info.n_nodes)
    return {"ok": bool(info.ok), "nodes": k, "hash": int(info.hash), "fail_event": int(info.fail_event),
            "kind": kind[:k], "begin": beg[:k], "end": end[:k], "dec": dec[:int(info.dec_bytes)].tobytes()}


def fnv(arr) -> int:
    a = np.ascontiguousarray(arr)
    return int(lib().oracle_fnv1a(FNV0, a.ctypes.data, a.nbytes))


def _u8(data) -> np.ndarray:
    if isinstance(data, (bytes, bytearray)):
        return np.frombuffer(bytes(data), dtype=np.uint8)
    return np.ascontiguousarray(data, dtype=np.uint8)


def classifier(plan: bytes) -> np.ndarray:
    out = np.zeros(256, dtype=np.uint32)
    lib().oracle_classifier(plan, out.ctypes.data)
    return out


def scan(plan: bytes, data, cap: int | None = None, bitmaps: bool = False) -> dict:
    """model::scanner<S>::scan (scan.hpp:27-73) + the per-byte masked/structural bitmaps."""
    d = _u8(data)
    n = d.size
    cap = n if cap is None else cap
    pos = np.empty(cap, np.uint32)
    ch = np.empty(cap, np.uint8)
    bpos = np.empty(cap, np.uint32)
    bch = np.empty(cap, np.uint8)
    counts = np.zeros(2, np.uint64)
    words = (n + 63) // 64
    mb = np.zeros(words, np.uint64) if bitmaps else None
    sb = np.zeros(words, np.uint64) if bitmaps else None
    unclosed = lib().oracle_scan(
        plan, d.ctypes.data, n, pos.ctypes.data, ch.ctypes.data, bpos.ctypes.data, bch.ctypes.data, cap,
        counts.ctypes.data, mb.ctypes.data if bitmaps else None, sb.ctypes.data if bitmaps else None,
    )
    k, kb = int(counts[0]), int(counts[1])
    if k > cap or kb > cap:
        raise ValueError("output capacity too small")
    r = {"positions": pos[:k].copy(), "chars": ch[:k].copy(), "bracket_pos": bpos[:kb].copy(),
         "bracket_char": bch[:kb].copy(), "unclosed_mode": bool(unclosed)}
    if bitmaps:
        r["masked_bitmap"] = mb
        r["struct_bitmap"] = sb
    return r


def pair(plan: bytes, bracket_char) -> dict:
    """model::pair_brackets<S> (nest.hpp:45-92)."""
    b = _u8(bracket_char)
    out = np.empty(b.size, np.int32)
    flags = np.zeros(3, np.int32)
    if lib().oracle_pair(plan, b.ctypes.data, b.size, out.ctypes.data, flags.ctypes.data) != 0:
        raise MemoryError("oracle_pair")
    return {"pair": out, "balanced": bool(flags[0]), "well_formed": bool(flags[1]), "max_depth": int(flags[2])}


def full(plan: bytes, data) -> dict:
    r = scan(plan, data)
    r.update(pair(plan, r["bracket_char"]))
    return r


def final_states(plan: bytes, data, in_states=None) -> np.ndarray:
    d = _u8(data)
    out = np.zeros(8, np.uint8)
    ins = None if in_states is None else np.ascontiguousarray(in_states, np.uint8)
    lib().oracle_final_states(plan, d.ctypes.data, d.size, ins.ctypes.data if ins is not None else None,
                              out.ctypes.data)
    return out


def pair_pos_device(positions: np.ndarray, bracket_pos: np.ndarray, pair_cpu: np.ndarray) -> np.ndarray:
    """The reference GPU convention (pipeline.cuh:492-500): per structural index, the
    partner's structural index or -1, derived from the CPU pair table."""
    k = positions.size
    out = np.full(k, -1, np.int32)
    if bracket_pos.size:
        sidx = np.searchsorted(positions, bracket_pos).astype(np.int64)
        ok = pair_cpu >= 0
        out[sidx[ok]] = sidx[pair_cpu[ok]].astype(np.int32)
    return out


# ---------------------------------------------------------------------------
# the reference's own CPU backend (built from its sources into oracle/_ref/)
# ---------------------------------------------------------------------------
def ref_available() -> bool:
    return os.path.exists(REF)


def ref_full(grammar: str, data) -> dict:
    d = _u8(data)
    with tempfile.TemporaryDirectory() as td:
        src, dst = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        d.tofile(src)
        subprocess.run([REF, "scan", grammar, src, dst], check=True)
        raw = open(dst, "rb").read()
    if raw[:8] != b"PRSREF01":
        raise ValueError("bad ref output")
    np_, nb = np.frombuffer(raw, np.uint64, 2, 8)
    np_, nb = int(np_), int(nb)
    hdr = np.frombuffer(raw, np.int32, 4, 24)
    o = 40
    pos = np.frombuffer(raw, np.uint32, np_, o); o += 4 * np_
    ch = np.frombuffer(raw, np.uint8, np_, o); o += np_
    bpos = np.frombuffer(raw, np.uint32, nb, o); o += 4 * nb
    bch = np.frombuffer(raw, np.uint8, nb, o); o += nb
    pr = np.frombuffer(raw, np.int32, nb, o)
    return {"positions": pos.copy(), "chars": ch.copy(), "bracket_pos": bpos.copy(), "bracket_char": bch.copy(),
            "pair": pr.copy(), "unclosed_mode": bool(hdr[0]), "balanced": bool(hdr[1]),
            "well_formed": bool(hdr[2]), "max_depth": int(hdr[3])}


def ref_time(grammar: str, path: str, max_bytes: int) -> dict:
    r = subprocess.run([REF, "time", grammar, path, str(max_bytes)], check=True, stdout=subprocess.PIPE, text=True)
    return _json.loads(r.stdout)


def batch_full(plan: bytes, buf: np.ndarray, spans) -> dict:
    """model::scanner::scan + pair_brackets run on every document of a packed
    batch alone (spans: (offset, size) pairs); per-document outputs concatenated
    in document order with their start indices, as pars_hip_scan_batch lays
    them out (positions document-local, pairs as document-local bracket indices)."""
    nd = len(spans)
    st, bst = np.zeros(nd + 1, np.uint64), np.zeros(nd + 1, np.uint64)
    pos, bpos, bch, pr = [], [], [], []
    flags = np.zeros(nd, np.uint8)
    depth = np.zeros(nd, np.int32)
    for d in range(nd):
        o, n = int(spans[d].offset), int(spans[d].size)
        r = full(plan, buf[o:o + n])
        pos.append(r["positions"])
        bpos.append(r["bracket_pos"])
        bch.append(r["bracket_char"])
        pr.append(r["pair"])
        st[d + 1] = st[d] + r["positions"].size
        bst[d + 1] = bst[d] + r["bracket_pos"].size
        flags[d] = (1 if r["balanced"] else 0) | (2 if r["well_formed"] else 0) | (4 if r["unclosed_mode"] else 0)
        depth[d] = r["max_depth"]
    cat = lambda xs, t: np.concatenate(xs).astype(t) if xs else np.zeros(0, t)  # noqa: E731
    return {"positions": cat(pos, np.uint32), "struct_start": st, "bracket_pos": cat(bpos, np.uint32),
            "bracket_char": cat(bch, np.uint8), "bracket_start": bst, "pair": cat(pr, np.int32),
            "flags": flags, "max_depth": depth}


# ---------------------------------------------------------------------------
# the timed CPU baseline (baseline.c)
# ---------------------------------------------------------------------------
class Run(C.Structure):
    _fields_ = [("seconds", C.c_double), ("bytes", C.c_uint64), ("n_struct", C.c_uint64),
                ("n_brackets", C.c_uint64), ("hash_pos", C.c_uint64), ("hash_pair", C.c_uint64),
                ("balanced", C.c_int32), ("well_formed", C.c_int32), ("max_depth", C.c_int32),
                ("unclosed", C.c_int32)]


SOURCES = ["pars_oracle.c", "json_dom.c", "baseline.c"]
NATIVE_FLAGS = ["-O3", "-march=native", "-std=gnu99", "-fPIC", "-shared"]


def native_lib(out_dir: str):
    """The restatement built -O3 -march=native for the host that runs it
    (gcc on that host, into out_dir).  Returns (CDLL, flags string), or the
    prebuilt portable liboracle.so and its flags when gcc is unavailable."""
    dst = os.path.join(out_dir, "liboracle_native.so")
    try:
        subprocess.run(["gcc", *NATIVE_FLAGS, "-o", dst] + [os.path.join(HERE, s) for s in SOURCES], check=True,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=120)
        L, flags = C.CDLL(dst), " ".join(NATIVE_FLAGS[:2])
    except (OSError, subprocess.SubprocessError):
        L, flags = C.CDLL(LIB), "-O2 (prebuilt; gcc unavailable on this host)"
    L.oracle_time_docs.restype = C.c_int
    L.oracle_time_docs.argtypes = [C.c_char_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p,
                                   C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(Run)]
    return L, flags


def host_info() -> dict:
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0))}


def time_docs(L, plan: bytes, buf: np.ndarray, offs, sizes, per_doc: bool = False, core: int | None = None) -> dict:
    """Run the baseline over documents buf[offs[d], +sizes[d]) pinned to one
    core (the first the process may use, unless given); returns the timing,
    totals, hashes, and per-document counts / flags / depths when asked."""
    buf = np.ascontiguousarray(buf, dtype=np.uint8)
    offs = np.ascontiguousarray(offs, dtype=np.uint64)
    sizes = np.ascontiguousarray(sizes, dtype=np.uint64)
    nd = offs.size
    arrs = None
    if per_doc:
        arrs = (np.zeros(nd, np.uint32), np.zeros(nd, np.uint32), np.zeros(nd, np.uint8), np.zeros(nd, np.int32))
    prev = os.sched_getaffinity(0)
    core = sorted(prev)[0] if core is None else core
    os.sched_setaffinity(0, {core})
    try:
        r = Run()
        rc = L.oracle_time_docs(plan, buf.ctypes.data, offs.ctypes.data, sizes.ctypes.data, nd,
                                *(a.ctypes.data for a in arrs) if arrs else (None, None, None, None), C.byref(r))
    finally:
        os.sched_setaffinity(0, prev)
    if rc:
        raise MemoryError("oracle_time_docs")
    out = {"seconds": r.seconds, "bytes": int(r.bytes), "n_struct": int(r.n_struct), "n_brackets": int(r.n_brackets),
           "hash_pos": int(r.hash_pos), "hash_pair": int(r.hash_pair), "balanced": bool(r.balanced),
           "well_formed": bool(r.well_formed), "max_depth": int(r.max_depth), "unclosed": bool(r.unclosed),
           "core": core}
    if arrs:
        out.update({"doc_struct": arrs[0], "doc_brackets": arrs[1], "doc_flags": arrs[2], "doc_depth": arrs[3]})
    return out
