"""Benchmark: the plain JSON pipeline on one seeded 2 GiB document, one MI355X,
and beside it the 1 GiB six-mode C-like document, the 100k-document batch and
the serving loop.

A step is one device-resident scan of the whole document through the C ABI
(pars_hip_scan_device: R1 classify .. R5 compaction in one kernel, then R6
pairing; the reference's N1b tier, bench/bench_boundary.cu:252-260), with the
input already in HBM.  `value` = input bytes / step time / 1e9 (the
reference's GB/s, bench/bench_boundary.cu:31-33).

roofline: the dominant kernel is the fused R1-R5 scan; its duration comes from
HIP events recorded by the library on the step's own stream around that
kernel.  Algorithmic bytes per launch (per document):
    B_alg = n (input) + n/8 (structural bitmap) + 4 k (uint32 positions)
            + 4 k_b (bracket-pair array)
with k, k_b the document's structural and bracket counts.  peak = 8 TB/s.
`traffic` is the HBM bytes one scan launch moves as measured by rocprofv3
counters (FETCH_SIZE + WRITE_SIZE, calibrated on a plain device copy); it is
read from the committed profile summary of this workload when present.

cpu_baseline: the C restatement of the reference's CPU path
(model::scanner::scan + model::pair_brackets, oracle/pars_oracle.c) built
-O3 -march=native for this host, pinned to one core, timed over the whole
document (C2, C3), every document of the batch (C4) and every request of the
serving stream (C5).  The GPU's outputs are compared with what it produced:
counts, flags and FNV-1a hashes of the positions and pair tables.

    python bench.py [--gpus 1] [--steps 10] [--warmup 2]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "input GB/s + % HBM roofline (2 GiB JSON); batched/served vs plain; p99 latency"
HBM_PEAK_GBS = 8000.0
PROFILES = os.path.join(ROOT, "profiles")


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def measured_traffic(name: str, n: int) -> int | None:
    """HBM bytes per scan launch from the committed counter summary (same workload size)."""
    try:
        with open(os.path.join(PROFILES, name)) as f:
            fields = dict(
                line.strip().lstrip("- ").split(": ", 1) for line in f if line.strip().lstrip("- ").startswith("traffic_")
            )
        if int(fields.get("traffic_input_bytes", -1)) != n:
            return None
        return int(fields["traffic_bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


# ---------------------------------------------------------------------------
# CPU baseline
# ---------------------------------------------------------------------------
class Baseline:
    """The restatement built for this host (-O3 -march=native), timed pinned to one core."""

    def __init__(self):
        from oracle import oracle as O

        self.O = O
        self.dir = tempfile.mkdtemp(prefix="pars_cpu_")
        self.lib, self.flags = O.native_lib(self.dir)
        self.host = O.host_info()

    def run(self, plan: str, buf, offs, sizes, what: str, per_doc: bool = False) -> tuple[dict, dict]:
        from pars_amd import Plan

        r = self.O.time_docs(self.lib, Plan.load(plan).blob, buf, offs, sizes, per_doc=per_doc)
        line = {
            "value": round(r["bytes"] / r["seconds"] / 1e9, 4),
            "unit": "GB/s",
            "cores": 1,
            "kind": "port",
            "sample": f"{what}: {r['bytes']} bytes, scan + pair_brackets, {r['seconds']:.2f} s",
            "build": self.flags,
            "pinned_core": r["core"],
            **self.host,
        }
        return line, r


def fnv(arr) -> int:
    from oracle import oracle as O

    return O.fnv(arr)


_hip = None


def d2h(ptr: int, count: int, dtype):
    """Device-to-host copy through the HIP runtime torch loaded (for the parity check)."""
    import numpy as np

    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so.7")
        _hip.hipMemcpy.restype = C.c_int
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    out = np.empty(count, dtype)
    if count:
        assert _hip.hipMemcpy(out.ctypes.data, C.c_void_p(ptr), out.nbytes, 2) == 0, "hipMemcpy"
    return out


def parity_single(r, cpu: dict) -> dict:
    """The GPU's outputs of one document against the baseline's."""
    import numpy as np

    pos = d2h(r.structural, r.n_struct, np.uint32)
    pair = d2h(r.bracket_pair, r.n_brackets, np.int32) if r.bracket_pair else np.zeros(0, np.int32)
    checks = {
        "counts": (int(r.n_struct), int(r.n_brackets)) == (cpu["n_struct"], cpu["n_brackets"]),
        "flags": (bool(r.balanced), bool(r.well_formed), int(r.max_depth), bool(r.unclosed_mode))
        == (cpu["balanced"], cpu["well_formed"], cpu["max_depth"], cpu["unclosed"]),
        "positions_fnv": fnv(pos) == cpu["hash_pos"],
        "pairs_fnv": fnv(pair) == cpu["hash_pair"],
    }
    checks["equal"] = all(checks.values())
    return checks


# ---------------------------------------------------------------------------
# one document (C2, C3)
# ---------------------------------------------------------------------------
def bench_document(torch, plan: str, d_ptr: int, n: int, steps: int, warmup: int, knobs: str, tune: bool,
                   traffic_file: str, kernel: str):
    """Timed steps of pars_hip_scan_device on a device-resident document."""
    import numpy as np

    from pars_amd import Plan, Scanner
    from pars_amd import _native as N

    sc = Scanner(Plan.load(plan))
    sc.set_option(N.PARS_OPT_WANT_BITMAP, 1)
    stream = torch.cuda.Stream()
    sp = stream.cuda_stream
    st = N.StageTimes()
    tuned = None
    if knobs:
        opts = {"rows": N.PARS_OPT_ROWS_PER_TILE, "lb_window": N.PARS_OPT_LOOKBACK_WINDOW,
                "blocks_per_cu": N.PARS_OPT_BLOCKS_PER_CU, "radix_bits": N.PARS_OPT_RADIX_BITS,
                "waves": N.PARS_OPT_WAVES_PER_BLOCK}
        tuned = {}
        for kv in knobs.split(","):
            k, v = kv.split("=")
            sc.set_option(opts[k.strip()], int(v))
            tuned[k.strip()] = int(v)
        log(f"{plan}: fixed knobs {tuned}")
    elif tune:
        t0 = time.perf_counter()
        tuned = sc.autotune(d_ptr, n, stream=sp, reps=3)
        log(f"{plan}: autotune ({time.perf_counter() - t0:.1f} s): {tuned}")
    for i in range(warmup):
        r = sc.scan_device(d_ptr, n, sp, want_pairs=False, run_pairs=True, stage_times=st)
        log(f"{plan} warmup {i}: {r}")
    scan_ms, pair_ms, counts = [], [], set()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        r = sc.scan_device(d_ptr, n, sp, want_pairs=False, run_pairs=True, stage_times=st)
        scan_ms.append(st.p1)
        pair_ms.append(st.p4)
        counts.add((r.n_struct, r.n_brackets, r.balanced, r.well_formed))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    assert len(counts) == 1, f"steps disagree: {counts}"
    ms_step = wall / steps * 1e3
    k, kb = r.n_struct, r.n_brackets
    b_alg = n + n / 8 + 4 * k + 4 * kb
    t_scan = float(np.mean(scan_ms)) / 1e3
    achieved = b_alg / t_scan / 1e9
    # the kernel as launched: rows per tile and waves per block as set or tuned
    kernel = kernel.format(rows=sc.get_option(N.PARS_OPT_ROWS_PER_TILE) or 4,
                           waves=sc.get_option(N.PARS_OPT_WAVES_PER_BLOCK) or 4)
    out = {
        "ms_per_step": ms_step,
        "value": n / (ms_step / 1e3) / 1e9,
        "n_struct": int(k),
        "n_brackets": int(kb),
        "max_depth": int(r.max_depth),
        "tuned": tuned,
        "roofline": {
            "bound": "hbm",
            "kernel": kernel,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(traffic_file, n),
            "b_alg_bytes": int(b_alg),
            "kernel_ms": round(t_scan * 1e3, 4),
        },
        "e2e_roofline_frac": round(b_alg / (ms_step / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
        "stage_ms": {"r1_r5": round(float(np.mean(scan_ms)), 4), "r6": round(float(np.mean(pair_ms)), 4)},
    }
    log(f"{plan}: {ms_step:.3f} ms/step; scan {out['stage_ms']['r1_r5']} ms, R6 {out['stage_ms']['r6']} ms")
    return sc, r, out, sp


# paper/plots/data/pars.csv, 512 MB rows (A100): N1a = R1-R5, N1b = R1-R6 GB/s, and tokens
REF_FORMATS_512 = {
    "json": (40.181, 19.997, 165941931), "csv": (141.189, 141.192, 68536716),
    "toml": (36.215, 30.604, 52123390), "clike": (49.299, 49.500, 29020050),
    "ini": (61.932, 61.937, 103244410), "xml": (34.601, 42.457, 89478487),
}
# the same rows' N2 tier (host input: H2D + R1-R6, GB/s, pars.csv:184-209), and
# cuJSON's R1-R6 with the input transfer at 512 MB (baselines.csv:50)
REF_N2_512 = {"json": 6.983, "csv": 11.896, "toml": 8.628, "clike": 10.245, "ini": 10.643, "xml": 10.087}
CUJSON_N2_512 = 10.876


def bench_formats(torch, mib: int, reps: int) -> dict:
    """The reference's per-format documents (bench_boundary.cu:312-354: a
    record repeated to the target size) through the device-resident tiers:
    N1a = R1-R5 (run_pairs off), N1b = R1-R6, and N2 = the same from host
    memory (pars_hip_scan_stream: PCIe upload overlapped with the scan).
    Tokens must equal the reference's published count for the same document
    (512 MiB)."""
    import numpy as np

    from pars_amd import Plan, Scanner
    from pars_amd import workloads as W

    out = {}
    stream = torch.cuda.Stream()
    sp = stream.cuda_stream
    for fmt, (ref_a, ref_b, ref_tok) in REF_FORMATS_512.items():
        doc = W.format_doc(fmt, mib << 20)
        d = torch.from_numpy(doc).to("cuda")
        torch.cuda.synchronize()
        sc = Scanner(Plan.load(fmt))
        sc.autotune(d.data_ptr(), doc.size, stream=sp, reps=2)
        line = {"bytes": int(doc.size)}
        for tier, pairs in (("n1a", False), ("n1b", True)):
            r = sc.scan_device(d.data_ptr(), doc.size, sp, want_pairs=False, run_pairs=pairs)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                r = sc.scan_device(d.data_ptr(), doc.size, sp, want_pairs=False, run_pairs=pairs)
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) / reps * 1e3
            line[tier + "_ms"] = round(ms, 4)
            line[tier + "_gb_s"] = round(doc.size / (ms / 1e3) / 1e9, 1)
        line["tokens"] = int(r.n_struct)
        if mib == 512:
            line["tokens_equal_reference"] = int(r.n_struct) == ref_tok
        # N2: the document handed over in host memory (H2D + R1-R6) through the
        # streaming boundary (64 MiB pinned chunks, 8 copy threads)
        walls = []
        for i in range(reps + 1):
            t0 = time.perf_counter()
            r2 = sc.scan_stream(doc, chunk_bytes=64 << 20, copy_threads=8, stream=sp)
            if i:
                walls.append(time.perf_counter() - t0)
        assert int(r2.n_struct) == int(r.n_struct)
        w = float(np.median(walls))
        line["n2_ms"] = round(w * 1e3, 3)
        line["n2_gb_s"] = round(doc.size / w / 1e9, 2)
        line["reference_a100_gb_s"] = {"n1a": ref_a, "n1b": ref_b, "n2": REF_N2_512[fmt]}
        if fmt == "json":
            line["cujson_a100_n2_gb_s"] = CUJSON_N2_512
        out[fmt] = line
        log(f"formats: {fmt} {line}")
        sc.close()
        del d
    return out


# The reference's full-parse rates for these documents at 512 MiB (GB/s, mean
# over its iterations; paper/plots/data/parse.csv:33-34, variant P3): its host
# hook replaying a structural index already on the host, the scan excluded.
REF_KV_P3_512 = {"toml": 0.397, "ini": 0.149}


def bench_kv(torch, mib: int, reps: int, no_cpu: bool) -> dict:
    """The INI / TOML stage-2 hooks (bench_parse_host.cu:69-134, 202-248) as one
    device call, pars_hip_kv_parse: R1-R5 and the key-value table of the
    reference benchmark's own documents (its record repeated to the target,
    bench_parse_host.cu:305-312), device-resident.  The table is checked
    record by record against the restatement.  CPU baseline: the restatement
    (oracle/kv_hooks.py, pure Python, one core) replaying the oracle's
    structural index of the document's first MiB (the replay + hook stage)."""
    import numpy as np

    from pars_amd import Plan, Scanner
    from pars_amd import _native as N
    from pars_amd import workloads as W

    out = {"call": "pars_hip_kv_parse (R1-R5 + key-value table, device-resident)"}
    stream = torch.cuda.Stream()
    sp = stream.cuda_stream
    for fmt in ("ini", "toml"):
        doc = W.format_doc(fmt, mib << 20)
        d = torch.from_numpy(doc).to("cuda")
        torch.cuda.synchronize()
        sc = Scanner(Plan.load(fmt))
        st, walls, parts, r = N.StageTimes(), [], [], None
        for i in range(reps + 1):
            t0 = time.perf_counter()
            r = sc.kv_parse(d.data_ptr(), doc.size, fmt, sp, st)
            if i:
                walls.append(time.perf_counter() - t0)
                parts.append((st.p1, st.p2))
        p = np.median(np.array(parts), axis=0)
        dev = float(p[0] + p[1])
        line = {"bytes": int(doc.size), "tokens": int(r.n_tokens), "n_role": r.n_role, "n_struct": int(r.n_struct),
                "device_ms": round(dev, 3), "gb_per_s": round(doc.size / dev / 1e6, 2),
                "stage_ms": {"r1_r5": round(float(p[0]), 3), "table": round(float(p[1]), 3)},
                "table_gb_per_s": round(doc.size / float(p[1]) / 1e6, 2),
                "call_ms": round(float(np.median(walls)) * 1e3, 3)}
        # the document is the record repeated: the table must be the record's, shifted
        from oracle import kv_hooks as K
        from oracle import oracle as O

        rec = W.FORMAT_RECORDS[fmt]
        one = K.tokens(fmt, rec, O.scan(Plan.load(fmt).blob, np.frombuffer(rec, np.uint8))["positions"])
        reps_ = doc.size // len(rec)
        torch.cuda.synchronize()
        role = d2h(r.role, r.n_tokens, np.uint8)
        beg = d2h(r.begin, r.n_tokens, np.uint32)
        shift = (np.arange(reps_, dtype=np.int64) * len(rec))[:, None]
        line["parity"] = bool(
            r.n_tokens == reps_ * len(one)
            and np.array_equal(role.reshape(reps_, -1), np.broadcast_to(np.array([t[0] for t in one], np.uint8),
                                                                        (reps_, len(one))))
            and np.array_equal(beg.reshape(reps_, -1).astype(np.int64),
                               np.array([t[1] for t in one], np.int64) + shift))
        if mib == 512:
            line["reference_p3_gb_s"] = REF_KV_P3_512[fmt]
            line["vs_reference_p3"] = round(line["gb_per_s"] / REF_KV_P3_512[fmt], 1)
        if not no_cpu:
            sample = np.ascontiguousarray(doc[: 1 << 20])
            pos = O.scan(Plan.load(fmt).blob, sample)["positions"]
            raw = sample.tobytes()
            prev = os.sched_getaffinity(0)
            os.sched_setaffinity(0, {sorted(prev)[0]})
            try:
                t0 = time.perf_counter()
                toks = K.tokens(fmt, raw, pos)
                dt = time.perf_counter() - t0
            finally:
                os.sched_setaffinity(0, prev)
            line["cpu_baseline"] = {"value": round(sample.size / dt / 1e9, 5), "unit": "GB/s", "cores": 1,
                                    "kind": "port",
                                    "stage": "replay + hook over a given structural index (oracle/kv_hooks.py, "
                                             "pure Python)",
                                    "sample": f"first {sample.size} bytes of the document, {len(toks)} tokens"}
        out[fmt] = line
        log(f"kv: {fmt} {line}")
        sc.close()
        del d
        torch.cuda.empty_cache()
    return out


def bench_stream(sc, doc, stream: int, reps: int = 3) -> dict:
    """The host boundary: the same document handed over in host memory and
    streamed through pinned staging (64 MiB chunks, 8 copy threads), uploads and
    scans overlapped; results on the device as in the main measurement."""
    import numpy as np

    from pars_amd import _native as N

    st = N.StageTimes()
    walls = []
    for i in range(reps + 1):
        t0 = time.perf_counter()
        r = sc.scan_stream(doc, chunk_bytes=64 << 20, copy_threads=8, stream=stream, stage_times=st)
        if i:
            walls.append(time.perf_counter() - t0)
    w = float(np.median(walls))
    return {
        "boundary": "host bytes -> pinned chunks -> PCIe -> scan (pars_hip_scan_stream)",
        "chunk_bytes": 64 << 20,
        "copy_threads": 8,
        "ms": round(w * 1e3, 3),
        "gb_per_s": round(doc.size / w / 1e9, 3),
        "n_struct": int(r.n_struct),
    }


def bench_tape(d_ptr: int, doc, stream: int, reps: int, cpu_sample: int, no_cpu: bool) -> dict:
    """The stage-2 JSON hook (bench_parse.cu:336-348: scan, replay into the JSON
    DOM builder) as one device call, pars_hip_json_tape: R1-R6 and the tape of
    a device-resident document.  Measured on the 2 GiB document and on the
    reference bench's own payload (its record repeated to 64 MiB,
    bench_parse.cu:403-412).  CPU baseline: the hook restated in C
    (oracle/json_dom.c), one core, replaying the oracle's structural index of
    the document's first `cpu_sample` bytes (the replay + hook stage only)."""
    import numpy as np
    import torch

    from pars_amd import Plan, Scanner
    from pars_amd import _native as N

    def run(sc, ptr, n):
        st, walls, last = N.StageTimes(), [], None
        parts = []
        for i in range(reps + 1):
            t0 = time.perf_counter()
            last = sc.json_tape(ptr, n, stream, st)
            if i:
                walls.append(time.perf_counter() - t0)
                parts.append((st.p1, st.p4, st.p2))
        dev = float(np.median([a + b + c for a, b, c in parts]))
        p = np.median(np.array(parts), axis=0)
        return {"bytes": n, "nodes": int(last.n_nodes), "ok": bool(last.ok),
                "device_ms": round(dev, 3), "gb_per_s": round(n / dev / 1e6, 2),
                "stage_ms": {"r1_r5": round(float(p[0]), 3), "r6": round(float(p[1]), 3), "tape": round(float(p[2]), 3)},
                "call_ms": round(float(np.median(walls)) * 1e3, 3)}

    sc = Scanner("json")
    out = {"call": "pars_hip_json_tape (R1-R6 + tape, device-resident)"}
    out["json_2gib"] = run(sc, d_ptr, doc.size)
    rec = b'{"id":123,"name":"x \\"y\\"","a":[1,2,3],"n":{"z":null}}'
    k = (64 << 20) // (len(rec) + 1)
    pay = np.frombuffer(b"[" + b",".join([rec] * k) + b"]", np.uint8)
    t = torch.from_numpy(pay.copy()).to("cuda")
    torch.cuda.synchronize()
    out["record_payload_64mib"] = run(sc, t.data_ptr(), pay.size)
    del t
    sc.close()
    if not no_cpu:
        from oracle import oracle as O

        sample = np.ascontiguousarray(doc[:cpu_sample])
        pos = O.scan(Plan.load("json").blob, sample)["positions"]
        prev = os.sched_getaffinity(0)
        os.sched_setaffinity(0, {sorted(prev)[0]})
        try:
            t0 = time.perf_counter()
            r = O.json_tape(sample, pos)
            dt = time.perf_counter() - t0
        finally:
            os.sched_setaffinity(0, prev)
        out["cpu_baseline"] = {"value": round(sample.size / dt / 1e9, 4), "unit": "GB/s", "cores": 1, "kind": "port",
                               "stage": "replay + JSON hook over a given structural index (oracle/json_dom.c)",
                               "sample": f"first {sample.size} bytes of the 2 GiB document, {r['nodes']} nodes"}
    return out


# ---------------------------------------------------------------------------
# C4: the batch
# ---------------------------------------------------------------------------
def bench_batch(torch, n_docs: int, iters: int, plain_sample: int, base: Baseline | None, tune: bool = True) -> dict:
    """The 100k-document batch: 70% JSON / 30% CSV, log-normal sizes (median
    4 KiB, sigma 1.85, clipped to [200 B, 4 MiB]), one batched launch chain per
    plan (graph-replayed), against one plain launch chain per document."""
    import numpy as np

    from pars_amd import Scanner
    from pars_amd import _native as N
    from pars_amd import workloads as W

    t0 = time.perf_counter()
    b = W.batch(4, n_docs, {"json": 70, "csv": 30}, median=4096.0, sigma=1.85)
    total = int(b["sizes"].sum())
    log(f"batch: {n_docs} documents, {total} bytes, generated in {time.perf_counter() - t0:.1f} s")
    cpu, cpu_runs = None, {}
    if base is not None:
        secs, nbytes = 0.0, 0
        for kind in ("json", "csv"):
            buf, spans, idx = b[kind]
            offs = np.array([s.offset for s in spans], np.uint64)
            sz = np.array([s.size for s in spans], np.uint64)
            line, r = base.run(kind, buf, offs, sz, kind, per_doc=True)
            cpu_runs[kind] = r
            secs += r["seconds"]
            nbytes += r["bytes"]
        cpu = dict(line)
        cpu["value"] = round(nbytes / secs / 1e9, 4)
        cpu["docs_per_s"] = round(n_docs / secs, 1)
        cpu["sample"] = (f"every document of the batch ({n_docs} documents, {nbytes} bytes) scanned and paired "
                         f"alone, one after the other: {secs:.2f} s")
        log(f"batch cpu baseline: {cpu}")
    stream = torch.cuda.Stream()
    sp = stream.cuda_stream
    groups = []
    for kind in ("json", "csv"):
        buf, spans, idx = b[kind]
        d = torch.from_numpy(buf).to("cuda")
        sc = Scanner(kind)
        sc.set_option(N.PARS_OPT_GRAPH, 1)
        groups.append((kind, d, spans, sc, len(idx)))
    torch.cuda.synchronize()
    tuned = {}
    if tune:
        for kind, d, spans, sc, nd in groups:
            t0 = time.perf_counter()
            tuned[kind] = sc.autotune(d.data_ptr(), d.numel(), spans, stream=sp, reps=2)
            log(f"batch autotune {kind} ({time.perf_counter() - t0:.1f} s): {tuned[kind]}")

    def one_pass():
        out = []
        for kind, d, spans, sc, nd in groups:
            r = sc.scan_batch(d.data_ptr(), d.numel(), spans, sp, want_pairs=False, run_pairs=True)
            out.append((r.n_struct, r.n_brackets))
        return out

    first = one_pass()  # eager
    for _ in range(2):  # capture, then replay
        assert one_pass() == first
    t0 = time.perf_counter()
    for _ in range(iters):
        one_pass()
    t_batch = (time.perf_counter() - t0) / iters
    # parity of the last replay with the baseline's per-document outputs
    parity = None
    if cpu_runs:
        ok = True
        for kind, d, spans, sc, nd in groups:
            r = sc.scan_batch(d.data_ptr(), d.numel(), spans, sp, want_pairs=False, run_pairs=True)
            c = cpu_runs[kind]
            ss = d2h(r.struct_start, nd + 1, np.uint32).astype(np.int64)
            bs = d2h(r.bracket_start, nd + 1, np.uint32).astype(np.int64)
            ok = ok and np.array_equal(np.diff(ss), c["doc_struct"].astype(np.int64))
            ok = ok and np.array_equal(np.diff(bs), c["doc_brackets"].astype(np.int64))
            ok = ok and np.array_equal(d2h(r.doc_flags, nd, np.uint8), c["doc_flags"])
            ok = ok and np.array_equal(d2h(r.max_depth, nd, np.int32), c["doc_depth"])
            ok = ok and fnv(d2h(r.structural, r.n_struct, np.uint32)) == c["hash_pos"]
            ok = ok and fnv(d2h(r.bracket_pair, r.n_brackets, np.int32)) == c["hash_pair"]
        parity = {"checked": "per-document counts, flags and max depth; hashes of every position and pair",
                  "equal": bool(ok)}
        log(f"batch parity: {parity}")

    # the same documents through one plain launch chain each (bounded sample, in document order)
    plain_t, plain_docs, plain_bytes = 0.0, 0, 0
    for kind, d, spans, sc, nd in groups:
        k = max(1, plain_sample * nd // n_docs)
        sc1 = Scanner(kind)
        base_ptr = d.data_ptr()
        sc1.scan_device(base_ptr + spans[0].offset, spans[0].size, sp)  # warm
        t0 = time.perf_counter()
        for i in range(min(k, nd)):
            sc1.scan_device(base_ptr + spans[i].offset, spans[i].size, sp, want_pairs=False, run_pairs=True)
        plain_t += time.perf_counter() - t0
        plain_docs += min(k, nd)
        plain_bytes += int(sum(spans[i].size for i in range(min(k, nd))))
        sc1.close()
    for g in groups:
        g[3].close()
    docs_s = n_docs / t_batch
    plain_docs_s = plain_docs / plain_t
    n_struct = int(sum(x[0] for x in first))
    n_br = int(sum(x[1] for x in first))
    b_alg = total + total / 8 + 4 * n_struct + 4 * n_br
    out = {
        "workload": "100k uneven documents, 70% JSON / 30% CSV, log-normal median 4 KiB (seed 4)",
        "docs": n_docs,
        "bytes": total,
        "ms_per_batch": round(t_batch * 1e3, 3),
        "docs_per_s": round(docs_s, 1),
        "gb_per_s": round(total / t_batch / 1e9, 3),
        "plain_docs_per_s": round(plain_docs_s, 1),
        "plain_gb_per_s": round(plain_bytes / plain_t / 1e9, 4),
        "plain_sample_docs": plain_docs,
        "speedup_docs_per_s": round(docs_s / plain_docs_s, 1),
        "n_struct": n_struct,
        "n_brackets": n_br,
        "graph": True,
        "tuned": tuned or None,
        "roofline": {
            "bound": "hbm",
            "kernel": "the whole batch chain per plan (prep, scan, R6, flags; graph replay), wall clock",
            "achieved": round(b_alg / t_batch / 1e9, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(b_alg / t_batch / 1e9 / HBM_PEAK_GBS, 4),
            "traffic": None,
            "b_alg_bytes": int(b_alg),
        },
    }
    if cpu is not None:
        out["cpu_baseline"] = cpu
        out["parity"] = parity
    return out


# ---------------------------------------------------------------------------
# C5: serving
# ---------------------------------------------------------------------------
def bench_serve(n_req: int, base: Baseline | None, seed: int = 5) -> dict:
    """Serving: Poisson arrivals of JSON / CSV / C-like requests (60/25/15,
    log-normal sizes, median 4 KiB, sigma 1.85, clipped to [200 B, 4 MiB])
    through a 4 GiB device ring under a 64 GiB HBM budget.  Saturation is
    measured first (closed loop: every request offered at once; after one
    whole pass, the median of three); then offered
    loads of 20%..100% of it in 10% steps.  Latency = scheduled arrival to
    outputs complete on the device."""
    import numpy as np

    from pars_amd import Server
    from pars_amd import workloads as W

    plans = ["json", "csv", "clike"]
    t0 = time.perf_counter()
    kinds = W.doc_kinds(seed, n_req, {"json": 60, "csv": 25, "clike": 15})
    sz = W.sizes(seed, n_req, median=4096.0, sigma=1.85)
    docs = [W.make_doc(k, W.doc_seed(seed, i), int(sz[i])) for i, k in enumerate(kinds)]
    pidx = [plans.index(k) for k in kinds]
    total = int(sz.sum())
    log(f"serve: {n_req} requests, {total} bytes, generated in {time.perf_counter() - t0:.1f} s")
    cpu = None
    if base is not None:
        secs, nbytes = 0.0, 0
        for p in plans:
            sel = [i for i, k in enumerate(kinds) if k == p]
            if not sel:
                continue
            buf = np.concatenate([docs[i] for i in sel])
            sizes = np.array([docs[i].size for i in sel], np.uint64)
            offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.uint64)
            line, r = base.run(p, buf, offs, sizes, p)
            secs += r["seconds"]
            nbytes += r["bytes"]
        cpu = dict(line)
        cpu["value"] = round(nbytes / secs / 1e9, 4)
        cpu["us_per_request"] = round(secs / n_req * 1e6, 2)
        cpu["requests_per_s"] = round(n_req / secs, 1)
        cpu["sample"] = (f"every request of the stream ({n_req} requests, {nbytes} bytes) scanned and paired alone, "
                         f"one after the other: {secs:.2f} s")
        log(f"serve cpu baseline: {cpu}")
    sv = Server(plans)  # 4 GiB ring, 1 GiB staging, 64 GiB budget, 64 MiB batches
    sv.replay(pidx[:200], docs[:200], np.zeros(200, np.uint64))  # warm every plan's kernels
    sv.replay(pidx, docs, np.zeros(n_req, np.uint64))  # one whole pass first: ring and staging in use
    els = sorted(sv.replay(pidx, docs, np.zeros(n_req, np.uint64))[1] for _ in range(3))
    el = els[1]  # the median of three closed-loop passes
    sat = n_req / (el / 1e9)
    log(f"serve: saturation {sat:.0f} req/s, {total / el:.1f} GB/s")
    rng = np.random.default_rng(seed)
    points = []
    for pct in range(20, 101, 10):
        rate = sat * pct / 100.0
        gaps = rng.exponential(1e9 / rate, n_req)
        arrival = np.cumsum(gaps)
        arrival = (arrival - arrival[0]).astype(np.uint64)
        lat, el = sv.replay(pidx, docs, arrival)
        pt = {
            "load_pct": pct,
            "offered_req_s": round(rate, 1),
            "p50_us": round(float(np.percentile(lat, 50)) / 1e3, 1),
            "p99_us": round(float(np.percentile(lat, 99)) / 1e3, 1),
            "gb_per_s": round(total / el, 3),
        }
        points.append(pt)
        log(f"serve: {pt}")
    st = sv.stats()
    sv.close()
    at80 = next(p for p in points if p["load_pct"] == 80)
    out = {
        "workload": "Poisson requests, 60% JSON / 25% CSV / 15% C-like, log-normal median 4 KiB (seed 5)",
        "requests_per_point": n_req,
        "bytes_per_point": total,
        "saturation_req_s": round(sat, 1),
        "saturation_gb_s": round(total / (n_req / sat) / 1e9, 3),
        "p50_us_at_80": at80["p50_us"],
        "p99_us_at_80": at80["p99_us"],
        "gb_per_s_at_80": at80["gb_per_s"],
        "points": points,
        "batches": int(st.batches),
        "ring_wraps": int(st.ring_wraps),
        "staging_wraps": int(st.staging_wraps),
        "device_bytes": int(st.device_bytes),
    }
    if cpu is not None:
        out["cpu_baseline"] = cpu
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size-mib", type=int, default=2048)
    ap.add_argument("--no-cpu", action="store_true", help="skip every CPU baseline")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain device copy of the document (counter calibration under rocprofv3)")
    ap.add_argument("--no-clike", action="store_true", help="skip the 1 GiB six-mode C-like document")
    ap.add_argument("--only-clike", action="store_true", help="run only the 1 GiB six-mode C-like document")
    ap.add_argument("--clike-mib", type=int, default=1024)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-documents measurement")
    ap.add_argument("--only-batch", action="store_true", help="run only the batched-documents measurement")
    ap.add_argument("--batch-docs", type=int, default=100_000)
    ap.add_argument("--batch-iters", type=int, default=5)
    ap.add_argument("--plain-sample", type=int, default=3000, help="documents timed through plain launches")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving measurement")
    ap.add_argument("--only-serve", action="store_true", help="run only the serving measurement")
    ap.add_argument("--serve-requests", type=int, default=20000)
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch knobs (no autotuner)")
    ap.add_argument("--no-stream", action="store_true", help="skip the host-boundary (streaming) measurement")
    ap.add_argument("--no-tape", action="store_true", help="skip the JSON tape (stage-2 hook) measurement")
    ap.add_argument("--no-formats", action="store_true",
                    help="skip the reference's six per-format documents (512 MiB each)")
    ap.add_argument("--formats-mib", type=int, default=512)
    ap.add_argument("--only-formats", action="store_true", help="run only the six per-format documents")
    ap.add_argument("--tape-cpu-mib", type=int, default=256, help="CPU baseline sample of the JSON tape")
    ap.add_argument("--no-kv", action="store_true", help="skip the INI / TOML key-value tables")
    ap.add_argument("--only-kv", action="store_true", help="run only the INI / TOML key-value tables")
    ap.add_argument("--kv-mib", type=int, default=512)
    ap.add_argument("--knobs", default="",
                    help="fixed launch knobs for the document scans instead of tuning, e.g. lb_window=16,rows=4 "
                         "(rows, lb_window, blocks_per_cu, radix_bits, waves); profiles use this")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for {args.gpus})", file=sys.stderr)
        return 2

    import numpy as np

    from pars_amd import workloads as W

    base = None if args.no_cpu else Baseline()
    if base is not None:
        log(f"cpu baseline build: {base.flags}; host {base.host}")

    if args.only_formats:
        import torch

        torch.cuda.init()
        print(json.dumps({"formats": bench_formats(torch, args.formats_mib, max(args.steps // 2, 3))}), flush=True)
        return 0

    if args.only_kv:
        import torch

        torch.cuda.init()
        print(json.dumps({"kv_tables": bench_kv(torch, args.kv_mib, max(args.steps // 2, 3), args.no_cpu)}),
              flush=True)
        return 0

    if args.only_serve or args.only_batch:
        import torch

        torch.cuda.init()
        if args.only_serve:
            print(json.dumps({"served": bench_serve(args.serve_requests, base)}), flush=True)
        else:
            r = bench_batch(torch, args.batch_docs, args.batch_iters, args.plain_sample, base, tune=not args.no_tune)
            print(json.dumps({"batched": r}), flush=True)
        return 0

    def clike_line(torch) -> dict:
        n3 = args.clike_mib << 20
        log(f"generating {n3} bytes of six-mode C-like source (seed 3, depth 24)")
        src = W.clike(3, n3, 24)
        cpu3 = raw3 = None
        if base is not None:
            cpu3, raw3 = base.run("clike6", src, [0], [src.size], f"the whole {n3 >> 20} MiB C-like document")
            log(f"clike6 cpu baseline: {cpu3}")
        d3 = torch.from_numpy(src).to("cuda")
        torch.cuda.synchronize()
        sc3, r3, m3, _ = bench_document(torch, "clike6", d3.data_ptr(), n3, max(args.steps // 2, 3), args.warmup,
                                        args.knobs, not args.no_tune, "k_scan_clike6_1gib.md",
                                        "k_scan<EventEval<5, true>, 2, false, 4> (R1-R5)")
        line3 = {
            "workload": f"clike6-{n3 >> 20}MiB plain pipeline (R1-R6, device-resident, N1b; six modes, 5 descriptors)",
            "bytes": n3,
            "seed": 3,
            "metric": "input GB/s",
            "value": round(m3["value"], 3),
            "unit": "GB/s",
            "ms_per_step": round(m3["ms_per_step"], 4),
            "max_depth": m3["max_depth"],
            "n_struct": m3["n_struct"],
            "n_brackets": m3["n_brackets"],
            "tuned": m3["tuned"],
            "roofline": m3["roofline"],
            "e2e_roofline_frac": m3["e2e_roofline_frac"],
            "stage_ms": m3["stage_ms"],
        }
        if cpu3 is not None:
            line3["cpu_baseline"] = cpu3
            line3["parity"] = parity_single(r3, raw3)
            log(f"clike6 parity: {line3['parity']}")
        sc3.close()
        del d3
        torch.cuda.empty_cache()
        return line3

    if args.only_clike:
        import torch

        torch.cuda.init()
        print(json.dumps({"clike6": clike_line(torch)}), flush=True)
        return 0

    n = args.size_mib << 20
    log(f"generating {n} bytes of JSON (seed 2, depth 32, 30% strings)")
    t0 = time.perf_counter()
    doc = W.json(2, n, 32, 0.30)
    str_frac = W.json.last_string_bytes / n
    log(f"generated in {time.perf_counter() - t0:.1f} s, string fraction {str_frac:.4f}")
    cpu = raw = None
    if base is not None:
        cpu, raw = base.run("json", doc, [0], [n], f"the whole {n >> 20} MiB document")
        log(f"cpu baseline: {cpu}")

    import torch

    torch.cuda.init()
    d_doc = torch.from_numpy(doc).to("cuda")
    torch.cuda.synchronize()
    if args.copy_calibration:
        for _ in range(2):
            d_copy = d_doc.clone()
            torch.cuda.synchronize()
            del d_copy
        log(f"copy calibration: 2 device copies of {n} bytes")

    sc, r, m, sp = bench_document(torch, "json", d_doc.data_ptr(), n, args.steps, args.warmup, args.knobs,
                                  not args.no_tune, "k_scan_json_2gib.md", "k_scan<EscapeEval, {rows}, false, {waves}> (R1-R5)")
    line = {
        "metric": METRIC,
        "value": round(m["value"], 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(m["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {
            "workload": "json-2GiB plain pipeline (R1-R6, device-resident, N1b)",
            "bytes": n,
            "seed": 2,
            "max_depth": m["max_depth"],
            "string_fraction": round(str_frac, 4),
            "n_struct": m["n_struct"],
            "n_brackets": m["n_brackets"],
            "parallelism": "single GPU",
            "tuned": m["tuned"],
        },
        "roofline": m["roofline"],
        "e2e_roofline_frac": m["e2e_roofline_frac"],
        "stage_ms": m["stage_ms"],
    }
    if cpu is not None:
        line["cpu_baseline"] = cpu
        line["parity"] = parity_single(r, raw)
        log(f"parity: {line['parity']}")
    if not args.no_stream:
        line["host_stream"] = bench_stream(sc, doc, sp)
        assert line["host_stream"]["n_struct"] == m["n_struct"], "streamed scan disagrees with the device-resident one"
        log(f"host stream: {line['host_stream']}")
    if not args.no_tape:
        line["json_tape"] = bench_tape(d_doc.data_ptr(), doc, sp, 5, args.tape_cpu_mib << 20, args.no_cpu)
        log(f"json tape: {line['json_tape']}")
    sc.close()
    del d_doc
    torch.cuda.empty_cache()
    if not args.no_clike:
        line["clike6"] = clike_line(torch)
    if not args.no_formats:
        torch.cuda.empty_cache()
        line["formats"] = bench_formats(torch, args.formats_mib, max(args.steps // 2, 3))
    if not args.no_kv:
        torch.cuda.empty_cache()
        line["kv_tables"] = bench_kv(torch, args.kv_mib, max(args.steps // 2, 3), args.no_cpu)
    if not args.no_batch:
        line["batched"] = bench_batch(torch, args.batch_docs, args.batch_iters, args.plain_sample, base,
                                      tune=not args.no_tune)
        log(f"batched: { {k: v for k, v in line['batched'].items() if k != 'tuned'} }")
    if not args.no_serve:
        torch.cuda.empty_cache()
        line["served"] = bench_serve(args.serve_requests, base)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
