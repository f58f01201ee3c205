"""Benchmark: the plain JSON pipeline on one seeded 2 GiB document, one MI355X.

A step is one device-resident scan of the whole document through the C ABI
(pars_hip_scan_device: R1 classify .. R5 compaction in one kernel, then R6
pairing; the reference's N1b tier, bench/bench_boundary.cu:252-260), with
the input already in HBM.  `value` = input bytes / step time / 1e9 (the
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

cpu_baseline: the reference's own CPU backend (model::scanner + pair_brackets,
compiled from its sources into oracle/_ref/) on a bounded prefix of the same
document, one core; the C restatement under oracle/ stands in ("port") where
that binary was not built.

    python bench.py [--gpus 1] [--steps 10] [--warmup 2]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "input GB/s + % HBM roofline (2 GiB JSON); batched/served vs plain; p99 latency"
HBM_PEAK_GBS = 8000.0
SCAN_KERNEL = "k_scan<EscapeEval, 4> (R1-R5)"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_scan_json_2gib.md")


def measured_traffic(n: int) -> int | None:
    """HBM bytes per scan launch from the committed counter summary (same workload size)."""
    try:
        with open(TRAFFIC_FILE) as f:
            fields = dict(
                line.strip().lstrip("- ").split(": ", 1) for line in f if line.strip().lstrip("- ").startswith("traffic_")
            )
        if int(fields.get("traffic_input_bytes", -1)) != n:
            return None
        return int(fields["traffic_bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(doc, sample_bytes: int) -> dict:
    """Reference CPU path on the first `sample_bytes` of the document, one core."""
    from oracle import oracle as O

    sample = doc[:sample_bytes]
    cores = 1
    if O.ref_available():
        with tempfile.NamedTemporaryFile(suffix=".json", delete=False) as f:
            path = f.name
        try:
            sample.tofile(path)
            cmd = [O.REF, "time", "json", path, str(sample.size)]
            r = json.loads(subprocess.run(cmd, check=True, stdout=subprocess.PIPE, text=True).stdout)
        finally:
            os.unlink(path)
        secs = r["scan_s"] + r["pair_s"]
        kind = "reference"
    else:
        from pars_amd import Plan

        plan = Plan.load("json").blob
        t0 = time.perf_counter()
        s = O.scan(plan, sample)
        O.pair(plan, s["bracket_char"])
        secs = time.perf_counter() - t0
        kind = "port"
    return {
        "value": round(sample.size / secs / 1e9, 4),
        "unit": "GB/s",
        "cores": cores,
        "kind": kind,
        "sample": f"first {sample.size >> 20} MiB of the 2 GiB document, scan + pair_brackets, {secs:.2f} s",
    }


def bench_batch(torch, n_docs: int, iters: int, plain_sample: int, tune: bool = True) -> dict:
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

    # the same documents through one plain launch chain each (bounded sample, in document order)
    plain_t, plain_docs, plain_bytes = 0.0, 0, 0
    for kind, d, spans, sc, nd in groups:
        k = max(1, plain_sample * nd // n_docs)
        sc1 = Scanner(kind)
        base = d.data_ptr()
        sc1.scan_device(base + spans[0].offset, spans[0].size, sp)  # warm
        t0 = time.perf_counter()
        for i in range(min(k, nd)):
            sc1.scan_device(base + spans[i].offset, spans[i].size, sp, want_pairs=False, run_pairs=True)
        plain_t += time.perf_counter() - t0
        plain_docs += min(k, nd)
        plain_bytes += int(sum(spans[i].size for i in range(min(k, nd))))
        sc1.close()
    for g in groups:
        g[3].close()
    docs_s = n_docs / t_batch
    plain_docs_s = plain_docs / plain_t
    return {
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
        "n_struct": int(sum(x[0] for x in first)),
        "n_brackets": int(sum(x[1] for x in first)),
        "graph": True,
        "tuned": tuned or None,
    }


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

    from pars_amd import Scanner
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
        pos = O.scan(Plan_json(), sample)["positions"]
        t0 = time.perf_counter()
        r = O.json_tape(sample, pos)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(sample.size / dt / 1e9, 4), "unit": "GB/s", "cores": 1, "kind": "port",
                               "stage": "replay + JSON hook over a given structural index (oracle/json_dom.c)",
                               "sample": f"first {sample.size} bytes of the 2 GiB document, {r['nodes']} nodes"}
    return out


def Plan_json() -> bytes:
    from pars_amd import Plan

    return Plan.load("json").blob


def bench_serve(n_req: int, seed: int = 5) -> dict:
    """Serving: Poisson arrivals of JSON / CSV / C-like requests (60/25/15,
    log-normal sizes, median 4 KiB, sigma 1.85, clipped to [200 B, 4 MiB])
    through a 4 GiB device ring under a 64 GiB HBM budget.  Saturation is
    measured first (closed loop: every request offered at once); then offered
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
    sv = Server(plans)  # 4 GiB ring, 1 GiB staging, 64 GiB budget, 64 MiB batches
    sv.replay(pidx[:200], docs[:200], np.zeros(200, np.uint64))  # warm every plan's kernels
    lat, el = sv.replay(pidx, docs, np.zeros(n_req, np.uint64))
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
    return {
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
        "device_bytes": int(st.device_bytes),
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size-mib", type=int, default=2048)
    ap.add_argument("--cpu-sample-mib", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain device copy of the document (counter calibration under rocprofv3)")
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
    ap.add_argument("--tape-cpu-mib", type=int, default=256, help="CPU baseline sample of the JSON tape")
    ap.add_argument("--knobs", default="",
                    help="fixed launch knobs for the document scan instead of tuning, e.g. lb_window=16,rows=4 "
                         "(rows, lb_window, blocks_per_cu, radix_bits); profiles use this")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for {args.gpus})", file=sys.stderr)
        return 2

    import numpy as np

    from pars_amd import Plan, Scanner
    from pars_amd import _native as N
    from pars_amd import workloads as W

    if args.only_serve:
        import torch

        torch.cuda.init()
        print(json.dumps({"served": bench_serve(args.serve_requests)}), flush=True)
        return 0
    if args.only_batch:
        import torch

        torch.cuda.init()
        r = bench_batch(torch, args.batch_docs, args.batch_iters, args.plain_sample, tune=not args.no_tune)
        print(json.dumps({"batched": r}), flush=True)
        return 0

    n = args.size_mib << 20
    log(f"generating {n} bytes of JSON (seed 2, depth 32, 30% strings)")
    t0 = time.perf_counter()
    doc = W.json(2, n, 32, 0.30)
    str_frac = W.json.last_string_bytes / n
    log(f"generated in {time.perf_counter() - t0:.1f} s, string fraction {str_frac:.3f}")
    cpu = None
    if not args.no_cpu:  # before the GPU is initialised: the baseline runs as a child process
        log("cpu baseline")
        cpu = cpu_baseline(doc, min(n, args.cpu_sample_mib << 20))
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

    sc = Scanner(Plan.load("json"))
    sc.set_option(N.PARS_OPT_WANT_BITMAP, 1)
    stream = torch.cuda.Stream()
    sp = stream.cuda_stream
    st = N.StageTimes()
    ptr = d_doc.data_ptr()
    tuned = None
    if args.knobs:
        opts = {"rows": N.PARS_OPT_ROWS_PER_TILE, "lb_window": N.PARS_OPT_LOOKBACK_WINDOW,
                "blocks_per_cu": N.PARS_OPT_BLOCKS_PER_CU, "radix_bits": N.PARS_OPT_RADIX_BITS}
        tuned = {}
        for kv in args.knobs.split(","):
            k, v = kv.split("=")
            sc.set_option(opts[k.strip()], int(v))
            tuned[k.strip()] = int(v)
        log(f"fixed knobs: {tuned}")
    elif not args.no_tune:
        t0 = time.perf_counter()
        tuned = sc.autotune(ptr, n, stream=sp, reps=3)
        log(f"autotune ({time.perf_counter() - t0:.1f} s): {tuned}")

    for i in range(args.warmup):
        r = sc.scan_device(ptr, n, sp, want_pairs=False, run_pairs=True, stage_times=st)
        log(f"warmup {i}: {r}")

    scan_ms, pair_ms, counts = [], [], set()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        r = sc.scan_device(ptr, n, sp, want_pairs=False, run_pairs=True, stage_times=st)
        scan_ms.append(st.p1)
        pair_ms.append(st.p4)
        counts.add((r.n_struct, r.n_brackets, r.balanced, r.well_formed))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = wall / args.steps * 1e3
    assert len(counts) == 1, f"steps disagree: {counts}"
    k, kb = r.n_struct, r.n_brackets
    log(f"steps done: {ms_step:.3f} ms/step; scan kernel {np.mean(scan_ms):.3f} ms, pairing {np.mean(pair_ms):.3f} ms")

    b_alg = n + n / 8 + 4 * k + 4 * kb
    t_scan = float(np.mean(scan_ms)) / 1e3
    achieved = b_alg / t_scan / 1e9
    value = n / (ms_step / 1e3) / 1e9
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {
            "workload": "json-2GiB plain pipeline (R1-R6, device-resident, N1b)",
            "bytes": n,
            "seed": 2,
            "max_depth": int(r.max_depth),
            "string_fraction": round(str_frac, 4),
            "n_struct": int(k),
            "n_brackets": int(kb),
            "parallelism": "single GPU",
            "tuned": tuned,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": SCAN_KERNEL,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(n),
            "b_alg_bytes": int(b_alg),
            "kernel_ms": round(t_scan * 1e3, 4),
        },
        "e2e_roofline_frac": round(b_alg / (ms_step / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
        "stage_ms": {"r1_r5": round(float(np.mean(scan_ms)), 4), "r6": round(float(np.mean(pair_ms)), 4)},
    }
    if cpu is not None:
        line["cpu_baseline"] = cpu
    if not args.no_stream:
        line["host_stream"] = bench_stream(sc, doc, sp)
        assert line["host_stream"]["n_struct"] == k, "streamed scan disagrees with the device-resident one"
        log(f"host stream: {line['host_stream']}")
    if not args.no_tape:
        line["json_tape"] = bench_tape(d_doc.data_ptr(), doc, sp, 5, args.tape_cpu_mib << 20, args.no_cpu)
        log(f"json tape: {line['json_tape']}")
    sc.close()
    del d_doc
    torch.cuda.empty_cache()
    if not args.no_batch:
        line["batched"] = bench_batch(torch, args.batch_docs, args.batch_iters, args.plain_sample,
                                      tune=not args.no_tune)
        log(f"batched: {line['batched']}")
    if not args.no_serve:
        torch.cuda.empty_cache()
        line["served"] = bench_serve(args.serve_requests)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
