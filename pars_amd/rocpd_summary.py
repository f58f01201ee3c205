"""Summarise a rocprofv3 result database (rocpd SQLite) as a short text table.

    python -m pars_amd.rocpd_summary <results.db> [more.db ...]

Kernel-trace databases give per-kernel call counts and durations (the rocpd
tables hold microseconds; printed here in milliseconds); counter
databases (rocprofv3 --pmc) give the per-dispatch average of each collected
counter (FETCH_SIZE / WRITE_SIZE are in KiB per dispatch).
"""
from __future__ import annotations

import os
import sqlite3
import sys
from collections import defaultdict


def short(name: str, width: int = 70) -> str:
    name = name.replace("pars_hip::", "").replace("(anonymous namespace)::", "")
    return name if len(name) <= width else name[: width - 3] + "..."


def summarize(path: str) -> str:
    if not os.path.isfile(path):
        raise SystemExit(f"{path}: no such result database")
    con = sqlite3.connect(path)
    out = [f"# {path}"]
    names = {r[0] for r in con.execute("select name from sqlite_master where type in ('table', 'view')")}
    kernels = []
    if "top_kernels" in names:
        kernels = list(con.execute("select name, total_calls, total_duration, average, percentage from top_kernels"))
    if kernels:
        out.append(f"{'kernel':70s} {'calls':>6s} {'avg_ms':>10s} {'total_ms':>11s} {'pct':>6s}")
        for name, calls, tot, avg, pct in kernels:
            out.append(f"{short(name):70s} {calls:6d} {avg / 1e3:10.4f} {tot / 1e3:11.3f} {pct:6.2f}")
    rows = []
    if "counters_collection" in names:
        rows = list(con.execute("select kernel_name, counter_name, value, duration from counters_collection"))
    if rows:
        acc = defaultdict(list)
        for kname, cname, value, _dur in rows:
            acc[(kname, cname)].append(value)
        out.append(f"{'kernel':70s} {'counter':>12s} {'dispatches':>10s} {'avg_value':>14s}")
        for (kname, cname), vals in sorted(acc.items()):
            out.append(f"{short(kname):70s} {cname:>12s} {len(vals):10d} {sum(vals) / len(vals):14.1f}")
    return "\n".join(out)


if __name__ == "__main__":
    for p in sys.argv[1:]:
        print(summarize(p))
        print()
