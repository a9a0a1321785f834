"""Summaries of rocprofv3 result databases (rocpd SQLite) as short text.

Usage:
  python -m genasis_amd.profiling trace  <kt_results.db>
  python -m genasis_amd.profiling counter <results.db> <COUNTER>

``trace`` prints per-kernel calls, total and average duration (microseconds,
as the rocpd ``top_kernels`` view gives them) from a
``rocprofv3 --kernel-trace --stats`` run; ``counter`` prints the per-dispatch
mean of one PMC counter per kernel from a ``rocprofv3 --pmc`` run.
"""
from __future__ import annotations

import collections
import sqlite3
import sys


def kernel_trace(db_path: str, top: int = 12) -> str:
    db = sqlite3.connect(db_path)
    rows = db.execute(
        "select name, total_calls, total_duration, average, percentage "
        "from top_kernels").fetchall()
    # top_kernels divides the nanosecond timestamps by 1000
    out = [f"{'kernel':64s} {'calls':>6s} {'total_us':>14s} {'avg_us':>12s} {'pct':>6s}"]
    for name, calls, total, avg, pct in rows[:top]:
        out.append(f"{name[:64]:64s} {calls:6d} {total:14.1f} {avg:12.1f} {pct:6.2f}")
    return "\n".join(out)


def counter_means(db_path: str, counter: str) -> dict:
    db = sqlite3.connect(db_path)
    acc = collections.defaultdict(list)
    for name, val in db.execute(
            "select kernel_name, value from counters_collection where counter_name=?",
            (counter,)):
        acc[name].append(val)
    return {k: (sum(v) / len(v), len(v)) for k, v in acc.items()}


def counter_table(db_path: str, counter: str) -> str:
    out = [f"{'kernel':64s} {'dispatches':>10s} {counter + '_KiB_mean':>20s}"]
    for name, (mean, n) in sorted(counter_means(db_path, counter).items(),
                                  key=lambda kv: -kv[1][0]):
        out.append(f"{name[:64]:64s} {n:10d} {mean:20.1f}")
    return "\n".join(out)


if __name__ == "__main__":
    if len(sys.argv) >= 3 and sys.argv[1] == "trace":
        print(kernel_trace(sys.argv[2]))
    elif len(sys.argv) >= 4 and sys.argv[1] == "counter":
        print(counter_table(sys.argv[2], sys.argv[3]))
    else:
        print(__doc__)
        sys.exit(2)
