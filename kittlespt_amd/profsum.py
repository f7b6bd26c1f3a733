"""Summaries of rocprofv3 result databases (rocpd SQLite) as short text.

    python -m kittlespt_amd.profsum stats <run_results.db>     # kernel stats
    python -m kittlespt_amd.profsum pmc <run_results.db> [kernel-substring]

`stats` prints the per-kernel totals of the kernel trace (the --stats summary:
calls, total and average duration); `pmc` prints per-kernel sums and per
dispatch averages of the collected counters (FETCH_SIZE / WRITE_SIZE in KiB).
"""
from __future__ import annotations

import sqlite3
import sys
from collections import defaultdict


def short(name, n=90):
    return name if len(name) <= n else name[: n - 3] + "..."


def stats(db):
    cur = sqlite3.connect(db).cursor()
    rows = list(cur.execute("select name, total_calls, total_duration, average, percentage from top_kernels"))
    # the top_kernels view reports durations in microseconds
    out = [f"{'calls':>7} {'total_ms':>11} {'avg_ms':>11} {'pct':>6}  kernel"]
    for name, calls, tot, avg, pct in rows:
        out.append(f"{calls:7d} {tot / 1e3:11.3f} {avg / 1e3:11.4f} {pct:6.2f}  {short(name)}")
    return "\n".join(out)


def pmc(db, needle=None):
    cur = sqlite3.connect(db).cursor()
    acc = defaultdict(lambda: [0, 0.0, set()])
    for name, counter, value, disp in cur.execute(
            "select e.name, e.counter_name, e.counter_value, e.dispatch_id from pmc_events e"):
        if needle and needle not in name:
            continue
        a = acc[(name, counter)]
        a[1] += value
        a[2].add(disp)
    out = [f"{'dispatches':>10} {'sum':>16} {'per_dispatch':>16}  counter  kernel"]
    for (name, counter), (_, total, disps) in sorted(acc.items(), key=lambda kv: -kv[1][1]):
        n = len(disps)
        out.append(f"{n:10d} {total:16.1f} {total / max(n, 1):16.1f}  {counter}  {short(name)}")
    return "\n".join(out)


if __name__ == "__main__":
    if len(sys.argv) < 3 or sys.argv[1] not in ("stats", "pmc"):
        print(__doc__)
        sys.exit(2)
    if sys.argv[1] == "stats":
        print(stats(sys.argv[2]))
    else:
        print(pmc(sys.argv[2], sys.argv[3] if len(sys.argv) > 3 else None))
