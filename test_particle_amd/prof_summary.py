"""Summarise rocprofv3 SQLite outputs (``*_results.db``) as text.

    python -m test_particle_amd.prof_summary trace DB [DB...]
        per-kernel calls, total and average duration
    python -m test_particle_amd.prof_summary pmc DB [--kernel SUBSTR]
        per-kernel average of each collected counter per dispatch
"""
import argparse
import sqlite3
import sys
from collections import defaultdict


def trace(dbs):
    rows = defaultdict(lambda: [0, 0.0])
    for db in dbs:
        con = sqlite3.connect(db)
        for name, dur in con.execute("select name, duration from kernels"):
            r = rows[name]
            r[0] += 1
            r[1] += dur
    tot = sum(v[1] for v in rows.values()) or 1.0
    print(f"{'calls':>7} {'total_ms':>10} {'avg_us':>10} {'pct':>6}  kernel")
    for name, (n, d) in sorted(rows.items(), key=lambda kv: -kv[1][1]):
        print(f"{n:7d} {d / 1e6:10.3f} {d / n / 1e3:10.2f} {100 * d / tot:6.2f}  {name[:110]}")


def pmc(dbs, kernel=None):
    acc = defaultdict(lambda: [0, 0.0])
    for db in dbs:
        con = sqlite3.connect(db)
        q = "select kernel_name, counter_name, value, dispatch_id from counters_collection"
        for name, cname, value, _ in con.execute(q):
            if kernel and kernel not in name:
                continue
            a = acc[(name, cname)]
            a[0] += 1
            a[1] += value
    print(f"{'dispatches':>10} {'mean_per_dispatch':>18}  counter  kernel")
    for (name, cname), (n, v) in sorted(acc.items()):
        print(f"{n:10d} {v / n:18.1f}  {cname}  {name[:100]}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["trace", "pmc"])
    ap.add_argument("dbs", nargs="+")
    ap.add_argument("--kernel")
    a = ap.parse_args()
    if a.what == "trace":
        trace(a.dbs)
    else:
        pmc(a.dbs, a.kernel)
    sys.stdout.flush()


if __name__ == "__main__":
    main()
