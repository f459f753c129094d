"""Summarises a rocprofv3 results database (rocpd SQLite) as text.

    python -m mechanic_amd.rocprof_summary RESULTS.db [--top 30] [--window KERNEL FIRST COUNT]
                                              [--last-ticks N --total-ticks T]

Prints, per kernel: calls, average / total duration and, when the run
collected hardware counters (--pmc), the average value of each counter per
dispatch. --window also prints the mean duration of dispatches FIRST ..
FIRST+COUNT-1 (in dispatch order) of the kernels whose name contains KERNEL.
--last-ticks N --total-ticks T averages each counter over the dispatches of
the last N of T ticks only (a kernel's calls per tick: its calls / T).
"""
from __future__ import annotations

import argparse
import sqlite3
import sys
from collections import defaultdict


def kernel_stats(cur):
    rows = cur.execute(
        "select name, count(*), avg(end - start), sum(end - start) from kernels group by name "
        "order by sum(end - start) desc").fetchall()
    return rows


def counter_stats(cur):
    try:
        rows = cur.execute(
            "select kernel_name, counter_name, avg(value), count(*) from counters_collection "
            "group by kernel_name, counter_name").fetchall()
    except sqlite3.Error:
        return {}
    out = defaultdict(dict)
    for kernel, counter, value, n in rows:
        out[kernel][counter] = (value, n)
    return out


def counter_tail_average(db: str, kernel_part: str, counter: str, last: int):
    """Mean per-dispatch value of `counter` over the last `last` dispatches of the
    kernels whose name contains `kernel_part` (dispatch order), or None."""
    con = sqlite3.connect(db)
    cur = con.cursor()
    try:
        cols = [c[1] for c in cur.execute("pragma table_info(counters_collection)").fetchall()]
        order = next((c for c in ("dispatch_id", "correlation_id", "start") if c in cols), "rowid")
        rows = cur.execute(
            f"select value from counters_collection where counter_name = ? and kernel_name like ? "
            f"order by {order}", (counter, f"%{kernel_part}%")).fetchall()
    except sqlite3.Error:
        return None
    finally:
        con.close()
    vals = [r[0] for r in rows][-last:]
    return sum(vals) / len(vals) if vals else None


def tail_counters(cur, last_ticks: int, total_ticks: int):
    """Per kernel and counter: the mean over the dispatches of the last
    `last_ticks` of `total_ticks` ticks, and the dispatches averaged."""
    try:
        cols = [c[1] for c in cur.execute("pragma table_info(counters_collection)").fetchall()]
        order = next((c for c in ("dispatch_id", "correlation_id", "start") if c in cols), "rowid")
        rows = cur.execute(f"select kernel_name, counter_name, value from counters_collection "
                           f"order by {order}").fetchall()
    except sqlite3.Error:
        return {}
    series = defaultdict(list)
    for kernel, counter, value in rows:
        series[(kernel, counter)].append(value)
    out = defaultdict(dict)
    for (kernel, counter), vals in series.items():
        per_tick = max(1, round(len(vals) / total_ticks))
        tail = vals[-last_ticks * per_tick:]
        out[kernel][counter] = (sum(tail) / len(tail), len(tail))
    return out


def window_average(cur, kernel_part: str, first: int, count: int):
    """Mean duration (ns) and count of dispatches first..first+count-1 (dispatch
    order) of the kernels whose name contains `kernel_part`."""
    rows = cur.execute("select end - start from kernels where name like ? order by dispatch_id",
                       (f"%{kernel_part}%",)).fetchall()
    vals = [r[0] for r in rows][first:first + count]
    return (sum(vals) / len(vals) if vals else None), len(vals), len(rows)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("db")
    ap.add_argument("--top", type=int, default=30)
    ap.add_argument("--window", nargs=3, metavar=("KERNEL", "FIRST", "COUNT"))
    ap.add_argument("--last-ticks", type=int, default=0)
    ap.add_argument("--total-ticks", type=int, default=0)
    args = ap.parse_args(argv)
    cur = sqlite3.connect(args.db).cursor()
    counters = counter_stats(cur)
    stats = kernel_stats(cur)
    total = sum(r[3] for r in stats) or 1
    print(f"{'kernel':64s} {'calls':>6s} {'avg_us':>9s} {'total_ms':>9s} {'pct':>6s}")
    for name, n, avg, tot in stats[: args.top]:
        print(f"{name[:64]:64s} {n:6d} {avg / 1e3:9.2f} {tot / 1e6:9.3f} {100 * tot / total:6.1f}")
    if counters:
        print("\ncounters (average per dispatch)")
        for kernel in sorted(counters, key=lambda k: -sum(v[0] for v in counters[k].values())):
            vals = ", ".join(f"{c}={v:.4g}" for c, (v, _) in sorted(counters[kernel].items()))
            print(f"{kernel[:64]:64s} {vals}")
    if args.last_ticks and args.total_ticks:
        tail = tail_counters(cur, args.last_ticks, args.total_ticks)
        print(f"\ncounters (average per dispatch over the last {args.last_ticks} of {args.total_ticks} ticks)")
        for kernel in sorted(tail, key=lambda k: -sum(v[0] for v in tail[k].values())):
            vals = ", ".join(f"{c}={v:.4g} (n={n})" for c, (v, n) in sorted(tail[kernel].items()))
            print(f"{kernel[:64]:64s} {vals}")
    if args.window:
        name, first, count = args.window[0], int(args.window[1]), int(args.window[2])
        avg, n, total_n = window_average(cur, name, first, count)
        if avg is None:
            print(f"\nwindow {name} [{first}, {first + count}): no dispatches (of {total_n})")
        else:
            print(f"\nwindow {name} dispatches [{first}, {first + n}) of {total_n}: mean {avg / 1e3:.2f} us")
    return 0


if __name__ == "__main__":
    sys.exit(main())
