"""Summaries of rocprofv3 result databases (the rocpd SQLite files written under -d DIR).

    python -m blochlib_amd.profsum DIR_OR_DB [--command "..."] [--out FILE]

Kernel table: per kernel (template arguments kept, parameter list dropped) the dispatch count,
total and average duration from the `kernels` view (start/end in ns).  Counter table (databases
of --pmc runs): per kernel and counter the value summed over each dispatch's instances, averaged
over dispatches and its maximum, with the dispatch count.  FETCH_SIZE and WRITE_SIZE are in KiB
per dispatch.
"""
from __future__ import annotations

import argparse
import glob
import os
import sqlite3
from collections import defaultdict


def short_name(name: str) -> str:
    s = name.split("(")[0]
    return s[5:] if s.startswith("void ") else s


def find_db(path: str) -> str:
    if os.path.isfile(path):
        return path
    hits = sorted(glob.glob(os.path.join(path, "**", "*.db"), recursive=True))
    if not hits:
        raise SystemExit(f"no rocpd database under {path}")
    return hits[0]


def kernel_table(con) -> list:
    """[(kernel, calls, total_ms, avg_ms)] sorted by total time."""
    agg = defaultdict(lambda: [0, 0.0])
    for name, start, end in con.execute("select name, start, end from kernels"):
        a = agg[short_name(name)]
        a[0] += 1
        a[1] += (end - start) * 1e-6
    rows = [(k, n, t, t / n) for k, (n, t) in agg.items()]
    return sorted(rows, key=lambda r: -r[2])


def counter_table(con) -> list:
    """[(kernel, counter, dispatches, mean value per dispatch, max value per dispatch)]."""
    try:
        cur = con.execute("select kernel_name, dispatch_id, counter_name, value "
                          "from counters_collection")
    except sqlite3.OperationalError:
        return []
    per = defaultdict(float)
    for kname, did, cname, val in cur:
        per[(short_name(kname), cname, did)] += float(val)
    acc = defaultdict(lambda: [0, 0.0, 0.0])
    for (k, cname, _), v in per.items():
        a = acc[(k, cname)]
        a[0] += 1
        a[1] += v
        a[2] = max(a[2], v)
    return sorted(((k, c, n, s / n, mx) for (k, c), (n, s, mx) in acc.items()))


def summarize(path: str, command: str = "") -> str:
    db = find_db(path)
    con = sqlite3.connect(db)
    out = []
    if command:
        out.append(f"# command (repository root): {command}")
    kt = kernel_table(con)
    if kt:
        tot = sum(r[2] for r in kt)
        out.append(f"# kernel dispatches: {sum(r[1] for r in kt)}, summed kernel time "
                   f"{tot:.1f} ms")
        out.append(f"{'kernel':<34s} {'calls':>6s} {'total_ms':>11s} {'avg_ms':>10s} {'pct':>6s}")
        for k, n, t, a in kt:
            out.append(f"{k:<34s} {n:>6d} {t:>11.2f} {a:>10.4f} {100 * t / tot:>6.2f}")
    ct = counter_table(con)
    if ct:
        out.append("")
        out.append(f"{'kernel':<34s} {'counter':<26s} {'dispatches':>10s} {'mean/dispatch':>16s} "
                   f"{'max/dispatch':>16s}")
        for k, c, n, v, mx in ct:
            out.append(f"{k:<34s} {c:<26s} {n:>10d} {v:>16.1f} {mx:>16.1f}")
    return "\n".join(out) + "\n"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("path")
    ap.add_argument("--command", default="")
    ap.add_argument("--out", default="")
    a = ap.parse_args()
    text = summarize(a.path, a.command)
    if a.out:
        with open(a.out, "w") as f:
            f.write(text)
    print(text, end="")


if __name__ == "__main__":
    main()
