"""Summarise a rocprofv3 kernel trace (rocpd SQLite output) as a text table.

Usage: python profiles/make_summary.py <run_results.db> "<command that made it>"
Prints per-kernel dispatch count, average / min / max duration and share.
"""
import sqlite3
import sys

db, cmd = sys.argv[1], sys.argv[2]
c = sqlite3.connect(db)
rows = list(c.execute("select name, count(*), avg(duration), min(duration), max(duration), "
                      "sum(duration), max(vgpr_count), max(sgpr_count), max(lds_size) "
                      "from kernels group by name order by 6 desc"))
durs = {}
for name, d in c.execute("select name, duration from kernels"):
    durs.setdefault(name, []).append(d)


def working_avg(name):
    """Average over launches that did work (a finished solve's kernels return
    at once; those launches take < 10% of the longest)."""
    d = sorted(durs[name])
    keep = [x for x in d if x >= 0.1 * d[-1]]
    return sum(keep) / len(keep), len(keep)
tot = sum(r[5] for r in rows)
print(f"command: {cmd}")
print(f"{'kernel':34s} {'calls':>6s} {'avg_us':>9s} {'work_n':>6s} {'work_avg_us':>11s} "
      f"{'min_us':>8s} {'max_us':>8s} {'total_ms':>9s} {'share':>6s} {'vgpr':>4s} {'sgpr':>4s} "
      f"{'lds':>6s}")
for name, n, avg, mn, mx, s, vg, sg, lds in rows:
    short = name.replace("void ", "").replace("odehip::(anonymous namespace)::", "")
    short = short.split("(")[0]
    wavg, wn = working_avg(name)
    print(f"{short[:34]:34s} {n:6d} {avg / 1e3:9.2f} {wn:6d} {wavg / 1e3:11.2f} {mn / 1e3:8.2f} "
          f"{mx / 1e3:8.2f} {s / 1e6:9.2f} {100 * s / tot:5.1f}% {vg:4d} {sg:4d} {lds:6d}")
