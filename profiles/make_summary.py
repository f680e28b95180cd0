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
tot = sum(r[5] for r in rows)
print(f"command: {cmd}")
print(f"{'kernel':60s} {'calls':>6s} {'avg_us':>10s} {'min_us':>9s} {'max_us':>9s} "
      f"{'total_ms':>9s} {'share':>6s} {'vgpr':>4s} {'sgpr':>4s} {'lds':>6s}")
for name, n, avg, mn, mx, s, vg, sg, lds in rows:
    short = name.split("(")[0].replace("odehip::(anonymous namespace)::", "")
    print(f"{short[:60]:60s} {n:6d} {avg / 1e3:10.2f} {mn / 1e3:9.2f} {mx / 1e3:9.2f} "
          f"{s / 1e6:9.2f} {100 * s / tot:5.1f}% {vg:4d} {sg:4d} {lds:6d}")
