"""Summarise a rocprofv3 kernel trace as a text table.

Usage: python -m ode_prototyping_amd.profile_summary <run_results.db | kernel_trace.csv> \
       "<command>" [--last=K]
Prints per-kernel dispatch count, average / min / max duration and share, and
the average over the launches that did most work (>= 50 % of the longest:
the launches of a solve in which every system is still running -- a finished
system's blocks return at once).  With --last=K, also the average of each
kernel's K latest launches (by start time): for a program that ends with K
timed back-to-back launches of one kernel, the figure its own events measure.
"""
import csv
import sqlite3
import sys

src, cmd = sys.argv[1], sys.argv[2]
last = 0
for a in sys.argv[3:]:
    if a.startswith("--last="):
        last = int(a.split("=", 1)[1])
durs, starts, regs = {}, {}, {}
if src.endswith(".csv"):
    for r in csv.DictReader(open(src)):
        name = r["Kernel_Name"]
        t0 = int(r["Start_Timestamp"])
        durs.setdefault(name, []).append(int(r["End_Timestamp"]) - t0)
        starts.setdefault(name, []).append(t0)
        regs[name] = (int(r["VGPR_Count"]), int(r["SGPR_Count"]), int(r["LDS_Block_Size"]))
else:
    c = sqlite3.connect(src)
    for name, t0, d, vg, sg, lds in c.execute(
            "select name, start, duration, vgpr_count, sgpr_count, lds_size from kernels"):
        durs.setdefault(name, []).append(d)
        starts.setdefault(name, []).append(t0)
        regs[name] = (vg, sg, lds)

tot = sum(sum(v) for v in durs.values())
print(f"command: {cmd}")
print(f"{'kernel':42s} {'calls':>6s} {'avg_us':>9s} {'full_n':>6s} {'full_avg_us':>11s} "
      f"{'min_us':>8s} {'max_us':>8s} {'total_ms':>9s} {'share':>6s} {'vgpr':>4s} {'sgpr':>4s} "
      f"{'lds':>6s}" + (f" {f'last{last}_us':>10s}" if last else ""))
for name in sorted(durs, key=lambda k: -sum(durs[k])):
    d = sorted(durs[name])
    by_start = [x for _, x in sorted(zip(starts[name], durs[name]))][-last:] if last else []
    full = [x for x in d if x >= 0.5 * d[-1]]
    short = name.replace("void ", "").replace("odehip::(anonymous namespace)::", "")
    short = short.replace("odehip::", "").split("(")[0]
    vg, sg, lds = regs[name]
    s = sum(d)
    print(f"{short[:42]:42s} {len(d):6d} {s / len(d) / 1e3:9.2f} {len(full):6d} "
          f"{sum(full) / len(full) / 1e3:11.2f} {d[0] / 1e3:8.2f} {d[-1] / 1e3:8.2f} "
          f"{s / 1e6:9.2f} {100 * s / tot:5.1f}% {vg:4d} {sg:4d} {lds:6d}"
          + (f" {sum(by_start) / len(by_start) / 1e3:10.2f}" if last else ""))
