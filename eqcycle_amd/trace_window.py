"""Per-kernel device time of the window that profile_window.py times, from a
rocprofv3 kernel trace.

    python -m eqcycle_amd.trace_window <run_kernel_trace.csv>

profile_window.py runs one 512 MiB copy (`convert_kernel<double, double>`)
right before its window; every dispatch after the last such copy belongs to
the window.  Prints the window's span and busy time, the idle time between dispatches
(host synchronisations), then per kernel the call count, total and average
duration, and, for the Newmark kernel, the gaps between consecutive launches
(host synchronisations of the dynamic batches).
"""
from __future__ import annotations

import collections
import csv
import sys


def main():
    path = sys.argv[1]
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"]))
    rows.sort()
    mk = max(i for i, r in enumerate(rows) if "convert_kernel<double, double>" in r[2] and r[1] - r[0] > 100000)
    w = rows[mk + 1:]
    span = (w[-1][1] - w[0][0]) / 1e3
    busy = sum(e - s for s, e, _ in w) / 1e3
    agg = collections.defaultdict(lambda: [0, 0.0])
    for s, e, n in w:
        k = n.split("(eq::Geom")[0].split("(long")[0].split("(eq::CgScal")[0][:90]
        agg[k][0] += 1
        agg[k][1] += (e - s) / 1e3
    print(f"window: {len(w)} dispatches, span {span:.1f} us, busy {busy:.1f} us ({100 * busy / span:.1f} %)")
    print("| kernel | calls | total us | avg us | % of busy |")
    print("|---|---|---|---|---|")
    for k, (c, t) in sorted(agg.items(), key=lambda x: -x[1][1])[:25]:
        print(f"| `{k}` | {c} | {t:.1f} | {t / c:.2f} | {100 * t / busy:.1f} |")
    idle = sorted(w[i + 1][0] - w[i][1] for i in range(len(w) - 1))
    big = [g for g in idle if g > 5000]
    print(f"idle between dispatches: {sum(g for g in idle if g > 0) / 1e3:.1f} us in total; "
          f"{len(big)} gaps over 5 us hold {sum(big) / 1e3:.1f} us")
    nm = [(s, e) for s, e, n in w if "newmark" in n]
    gaps = sorted(nm[i + 1][0] - nm[i][1] for i in range(len(nm) - 1))
    if gaps:
        print(f"Newmark launches: {len(nm)}, gap median {gaps[len(gaps) // 2] / 1e3:.2f} us, mean "
              f"{sum(gaps) / len(gaps) / 1e3:.2f} us; gaps > 20 us: {sum(1 for g in gaps if g > 20000)}, "
              f"{sum(g for g in gaps if g > 20000) / 1e3:.1f} us in total")


if __name__ == "__main__":
    main()
