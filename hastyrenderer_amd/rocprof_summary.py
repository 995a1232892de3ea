"""Markdown summaries of rocprofv3 CSV output for this library's kernels.

    python -m hastyrenderer_amd.rocprof_summary stats <dir>/run_kernel_stats.csv [<dir>/run_kernel_trace.csv]
    python -m hastyrenderer_amd.rocprof_summary pmc <dir>/run_counter_collection.csv COUNTER [match ...]

`stats` prints one row per kernel (calls, total and average duration, share,
and the register / LDS / scratch figures of the trace); `pmc` prints the
counter's mean value per dispatch for each kernel whose name contains one of
the `match` strings (all kernels when none is given).
"""
from __future__ import annotations

import csv
import sys
from collections import defaultdict


def short(name: str) -> str:
    n = name.replace("hasty::(anonymous namespace)::", "").replace("(anonymous namespace)::", "")
    return n.split("(")[0][:80]


def _col(row, *names):
    for n in names:
        if n in row and row[n] != "":
            return row[n]
    return None


def stats(stats_csv: str, trace_csv: str | None = None) -> str:
    res = {}
    if trace_csv:
        with open(trace_csv) as f:
            for r in csv.DictReader(f):
                k = _col(r, "Kernel_Name", "KernelName")
                res[k] = (_col(r, "VGPR_Count", "Arch_VGPR_Count"), _col(r, "SGPR_Count"),
                          _col(r, "LDS_Block_Size", "Lds_Size", "LDS_Size"),
                          _col(r, "Scratch_Size", "Private_Segment_Size"))
    out = ["| kernel | calls | total ms | avg us | % | VGPR | SGPR | LDS B | scratch B |",
           "|---|---|---|---|---|---|---|---|---|"]
    with open(stats_csv) as f:
        rows = list(csv.DictReader(f))
    for r in rows:
        name = _col(r, "Name", "KERNEL_NAME")
        calls = int(float(_col(r, "Calls", "CALLS") or 0))
        tot = float(_col(r, "TotalDurationNs", "TOTAL_DURATION") or 0)
        avg = float(_col(r, "AverageNs", "AVERAGE") or 0)
        pct = float(_col(r, "Percentage", "PERCENTAGE") or 0)
        v = res.get(name, ("", "", "", ""))
        out.append("| %s | %d | %.3f | %.1f | %.1f | %s | %s | %s | %s |" % (
            short(name), calls, tot / 1e6, avg / 1e3, pct, *[x if x is not None else "" for x in v]))
    return "\n".join(out)


def pmc(counter_csv: str, counter: str, matches=()) -> str:
    acc = defaultdict(list)
    with open(counter_csv) as f:
        for r in csv.DictReader(f):
            if _col(r, "Counter_Name") != counter:
                continue
            name = _col(r, "Kernel_Name")
            if matches and not any(m in name for m in matches):
                continue
            acc[short(name)].append(float(_col(r, "Counter_Value")))
    out = ["| kernel | dispatches | mean %s per dispatch | total |" % counter, "|---|---|---|---|"]
    for k, v in sorted(acc.items()):
        out.append("| %s | %d | %.1f | %.1f |" % (k, len(v), sum(v) / len(v), sum(v)))
    return "\n".join(out)


def main(argv):
    if len(argv) >= 2 and argv[0] == "stats":
        print(stats(argv[1], argv[2] if len(argv) > 2 else None))
    elif len(argv) >= 3 and argv[0] == "pmc":
        print(pmc(argv[1], argv[2], tuple(argv[3:])))
    else:
        print(__doc__)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
