"""Per-kernel statistics of the last dispatches in a rocprofv3 CSV file.

A whole bench run dispatches the pass kernels thousands of times (the sweep
comparison) before the timed matvecs of the real operator come last; this
summarises only the last N dispatches of the kernels whose names match a
pattern, grouped by kernel and grid size, so that the averages describe the
timed launches (and the autotuner's candidates dispatched just before them).

    python3 -m dmrg_amd.profile_tail <kernel_trace.csv | counter_collection.csv> N [name-regex] [--by-name]

Kernel traces give durations (us); counter collections give each counter's
mean value per dispatch.  With --by-name a kernel trace is grouped by kernel
only (all grid sizes together), with each kernel's share of the summed time.
"""
from __future__ import annotations

import collections
import csv
import re
import sys


def kernel_name(full: str) -> str:
    """The kernel's name with its template arguments, without the namespace
    and the parameter list ("k_gemm_bf3p<true, 1>")."""
    m = re.search(r"\b(k_\w+(?:<[^<>]*>)?)\(", full)
    if m:
        return m.group(1)
    return full.split("(")[0][-40:] if not full.startswith("void ") else full[5:45]


def summarise(path: str, n: int, pattern: str = ".", by_name: bool = False) -> str:
    rx = re.compile(pattern)
    rows = [r for r in csv.DictReader(open(path)) if rx.search(r["Kernel_Name"])]
    counters = "Counter_Name" in (rows[0] if rows else {})
    key_id = "Dispatch_Id" if counters else "Start_Timestamp"
    ids = sorted({int(r[key_id]) for r in rows})[-n:]
    keep = set(ids)
    rows = [r for r in rows if int(r[key_id]) in keep]
    out = [f"{path}: last {len(ids)} dispatches matching {pattern!r}"]
    if counters:
        agg = collections.defaultdict(lambda: collections.defaultdict(list))
        for r in rows:
            k = (kernel_name(r["Kernel_Name"]), r["Grid_Size"])
            agg[k][r["Counter_Name"]].append(float(r["Counter_Value"]))
        for (name, grid), cs in sorted(agg.items()):
            vals = ", ".join(f"{c} {sum(v) / len(v):.1f}" for c, v in sorted(cs.items()))
            out.append(f"{name:40s} grid {grid:>9s} dispatches {len(next(iter(cs.values()))):4d}: {vals}")
    else:
        agg = collections.defaultdict(list)
        for r in rows:
            gx = "all" if by_name else (r.get("Grid_Size_X") or r.get("Grid_Size", ""))
            agg[(kernel_name(r["Kernel_Name"]), gx)].append(
                (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1000)
        whole = sum(sum(v) for v in agg.values()) or 1.0
        for (name, grid), v in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
            share = f" ({100 * sum(v) / whole:4.1f} %)" if by_name else ""
            out.append(f"{name:40s} grid {grid:>9s} calls {len(v):4d} avg {sum(v) / len(v):8.1f} us "
                       f"min {min(v):8.1f} total {sum(v):9.1f}{share}")
    return "\n".join(out)


if __name__ == "__main__":
    args = [a for a in sys.argv[1:] if a != "--by-name"]
    if len(args) < 2:
        sys.exit(__doc__)
    print(summarise(args[0], int(args[1]), args[2] if len(args) > 2 else ".", "--by-name" in sys.argv[1:]))
