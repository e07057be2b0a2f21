"""Summaries of rocprofv3 output for the profiles/ notes.

``python -m gridfoam_amd.profiling kernels DIR [N]`` prints per-kernel call
counts and durations from the ``*_kernel_trace.csv`` files under DIR;
``python -m gridfoam_amd.profiling counters DIR`` prints, per kernel, the
mean value per dispatch of every counter in the ``*_counter_collection.csv``
files under DIR (run rocprofv3 with ``--output-format csv``);
``python -m gridfoam_amd.profiling gaps DIR [N]`` sums the GPU's idle time
between consecutive kernels per kernel transition.
"""

from __future__ import annotations

import collections
import csv
import pathlib
import re
import sys


def _short(name: str) -> str:
    name = re.sub(r"\(anonymous namespace\)::", "", name)
    name = re.sub(r"^void ", "", name)
    return re.sub(r"\(.*$", "", name)[:64]


def kernels(root: str, top: int = 25) -> str:
    agg = collections.defaultdict(list)
    for f in pathlib.Path(root).rglob("*_kernel_trace.csv"):
        with open(f, newline="") as fh:
            for row in csv.DictReader(fh):
                agg[_short(row["Kernel_Name"])].append(int(row["End_Timestamp"]) - int(row["Start_Timestamp"]))
    tot = sum(sum(v) for v in agg.values()) or 1
    lines = [f"{'kernel':64s} {'calls':>7s} {'total_ms':>9s} {'avg_us':>8s} {'min_us':>8s} {'max_us':>8s} {'%':>5s}"]
    for k, v in sorted(agg.items(), key=lambda kv: -sum(kv[1]))[:top]:
        lines.append(f"{k:64s} {len(v):7d} {sum(v) / 1e6:9.2f} {sum(v) / len(v) / 1e3:8.2f} {min(v) / 1e3:8.2f} "
                     f"{max(v) / 1e3:8.2f} {100 * sum(v) / tot:5.1f}")
    return "\n".join(lines)


def gaps(root: str, top: int = 20, max_gap_us: float = 5000.0, between: str = "") -> str:
    """Idle time of the GPU between consecutive kernels (all queues merged),
    summed per (previous kernel -> next kernel) transition; gaps longer than
    max_gap_us (set-up, compilation, host pauses between phases of the
    measurement) are left out.  With `between`, only the span from the
    second to the last launch of kernels whose name contains it counts."""
    ev = []
    for f in pathlib.Path(root).rglob("*_kernel_trace.csv"):
        with open(f, newline="") as fh:
            for row in csv.DictReader(fh):
                ev.append((int(row["Start_Timestamp"]), int(row["End_Timestamp"]), _short(row["Kernel_Name"])))
    ev.sort()
    if between:
        marks = [e[0] for e in ev if between in e[2]]
        if len(marks) >= 3:
            ev = [e for e in ev if marks[1] <= e[0] <= marks[-1]]
    agg = collections.defaultdict(list)
    busy_end, prev = None, None
    busy = idle = 0
    for st, en, name in ev:
        if busy_end is not None and st > busy_end:
            g = st - busy_end
            if g <= max_gap_us * 1e3:
                agg[(prev, name)].append(g)
                idle += g
        busy += en - st
        if busy_end is None or en > busy_end:
            busy_end, prev = en, name
    span = (ev[-1][1] - ev[0][0]) if ev else 0
    lines = [f"span {span / 1e6:.2f} ms, kernel time {busy / 1e6:.2f} ms, idle gaps <= {max_gap_us:.0f} us: "
             f"{idle / 1e6:.2f} ms",
             f"{'previous kernel -> next kernel':100s} {'gaps':>6s} {'total_ms':>9s} {'avg_us':>8s}"]
    for (a, b), v in sorted(agg.items(), key=lambda kv: -sum(kv[1]))[:top]:
        lines.append(f"{(a or '')[:48]:48s} -> {b[:48]:48s} {len(v):6d} {sum(v) / 1e6:9.2f} {sum(v) / len(v) / 1e3:8.2f}")
    return "\n".join(lines)


def counters(root: str) -> str:
    agg = collections.defaultdict(list)
    for f in pathlib.Path(root).rglob("*_counter_collection.csv"):
        with open(f, newline="") as fh:
            for row in csv.DictReader(fh):
                agg[(_short(row["Kernel_Name"]), row["Counter_Name"])].append(float(row["Counter_Value"]))
    lines = [f"{'kernel':64s} {'counter':>12s} {'dispatches':>10s} {'mean/dispatch':>14s}"]
    for (k, c), v in sorted(agg.items()):
        lines.append(f"{k:64s} {c:>12s} {len(v):10d} {sum(v) / len(v):14.1f}")
    return "\n".join(lines)


if __name__ == "__main__":
    mode, root = sys.argv[1], sys.argv[2]
    if mode == "kernels":
        print(kernels(root, int(sys.argv[3]) if len(sys.argv) > 3 else 25))
    elif mode == "counters":
        print(counters(root))
    elif mode == "gaps":
        print(gaps(root, int(sys.argv[3]) if len(sys.argv) > 3 else 20,
                   between=sys.argv[4] if len(sys.argv) > 4 else ""))
    else:
        raise SystemExit(f"unknown mode {mode!r}")
