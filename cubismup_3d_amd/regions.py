"""Per-region kernel times from a rocprofv3 database: each kernel dispatch is filed under the
roctx range that enclosed its HIP launch call, i.e. under the reference's profiler region names
(SimulationData::startProfiler, SimulationData.cpp:293-305) that the host layer and the fused
step's stages open. Needs a run with --marker-trace --hip-runtime-trace --kernel-trace:

    rocprofv3 --marker-trace --hip-runtime-trace --kernel-trace -d out -o run -- <program>
    python -m cubismup_3d_amd.regions out/run_results.db

In the database a kernel's stack_id is that of its launch call, and the call's parent_stack_id
that of the innermost range open on the launching thread. Prints a markdown table per database:
region, ranges opened, kernel, launches, mean and total GPU time."""
from __future__ import annotations

import collections
import json
import sqlite3
import sys


def summarize(db: str):
    """{region: {kernel: (launches, total ns)}} in order of first launch, and ranges per region."""
    c = sqlite3.connect(db)
    marks, opened = {}, collections.Counter()
    for sid, x in c.execute("select stack_id, extdata from regions where category like 'MARKER%'"):
        marks[sid] = json.loads(x).get("message", "")
        opened[marks[sid]] += 1
    parent = dict(c.execute("select stack_id, parent_stack_id from regions where category like 'HIP%'"))
    per = collections.OrderedDict()
    for name, sid, dur in c.execute("select name, stack_id, duration from kernels order by start"):
        p = parent.get(sid)
        reg = "(no launch record)" if p is None else marks.get(p, "(outside any region)")
        k = name.split("(")[0].replace("void ", "")
        d = per.setdefault(reg, collections.OrderedDict())
        n, tot = d.get(k, (0, 0))
        d[k] = (n + 1, tot + dur)
    return per, opened


def table(db: str) -> str:
    per, opened = summarize(db)
    out = ["| region | ranges | kernel | launches | mean us | total ms |", "|---|---|---|---|---|---|"]
    for reg, ks in per.items():
        for k, (n, tot) in sorted(ks.items(), key=lambda kv: -kv[1][1]):
            out.append(f"| {reg} | {opened.get(reg, '-')} | `{k}` | {n} | {tot / n / 1e3:.1f} | {tot / 1e6:.3f} |")
    return "\n".join(out)


if __name__ == "__main__":
    for path in sys.argv[1:]:
        print(f"### {path}\n\n{table(path)}\n")
