"""Summaries of rocprofv3 results databases (rocpd SQLite) for this library.

    python -m opentissue_amd.rocprof_summary trace RESULTS.db [--match SUBSTR] [--split-us T]
        per-kernel dispatch count, average / min / max duration (us), total
        (ms), share of the traced kernel time, grid (workgroups), block, LDS;
        with --split-us, dispatches shorter than T us are listed in a row of
        their own (CG kernels issued after every body stopped exit at once)
    python -m opentissue_amd.rocprof_summary pmc RESULTS.db [...] [--drop-below F]
        per-kernel counter values averaged over dispatches (JSON lines); with
        --drop-below F, dispatches whose counter value is below F times that
        kernel's largest are left out of the average (CG launches issued after
        every body stopped read almost nothing and would dilute it)

Only reads the database; nothing is executed on the device.
"""
from __future__ import annotations

import collections
import json
import sqlite3
import statistics
import sys


def _names(cur):
    strs = {r[0]: r[1] for r in cur.execute("select id, string from rocpd_string")}
    ks = {r[0]: r[1] for r in cur.execute("select id, display_name from rocpd_info_kernel_symbol")}
    return {k: strs.get(v, v) for k, v in ks.items()}


def kernels(db):
    con = sqlite3.connect(db)
    cur = con.cursor()
    names = _names(cur)
    out = collections.defaultdict(list)
    q = ("select kernel_id, start, end, grid_size_x, workgroup_size_x, group_segment_size "
         "from rocpd_kernel_dispatch")
    for kid, st, en, gx, wx, lds in cur.execute(q):
        out[str(names.get(kid, kid))].append((en - st, gx // max(1, wx), wx, lds))
    return out


def pmc(db, drop_below=None):
    """{kernel: {counter: mean per dispatch}} (plus "dispatches": the number
    averaged).  With drop_below, a dispatch counts only if its value of the
    kernel's first counter is at least drop_below times the largest one."""
    con = sqlite3.connect(db)
    cur = con.cursor()
    names = _names(cur)
    pm = {r[0]: r[1] for r in cur.execute("select id, name from rocpd_info_pmc")}
    per = collections.defaultdict(lambda: collections.defaultdict(lambda: collections.defaultdict(float)))
    q = ("select d.kernel_id, e.pmc_id, e.value, d.id from rocpd_pmc_event e "
         "join rocpd_kernel_dispatch d on e.event_id = d.event_id")
    for kid, pid, val, did in cur.execute(q):
        per[str(names.get(kid, kid))][did][pm.get(pid)] += val
    out = {}
    for k, disp in per.items():
        rows = list(disp.values())
        if drop_below is not None and rows:
            first = sorted(rows[0])[0]
            top = max(r.get(first, 0.0) for r in rows)
            rows = [r for r in rows if r.get(first, 0.0) >= drop_below * top]
        if not rows:
            continue
        keys = sorted({c for r in rows for c in r})
        out[k] = {c: sum(r.get(c, 0.0) for r in rows) / len(rows) for c in keys}
        out[k]["dispatches"] = len(rows)
    return out


def trace_table(db, match=None, split_us=None):
    rows = []
    for n, v in kernels(db).items():
        if match and match not in n:
            continue
        groups = [(n, v)]
        if split_us is not None:
            short = [x for x in v if x[0] < split_us * 1e3]
            groups = [(n, [x for x in v if x[0] >= split_us * 1e3]),
                      (n + f" [< {split_us:g} us]", short)]
        for name, g in groups:
            if not g:
                continue
            d = [x[0] for x in g]
            rows.append((sum(d), name, len(g), statistics.mean(d) / 1e3, min(d) / 1e3,
                         max(d) / 1e3, g[0][1], g[0][2], g[0][3]))
    rows.sort(reverse=True)
    total = sum(r[0] for r in rows) or 1
    lines = ["| kernel | calls | avg us | min us | max us | total ms | % | grid | block | LDS B |",
             "|---|---|---|---|---|---|---|---|---|---|"]
    for r in rows:
        lines.append(f"| `{r[1][:90]}` | {r[2]} | {r[3]:.2f} | {r[4]:.2f} | {r[5]:.2f} | "
                     f"{r[0] / 1e6:.2f} | {100 * r[0] / total:.1f} | {r[6]} | {r[7]} | {r[8]} |")
    return "\n".join(lines)


def main(argv):
    if len(argv) < 2 or argv[0] not in ("trace", "pmc"):
        print(__doc__)
        return 2
    if argv[0] == "trace":
        match = split = None
        if "--match" in argv:
            match = argv[argv.index("--match") + 1]
        if "--split-us" in argv:
            split = float(argv[argv.index("--split-us") + 1])
        print(trace_table(argv[1], match, split))
    else:
        drop = None
        args = list(argv[1:])
        if "--drop-below" in args:
            i = args.index("--drop-below")
            drop = float(args[i + 1])
            del args[i:i + 2]
        for db in args:
            for k, d in pmc(db, drop).items():
                print(json.dumps({"db": db, "kernel": k, "counters": d}))
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
