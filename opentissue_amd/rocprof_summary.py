"""Summaries of rocprofv3 results databases (rocpd SQLite) for this library.

    python -m opentissue_amd.rocprof_summary trace RESULTS.db [--match SUBSTR] [--split-us T]
        per-kernel dispatch count, average / min / max duration (us), total
        (ms), share of the traced kernel time, grid (workgroups), block, LDS;
        with --split-us, dispatches shorter than T us are listed in a row of
        their own (CG kernels issued after every body stopped exit at once)
    python -m opentissue_amd.rocprof_summary pmc RESULTS.db [...]
        per-kernel counter values averaged over dispatches (JSON lines)

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


def pmc(db):
    con = sqlite3.connect(db)
    cur = con.cursor()
    names = _names(cur)
    pm = {r[0]: r[1] for r in cur.execute("select id, name from rocpd_info_pmc")}
    agg = collections.defaultdict(lambda: collections.defaultdict(float))
    cnt = collections.defaultdict(set)
    q = ("select d.kernel_id, e.pmc_id, e.value, d.id from rocpd_pmc_event e "
         "join rocpd_kernel_dispatch d on e.event_id = d.event_id")
    for kid, pid, val, did in cur.execute(q):
        n = str(names.get(kid, kid))
        agg[n][pm.get(pid)] += val
        cnt[n].add(did)
    return {k: {c: v / len(cnt[k]) for c, v in d.items()} for k, d in agg.items()}


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
        for db in argv[1:]:
            for k, d in pmc(db).items():
                print(json.dumps({"db": db, "kernel": k, "counters": d}))
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
