"""Summaries of rocprofv3 result databases (rocpd SQLite) for profiles/.

    python parallel-computing_amd/rocprof_summary.py stats RUN.db OUT.txt --cmd "..."
        per-kernel dispatch count and average / min / max / total duration
    python parallel-computing_amd/rocprof_summary.py gaps RUN.db OUT.txt --cmd "..."
        per kernel: average duration and average idle time since the previous
        dispatch ended (graph-replayed iteration loops)
    python parallel-computing_amd/rocprof_summary.py counters RUN.db OUT.txt --cmd "..."
        per kernel: average of every collected counter per dispatch, and per
        wave for SQ_INSTS_* when SQ_WAVES was collected
    python parallel-computing_amd/rocprof_summary.py traffic FETCH.db WRITE.db OUT.json \\
        --kernel NAME --copy-bytes B --cmd "..." [--p P --nelem N]
        HBM bytes per launch of the kernel whose name contains NAME, from
        FETCH_SIZE and WRITE_SIZE (KiB per dispatch), each scaled by what the
        same counter reads for the largest device-copy dispatch of B bytes
        read and B bytes written in the same run.
"""
import argparse
import collections
import json
import sqlite3


def kernel_rows(db):
    cur = sqlite3.connect(db).cursor()
    cols = [r[1] for r in cur.execute("pragma table_info(kernels)")]
    name, start, end = cols.index("name"), cols.index("start"), cols.index("end")
    return [(r[name], (r[end] - r[start]) / 1e3) for r in cur.execute("select * from kernels")]


def counter_rows(db):
    cur = sqlite3.connect(db).cursor()
    cols = [r[1] for r in cur.execute("pragma table_info(counters_collection)")]
    ki, ci = cols.index("kernel_name"), cols.index("counter_name")
    vi, di = cols.index("value"), cols.index("dispatch_id")
    out = collections.defaultdict(lambda: collections.defaultdict(float))
    names = {}
    for r in cur.execute("select * from counters_collection"):
        out[r[di]][r[ci]] += r[vi]
        names[r[di]] = r[ki]
    return names, out


def short(name, n=120):
    return name if len(name) <= n else name[:n - 3] + "..."


def stats(a):
    agg = collections.defaultdict(list)
    for n, us in kernel_rows(a.db):
        agg[n].append(us)
    lines = [f"# rocprofv3 --kernel-trace --stats summary", f"# command: {a.cmd}", "",
             f"{'calls':>6} {'avg_us':>10} {'min_us':>10} {'max_us':>10} {'total_us':>12}  kernel"]
    for n, v in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
        lines.append(f"{len(v):6d} {sum(v) / len(v):10.2f} {min(v):10.2f} {max(v):10.2f} "
                     f"{sum(v):12.1f}  {short(n)}")
    open(a.out, "w").write("\n".join(lines) + "\n")
    print("\n".join(lines[:14]))


def kernel_spans(db):
    cur = sqlite3.connect(db).cursor()
    cols = [r[1] for r in cur.execute("pragma table_info(kernels)")]
    name, start, end = cols.index("name"), cols.index("start"), cols.index("end")
    return sorted((r[start], r[end], r[name]) for r in cur.execute("select * from kernels"))


def gaps(a):
    rows = kernel_spans(a.db)
    dur = collections.defaultdict(list)
    gap = collections.defaultdict(list)
    for (s0, e0, _), (s1, e1, n1) in zip(rows, rows[1:]):
        g = (s1 - e0) / 1e3
        if g < 50.0:  # within a replayed loop, not a host round trip
            dur[n1].append((e1 - s1) / 1e3)
            gap[n1].append(g)
    lines = ["# rocprofv3 --kernel-trace: per kernel, average duration and idle time "
             "before it (gaps under 50 us)", f"# command: {a.cmd}", "",
             f"{'calls':>6} {'avg_us':>10} {'gap_us':>8}  kernel"]
    for n in sorted(dur, key=lambda k: -sum(dur[k])):
        v, g = dur[n], gap[n]
        lines.append(f"{len(v):6d} {sum(v) / len(v):10.2f} {sum(g) / len(g):8.2f}  "
                     f"{short(n.replace('(anonymous namespace)::', '').split('(')[0].replace('void ', ''), 90)}")
    open(a.out, "w").write("\n".join(lines) + "\n")
    print("\n".join(lines[:20]))


def counters(a):
    names, vals = counter_rows(a.db)
    per = collections.defaultdict(lambda: collections.defaultdict(list))
    for d, n in names.items():
        for c, v in vals[d].items():
            per[n][c].append(v)
    lines = ["# rocprofv3 --pmc: per kernel, counter averages per dispatch (per wave for "
             "SQ_INSTS_*)", f"# command: {a.cmd}", ""]
    for n in sorted(per):
        cs = per[n]
        waves = sum(cs["SQ_WAVES"]) / len(cs["SQ_WAVES"]) if "SQ_WAVES" in cs else 0.0
        lines.append(short(n.replace("(anonymous namespace)::", "").split("(")[0]
                           .replace("void ", ""), 110))
        for c in sorted(cs):
            avg = sum(cs[c]) / len(cs[c])
            extra = f"  ({avg / waves:.1f} per wave)" if waves and c.startswith("SQ_INSTS") else ""
            lines.append(f"    {c:24s} {avg:14.6g}  [{len(cs[c])} dispatches]{extra}")
    open(a.out, "w").write("\n".join(lines) + "\n")
    print("\n".join(lines))


def base(name):
    """A kernel's name without its argument list (and without the
    anonymous-namespace qualifier, whose parenthesis would cut it short)."""
    return name.replace("(anonymous namespace)::", "").split("(")[0]


def pick(names, vals, counter, pred):
    return [(d, vals[d][counter]) for d, n in names.items() if pred(n) and counter in vals[d]]


def traffic(a):
    res = {"kernel_prefix": a.kernel, "command": a.cmd, "copy_bytes": a.copy_bytes}
    per = {}
    for counter, db in (("FETCH_SIZE", a.fetch), ("WRITE_SIZE", a.write)):
        names, vals = counter_rows(db)
        copies = pick(names, vals, counter, lambda n: "copy" in n.lower())
        kern = pick(names, vals, counter, lambda n: a.kernel in base(n))
        if not copies or not kern:
            raise SystemExit(f"{counter}: no copy ({len(copies)}) or kernel ({len(kern)}) dispatch")
        copy_kib = max(v for _, v in copies)
        scale = a.copy_bytes / (copy_kib * 1024.0)
        kib = sum(v for _, v in kern) / len(kern)
        per[counter] = {"kernel_kib_per_dispatch": kib, "copy_kib": copy_kib, "scale": scale,
                        "bytes_per_launch": kib * 1024.0 * scale, "dispatches": len(kern)}
        res["kernel"] = short(base(names[kern[0][0]]).replace("void igak::", "").replace(
            "void ", ""), 200)
    res.update(per)
    res["bytes_per_launch"] = per["FETCH_SIZE"]["bytes_per_launch"] + \
        per["WRITE_SIZE"]["bytes_per_launch"]
    if a.p:
        res["p"] = a.p
    if a.nelem:
        res["nelem"] = a.nelem
    json.dump(res, open(a.out, "w"), indent=1)
    print(json.dumps(res, indent=1))


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="what", required=True)
    s = sub.add_parser("stats")
    s.add_argument("db")
    s.add_argument("out")
    s.add_argument("--cmd", default="")
    c = sub.add_parser("counters")
    c.add_argument("db")
    c.add_argument("out")
    c.add_argument("--cmd", default="")
    g = sub.add_parser("gaps")
    g.add_argument("db")
    g.add_argument("out")
    g.add_argument("--cmd", default="")
    t = sub.add_parser("traffic")
    t.add_argument("fetch")
    t.add_argument("write")
    t.add_argument("out")
    t.add_argument("--kernel", required=True)
    t.add_argument("--copy-bytes", type=float, required=True)
    t.add_argument("--cmd", default="")
    t.add_argument("--p", type=int, default=0)
    t.add_argument("--nelem", type=int, default=0)
    a = ap.parse_args()
    {"stats": stats, "gaps": gaps, "counters": counters, "traffic": traffic}[a.what](a)


if __name__ == "__main__":
    main()
