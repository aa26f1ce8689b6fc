"""Summaries of rocprofv3 result databases (rocpd SQLite, `*_results.db`).

    python -m vfb_amd.profsum trace <db> [kernel-substring]
        per-kernel calls / total / average / min / max duration
    python -m vfb_amd.profsum pmc <db> [kernel-substring]
        per-kernel average of each collected counter per dispatch, and the
        sum over all dispatches

A measurement aid for the committed profile summaries; it reads only the
database rocprofv3 wrote.
"""
import collections
import sqlite3
import sys


def _q(db, sql, args=()):
    con = sqlite3.connect(db)
    try:
        return con.execute(sql, args).fetchall()
    finally:
        con.close()


def kernel_stats(db, like="%"):
    rows = _q(db, "select name, duration from kernels where name like ?", (like,))
    acc = collections.OrderedDict()
    for name, dur in rows:
        short = name.split("(")[0]
        acc.setdefault(short, []).append(dur)
    out = []
    for k, ds in acc.items():
        out.append({"kernel": k, "calls": len(ds), "total_ns": sum(ds), "avg_ns": sum(ds) / len(ds),
                    "min_ns": min(ds), "max_ns": max(ds)})
    return sorted(out, key=lambda r: -r["total_ns"])


def pmc_stats(db, like="%"):
    rows = _q(db, "select dispatch_id, name, counter_name, sum(counter_value), max(duration) from pmc_events "
                  "where name like ? group by dispatch_id, counter_name", (like,))
    per = collections.OrderedDict()
    for d, name, cname, val, dur in rows:
        short = name.split("(")[0]
        e = per.setdefault(short, {"dispatches": set(), "sum": collections.Counter(), "ns": {}})
        e["dispatches"].add(d)
        e["sum"][cname] += val
        e["ns"][d] = dur
    out = []
    for k, e in per.items():
        n = len(e["dispatches"])
        out.append({"kernel": k, "dispatches": n, "total_ns": sum(e["ns"].values()),
                    "avg": {c: v / n for c, v in e["sum"].items()}, "sum": dict(e["sum"])})
    return sorted(out, key=lambda r: -r["total_ns"])


def main(argv):
    if len(argv) < 3 or argv[1] not in ("trace", "pmc"):
        print(__doc__)
        return 2
    like = f"%{argv[3]}%" if len(argv) > 3 else "%"
    if argv[1] == "trace":
        print("| kernel | calls | total ms | avg us | min us | max us |\n|---|---|---|---|---|---|")
        for r in kernel_stats(argv[2], like):
            print(f"| `{r['kernel']}` | {r['calls']} | {r['total_ns'] / 1e6:.2f} | {r['avg_ns'] / 1e3:.1f} | "
                  f"{r['min_ns'] / 1e3:.1f} | {r['max_ns'] / 1e3:.1f} |")
    else:
        for r in pmc_stats(argv[2], like):
            print(f"{r['kernel']}: {r['dispatches']} dispatches, {r['total_ns'] / 1e6:.2f} ms")
            for c in sorted(r["avg"]):
                print(f"    {c}: avg {r['avg'][c]:.6g} per dispatch, sum {r['sum'][c]:.6g}")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
