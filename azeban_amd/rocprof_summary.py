"""Summaries of rocprofv3 result databases (kernel trace and hardware counters).

    python -m azeban_amd.rocprof_summary trace  <kt_results.db>  --title T --cmd C
    python -m azeban_amd.rocprof_summary counters --fetch F.db --write W.db [--sq S.db ...]
                                          --title T --cmd C --json out.json
    python -m azeban_amd.rocprof_summary dispatches <pmc_results.db> --title T --cmd C
                                          [--match azb:: --limit 12]

Counter runs hold one pass per database (FETCH_SIZE and WRITE_SIZE need their own
runs).  The HBM bytes of a dispatch are FETCH_SIZE x fetch_scale + WRITE_SIZE x
write_scale (both in KiB); the scales come from a device-to-device copy of known
size in the same program (the first `__amd_rocclr_copyBuffer` dispatch), so that
the counters' own accounting (FETCH_SIZE reads half the copied bytes on this
part) is measured rather than assumed.
"""
from __future__ import annotations

import argparse
import collections
import json
import sqlite3
import sys


def _kernel_short(name: str) -> str:
    n = name.split("(")[0].replace("void ", "").strip()
    return n


def trace_table(db: str) -> list[tuple]:
    con = sqlite3.connect(db)
    rows = con.execute(
        "select name, total_calls, total_duration, average, percentage from top_kernels "
        "order by total_duration desc").fetchall()
    return rows


def counter_events(db: str) -> list[dict]:
    con = sqlite3.connect(db)
    rows = con.execute(
        "select dispatch_id, name, duration, counter_name, sum(counter_value) from pmc_events "
        "group by dispatch_id, counter_name order by dispatch_id").fetchall()
    out = collections.OrderedDict()
    for did, name, dur, cn, v in rows:
        out.setdefault(did, {"name": name, "dur_ns": dur})[cn] = v
    return list(out.values())


def calibration(events: list[dict], counter: str, copy_bytes: float) -> float:
    for e in events:
        if "copyBuffer" in e["name"] and counter in e:
            return copy_bytes / (e[counter] * 1024.0)
    raise ValueError(f"no calibration copy dispatch with {counter} found")


def per_kernel(events: list[dict], counters: list[str]) -> dict:
    agg = collections.OrderedDict()
    for e in events:
        k = _kernel_short(e["name"])
        a = agg.setdefault(k, {"dispatches": 0, "dur_ns": 0.0, **{c: 0.0 for c in counters}})
        a["dispatches"] += 1
        a["dur_ns"] += e["dur_ns"]
        for c in counters:
            a[c] += e.get(c, 0.0)
    for a in agg.values():
        n = a["dispatches"]
        a["dur_ns"] /= n
        for c in counters:
            a[c] /= n
    return agg


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="mode", required=True)
    t = sub.add_parser("trace")
    t.add_argument("db")
    t.add_argument("--title", required=True)
    t.add_argument("--cmd", required=True)
    c = sub.add_parser("counters")
    c.add_argument("--fetch", required=True)
    c.add_argument("--write", required=True)
    c.add_argument("--sq", nargs="*", default=[])
    c.add_argument("--copy-bytes", type=float, default=2.0 * 2**30)
    c.add_argument("--title", required=True)
    c.add_argument("--cmd", required=True)
    c.add_argument("--json")
    c.add_argument("--match", default="azb::", help="only kernels whose name contains this")
    d = sub.add_parser("dispatches")
    d.add_argument("db")
    d.add_argument("--title", required=True)
    d.add_argument("--cmd", required=True)
    d.add_argument("--match", default="azb::", help="only kernels whose name contains this")
    d.add_argument("--limit", type=int, default=12, help="first dispatches shown")
    a = ap.parse_args(argv)

    if a.mode == "dispatches":
        # one row per dispatch, in launch order: tells apart kernels that share one
        # template instantiation but run different passes
        ev = [e for e in counter_events(a.db) if a.match in e["name"]][: a.limit]
        names = sorted({k for e in ev for k in e if k not in ("name", "dur_ns")})
        print(f"# {a.title}\n")
        print("Command (from the repository root, one MI355X):\n")
        print(f"```\n{a.cmd}\n```\n")
        print("Per dispatch, in launch order (counters summed over the device):\n")
        print("| # | kernel | us | " + " | ".join(names) + " |")
        print("|---|---|---|" + "---|" * len(names))
        for i, e in enumerate(ev):
            vals = " | ".join(f"{e.get(n, 0):.0f}" for n in names)
            print(f"| {i} | `{_kernel_short(e['name'])}` | {e['dur_ns'] / 1e3:.1f} | {vals} |")
        return 0

    if a.mode == "trace":
        print(f"# {a.title}\n")
        print("Command (from the repository root, one MI355X):\n")
        print(f"```\n{a.cmd}\n```\n")
        print("`rocprofv3 --kernel-trace --stats` summary (durations in milliseconds):\n")
        print("| kernel | calls | total ms | avg ms | % |")
        print("|---|---|---|---|---|")
        for n, calls, tot, avg, pct in trace_table(a.db):
            print(f"| `{_kernel_short(n)}` | {calls} | {tot / 1e3:.2f} | {avg / 1e3:.3f} | {pct:.2f} |")
        return 0

    fe, we = counter_events(a.fetch), counter_events(a.write)
    fs = calibration(fe, "FETCH_SIZE", a.copy_bytes)
    ws = calibration(we, "WRITE_SIZE", a.copy_bytes)
    kf = per_kernel([e for e in fe if a.match in e["name"]], ["FETCH_SIZE"])
    kw = per_kernel([e for e in we if a.match in e["name"]], ["WRITE_SIZE"])
    sq_counters = []
    ks = collections.OrderedDict()
    for db in a.sq:
        ev = [e for e in counter_events(db) if a.match in e["name"]]
        names = sorted({k for e in ev for k in e if k not in ("name", "dur_ns")})
        sq_counters += [n for n in names if n not in sq_counters]
        for k, v in per_kernel(ev, names).items():
            ks.setdefault(k, {}).update({n: v[n] for n in names})
    rows = {}
    for k in kf:
        if k not in kw:
            continue
        fb = kf[k]["FETCH_SIZE"] * 1024.0 * fs
        wb = kw[k]["WRITE_SIZE"] * 1024.0 * ws
        rows[k] = {"dispatches": kf[k]["dispatches"], "avg_us": kf[k]["dur_ns"] / 1e3,
                   "fetch_bytes": fb, "write_bytes": wb, "hbm_bytes": fb + wb,
                   **{n: ks.get(k, {}).get(n) for n in sq_counters}}
    print(f"# {a.title}\n")
    print("Commands (from the repository root, one MI355X; one counter pass per run):\n")
    print(f"```\n{a.cmd}\n```\n")
    print(f"Calibration: a {a.copy_bytes / 2**30:.0f} GiB device-to-device copy in the same program reads "
          f"FETCH_SIZE x {fs:.3f} and WRITE_SIZE x {ws:.3f} bytes per KiB counted; HBM bytes below "
          "are the counters times these factors (per dispatch, averaged over the dispatches of a "
          "kernel).\n")
    print("| kernel | dispatches | avg us | HBM read GB | HBM write GB | HBM total GB | GB/s |")
    print("|---|---|---|---|---|---|---|")
    for k, r in rows.items():
        gbs = r["hbm_bytes"] / (r["avg_us"] * 1e-6) / 1e9
        print(f"| `{k}` | {r['dispatches']} | {r['avg_us']:.1f} | {r['fetch_bytes'] / 1e9:.3f} | "
              f"{r['write_bytes'] / 1e9:.3f} | {r['hbm_bytes'] / 1e9:.3f} | {gbs:.0f} |")
    if sq_counters:
        print("\nSQ counters (per dispatch, summed over the device):\n")
        print("| kernel | " + " | ".join(sq_counters) + " |")
        print("|---|" + "---|" * len(sq_counters))
        for k, r in rows.items():
            vals = [r.get(n) for n in sq_counters]
            print(f"| `{k}` | " + " | ".join("" if v is None else f"{v:.4g}" for v in vals) + " |")
    if a.json:
        with open(a.json, "w") as f:
            json.dump({"title": a.title, "command": a.cmd, "fetch_scale": fs, "write_scale": ws,
                       "kernels": rows}, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
