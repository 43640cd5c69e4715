"""Summarise rocprofv3 databases (rocpd SQLite): kernel-trace statistics and HBM traffic per dispatch.

    python -m mppi_generic_vendor_amd.rocprof_summary TRACE_DB [FETCH_DB WRITE_DB] [--calibration-mib 256]
                                                     [--json OUT.json]
    python -m mppi_generic_vendor_amd.rocprof_summary --counters NAME[,NAME...] PMC_DB [--json OUT.json]

TRACE_DB comes from `rocprofv3 --kernel-trace --stats`, FETCH_DB from `rocprofv3 --pmc FETCH_SIZE --kernel-trace`,
WRITE_DB from `rocprofv3 --pmc WRITE_SIZE --kernel-trace`, all of a run that also performs one device-to-device copy
of `calibration-mib` MiB (bench.py does). The copy reads and writes exactly that many bytes, so the counter value it
produces divided by its size is the factor that converts the counter into HBM bytes; every kernel's median counter
value is divided by that factor.

With --counters, PMC_DB comes from `rocprofv3 --pmc NAME ... --kernel-trace` and the table lists, per kernel, the median
over its dispatches of each named counter (raw or derived, e.g. VALUBusy, VALUUtilization, MfmaUtil).
"""
from __future__ import annotations

import argparse
import json
import sqlite3
import subprocess
from collections import defaultdict

COPY_KERNEL = "__amd_rocclr_copyBuffer"


def short_name(n: str) -> str:
    """Kernel name without return type, namespace and parameter list; mangled names (some traces keep them) are
    demangled with c++filt first when it is available."""
    if n.startswith("_Z"):
        try:
            n = subprocess.run(["c++filt", "-p", n], capture_output=True, text=True, check=True).stdout.strip() or n
        except (OSError, subprocess.CalledProcessError):
            pass
    return n.split("(")[0].replace("void ", "").replace("mppi_amd::", "")


def kernel_stats(path: str):
    db = sqlite3.connect(path)
    return [(short_name(n), int(c), float(a), float(p))
            for n, c, a, p in db.execute("select name, total_calls, average, percentage from top_kernels")]


def counter_values(path: str, counter: str):
    db = sqlite3.connect(path)
    acc = defaultdict(list)
    for name, v in db.execute("select name, counter_value from pmc_events where counter_name=?", (counter,)):
        acc[short_name(name)].append(float(v))
    return acc


def summarise(trace_db: str, fetch_db: str, write_db: str, calibration_mib: int):
    stats = kernel_stats(trace_db)
    f = counter_values(fetch_db, "FETCH_SIZE")
    w = counter_values(write_db, "WRITE_SIZE")
    cal_kib = calibration_mib * 1024.0
    cf = max(f.get(COPY_KERNEL, [0.0])) / cal_kib
    cw = max(w.get(COPY_KERNEL, [0.0])) / cal_kib
    if cf <= 0 or cw <= 0:
        raise SystemExit("the calibration copy is missing from the counter runs")
    traffic = {}
    for k in sorted(set(f) | set(w)):
        if k == COPY_KERNEL:
            continue
        fv = sorted(f.get(k, [0.0]))
        wv = sorted(w.get(k, [0.0]))
        fetch = fv[len(fv) // 2] / cf * 1024.0
        write = wv[len(wv) // 2] / cw * 1024.0
        traffic[k] = {"fetch_bytes": round(fetch), "write_bytes": round(write),
                      "total_bytes": round(fetch + write), "dispatches": len(fv)}
    return {"kernels": stats, "fetch_factor": cf, "write_factor": cw, "calibration_mib": calibration_mib,
            "traffic": traffic}


def counter_table(path: str, names):
    """{kernel: {counter: (median over dispatches, dispatches)}} for the named counters."""
    out = defaultdict(dict)
    for c in names:
        for k, vals in counter_values(path, c).items():
            v = sorted(vals)
            out[k][c] = (v[len(v) // 2], len(v))
    return dict(out)


def render(s) -> str:
    out = ["kernel-trace statistics (average over all dispatches of the run):"]
    for name, calls, avg, pct in s["kernels"]:
        out.append(f"  {avg:10.3f} us  calls {calls:6d}  {pct:5.1f}%  {name}")
    out.append(f"calibration copy of {s['calibration_mib']} MiB: FETCH_SIZE factor {s['fetch_factor']:.4f}, "
               f"WRITE_SIZE factor {s['write_factor']:.4f} (counter KiB per copied KiB)")
    out.append("HBM traffic per dispatch (median over dispatches, scaled by the factors):")
    for k, t in s["traffic"].items():
        out.append(f"  fetch {t['fetch_bytes'] / 1024:10.1f} KiB  write {t['write_bytes'] / 1024:10.1f} KiB  "
                   f"total {t['total_bytes'] / 1024:10.1f} KiB  ({t['dispatches']} dispatches)  {k}")
    return "\n".join(out)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--counters", default=None, help="comma-separated counter names: print their per-kernel medians")
    ap.add_argument("trace_db")
    ap.add_argument("fetch_db", nargs="?")
    ap.add_argument("write_db", nargs="?")
    ap.add_argument("--calibration-mib", type=int, default=256)
    ap.add_argument("--json", default=None, help="also write the traffic table as JSON")
    a = ap.parse_args()
    if a.counters:
        names = [c for c in a.counters.split(",") if c]
        table = counter_table(a.trace_db, names)
        print("per-kernel median over dispatches: " + ", ".join(names))
        for k in sorted(table):
            cells = "  ".join(f"{c} {table[k][c][0]:10.3f}" if c in table[k] else f"{c} {'-':>10}" for c in names)
            n = max(v[1] for v in table[k].values())
            print(f"  {cells}  ({n} dispatches)  {k}")
        if a.json:
            with open(a.json, "w") as fh:
                json.dump({k: {c: v[0] for c, v in t.items()} for k, t in table.items()}, fh, indent=1)
        return
    if a.fetch_db is None or a.write_db is None:
        print("kernel-trace statistics (average over all dispatches of the run):")
        for name, calls, avg, pct in kernel_stats(a.trace_db):
            print(f"  {avg:10.3f} us  calls {calls:6d}  {pct:5.1f}%  {name}")
        return
    s = summarise(a.trace_db, a.fetch_db, a.write_db, a.calibration_mib)
    print(render(s))
    if a.json:
        with open(a.json, "w") as fh:
            json.dump({"fetch_factor": s["fetch_factor"], "write_factor": s["write_factor"],
                       "calibration_mib": s["calibration_mib"], "traffic": s["traffic"]}, fh, indent=1)


if __name__ == "__main__":
    main()
