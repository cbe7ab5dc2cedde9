"""Helpers for the rocprofv3 measurements behind bench.py's roofline numbers.

  python -m snncuda_amd.profiling copy [--mib N] [--reps R]
      A plain device copy of N MiB (torch), repeated R times: the calibration workload for
      FETCH_SIZE / WRITE_SIZE (a copy moves exactly N MiB in and N MiB out).
  python -m snncuda_amd.profiling summary --trace T.db [--fetch F.db --write W.db] [--skip K]
                                          [--kernel NAME --events-per-launch X] --cmd "..." --out FILE
      Per-kernel duration statistics from a --kernel-trace --stats database and, when given, the
      per-dispatch FETCH_SIZE / WRITE_SIZE (KiB) of every kernel, skipping the first K dispatches of
      each kernel (ramp-up ticks). Writes a short text summary.
  python -m snncuda_amd.profiling calibrate --fetch F.db --write W.db --mib N --out FILE
      Ratio of the copy kernel's counter bytes to the bytes it moves.
  python -m snncuda_amd.profiling counters DB... [--skip K] [--kernels PREFIX...] --cmd "..." --out FILE
      Mean per dispatch of every counter in --pmc databases, per kernel.
"""
from __future__ import annotations

import argparse
import json
import re
import sqlite3
import statistics
import sys


def _short(name: str) -> str:
    name = name.replace("(anonymous namespace)::", "")
    name = re.sub(r"\(.*", "", name)
    return re.sub(r"^void ", "", name).strip()


def kernel_durations(db: str):
    c = sqlite3.connect(db).cursor()
    out = {}
    for name, dur in c.execute("select name, duration from kernels order by dispatch_id"):
        out.setdefault(_short(name), []).append(dur)
    return out


def counter_values(db: str, counter: str):
    c = sqlite3.connect(db).cursor()
    out = {}
    q = ("select k.name, p.counter_value from pmc_events p join kernels k on p.dispatch_id = k.dispatch_id "
         "where p.counter_name = ? order by k.dispatch_id")
    for name, v in c.execute(q, (counter,)):
        out.setdefault(_short(name), []).append(v)
    return out


def cmd_copy(a):
    import torch
    n = a.mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(a.reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(f"copied {a.mib} MiB x {a.reps}")


def cmd_calibrate(a):
    f = counter_values(a.fetch, "FETCH_SIZE")
    w = counter_values(a.write, "WRITE_SIZE")
    name = max(f, key=lambda k: statistics.mean(f[k]))
    fk = statistics.median(f[name][1:] or f[name])
    wk = statistics.median(w[name][1:] or w[name])
    moved_kib = a.mib * 1024
    fs, ws = moved_kib / fk, moved_kib / wk
    text = (f"copy calibration: kernel {name}, {a.mib} MiB read + {a.mib} MiB written per dispatch\n"
            f"FETCH_SIZE median {fk:.1f} KiB -> fetch scale (bytes read / counted) = {fs:.4f}\n"
            f"WRITE_SIZE median {wk:.1f} KiB -> write scale (bytes written / counted) = {ws:.4f}\n")
    open(a.out, "w").write(text)
    if a.json:
        open(a.json, "w").write(json.dumps({"fetch_scale": fs, "write_scale": ws, "mib": a.mib}, indent=1) + "\n")
    print(text)


def cmd_summary(a):
    dur = kernel_durations(a.trace)
    lines = [f"command: {a.cmd}", f"source: rocprofv3 --kernel-trace --stats; first {a.skip} dispatches of each "
             "kernel skipped (network ramp-up)", "",
             f"{'kernel':58s} {'calls':>6s} {'mean us':>10s} {'median us':>10s} {'total ms':>10s}"]
    rows = []
    for k, v in dur.items():
        v = v[a.skip:] if len(v) > a.skip else v
        rows.append((sum(v), k, v))
    for tot, k, v in sorted(rows, reverse=True)[:16]:
        lines.append(f"{k[:58]:58s} {len(v):6d} {statistics.mean(v)/1e3:10.2f} {statistics.median(v)/1e3:10.2f} {tot/1e6:10.3f}")
    result = {}
    if a.fetch and a.write:
        f = counter_values(a.fetch, "FETCH_SIZE")
        w = counter_values(a.write, "WRITE_SIZE")
        lines += ["", f"source: rocprofv3 --pmc FETCH_SIZE / --pmc WRITE_SIZE (KiB per dispatch), "
                      f"first {a.skip} dispatches skipped",
                  f"{'kernel':58s} {'FETCH KiB':>12s} {'WRITE KiB':>12s}"]
        for k in sorted(f, key=lambda k: -statistics.mean(f[k][a.skip:] or f[k]))[:10]:
            fv = statistics.mean(f[k][a.skip:] or f[k])
            wv = statistics.mean(w.get(k, [0])[a.skip:] or w.get(k, [0]))
            lines.append(f"{k[:58]:58s} {fv:12.1f} {wv:12.1f}")
            if a.kernel and k.startswith(a.kernel):
                result = {"kernel": k, "fetch_kib": fv, "write_kib": wv, "bytes_per_launch": (fv + wv) * 1024}
    if result and a.calibration:
        cal = json.load(open(a.calibration))
        result["fetch_scale"], result["write_scale"] = cal["fetch_scale"], cal["write_scale"]
        result["bytes_per_launch"] = (result["fetch_kib"] * cal["fetch_scale"] + result["write_kib"] * cal["write_scale"]) * 1024
        lines += ["", f"scaled by the copy calibration: fetch x{cal['fetch_scale']:.3f}, write x{cal['write_scale']:.3f}"]
    if a.kernel and a.events_per_launch and result:
        result["events_per_launch"] = a.events_per_launch
        result["bytes_per_event"] = result["bytes_per_launch"] / a.events_per_launch
        lines += ["", f"{result['kernel']}: {result['bytes_per_launch']/1e6:.1f} MB per launch, "
                      f"{result['bytes_per_event']:.1f} B per delivered event"]
    text = "\n".join(lines) + "\n"
    open(a.out, "w").write(text)
    if a.json and result:
        result["stdp"] = a.stdp
        open(a.json, "w").write(json.dumps(result, indent=1) + "\n")
    print(text)


def cmd_counters(a):
    """Mean per dispatch of every counter in the given --pmc databases, per kernel (first --skip
    dispatches of each kernel left out), for the kernels named by --kernels prefixes."""
    per = {}
    for db in a.dbs:
        c = sqlite3.connect(db).cursor()
        names = [r[0] for r in c.execute("select distinct counter_name from pmc_events")]
        for cn in names:
            for k, v in counter_values(db, cn).items():
                if a.kernels and not any(k.startswith(p) for p in a.kernels):
                    continue
                v = v[a.skip:] if len(v) > a.skip else v
                per.setdefault(k, {})[cn] = statistics.mean(v)
    lines = [f"command: {a.cmd}", f"source: rocprofv3 --pmc (values per dispatch as rocprofv3 reports them), "
             f"mean over dispatches after the first {a.skip} of each kernel", ""]
    for k in sorted(per):
        lines.append(k)
        for cn in sorted(per[k]):
            lines.append(f"   {cn:28s} {per[k][cn]:16.1f}")
    text = "\n".join(lines) + "\n"
    open(a.out, "w").write(text)
    print(text)


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd_name", required=True)
    c = sub.add_parser("copy")
    c.add_argument("--mib", type=int, default=1024)
    c.add_argument("--reps", type=int, default=5)
    k = sub.add_parser("calibrate")
    k.add_argument("--fetch", required=True)
    k.add_argument("--write", required=True)
    k.add_argument("--mib", type=int, default=1024)
    k.add_argument("--out", required=True)
    k.add_argument("--json")
    s = sub.add_parser("summary")
    s.add_argument("--trace", required=True)
    s.add_argument("--fetch")
    s.add_argument("--write")
    s.add_argument("--skip", type=int, default=0)
    s.add_argument("--kernel")
    s.add_argument("--events-per-launch", type=float)
    s.add_argument("--stdp", action="store_true")
    s.add_argument("--cmd", required=True)
    s.add_argument("--out", required=True)
    s.add_argument("--json")
    s.add_argument("--calibration")
    n = sub.add_parser("counters")
    n.add_argument("dbs", nargs="+")
    n.add_argument("--skip", type=int, default=0)
    n.add_argument("--kernels", nargs="*")
    n.add_argument("--cmd", required=True)
    n.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    {"copy": cmd_copy, "calibrate": cmd_calibrate, "summary": cmd_summary, "counters": cmd_counters}[a.cmd_name](a)


if __name__ == "__main__":
    main()
