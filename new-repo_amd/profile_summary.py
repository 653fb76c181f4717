"""Summaries of rocprofv3 CSV output for the benchmark's evidence.

kernels DIR            per-kernel table from <DIR>/**/*kernel_stats.csv
traffic FETCH WRITE    HBM bytes per launch of a kernel family from two
                       counter passes (FETCH_SIZE, WRITE_SIZE, KiB per
                       dispatch), each scaled by what the same counter reads
                       for the copy-calibration kernel of known size
                       (bench.py --copy-calibration MIB, an elementwise
                       y = x * 1 over float32)

Usage:
  python new-repo_amd/profile_summary.py kernels gpurun_out/trace
  python new-repo_amd/profile_summary.py traffic gpurun_out/pmc_fetch gpurun_out/pmc_write \\
      --family cols --pattern k_cols --variant learned --grid 256,256,256 --copy-mib 256 \\
      --algorithmic-bytes B --command "..." --out profiles/pmc_traffic.json
"""
from __future__ import annotations

import argparse
import csv
import json
import pathlib
import sys


def _rows(d: str, suffix: str):
    files = sorted(pathlib.Path(d).rglob(f"*{suffix}"))
    if not files:
        raise SystemExit(f"no *{suffix} under {d}")
    out = []
    for f in files:
        with open(f, newline="") as fh:
            out += list(csv.DictReader(fh))
    return out


def _col(row, *names):
    low = {k.lower(): v for k, v in row.items()}
    for n in names:
        if n.lower() in low:
            return low[n.lower()]
    raise KeyError(f"none of {names} in {list(row)}")


def kernels(d: str, top: int = 25):
    rows = _rows(d, "kernel_stats.csv")
    tab = []
    for r in rows:
        tab.append((_col(r, "Name", "KernelName", "Kernel_Name"), int(_col(r, "Calls")),
                    float(_col(r, "TotalDurationNs")), float(_col(r, "AverageNs")), float(_col(r, "Percentage"))))
    tab.sort(key=lambda t: -t[2])
    lines = ["| kernel | calls | total ms | avg us | % |", "|---|---|---|---|---|"]
    for n, c, tot, avg, pct in tab[:top]:
        lines.append(f"| `{n[:110]}` | {c} | {tot / 1e6:.3f} | {avg / 1e3:.1f} | {pct:.1f} |")
    return "\n".join(lines), tab


def _per_dispatch(d: str, counter: str):
    rows = _rows(d, "counter_collection.csv")
    vals = {}
    for r in rows:
        if _col(r, "Counter_Name").strip() != counter:
            continue
        key = (_col(r, "Dispatch_Id", "Correlation_Id"), _col(r, "Kernel_Name"))
        vals[key] = vals.get(key, 0.0) + float(_col(r, "Counter_Value"))
    return [(k[1], v) for k, v in vals.items()]


def traffic(args):
    def avg(d, counter, pat):
        v = [x for n, x in _per_dispatch(d, counter) if pat in n]
        if not v:
            raise SystemExit(f"no {counter} dispatches matching {pat!r} in {d}")
        return sum(v) / len(v), len(v)
    copy_pat = args.copy_pattern
    exp_kib = args.copy_mib * 1024.0
    f_cal, nf_cal = avg(args.fetch, "FETCH_SIZE", copy_pat)
    w_cal, nw_cal = avg(args.write, "WRITE_SIZE", copy_pat)
    f_k, nf = avg(args.fetch, "FETCH_SIZE", args.pattern)
    w_k, nw = avg(args.write, "WRITE_SIZE", args.pattern)
    sf, sw = exp_kib / f_cal, exp_kib / w_cal
    tbytes = (f_k * sf + w_k * sw) * 1024.0
    out = {"variant": args.variant, "kernel_family": args.family, "kernel_pattern": args.pattern,
           "grid": [int(x) for x in args.grid.split(",")],
           "fetch_kib_per_dispatch": f_k, "write_kib_per_dispatch": w_k, "dispatches": [nf, nw],
           "copy_calibration": {"pattern": copy_pat, "mib": args.copy_mib, "fetch_kib": f_cal, "write_kib": w_cal,
                                "fetch_scale": sf, "write_scale": sw, "dispatches": [nf_cal, nw_cal]},
           "traffic_bytes_per_launch": tbytes,
           "algorithmic_bytes_per_launch": args.algorithmic_bytes,
           "traffic_over_algorithmic": tbytes / args.algorithmic_bytes if args.algorithmic_bytes else None,
           "commands": args.command}
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    k = sub.add_parser("kernels")
    k.add_argument("dir")
    k.add_argument("--top", type=int, default=25)
    t = sub.add_parser("traffic")
    t.add_argument("fetch")
    t.add_argument("write")
    t.add_argument("--family", required=True)
    t.add_argument("--pattern", required=True)
    t.add_argument("--copy-pattern", default="elementwise")
    t.add_argument("--copy-mib", type=int, required=True)
    t.add_argument("--variant", required=True)
    t.add_argument("--grid", required=True)
    t.add_argument("--algorithmic-bytes", type=float, default=0.0)
    t.add_argument("--command", action="append", default=[])
    t.add_argument("--out")
    a = ap.parse_args(argv)
    if a.cmd == "kernels":
        text, _ = kernels(a.dir, a.top)
        print(text)
    else:
        res = traffic(a)
        s = json.dumps(res, indent=1)
        if a.out:
            pathlib.Path(a.out).write_text(s + "\n")
        print(s)
    return 0


if __name__ == "__main__":
    sys.exit(main())
