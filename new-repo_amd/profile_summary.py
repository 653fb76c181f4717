"""Summaries of rocprofv3 CSV output for the benchmark's evidence.

kernels DIR            per-kernel table from <DIR>/**/*kernel_stats.csv
report DIR             markdown report of a kernel trace (see --help)
sq DIR                 per-kernel instruction counts of an SQ counter pass
traffic FETCH WRITE    HBM bytes per launch of each kernel family from two
                       counter passes (FETCH_SIZE, WRITE_SIZE, KiB per
                       dispatch), each scaled by what the same counter reads
                       for the copy-calibration kernel of known size
                       (bench.py --copy-calibration MIB, an elementwise
                       y = x * 1 over float32), against the algorithmic bytes
                       of the same dispatches (bench.py --steps 0 output)

Usage:
  python new-repo_amd/profile_summary.py kernels out/trace
  python new-repo_amd/profile_summary.py traffic out/pmc_fetch out/pmc_write \\
      --bench-json out/pmc_fetch.json --variant learned --copy-mib 256 \\
      --command "..." --out profiles/pmc_traffic_learned.json
"""
from __future__ import annotations

import argparse
import csv
import json
import pathlib
import re
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
    """(kernel name, counter value) per dispatch, in dispatch order."""
    rows = _rows(d, "counter_collection.csv")
    vals = {}
    for r in rows:
        if _col(r, "Counter_Name").strip() != counter:
            continue
        key = (int(_col(r, "Dispatch_Id", "Correlation_Id")), _col(r, "Kernel_Name"))
        vals[key] = vals.get(key, 0.0) + float(_col(r, "Counter_Value"))
    return [(k[1], vals[k]) for k in sorted(vals)]


def _last(vals: list, n: int) -> list:
    """The last n dispatches: the HIP-event cycle that closes the counter
    run (its warm-up steps are profiled too, and may run other work, such
    as the first, full-column dynamic update)."""
    return vals[-n:] if n > 0 else vals


def rocprof_key(name: str) -> str:
    """rocprofv3's kernel name -> the key of the library's per-kernel
    statistics: the kernel name with its floating-point template arguments
    (at most two), without other template arguments or the parameter list:
    'void k_cols_c16<float, 1, 2, 2>(ColArgs<float>, ...)' ->
    'k_cols_c16<float>', 'k_proj(ProjArgs)' -> 'k_proj'."""
    n = name.strip()
    if n.startswith("void "):
        n = n[5:]
    depth, cut = 0, len(n)
    for i, ch in enumerate(n):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            cut = i
            break
    n = n[:cut]
    if "<" not in n:
        return n
    base, args = n.split("<", 1)
    parts, depth, cur = [], 0, ""
    for ch in args[:-1]:
        if ch == "," and depth == 0:
            parts.append(cur.strip())
            cur = ""
            continue
        depth += ch == "<"
        depth -= ch == ">"
        cur += ch
    parts.append(cur.strip())
    types = [a for a in parts if a in ("float", "double")][:2]
    return base + (f"<{', '.join(types)}>" if types else "")


# kernel-name patterns of the benchmark's kernel families
FAMILY_KERNELS = {"rows_r2c": r"\bk_rows_r2c<", "rows_c2r": r"\bk_rows_c2r(_div<|_proj\(|<)",
                  "cols": r"\bk_cols(_256|_c16)?<", "rows_resample": r"\bk_rows_resample<",
                  "cols_resample": r"\bk_(cols_resample|colres_c16)<", "tridiag": r"\bk_tridiag",
                  "lagrangian": r"\bk_lagr<"}


def traffic(args):
    """Per family: counter bytes summed over its dispatches (scaled by the
    copy calibration) against the algorithmic bytes the library accumulated
    for the same dispatches (bench.py --steps 0 output of the counter run)."""
    lines = [ln for ln in pathlib.Path(args.bench_json).read_text().splitlines() if ln.startswith("{")]
    bench = json.loads(lines[-1])
    fam = bench["variants"][args.variant]["families"]
    fd, wd = _per_dispatch(args.fetch, "FETCH_SIZE"), _per_dispatch(args.write, "WRITE_SIZE")

    def cal(d, counter):
        v = [x for n, x in d if args.copy_pattern in n]
        if not v:
            raise SystemExit(f"no {counter} dispatches of the copy kernel ({args.copy_pattern!r})")
        return args.copy_mib * 1024.0 / (sum(v) / len(v)), len(v)

    sf, ncf = cal(fd, "FETCH_SIZE")
    sw, ncw = cal(wd, "WRITE_SIZE")
    out = {"variant": args.variant, "grid": bench["grid"],
           "copy_calibration": {"pattern": args.copy_pattern, "mib": args.copy_mib, "fetch_scale": sf,
                                "write_scale": sw, "dispatches": [ncf, ncw]},
           "commands": args.command, "families": {}}
    for name, pat in FAMILY_KERNELS.items():
        if name not in fam:
            continue
        calls = fam[name]["calls"]
        f = _last([x for n, x in fd if re.search(pat, n)], calls)
        w = _last([x for n, x in wd if re.search(pat, n)], calls)
        if not f or not w:
            continue
        tb = (sum(f) * sf + sum(w) * sw) * 1024.0
        out["families"][name] = {
            "kernel_pattern": pat, "dispatches": [len(f), len(w)], "calls": calls,
            "fetch_bytes_per_launch": sum(f) * sf * 1024.0 / len(f),
            "write_bytes_per_launch": sum(w) * sw * 1024.0 / len(w),
            "traffic_bytes_per_launch": tb / len(f),
            "algorithmic_bytes_per_launch": fam[name]["bytes"] / calls,
            "traffic_over_algorithmic": tb / fam[name]["bytes"]}
    # per kernel: the library's per-kernel statistics of the same launches,
    # aggregated by key (all variants of one kernel at one precision)
    lib = {}
    for name, v in bench["variants"][args.variant].get("kernels", {}).items():
        if not name.startswith("k_"):
            continue
        key = name.split(" [")[0]
        e = lib.setdefault(key, {"calls": 0, "bytes": 0.0, "ms": 0.0})
        e["calls"] += v["calls"]
        e["bytes"] += v["bytes"]
        e["ms"] += v["ms"]
    out["kernels"] = {}
    for key, e in sorted(lib.items(), key=lambda kv: -kv[1]["ms"]):
        f = _last([x for n, x in fd if rocprof_key(n) == key], e["calls"])
        w = _last([x for n, x in wd if rocprof_key(n) == key], e["calls"])
        if not f or not w:
            continue
        fb, wb = sum(f) * sf * 1024.0, sum(w) * sw * 1024.0
        out["kernels"][key] = {
            "dispatches": [len(f), len(w)], "calls": e["calls"],
            "fetch_bytes_per_launch": fb / len(f), "write_bytes_per_launch": wb / len(w),
            "traffic_bytes_per_launch": fb / len(f) + wb / len(w),
            "algorithmic_bytes_per_launch": e["bytes"] / e["calls"],
            "traffic_over_algorithmic": (fb / len(f) + wb / len(w)) / (e["bytes"] / e["calls"])}
    # no rates: counter collection slows every dispatch, so the times of
    # these runs are not the kernels' (those are in the trace summaries)
    out["note"] = ("bytes per launch only; kernel times and GB/s come from the runs without counters "
                   "(kernel-trace summaries and the benchmark line)")
    return out


def report(args):
    """Markdown kernel-trace report: top kernels, the HIP-event per-kernel
    averages of the traced process beside rocprofv3's (kernels keyed by name
    and floating-point template arguments), and the step figures of a default
    benchmark line."""
    text, tab = kernels(args.trace, args.top)
    lines = [ln for ln in pathlib.Path(args.trace_json).read_text().splitlines() if ln.startswith("{")]
    tj = json.loads(lines[-1])
    kern = tj.get("kernels") or tj.get("variants", {}).get(args.variant, {}).get("kernels", {})
    bb = json.loads(pathlib.Path(args.bench_json).read_text().splitlines()[-1])
    lib = {}
    for name, v in kern.items():
        if not name.startswith("k_"):
            continue
        e = lib.setdefault(name.split(" [")[0], {"ms": 0.0, "calls": 0})
        e["ms"] += v["ms"]
        e["calls"] += v["calls"]
    roc = {}
    for n, c, tot, avg, pct in tab:
        e = roc.setdefault(rocprof_key(n), {"ns": 0.0, "calls": 0})
        e["ns"] += tot
        e["calls"] += c
    rows = []
    for key, e in sorted(lib.items(), key=lambda kv: -kv[1]["ms"])[:args.top]:
        r = roc.get(key)
        ra = f"{r['ns'] / r['calls'] / 1e6:.4f}" if r else "-"
        rows.append(f"| `{key}` | {e['ms'] / e['calls']:.4f} | {e['calls']} | {ra} | {r['calls'] if r else '-'} |")
    steps = [f"| {k} | {v['ms_per_step']:.3f} | {v['mpts']:.0f} |" for k, v in bb["variants"].items() if "mpts" in v]
    r = bb["roofline"]
    out = [f"# {args.title}", "", "Commands (repository root, on the MI355X):", "", "```"] + args.command + ["```", "",
           args.workload, "", "## Top kernels (rocprofv3 --stats)", "", text, "",
           "## HIP-event per-kernel averages vs rocprofv3", "",
           "bench.py times every launch of a 5-step cycle with HIP events on the library's stream (one stream, each "
           "launch bracketed and synchronised).  rocprofv3 averages all launches of the same kernel over the whole "
           "trace.  Kernels are keyed by name and floating-point template arguments (all launch sites and "
           "variants of a kernel together).", "",
           "| kernel | HIP-event avg ms | launches (cycle) | rocprof avg ms | rocprof launches |",
           "|---|---|---|---|---|"] + rows + [
           "", "## Default benchmark line (`python3 bench.py`, separate run)", "",
           "| variant | ms/step | Mpt-steps/s |", "|---|---|---|"] + steps + [
           "", f"Dominant single kernel: `{r['kernel']}` at {r['achieved']:.0f} GB/s algorithmic "
           f"({r['frac']:.3f} of 8 TB/s), {r['avg_launch_ms']:.4f} ms per launch; whole step "
           f"{bb['step_roofline']['achieved_GBs']:.0f} GB/s ({bb['step_roofline']['frac']:.3f}) at "
           f"{bb['step_roofline']['bytes_per_point_step']} B per point-step.", ""]
    return "\n".join(out)


def sq(args):
    """Per-kernel instruction counts from a counter pass (SQ_INSTS_VALU,
    _SALU, _LDS, _VMEM_RD, _VMEM_WR, SQ_WAVES) and the VALU issue fraction
    they imply against the HIP-event launch time of a run without the
    profiler: VALU instructions x 4 cycles (a wave64 instruction on a
    16-lane SIMD) / (1024 SIMDs x launch time x clock)."""
    rows = _rows(args.dir, "counter_collection.csv")
    bench = json.loads(pathlib.Path(args.bench_json).read_text().splitlines()[-1])
    kern = {}
    for name, v in bench.get("kernels", {}).items():
        e = kern.setdefault(name.split(" [")[0], {"ms": 0.0, "calls": 0})
        e["ms"] += v["ms"]
        e["calls"] += v["calls"]
    # the counter run's last `calls` dispatches of each kernel: its HIP-event
    # cycle, the same steps of the cycle as the benchmark's
    ids = {}
    for r in rows:
        ids.setdefault(rocprof_key(_col(r, "Kernel_Name")), set()).add(int(_col(r, "Dispatch_Id", "Correlation_Id")))
    keep = {k: set(_last(sorted(v), kern.get(k, {}).get("calls", 0))) for k, v in ids.items()}
    agg, disp = {}, {}
    for r in rows:
        k = rocprof_key(_col(r, "Kernel_Name"))
        did = int(_col(r, "Dispatch_Id", "Correlation_Id"))
        if did not in keep[k]:
            continue
        a = agg.setdefault(k, {})
        n = _col(r, "Counter_Name").strip()
        a[n] = a.get(n, 0.0) + float(_col(r, "Counter_Value"))
        disp.setdefault(k, set()).add(did)
    lines = [f"# {args.title}", "", "Commands (repository root, on the MI355X):", "", "```"] + args.command + [
        "```", "", f"VALU issue fraction = SQ_INSTS_VALU x 4 / (1024 SIMDs x HIP-event launch time x "
        f"{args.clock_ghz} GHz); the HIP-event times are those of the default benchmark run (no profiler).", "",
        "| kernel | dispatches | VALU / launch | SALU / launch | LDS / launch | VMEM / launch | waves | "
        "HIP-event ms | VALU issue fraction |", "|---|---|---|---|---|---|---|---|---|"]
    order = sorted(agg, key=lambda k: -(kern.get(k, {}).get("ms", 0.0)))
    for k in order:
        if not k.startswith("k_") or k not in kern:
            continue
        c, nd = agg[k], len(disp[k])
        avg = kern[k]["ms"] / kern[k]["calls"]
        valu = c.get("SQ_INSTS_VALU", 0.0) / nd
        util = valu * 4 / (1024 * avg * 1e-3 * args.clock_ghz * 1e9)
        vmem = (c.get("SQ_INSTS_VMEM_RD", 0.0) + c.get("SQ_INSTS_VMEM_WR", 0.0)) / nd
        lines.append(f"| `{k}` | {nd} | {valu:.3e} | {c.get('SQ_INSTS_SALU', 0.0) / nd:.2e} | "
                     f"{c.get('SQ_INSTS_LDS', 0.0) / nd:.2e} | {vmem:.2e} | {c.get('SQ_WAVES', 0.0) / nd:.0f} | "
                     f"{avg:.4f} | {util:.2f} |")
    return "\n".join(lines) + "\n"


def _cycle_counters(path, calls):
    """Counter sums per kernel key over the last `calls[key]` dispatches of
    each kernel in a counter directory (the HIP-event cycle)."""
    rows = _rows(path, "counter_collection.csv")
    ids = {}
    for r in rows:
        ids.setdefault(rocprof_key(_col(r, "Kernel_Name")), set()).add(int(_col(r, "Dispatch_Id", "Correlation_Id")))
    keep = {k: set(_last(sorted(v), calls.get(k, 0))) for k, v in ids.items()}
    agg = {}
    for r in rows:
        k = rocprof_key(_col(r, "Kernel_Name"))
        if int(_col(r, "Dispatch_Id", "Correlation_Id")) not in keep[k]:
            continue
        a = agg.setdefault(k, {})
        n = _col(r, "Counter_Name").strip()
        a[n] = a.get(n, 0.0) + float(_col(r, "Counter_Value"))
    return agg


def stalls(args):
    """Where the waves of each kernel spend their time, from two counter
    passes: SQ_WAIT_ANY (waiting on memory counters, barriers or anything
    else), SQ_WAIT_INST_ANY (ready, waiting for issue) and SQ_WAIT_INST_LDS,
    each over SQ_WAVE_CYCLES (all in quad-cycles); and the LDS bank-conflict
    stall cycles over the conflict-free LDS cycles (SQ_LDS_BANK_CONFLICT /
    (SQ_LDS_IDX_ACTIVE - SQ_LDS_BANK_CONFLICT))."""
    bench = json.loads(pathlib.Path(args.bench_json).read_text().splitlines()[-1])
    calls, ms = {}, {}
    for name, v in bench.get("kernels", {}).items():
        k = name.split(" [")[0]
        calls[k] = calls.get(k, 0) + v["calls"]
        ms[k] = ms.get(k, 0.0) + v["ms"]
    a = _cycle_counters(args.dir_wait, calls)
    b = _cycle_counters(args.dir_lds, calls)
    lines = [f"# {args.title}", "", "Commands (repository root, on the MI355X):", "", "```"] + args.command + [
        "```", "", "Fractions of the waves' cycles (SQ_WAVE_CYCLES) over the launches of the benchmark's HIP-event "
        "cycle. Kernels in order of their HIP-event time per step.", "",
        "| kernel | waiting (SQ_WAIT_ANY) | ready, not issued (SQ_WAIT_INST_ANY) | waiting for LDS issue | "
        "LDS bank-conflict stalls / LDS cycles |", "|---|---|---|---|---|"]
    for k in sorted(a, key=lambda k: -ms.get(k, 0.0)):
        if not k.startswith("k_") or k not in ms:
            continue
        c = a[k]
        wc = c.get("SQ_WAVE_CYCLES", 0.0)
        if wc <= 0:
            continue
        d = b.get(k, {})
        act = d.get("SQ_LDS_IDX_ACTIVE", 0.0) - d.get("SQ_LDS_BANK_CONFLICT", 0.0)
        bank = f"{d.get('SQ_LDS_BANK_CONFLICT', 0.0) / act:.2f}" if act > 0 else "–"
        lines.append(f"| `{k}` | {c.get('SQ_WAIT_ANY', 0.0) / wc:.2f} | {c.get('SQ_WAIT_INST_ANY', 0.0) / wc:.2f} | "
                     f"{c.get('SQ_WAIT_INST_LDS', 0.0) / wc:.3f} | {bank} |")
    return "\n".join(lines) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    k = sub.add_parser("kernels")
    k.add_argument("dir")
    k.add_argument("--top", type=int, default=25)
    t = sub.add_parser("traffic")
    t.add_argument("fetch")
    t.add_argument("write")
    t.add_argument("--bench-json", required=True, help="stdout of the counter run (bench.py --steps 0)")
    t.add_argument("--copy-pattern", default="MulFunctor")
    t.add_argument("--copy-mib", type=int, required=True)
    t.add_argument("--variant", required=True)
    t.add_argument("--command", action="append", default=[])
    t.add_argument("--out")
    rp = sub.add_parser("report")
    rp.add_argument("trace")
    rp.add_argument("--trace-json", required=True, help="stdout of the traced bench.py run")
    rp.add_argument("--bench-json", required=True, help="stdout of a default bench.py run")
    rp.add_argument("--title", required=True)
    rp.add_argument("--workload", required=True)
    rp.add_argument("--command", action="append", default=[])
    rp.add_argument("--top", type=int, default=30)
    rp.add_argument("--variant", default="learned", help="variant of the traced run")
    rp.add_argument("--out")
    q = sub.add_parser("sq")
    q.add_argument("dir")
    q.add_argument("--bench-json", required=True, help="stdout of a default bench.py run")
    q.add_argument("--title", required=True)
    q.add_argument("--command", action="append", default=[])
    q.add_argument("--clock-ghz", type=float, default=2.4)
    q.add_argument("--out")
    st = sub.add_parser("stalls")
    st.add_argument("dir_wait", help="counter pass with SQ_WAVE_CYCLES, SQ_WAIT_ANY, SQ_WAIT_INST_ANY, SQ_WAIT_INST_LDS")
    st.add_argument("dir_lds", help="counter pass with SQ_LDS_BANK_CONFLICT, SQ_LDS_IDX_ACTIVE")
    st.add_argument("--bench-json", required=True, help="stdout of a default bench.py run")
    st.add_argument("--title", required=True)
    st.add_argument("--command", action="append", default=[])
    st.add_argument("--out")
    a = ap.parse_args(argv)
    if a.cmd in ("sq", "stalls"):
        text = sq(a) if a.cmd == "sq" else stalls(a)
        if a.out:
            pathlib.Path(a.out).write_text(text)
        print(text)
        return 0
    if a.cmd == "kernels":
        text, _ = kernels(a.dir, a.top)
        print(text)
    elif a.cmd == "report":
        text = report(a)
        if a.out:
            pathlib.Path(a.out).write_text(text)
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
