#!/usr/bin/env python3
"""Summaries of rocprofv3 CSV output for the solver's kernels.

    rocprof_summary.py stats  <run_kernel_stats.csv>                 [--title T] [--cmd C]
    rocprof_summary.py traffic <fetch counter_collection.csv> <write counter_collection.csv>
                               --workload W [--copy-mib M] [--cmd C]
    rocprof_summary.py busy   <run_kernel_trace.csv>                 [--title T] [--cmd C] [--top K]

`stats` prints a Markdown table of per-kernel calls and durations (from
`rocprofv3 --kernel-trace --stats --output-format csv`).  A kernel that ran
as several template variants (the autotuner's trial launches) is reported,
in both commands, as the variant with the most dispatches: the one the step
used.  `traffic` turns two
`rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE` runs into HBM bytes per
launch for each kernel, JSON on stdout.  Both counters are in KiB per
dispatch; each is scaled by what it reads for the runtime's device-to-device
copy kernel (`__amd_rocclr_copyBuffer`) of a known size M MiB that the
profiled program ran first (bench.py --copy-calibration-mb M): a copy of M MiB
fetches and writes M MiB, so factor = M MiB / sum of the counter over the copy
dispatches.  `busy` reports how much of a kernel trace's span the device had a
kernel running, and how much kernels of several streams overlapped.
"""
import argparse
import csv
import json
import re
import sys
from collections import defaultdict


def template_args(full, start):
    """Top-level template arguments of the name whose '<' is at `start`."""
    depth, cur, out = 0, "", []
    for ch in full[start:]:
        if ch == "<":
            depth += 1
            if depth == 1:
                continue
        elif ch == ">":
            depth -= 1
            if depth == 0:
                out.append(cur.strip())
                return out
        elif ch == "," and depth == 1:
            out.append(cur.strip())
            cur = ""
            continue
        cur += ch
    return out


def short_name(full):
    """Kernel name as bench.py reports it, or None for kernels outside the step."""
    m = re.search(r"dbec::(k_\w+)", full)
    if not m:
        return "copy" if "copyBuffer" in full else None
    k = m.group(1)
    targs = template_args(full, m.end()) if full[m.end():m.end() + 1] == "<" else []
    if k == "k_ypass":
        return "ypass_fwd" if targs[1] == "-1" else "ypass_inv"
    if k == "k_sweep_xy":  # template <typename S, int AXIS, int L, int M>
        return "sweep_x" if targs[1] == "0" else "sweep_y"
    return {
        "k_zfwd": "zfwd", "k_xconv": "xconv", "k_zinv_nu": "zinv_nu", "k_zinv": "zinv",
        "k_sweep_z": "sweep_z", "k_sweepz_zfwd": "sweep_z_zfwd", "k_norm_cols": "norm", "k_scale": "scale",
    }.get(k, k[2:])


def dominant(variants):
    """The variant (full kernel name) with the most dispatches.  A kernel that
    ran as several template instantiations (the autotuner's trial launches at
    context creation) is reported as the one the step used."""
    return max(variants.items(), key=lambda kv: kv[1][0])


def cmd_stats(args):
    rows = list(csv.DictReader(open(args.csv)))
    byn = defaultdict(dict)
    for r in rows:
        n = short_name(r["Name"]) or r["Name"][:40]
        c, t = byn[n].get(r["Name"], (0, 0.0))
        byn[n][r["Name"]] = (c + int(r["Calls"]), t + float(r["TotalDurationNs"]))
    agg = {}
    for n, variants in byn.items():
        full, (calls, tot) = dominant(variants)
        agg[n] = [calls, tot, full]
    total = sum(a[1] for a in agg.values())
    print(f"# {args.title}\n")
    if args.cmd:
        print(f"Command: `{args.cmd}`\n")
    print("| kernel | calls | avg us | share | full name |")
    print("|---|---:|---:|---:|---|")
    for n, (calls, tot, full) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        print(f"| {n} | {calls} | {tot / calls / 1e3:.1f} | {100 * tot / total:.1f}% | `{full[:110]}` |")


def first_targ(full):
    """The first template argument of a dbec kernel (its transform length or
    storage type), or ''."""
    m = re.search(r"dbec::(k_\w+)", full)
    if not m or full[m.end():m.end() + 1] != "<":
        return ""
    targs = template_args(full, m.end())
    return targs[0] if targs else ""


def cmd_busy(args):
    """Device occupancy of a kernel trace (rocprofv3 --kernel-trace
    run_kernel_trace.csv): the span from the first kernel's start to the last
    one's end, the union of the kernels' intervals (time with at least one
    kernel running), their summed durations (more than the union when
    kernels of several streams overlap), and the summed time per kernel and
    first template argument."""
    iv, per = [], defaultdict(lambda: [0, 0.0])
    for r in csv.DictReader(open(args.csv)):
        a, b = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
        iv.append((a, b))
        name = r["Kernel_Name"]
        key = (short_name(name) or name[:40], first_targ(name))
        per[key][0] += 1
        per[key][1] += b - a
    iv.sort()
    span = iv[-1][1] - iv[0][0]
    union, cur_a, cur_b = 0, iv[0][0], iv[0][1]
    for a, b in iv[1:]:
        if a > cur_b:
            union += cur_b - cur_a
            cur_a, cur_b = a, b
        else:
            cur_b = max(cur_b, b)
    union += cur_b - cur_a
    summed = sum(b - a for a, b in iv)
    print(f"# {args.title}\n")
    if args.cmd:
        print(f"Command: `{args.cmd}`\n")
    print(f"* kernels: {len(iv)}")
    print(f"* span, first kernel start to last kernel end: {span / 1e9:.3f} s")
    print(f"* time with at least one kernel running: {union / 1e9:.3f} s ({100 * union / span:.1f} % of the span)")
    print(f"* summed kernel durations: {summed / 1e9:.3f} s ({summed / union:.2f}x the busy time: kernels of "
          f"several streams overlapping)\n")
    print("| kernel | template arg | calls | total s | share of summed |")
    print("|---|---|---:|---:|---:|")
    for (n, t), (c, d) in sorted(per.items(), key=lambda kv: -kv[1][1])[: args.top]:
        print(f"| {n} | {t} | {c} | {d / 1e9:.3f} | {100 * d / summed:.1f}% |")


def counter_avg(path, counter):
    per = defaultdict(list)
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter:
            continue
        per[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return per


def cmd_traffic(args):
    fetch = counter_avg(args.fetch_csv, "FETCH_SIZE")
    write = counter_avg(args.write_csv, "WRITE_SIZE")
    mib = args.copy_mib

    def factor(per):
        vals = [v for k, vs in per.items() if "copyBuffer" in k for v in vs]
        if not vals:
            raise SystemExit("no copy dispatch found: run the program with --copy-calibration-mb")
        return mib * 1024.0 / sum(vals)

    ff, fw = factor(fetch), factor(write)
    out = {"workload": args.workload, "command": args.cmd, "copy_calibration_mib": mib,
           "fetch_factor": ff, "write_factor": fw, "kernels": {}}
    def dominant_per_short(per):
        variants = defaultdict(dict)  # short name -> full name -> (dispatches,)
        for k, vs in per.items():
            s = short_name(k)
            if s and s != "copy":
                variants[s][k] = (len(vs),)
        return {s: dominant(vv)[0] for s, vv in variants.items()}

    # the dominant variant is chosen per counter pass: the two passes are
    # separate runs, and a tuned launch choice may differ between them
    fdom, wdom = dominant_per_short(fetch), dominant_per_short(write)
    for s in sorted(set(fdom) | set(wdom)):
        if s not in fdom or s not in wdom:
            print(f"warning: {s} appears in only one counter pass; not reported", file=sys.stderr)
            continue
        f, w = fetch[fdom[s]], write[wdom[s]]
        rec = {
            "dispatches": len(f),
            "fetch_kib_raw": sum(f) / len(f),
            "write_kib_raw": sum(w) / len(w),
            "fetch_bytes": ff * 1024.0 * sum(f) / len(f),
            "write_bytes": fw * 1024.0 * sum(w) / len(w),
        }
        if fdom[s] != wdom[s]:
            print(f"warning: {s} ran different variants in the two passes: {fdom[s]} / {wdom[s]}", file=sys.stderr)
            rec["fetch_variant"], rec["write_variant"] = fdom[s], wdom[s]
        out["kernels"][s] = rec
    json.dump(out, sys.stdout, indent=1)
    print()


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="what", required=True)
    a = sub.add_parser("stats")
    a.add_argument("csv")
    a.add_argument("--title", default="rocprofv3 kernel statistics")
    a.add_argument("--cmd", default="")
    c = sub.add_parser("busy")
    c.add_argument("csv")
    c.add_argument("--title", default="Device occupancy of a kernel trace")
    c.add_argument("--cmd", default="")
    c.add_argument("--top", type=int, default=25)
    b = sub.add_parser("traffic")
    b.add_argument("fetch_csv")
    b.add_argument("write_csv")
    b.add_argument("--workload", required=True)
    b.add_argument("--copy-mib", type=float, default=512)
    b.add_argument("--cmd", default="")
    args = ap.parse_args()
    {"stats": cmd_stats, "traffic": cmd_traffic, "busy": cmd_busy}[args.what](args)


if __name__ == "__main__":
    main()
