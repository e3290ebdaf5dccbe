"""Summaries of rocprofv3 output for the benchmark's kernels, and the copy that calibrates the HBM counters.

    python -m o_o_solver_amd.profiling copy
        A device-to-device copy of 2 GiB (float64), nothing else: run under `rocprofv3 --pmc FETCH_SIZE` and
        `--pmc WRITE_SIZE` it gives what each counter reads for a known number of bytes.
    python -m o_o_solver_amd.profiling stats KERNEL_STATS_CSV OUT [--command CMD]
        The per-kernel table of a `rocprofv3 --kernel-trace --stats -f csv` run (calls, total and average duration).
    python -m o_o_solver_amd.profiling traffic --fetch CSV --write CSV --copy-fetch CSV --copy-write CSV OUT
        Per kernel, the average FETCH_SIZE and WRITE_SIZE per dispatch (KiB, as the counters read), each scaled by
        copy bytes / counter reading of the calibration copy, and their sum: the HBM bytes per launch.

    python -m o_o_solver_amd.profiling mfma COUNTERS_CSV OUT [--command CMD] [--stats STATS_SUMMARY]
        Per kernel, the MFMA flops per dispatch from SQ_INSTS_VALU_MFMA_MOPS_F64 and _F32 (units of 512 flops), and
        with a statistics summary their rate over the kernel's average duration and its fraction of the MFMA peak.

    python -m o_o_solver_amd.profiling levels KERNEL_TRACE_CSV OUT [--command CMD]
        From a `rocprofv3 --kernel-trace -f csv` run of the benchmark: for the last combined factorisation and solve
        of each precision in the trace (a factorisation with the solve's forward sweep beside it, as the benchmark's
        timed steps run; the fastest factorisation when there is none), the time of each level (from its assembly
        launch to the next level's, the small fronts before the first assembly apart), and the kernel time of the
        solve that follows it.

read_traffic(path, kernel) returns the calibrated bytes per launch of one kernel from a traffic summary.
"""

import argparse
import csv
import collections
import os
import re
import sys

COPY_BYTES = 1 << 31  # the calibration copy: 2^28 float64 values read and written once


def kernel_name(full):
    """'void oosg::extendTilesKernel<double, 64>(...)' → 'extendTilesKernel<double,64>'."""
    name = full.strip().replace("(anonymous namespace)::", "")
    if name.startswith("void "):
        name = name[5:]
    depth, end = 0, len(name)
    for i, ch in enumerate(name):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            end = i
            break
    name = name[:end].replace(", ", ",")  # one token: the summaries split their lines on spaces
    return re.sub(r"^(oosg::)+", "", name)


def counter_per_kernel(path, counter):
    """{kernel: [value per dispatch]} of one counter of a counter_collection.csv."""
    values = collections.defaultdict(list)
    with open(path) as f:
        for row in csv.DictReader(f):
            if row["Counter_Name"] == counter:
                values[kernel_name(row["Kernel_Name"])].append(float(row["Counter_Value"]))
    return values


def copy_factor(path, counter):
    """COPY_BYTES / (the counter's reading, in bytes, for the calibration copy's kernel)."""
    readings = counter_per_kernel(path, counter)
    copies = [v for k, vs in readings.items() if "copyBuffer" in k for v in vs]
    if len(copies) != 1:
        raise ValueError(f"{path}: expected one copy kernel dispatch, found {len(copies)}")
    return COPY_BYTES / (copies[0] * 1024.0)


def stats(csv_path, out, command):
    with open(csv_path) as f:
        rows = list(csv.DictReader(f))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    total = sum(float(r["TotalDurationNs"]) for r in rows)
    lines = [f"# rocprofv3 kernel statistics", f"# command: {command}", "#",
             f"# {'kernel':<44} {'calls':>7} {'total ms':>10} {'avg us':>10} {'min us':>9} {'max us':>9} {'share':>6}"]
    for r in rows:
        lines.append(f"  {kernel_name(r['Name']):<44} {int(r['Calls']):>7} {float(r['TotalDurationNs']) / 1e6:>10.3f} "
                     f"{float(r['AverageNs']) / 1e3:>10.2f} {float(r['MinNs']) / 1e3:>9.2f} "
                     f"{float(r['MaxNs']) / 1e3:>9.2f} {float(r['TotalDurationNs']) / total:>6.3f}")
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")


def traffic(fetch, write, copy_fetch, copy_write, out, commands):
    kf = copy_factor(copy_fetch, "FETCH_SIZE")
    kw = copy_factor(copy_write, "WRITE_SIZE")
    f = counter_per_kernel(fetch, "FETCH_SIZE")
    w = counter_per_kernel(write, "WRITE_SIZE")
    lines = ["# HBM traffic per dispatch: rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE (KiB per dispatch), each",
             "# scaled by copy bytes / reading of a plain 2 GiB device copy under the same counter.",
             f"# calibration: FETCH_SIZE x {kf:.4f}, WRITE_SIZE x {kw:.4f}"]
    lines += [f"# command: {c}" for c in commands]
    lines += ["#", f"# {'kernel':<40} {'dispatches':>10} {'fetch KiB':>12} {'write KiB':>12} {'traffic bytes':>16}"]
    for k in sorted(set(f) & set(w)):
        fa = sum(f[k]) / len(f[k])
        wa = sum(w[k]) / len(w[k])
        bytes_ = (fa * kf + wa * kw) * 1024.0
        lines.append(f"  {k:<40} {len(f[k]):>10} {fa:>12.1f} {wa:>12.1f} {bytes_:>16.0f}")
    with open(out, "w") as fo:
        fo.write("\n".join(lines) + "\n")


MFMA_COUNTERS = ["SQ_INSTS_VALU_MFMA_MOPS_F64", "SQ_INSTS_VALU_MFMA_MOPS_F32"]
# MI355X dense matrix-core peaks (FLOP/s).
MFMA_PEAK = {"f64": 78.6e12, "f32": 157.3e12}


def read_stats(path):
    """{kernel: average duration in seconds} of a kernel statistics summary (stats())."""
    out = {}
    with open(path) as f:
        for line in f:
            if line.startswith("#"):
                continue
            parts = line.split()
            if len(parts) == 7:
                out[parts[0]] = float(parts[3]) * 1e-6
    return out


def mfma(csv_path, out, command, stats_path=None):
    per = {c: counter_per_kernel(csv_path, c) for c in MFMA_COUNTERS}
    durations = read_stats(stats_path) if stats_path else {}
    kernels = sorted(set().union(*[set(v) for v in per.values()]))
    lines = ["# Matrix-core work per dispatch (rocprofv3 --pmc SQ_INSTS_VALU_MFMA_MOPS_F64 / _F32: MFMA operations in",
             "# units of 512 flops, all padding included), and its rate over the kernel's average duration in the",
             f"# statistics summary {os.path.basename(stats_path) if stats_path else '(none)'}, against the dense MFMA",
             "# peak of its precision (FP64 78.6, FP32 157.3 TFLOP/s).",
             f"# command: {command}", "#",
             f"# {'kernel':<40} {'dispatches':>10} {'mfma flops':>12} {'avg us':>9} {'TFLOP/s':>8} {'of peak':>8}"]
    for k in kernels:
        f64 = per["SQ_INSTS_VALU_MFMA_MOPS_F64"].get(k, [])
        f32 = per["SQ_INSTS_VALU_MFMA_MOPS_F32"].get(k, [])
        count = max(len(f64), len(f32))
        a64 = sum(f64) / len(f64) if f64 else 0.0
        a32 = sum(f32) / len(f32) if f32 else 0.0
        flops = 512.0 * (a64 + a32)
        if flops == 0.0:
            continue
        peak = MFMA_PEAK["f32" if a32 > a64 else "f64"]
        t = durations.get(k)
        rate = flops / t if t else float("nan")
        lines.append(f"  {k:<40} {count:>10} {flops:>12.4g} {t * 1e6 if t else float('nan'):>9.2f} "
                     f"{rate / 1e12:>8.2f} {rate / peak:>8.3f}")
    with open(out, "w") as fo:
        fo.write("\n".join(lines) + "\n")


def read_traffic(path, kernel):
    """The calibrated HBM bytes per launch of `kernel` in a traffic summary, or None."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("#"):
                continue
            parts = line.split()
            if len(parts) == 5 and parts[0] == kernel:
                return float(parts[4])
    return None


def factorisations(csv_path):
    """The factorisations of a kernel trace: per factorisation its precision, its kernels (name, start, end in ns)
    from its first small-front or assembly launch to its pivot statistics, and the kernels of the solve after it."""
    with open(csv_path, newline="") as f:
        rows = [(kernel_name(r["Kernel_Name"]), int(r["Start_Timestamp"]), int(r["End_Timestamp"]))
                for r in csv.DictReader(f)]
    rows.sort(key=lambda r: r[1])
    found, current, after = [], None, None
    for r in rows:
        name = r[0]
        starts = name.startswith("smallFrontKernel") or name.startswith("assembleKernel")
        if current is None and starts:
            current, after = [], None
        if current is not None:
            current.append(r)
            if name.startswith("pivotStatisticsKernel"):
                precision = next((k.split("<")[1].split(">")[0] for k, _, _ in current if "<" in k
                                  and (k.startswith("assembleKernel") or k.startswith("smallFrontKernel"))), "?")
                after = []
                found.append({"precision": precision, "kernels": current, "solve": after})
                current = None
        elif after is not None:
            after.append(r)
    return found


def levels(csv_path, out, command):
    lines = [f"# levels of the last combined factorisation and solve of each precision, from "
             f"{os.path.basename(csv_path)}",
             f"# command: {command}" if command else "#",
             "# level: from its assembly launch to the next level's (the last: to the pivot statistics); busy: the",
             "# kernel time started in that window (the streams' kernels overlap, so busy can exceed the span)"]
    found = factorisations(csv_path)
    for precision in sorted({f["precision"] for f in found}):
        runs = [f for f in found if f["precision"] == precision]
        span = lambda f: f["kernels"][-1][1] - f["kernels"][0][1]
        combined = [f for f in runs if any(k[0].startswith("tinyForwardKernel") for k in f["kernels"])]
        best = combined[-1] if combined else min(runs, key=span)
        ks = best["kernels"]
        t0 = ks[0][1]
        marks = [k[1] for k in ks if k[0].startswith("assembleKernel")] + [ks[-1][1]]
        lines.append(f"\n{precision}: {len(runs)} factorisations in the trace, {len(combined)} combined with the solve; "
                     f"{'the last combined' if combined else 'the fastest'} {span(best) / 1e6:.3f} ms "
                     f"(first launch to pivot statistics)")
        lines.append(f"  {'level':>5s} {'span ms':>9s} {'busy ms':>9s}")
        windows = [("small", t0, marks[0])] + [(str(i), marks[i], marks[i + 1]) for i in range(len(marks) - 1)]
        for label, a, b in windows:
            busy = sum(e - s for _, s, e in ks if a <= s < b)
            lines.append(f"  {label:>5s} {(b - a) / 1e6:9.3f} {busy / 1e6:9.3f}")
        solve = best["solve"]
        if solve:
            lines.append(f"  solve after it: {len(solve)} kernels, span {(solve[-1][2] - solve[0][1]) / 1e6:.3f} ms, "
                         f"kernel time {sum(e - s for _, s, e in solve) / 1e6:.3f} ms")
            per = collections.defaultdict(lambda: [0, 0])
            for name, s0, e0 in solve:
                per[name][0] += 1
                per[name][1] += e0 - s0
            for name, (count, ns) in sorted(per.items(), key=lambda x: -x[1][1])[:12]:
                lines.append(f"    {ns / 1e6:8.3f} ms {count:4d}  {name}")
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


def copy():
    import torch
    n = COPY_BYTES // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda:0")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    print(f"copied {COPY_BYTES} bytes", flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="what", required=True)
    sub.add_parser("copy")
    s = sub.add_parser("stats")
    s.add_argument("csv")
    s.add_argument("out")
    s.add_argument("--command", default="")
    t = sub.add_parser("traffic")
    for name in ("fetch", "write", "copy-fetch", "copy-write"):
        t.add_argument("--" + name, required=True)
    t.add_argument("--command", action="append", default=[])
    t.add_argument("out")
    m = sub.add_parser("mfma")
    m.add_argument("csv")
    m.add_argument("out")
    m.add_argument("--command", default="")
    m.add_argument("--stats", default=None)
    v = sub.add_parser("levels")
    v.add_argument("csv")
    v.add_argument("out")
    v.add_argument("--command", default="")
    args = ap.parse_args(argv)
    if args.what == "copy":
        copy()
    elif args.what == "stats":
        stats(args.csv, args.out, args.command)
    elif args.what == "mfma":
        mfma(args.csv, args.out, args.command, args.stats)
    elif args.what == "levels":
        levels(args.csv, args.out, args.command)
    else:
        traffic(args.fetch, args.write, args.copy_fetch, args.copy_write, args.out, args.command)


if __name__ == "__main__":
    sys.exit(main())
