"""Summarize rocprofv3 output of bench.py into the text files under profiles/.

Inputs are the CSV directories of three rocprofv3 runs of the same bench
command (``--kernel-trace --stats``, ``--pmc FETCH_SIZE``, ``--pmc WRITE_SIZE``).
Byte counters are scaled by what they read for the calibration kernel that
bench.py runs on a known size (an elementwise multiply reading and writing
16 Mi floats, 67.1 MB each way), so the figures are HBM bytes per launch.

    python -m waveletpotentials_amd.rocprof_summary TRACE_DIR FETCH_DIR WRITE_DIR OUT_PREFIX \\
        --command "..."

writes OUT_PREFIX.md (per-kernel durations and bytes) and OUT_PREFIX.json
(per-kernel fetch/write bytes, read by bench.py for ``roofline.traffic``).
With ``--trace-only TRACE_DIR OUT_PREFIX`` it writes OUT_PREFIX.md with the
per-kernel durations of one ``--kernel-trace --stats`` run.
"""
import argparse
import collections
import csv
import glob
import json
import os
import re
import statistics

CAL_BYTES = 16 * 1024 * 1024 * 4  # calibration kernel: bytes read and bytes written
CAL_NAME = "vectorized_elementwise_kernel"
CAL_OP = "MulFunctor<float>"  # torch.mul(src, 1.0, out=dst) on FP32 tensors

_WT = {"Cdf37": "cdf3.7 n", "Cdf46": "cdf4.6 d"}


def kernel_label(name):
    """The label the library's stage trace uses for a kernel symbol."""
    m = re.search(r"lp_kernel.*Up(Cdf\d\d)E.*ELi(\d)ELi(\d)ELb([01])ELb([01])E", name)
    if m:
        return (f"lp_kernel<{_WT.get(m.group(1), m.group(1))},L={m.group(2)},"
                f"{'fwd' if m.group(4) == '1' else 'inv'},"
                f"{'contiguous' if m.group(5) == '1' else 'strided'}>")
    m = re.search(r"lp_kernel<.*Up(Cdf\d\d),.*>, (\d), (\d), (true|false), (true|false)>", name)
    if m:
        return (f"lp_kernel<{_WT.get(m.group(1), m.group(1))},L={m.group(2)},"
                f"{'fwd' if m.group(4) == 'true' else 'inv'},"
                f"{'contiguous' if m.group(5) == 'true' else 'strided'}>")
    m = re.search(r"line_tile_kernel.*Up(Cdf\d\d)E.*EE([fd])Lb([01])ELb([01])E", name)
    if m:
        return (f"line_tile_kernel<{_WT.get(m.group(1), m.group(1))},"
                f"{'f32' if m.group(2) == 'f' else 'f64'},{'fwd' if m.group(3) == '1' else 'inv'},"
                f"{'strided' if m.group(4) == '1' else 'contiguous'}>")
    m = re.search(r"unet\d+(conv3_wgrad_kernel|conv3_kernel)I((?:L[ib]\d+E)+)E", name)
    if m:
        args = [v if k == "i" else ("true" if v == "1" else "false")
                for k, v in re.findall(r"L([ib])(\d+)E", m.group(2))]
        return f"unet::{m.group(1)}<{', '.join(args)}>"
    m = re.match(r"(?:void )?whhd::((?:\w+::)*)(?:\(anonymous namespace\)::)?(\w+(?:<[^>]*>)?)",
                 name)
    if m:
        ns = "unet::" if m.group(1).startswith("unet") else ""
        return ns + m.group(2)
    return name.split("(")[0][:80]


def is_library(name):
    """A kernel of this library: demangled under whhd::, or a mangled whhd symbol
    (rocprofv3 leaves some long template names mangled)."""
    return "whhd::" in name or name.startswith("_ZN4whhd")


def _one(pattern):
    hits = glob.glob(pattern)
    if not hits:
        raise FileNotFoundError(pattern)
    return hits[0]


def counters(d, by_grid=False):
    out = collections.defaultdict(list)
    with open(_one(os.path.join(d, "*counter_collection.csv"))) as f:
        for r in csv.DictReader(f):
            key = (r["Kernel_Name"], int(r["Grid_Size"])) if by_grid else r["Kernel_Name"]
            out[key].append(float(r["Counter_Value"]) * 1024.0)  # KiB -> bytes
    return out


def durations_by_grid(trace_dir):
    """(kernel name, total work-items) -> per-dispatch durations in ns."""
    out = collections.defaultdict(list)
    hits = glob.glob(os.path.join(trace_dir, "*kernel_trace.csv"))
    if not hits:
        return out
    with open(hits[0]) as f:
        for r in csv.DictReader(f):
            g = int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"])
            out[(r["Kernel_Name"], g)].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    return out


def summarize(trace_dir, fetch_dir, write_dir):
    stats = {}
    with open(_one(os.path.join(trace_dir, "*kernel_stats.csv"))) as f:
        for r in csv.DictReader(f):
            stats[r["Name"]] = r
    fetch, write = counters(fetch_dir), counters(write_dir)
    cal = [k for k in fetch if CAL_NAME in k and CAL_OP in k]
    if not cal:
        raise RuntimeError("calibration kernel not found in the counter runs")
    f_scale = CAL_BYTES / statistics.median(fetch[cal[0]])
    w_scale = CAL_BYTES / statistics.median(write[cal[0]])
    kernels = {}
    for name, r in stats.items():
        label = kernel_label(name)
        row = kernels.setdefault(label, {"calls": 0, "total_ns": 0.0, "library": is_library(name)})
        row["calls"] += int(r["Calls"])
        row["total_ns"] += float(r["TotalDurationNs"])
        if name in fetch:
            row["fetch_bytes"] = statistics.median(fetch[name]) * f_scale
        if name in write:
            row["write_bytes"] = statistics.median(write[name]) * w_scale
    for row in kernels.values():
        row["avg_us"] = row["total_ns"] / row["calls"] / 1000.0
    # library kernels launched at several sizes (one per multigrid level, say):
    # one row per launch size
    durs = durations_by_grid(trace_dir)
    fg, wg = counters(fetch_dir, True), counters(write_dir, True)
    sizes = collections.defaultdict(set)
    for (name, g) in durs:
        sizes[name].add(g)
    by_size = []
    for (name, g), d in durs.items():
        if not is_library(name) or len(sizes[name]) < 2:
            continue
        row = {"kernel": kernel_label(name), "grid": g, "calls": len(d),
               "avg_us": statistics.mean(d) / 1000.0}
        if (name, g) in fg:
            row["fetch_bytes"] = statistics.median(fg[(name, g)]) * f_scale
        if (name, g) in wg:
            row["write_bytes"] = statistics.median(wg[(name, g)]) * w_scale
        by_size.append(row)
    by_size.sort(key=lambda r: (r["kernel"], -r["grid"]))
    return kernels, f_scale, w_scale, by_size


def write_outputs(kernels, f_scale, w_scale, prefix, command, by_size=()):
    lib = {k: v for k, v in kernels.items() if v.get("library")}
    with open(prefix + ".json", "w") as f:
        json.dump({"command": command, "fetch_scale": f_scale, "write_scale": w_scale,
                   "kernels": {k: {"avg_us": round(v["avg_us"], 3), "calls": v["calls"],
                                   "fetch_bytes": round(v.get("fetch_bytes", 0.0)),
                                   "write_bytes": round(v.get("write_bytes", 0.0))}
                               for k, v in sorted(lib.items())}}, f, indent=1)
    lines = [f"# rocprofv3 summary: {os.path.basename(prefix)}", "",
             "Command (three runs; byte counters scaled by the calibration kernel, "
             f"FETCH_SIZE x{f_scale:.3f}, WRITE_SIZE x{w_scale:.3f}):", "", "```", command, "```",
             "", "HBM GB/s = (fetch + write bytes) / average duration; peak 8000 GB/s.", "",
             "| kernel | calls | avg us | HBM fetch MB/launch | HBM write MB/launch | HBM GB/s | % of peak |",
             "|---|---:|---:|---:|---:|---:|---:|"]
    for k, v in sorted(lib.items(), key=lambda kv: -kv[1]["total_ns"]):
        b = v.get("fetch_bytes", 0) + v.get("write_bytes", 0)
        gbs = b / (v["avg_us"] * 1e-6) / 1e9 if v["avg_us"] > 0 else 0.0
        lines.append(f"| `{k}` | {v['calls']} | {v['avg_us']:.2f} | "
                     f"{v.get('fetch_bytes', 0) / 1e6:.1f} | {v.get('write_bytes', 0) / 1e6:.1f} | "
                     f"{gbs:.0f} | {100 * gbs / 8000:.1f}% |")
    if by_size:
        lines += ["", "Kernels launched at several sizes, one row per launch size (work-items):", "",
                  "| kernel | work-items | calls | avg us | HBM fetch MB/launch | HBM write MB/launch | HBM GB/s | % of peak |",
                  "|---|---:|---:|---:|---:|---:|---:|---:|"]
        for v in by_size:
            b = v.get("fetch_bytes", 0) + v.get("write_bytes", 0)
            gbs = b / (v["avg_us"] * 1e-6) / 1e9 if v["avg_us"] > 0 else 0.0
            lines.append(f"| `{v['kernel']}` | {v['grid']} | {v['calls']} | {v['avg_us']:.2f} | "
                         f"{v.get('fetch_bytes', 0) / 1e6:.1f} | {v.get('write_bytes', 0) / 1e6:.1f} | "
                         f"{gbs:.0f} | {100 * gbs / 8000:.1f}% |")
    with open(prefix + ".md", "w") as f:
        f.write("\n".join(lines) + "\n")


def write_trace_only(trace_dir, prefix, command, top=30):
    rows = []
    with open(_one(os.path.join(trace_dir, "*kernel_stats.csv"))) as f:
        for r in csv.DictReader(f):
            rows.append((kernel_label(r["Name"]), int(r["Calls"]), float(r["TotalDurationNs"])))
    merged = {}
    for k, c, t in rows:
        m = merged.setdefault(k, [0, 0.0])
        m[0] += c
        m[1] += t
    total = sum(v[1] for v in merged.values())
    lines = [f"# rocprofv3 summary: {os.path.basename(prefix)}", "", "Command:", "", "```",
             command, "```", "", "| kernel | calls | avg us | total ms | share |",
             "|---|---:|---:|---:|---:|"]
    for k, (c, t) in sorted(merged.items(), key=lambda kv: -kv[1][1])[:top]:
        lines.append(f"| `{k}` | {c} | {t / c / 1000:.2f} | {t / 1e6:.3f} | {100 * t / total:.1f}% |")
    with open(prefix + ".md", "w") as f:
        f.write("\n".join(lines) + "\n")


def main():
    p = argparse.ArgumentParser()
    p.add_argument("dirs", nargs="+", help="TRACE FETCH WRITE OUT_PREFIX, or TRACE OUT_PREFIX "
                                           "with --trace-only")
    p.add_argument("--command", required=True)
    p.add_argument("--trace-only", action="store_true")
    a = p.parse_args()
    if a.trace_only:
        write_trace_only(a.dirs[0], a.dirs[1], a.command)
        return
    trace_dir, fetch_dir, write_dir, out_prefix = a.dirs
    kernels, fs, ws, by_size = summarize(trace_dir, fetch_dir, write_dir)
    write_outputs(kernels, fs, ws, out_prefix, a.command, by_size)


if __name__ == "__main__":
    main()
