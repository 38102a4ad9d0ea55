"""Summaries of rocprofv3 runs of the docking pipeline (kernel times, HBM traffic and SQ
counters).

    python -m megadock_amd.profiling trace DIR           # --kernel-trace --stats, -f csv
    python -m megadock_amd.profiling pmc --batch B --json OUT DIR [DIR ...]

DIR are rocprofv3 output folders written with `-f csv -o run`.  `pmc` reads the counter
passes of one command (FETCH_SIZE in one pass, WRITE_SIZE in another, SQ counters in
others), keeps for each kernel the dispatches of its largest grid (full batches of the
pipeline), averages each counter over them and writes, per kernel and per dispatch:
    hbm_bytes = FETCH_SIZE * 1024 * f_fetch + WRITE_SIZE * 1024 * f_write
where f_fetch and f_write are the factors a plain device-to-device copy of known size in
the same run (bench.py --copy-probe: one dispatch of `k_copy_probe` reading and writing
1 GiB) needs for its counters to read its true bytes; and every SQ counter as read.  bench.py reads the JSON
to report the measured HBM bytes and VALU instructions of its dominant kernel.
"""
import argparse
import collections
import csv
import json
import os
import re
import sys

COPY_KERNEL = "k_copy_probe"
PIPELINE = re.compile(r"kp_|k_vox|k_reduce|k_fwd|k_x_corr|k_inv|k_topt")


def short_name(kernel):
    k = kernel.replace("(anonymous namespace)::", "").replace("HIP_vector_type<float, 2u>", "float2")
    if k.startswith("void "):
        k = k[5:]
    return k.split("(")[0]


def kernel_stats(trace_dir):
    rows = []
    for r in csv.DictReader(open(os.path.join(trace_dir, "run_kernel_stats.csv"))):
        rows.append({"kernel": short_name(r["Name"]), "calls": int(r["Calls"]),
                     "total_ms": int(r["TotalDurationNs"]) / 1e6, "avg_us": float(r["AverageNs"]) / 1e3,
                     "percent": float(r["Percentage"])})
    return rows


def dispatch_stats(trace_dir):
    """Per (kernel, grid) duration statistics from run_kernel_trace.csv: full batches and
    single-rotation calls have different grids and are reported apart."""
    d = collections.defaultdict(list)
    for r in csv.DictReader(open(os.path.join(trace_dir, "run_kernel_trace.csv"))):
        grid = int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"])
        d[(short_name(r["Kernel_Name"]), grid)].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    rows = []
    for (k, g), v in d.items():
        v.sort()
        rows.append({"kernel": k, "grid": g, "calls": len(v), "avg_us": sum(v) / len(v) / 1e3,
                     "median_us": v[len(v) // 2] / 1e3, "total_ms": sum(v) / 1e6})
    rows.sort(key=lambda r: -r["total_ms"])
    return rows


def counter_values(pmc_dir):
    """{counter: {kernel: [(grid, value per dispatch), ...]}} of one counter pass."""
    vals = collections.defaultdict(lambda: collections.defaultdict(list))
    for r in csv.DictReader(open(os.path.join(pmc_dir, "run_counter_collection.csv"))):
        vals[r["Counter_Name"]][short_name(r["Kernel_Name"])].append((int(r["Grid_Size"]), float(r["Counter_Value"])))
    return vals


def largest_grid_mean(v):
    g = max(x[0] for x in v)
    sel = [x[1] for x in v if x[0] == g]
    return g, sum(sel) / len(sel), len(sel)


def pmc_summary(dirs, batch, copy_bytes=1 << 30, min_dispatches=4):
    counters = collections.defaultdict(dict)  # kernel -> counter -> (grid, mean, n)
    for d in dirs:
        for c, per_kernel in counter_values(d).items():
            for k, v in per_kernel.items():
                counters[k][c] = largest_grid_mean(v)
    factors = {}
    for c in ("FETCH_SIZE", "WRITE_SIZE"):
        got = counters.get(COPY_KERNEL, {}).get(c)
        factors[c] = (copy_bytes / 1024.0) / got[1] if got and got[1] > 0 else None
    # of the template variants of one kernel (autotuning runs them all), keep those
    # dispatched as often as the most dispatched one: the variants the timed passes ran
    most = collections.defaultdict(int)
    for k, cs in counters.items():
        most[k.split("<")[0]] = max(most[k.split("<")[0]], max(n for _, _, n in cs.values()))
    kernels = {}
    for k, cs in counters.items():
        n = max(n for _, _, n in cs.values())
        if not PIPELINE.match(k) or n < min_dispatches or n < most[k.split("<")[0]]:
            continue
        row = {"grid": max(g for g, _, _ in cs.values()), "dispatches": max(n for _, _, n in cs.values())}
        for c, (_, mean, _) in cs.items():
            row[c] = mean
        if "FETCH_SIZE" in cs and "WRITE_SIZE" in cs and all(factors.values()):
            row["fetch_bytes"] = cs["FETCH_SIZE"][1] * 1024 * factors["FETCH_SIZE"]
            row["write_bytes"] = cs["WRITE_SIZE"][1] * 1024 * factors["WRITE_SIZE"]
            row["hbm_bytes"] = row["fetch_bytes"] + row["write_bytes"]
        kernels[k] = row
    out = {"batch": batch, "copy_calibration": {"bytes": copy_bytes, "factors": factors}, "kernels": kernels}
    if kernels and all("hbm_bytes" in v for v in kernels.values()):
        out["pipeline_hbm_bytes"] = sum(v["hbm_bytes"] for v in kernels.values())
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("dir")
    p = sub.add_parser("pmc")
    p.add_argument("dirs", nargs="+")
    p.add_argument("--batch", type=int, required=True)
    p.add_argument("--copy-bytes", type=int, default=1 << 30)
    p.add_argument("--json", required=True)
    a = ap.parse_args(argv)
    if a.cmd == "trace":
        print("rocprofv3 --stats summary (all dispatches):\n")
        print("| kernel | calls | avg (us) | total (ms) | % |\n|---|---:|---:|---:|---:|")
        for r in kernel_stats(a.dir):
            print("| %s | %d | %.1f | %.2f | %.2f |" % (r["kernel"], r["calls"], r["avg_us"], r["total_ms"], r["percent"]))
        print("\nPer grid size, from the kernel trace (a full batch and a single rotation apart):\n")
        print("| kernel | grid (work-items) | calls | avg (us) | median (us) | total (ms) |\n|---|---:|---:|---:|---:|---:|")
        for r in dispatch_stats(a.dir):
            if r["total_ms"] < 0.05:
                continue
            print("| %s | %d | %d | %.1f | %.1f | %.2f |" % (r["kernel"], r["grid"], r["calls"], r["avg_us"],
                                                             r["median_us"], r["total_ms"]))
        return 0
    s = pmc_summary(a.dirs, a.batch, a.copy_bytes)
    with open(a.json, "w") as fh:
        json.dump(s, fh, indent=1)
    f = s["copy_calibration"]["factors"]
    print("Copy calibration (%d bytes copied): FETCH_SIZE x %s, WRITE_SIZE x %s\n"
          % (s["copy_calibration"]["bytes"], f["FETCH_SIZE"], f["WRITE_SIZE"]))
    cols = sorted({c for v in s["kernels"].values() for c in v if c.isupper() or c.startswith("SQ_")})
    print("Per full-batch dispatch (%d rotations), mean over the dispatches of the largest grid:\n" % a.batch)
    print("| kernel | dispatches | HBM MB | " + " | ".join(cols) + " |\n|---|---:|---:|" + "---:|" * len(cols))
    for k, v in sorted(s["kernels"].items(), key=lambda kv: -kv[1].get("hbm_bytes", 0)):
        hb = "%.1f" % (v["hbm_bytes"] / 1e6) if "hbm_bytes" in v else "-"
        print("| %s | %d | %s | " % (k, v["dispatches"], hb) + " | ".join("%.4g" % v[c] if c in v else "-" for c in cols) + " |")
    if "pipeline_hbm_bytes" in s:
        print("\nPipeline: %.1f MB per batch, %.2f MB per rotation"
              % (s["pipeline_hbm_bytes"] / 1e6, s["pipeline_hbm_bytes"] / a.batch / 1e6))
    return 0


if __name__ == "__main__":
    sys.exit(main())
