"""Summaries of rocprofv3 runs of the docking pipeline (kernel times and HBM traffic).

    python -m megadock_amd.profiling trace  DIR            # --kernel-trace --stats, -f csv
    python -m megadock_amd.profiling traffic FETCH_DIR WRITE_DIR [--copy-fetch-factor 2.0]
                                             [--copy-write-factor 1.0] --batch B --json OUT

DIR, FETCH_DIR and WRITE_DIR are rocprofv3 output folders written with `-f csv -o run`.
The traffic summary keeps, for each kernel, the dispatches of the largest grid (full
batches), averages FETCH_SIZE and WRITE_SIZE (KiB per dispatch) over them and scales each
by the factor a plain device-to-device copy of known size needs to read its true bytes.
bench.py reads the JSON to report the measured HBM bytes next to the algorithmic ones.
"""
import argparse
import collections
import csv
import json
import os
import sys


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


def counter_per_kernel(pmc_dir, counter):
    vals = collections.defaultdict(list)
    for r in csv.DictReader(open(os.path.join(pmc_dir, "run_counter_collection.csv"))):
        if r["Counter_Name"] != counter:
            continue
        vals[short_name(r["Kernel_Name"])].append((int(r["Grid_Size"]), float(r["Counter_Value"])))
    out = {}
    for k, v in vals.items():
        g = max(x[0] for x in v)
        sel = [x[1] for x in v if x[0] == g]
        out[k] = {"grid": g, "kib": sum(sel) / len(sel), "dispatches": len(sel)}
    return out


def traffic(fetch_dir, write_dir, batch, ff, wf, min_dispatches=4):
    """Per-batch HBM bytes of the kernels dispatched at least min_dispatches times at their
    largest grid (the per-batch pipeline; set-up kernels such as the receptor transform
    run once per context)."""
    fetch = counter_per_kernel(fetch_dir, "FETCH_SIZE")
    write = counter_per_kernel(write_dir, "WRITE_SIZE")
    kernels = {}
    for k in sorted(set(fetch) | set(write)):
        n = max(fetch.get(k, {}).get("dispatches", 0), write.get(k, {}).get("dispatches", 0))
        if k.startswith("__amd_rocclr") or n < min_dispatches:
            continue
        fb = fetch.get(k, {}).get("kib", 0.0) * 1024 * ff
        wb = write.get(k, {}).get("kib", 0.0) * 1024 * wf
        kernels[k] = {"fetch_bytes": round(fb), "write_bytes": round(wb),
                      "dispatches": fetch.get(k, write.get(k, {})).get("dispatches", 0)}
    total = sum(v["fetch_bytes"] + v["write_bytes"] for v in kernels.values())
    return {"batch": batch, "copy_calibration": {"FETCH_SIZE": ff, "WRITE_SIZE": wf},
            "bytes_per_batch": total, "bytes_per_rotation": total / batch, "kernels": kernels}


def counters(dirs, batch, kernels_like="kp_|k_vox|k_reduce|k_clear"):
    """Counter values per full-batch dispatch for every counter found in the given pmc
    output folders, divided by the batch (per rotation)."""
    import re
    table = collections.defaultdict(dict)
    for d in dirs:
        names = set(r["Counter_Name"] for r in csv.DictReader(open(os.path.join(d, "run_counter_collection.csv"))))
        for c in names:
            for k, v in counter_per_kernel(d, c).items():
                if re.match(kernels_like, k) and v["dispatches"] >= 4:
                    table[k][c] = v["kib"] / batch  # the value per dispatch (named kib by the helper)
    return table


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("dir")
    f = sub.add_parser("traffic")
    f.add_argument("fetch_dir")
    f.add_argument("write_dir")
    f.add_argument("--batch", type=int, required=True)
    f.add_argument("--copy-fetch-factor", type=float, default=2.0)
    f.add_argument("--copy-write-factor", type=float, default=1.0)
    f.add_argument("--json", required=True)
    c = sub.add_parser("counters")
    c.add_argument("dirs", nargs="+")
    c.add_argument("--batch", type=int, required=True)
    a = ap.parse_args(argv)
    if a.cmd == "counters":
        tab = counters(a.dirs, a.batch)
        cols = sorted(set(x for v in tab.values() for x in v))
        print("Per rotation (counter value of a full-batch dispatch / batch):\n")
        print("| kernel | " + " | ".join(cols) + " |\n|---|" + "---:|" * len(cols))
        for k in sorted(tab, key=lambda k: -tab[k].get("SQ_INSTS_VALU", 0)):
            print("| %s | " % k + " | ".join("%.4g" % tab[k].get(x, float("nan")) for x in cols) + " |")
        return 0
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
    else:
        s = traffic(a.fetch_dir, a.write_dir, a.batch, a.copy_fetch_factor, a.copy_write_factor)
        with open(a.json, "w") as fh:
            json.dump(s, fh, indent=1)
        print("| kernel | fetch (MB/batch) | write (MB/batch) |\n|---|---:|---:|")
        for k, v in s["kernels"].items():
            print("| %s | %.1f | %.1f |" % (k, v["fetch_bytes"] / 1e6, v["write_bytes"] / 1e6))
        print("| total | %.1f MB per batch of %d | %.2f MB per rotation |"
              % (s["bytes_per_batch"] / 1e6, a.batch, s["bytes_per_rotation"] / 1e6))


if __name__ == "__main__":
    sys.exit(main())
