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
    a = ap.parse_args(argv)
    if a.cmd == "trace":
        print("| kernel | calls | avg (us) | total (ms) | % |\n|---|---:|---:|---:|---:|")
        for r in kernel_stats(a.dir):
            print("| %s | %d | %.1f | %.2f | %.2f |" % (r["kernel"], r["calls"], r["avg_us"], r["total_ms"], r["percent"]))
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
