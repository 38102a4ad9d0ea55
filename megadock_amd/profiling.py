"""Summaries of rocprofv3 runs of the docking pipeline (kernel times, HBM traffic and SQ
counters).

    python -m megadock_amd.profiling trace DIR           # --kernel-trace --stats, -f csv
    python -m megadock_amd.profiling pmc --batch B --json OUT DIR [DIR ...]
    python -m megadock_amd.profiling pmc-md OUT           # the tables of a written summary

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


def busy(trace_dir, gap_ms=0.2, split_ms=50.0):
    """GPU-busy share of a kernel trace: the union of the dispatches' [start, end] intervals
    over the span from the first docking kernel to the last one (kernels named kp_ or k_),
    and the idle gaps longer than gap_ms.  Gaps longer than split_ms (a process ending and
    the next starting, e.g. a warm-up run before the timed one) split the trace into
    segments; the longest segment is reported apart ("main"), with its gaps by length and
    the kernels on either side of its longest gaps."""
    iv = []
    for r in csv.DictReader(open(os.path.join(trace_dir, "run_kernel_trace.csv"))):
        k = short_name(r["Kernel_Name"])
        if PIPELINE.match(k) or k.startswith("k_"):
            iv.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), k.split("<")[0]))
    if not iv:
        return None
    iv.sort()

    def scan(seg):
        covered, idle, gl = 0, 0, []
        cs, ce, last = seg[0][0], seg[0][1], seg[0][2]
        for s, e, k in seg[1:]:
            if s > ce:
                covered += ce - cs
                idle += s - ce
                gl.append((s - ce, last, k))
                cs, ce = s, e
            else:
                ce = max(ce, e)
            if e >= ce:
                last = k
        covered += ce - cs
        return covered, idle, ce - seg[0][0], gl

    covered, idle, span, gl = scan(iv)
    out = {"span_ms": span / 1e6, "busy_ms": covered / 1e6, "busy_frac": covered / span, "idle_ms": idle / 1e6,
           "gaps_over_%gms" % gap_ms: sum(g > gap_ms * 1e6 for g, _, _ in gl), "dispatches": len(iv)}
    # segments split at the long gaps; the longest is the main run
    segs, cur, end = [], [iv[0]], iv[0][1]
    for x in iv[1:]:
        if x[0] - end > split_ms * 1e6:
            segs.append(cur)
            cur = []
        cur.append(x)
        end = max(end, x[1])
    segs.append(cur)
    main = max(segs, key=lambda sg: max(e for _, e, _ in sg) - sg[0][0])
    mc, mi, ms, mg = scan(main)
    edges = [0.2, 0.5, 1, 2, 5, split_ms]
    hist = {}
    for lo, hi in zip(edges, edges[1:]):
        sel = [g for g, _, _ in mg if lo * 1e6 < g <= hi * 1e6]
        hist["%g-%g ms" % (lo, hi)] = {"gaps": len(sel), "ms": round(sum(sel) / 1e6, 2)}
    pairs = collections.defaultdict(lambda: [0, 0])
    for g, a, b in mg:
        if g > gap_ms * 1e6:
            pairs[a + " -> " + b][0] += 1
            pairs[a + " -> " + b][1] += g
    top = sorted(pairs.items(), key=lambda kv: -kv[1][1])[:8]
    out["segments"] = len(segs)
    out["main"] = {"span_ms": ms / 1e6, "busy_ms": mc / 1e6, "busy_frac": mc / ms, "idle_ms": mi / 1e6,
                   "dispatches": len(main), "gaps_by_length": hist,
                   "around_gaps": {k: {"gaps": v[0], "ms": round(v[1] / 1e6, 2)} for k, v in top}}
    return out


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
    out = {"batch": batch, "copy_calibration": {"bytes": copy_bytes, "factors": factors},
           "kernels": timed_pass_kernels(kernels)}
    finish_pipeline(out)
    return out


def timed_pass_kernels(kernels):
    """The kernels of the timed passes: those dispatched at least half as often as the most
    dispatched one.  The autotuner's trial batches run other forms of a stage (for example
    the two-pass inverse next to the fused one) a few times; they do not count."""
    if not kernels:
        return kernels
    top = max(v["dispatches"] for v in kernels.values())
    return {k: v for k, v in kernels.items() if v["dispatches"] * 2 >= top}


def finish_pipeline(out):
    ks = out["kernels"]
    out.pop("pipeline_hbm_bytes", None)
    if ks and all("hbm_bytes" in v for v in ks.values()):
        out["pipeline_hbm_bytes"] = sum(v["hbm_bytes"] for v in ks.values())


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("dir")
    t.add_argument("--busy", action="store_true", help="also the GPU-busy share of the docking span")
    p = sub.add_parser("pmc")
    p.add_argument("dirs", nargs="+")
    p.add_argument("--batch", type=int, required=True)
    p.add_argument("--copy-bytes", type=int, default=1 << 30)
    p.add_argument("--json", required=True)
    m = sub.add_parser("pmc-md", help="the tables of a summary JSON written by pmc")
    m.add_argument("json")
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
        if a.busy:
            b = busy(a.dir)
            print("\nGPU busy over the docking span (union of the kernels' intervals): %s" % json.dumps(b))
        return 0
    if a.cmd == "pmc-md":
        print_pmc(json.load(open(a.json)))
        return 0
    s = pmc_summary(a.dirs, a.batch, a.copy_bytes)
    with open(a.json, "w") as fh:
        json.dump(s, fh, indent=1)
    print_pmc(s)
    return 0


def print_pmc(s):
    """The markdown tables of a counter summary (pmc_summary's dict)."""
    batch = s["batch"]
    f = s["copy_calibration"]["factors"]
    print("Copy calibration (%d bytes copied): FETCH_SIZE x %s, WRITE_SIZE x %s\n"
          % (s["copy_calibration"]["bytes"], f["FETCH_SIZE"], f["WRITE_SIZE"]))
    cols = sorted({c for v in s["kernels"].values() for c in v if c.isupper() or c.startswith("SQ_")})
    print("Per full-batch dispatch (%d rotations), mean over the dispatches of the largest grid:\n" % batch)
    print("| kernel | dispatches | HBM MB | " + " | ".join(cols) + " |\n|---|---:|---:|" + "---:|" * len(cols))
    for k, v in sorted(s["kernels"].items(), key=lambda kv: -kv[1].get("hbm_bytes", 0)):
        hb = "%.1f" % (v["hbm_bytes"] / 1e6) if "hbm_bytes" in v else "-"
        print("| %s | %d | %s | " % (k, v["dispatches"], hb) + " | ".join("%.4g" % v[c] if c in v else "-" for c in cols) + " |")
    if "pipeline_hbm_bytes" in s:
        print("\nPipeline: %.1f MB per batch, %.2f MB per rotation"
              % (s["pipeline_hbm_bytes"] / 1e6, s["pipeline_hbm_bytes"] / batch / 1e6))


if __name__ == "__main__":
    sys.exit(main())
