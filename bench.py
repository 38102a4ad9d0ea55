#!/usr/bin/env python3
"""Rotations/s of the FFT docking rotation loop on one MI355X.

Workload (default c1): synthetic receptor of 600 residues and ligand of 250 residues,
144^3 FP32 complex grid (-F 144), the built-in 54000-rotation set, top 3 per rotation.
A step is one pass of the batched hot path over all 54000 rotations (voxelize, forward
FFT, spectral product, inverse FFT, per-rotation top-t), with the receptor spectrum,
ligand atoms and rotation table already resident in HBM.

Also reported on the same JSON line:
  plain      the same kernels with one rotation per call and a host round trip per
             rotation (the reference's GPU loop structure), on a sample of rotations
  roofline   bytes moved per rotation by the algorithmic yardstick B_rot = 32 nf^3
             (forward and inverse 3D FFT, one read and one write of the complex grid
             each) per second, against MI355X's 8 TB/s; per-stage times from HIP events
  cpu_baseline  the CPU oracle (a C restatement of the reference's OpenMP+FFTW loop,
             with its own FFT since FFTW is not installed) on a sample of rotations,
             timed on this host's cores; its results double as a parity spot check
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "rotations/s at 144³ (plain vs batched), %HBM peak, top-score rel err ≤1e-4"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


# Algorithmic HBM bytes per rotation, B_rot = 32 nf^3: the forward and the inverse 3D FFT
# each read and write the complex64 grid once (4 G, G = 8 nf^3).  Each of the six axis
# passes (z, y, x forward; x, y, z inverse) carries a sixth of it; voxelization, the
# spectral product (fused into the x passes) and the top-t reduction carry none.  A kernel
# that performs several passes carries their sum, so the shares of the kernels of one batch
# add up to B_rot.
STAGE_PASSES = {"voxelize": 0, "fwd_z": 1, "fwd_y": 1, "x_corr": 2, "inv_y": 1, "inv_z_top": 1, "reduce": 0}
STAGE_KERNELS = {"voxelize": "k_vox_lds / k_vox_atoms + k_vox_masks, k_vox_charge", "fwd_z": "kp_fwd_z",
                 "fwd_y": "kp_fwd_y", "x_corr": "kp_x_corr", "inv_y": "kp_inv_y", "inv_z_top": "kp_inv_z_top",
                 "reduce": "k_reduce_top"}
VALU_ISSUE_PEAK = 256 * 2.4e9  # wave64 VALU instructions per second: one per CU per cycle, 256 CUs at 2.4 GHz


def kernel_table(stage_ms, nfull, nf, batch):
    """Per-kernel launch durations (HIP events around each stage of every full batch of the
    timed passes) with their algorithmic bytes.  A stage whose kernel is fused into a
    neighbour has an empty event interval: its passes move to that neighbour (forward y
    into the fused forward z + y kernel, inverse y into the fused inverse y + z + top-t)."""
    b_rot = 32.0 * nf ** 3
    ms = {k: v / max(nfull, 1) for k, v in stage_ms.items()}
    passes = dict(STAGE_PASSES)
    names = dict(STAGE_KERNELS)
    empty = lambda k: ms[k] < 0.02 * max(ms.values())
    if empty("fwd_y"):
        passes["fwd_z"] += passes.pop("fwd_y")
        names["fwd_z"] = ("kp_fwd_half (forward z + y over half planes)" if nf == 256
                          else "kp_fwd_zy (forward z + y)")
    if empty("inv_y"):
        passes["inv_z_top"] += passes.pop("inv_y")
        names["inv_z_top"] = ("kp_inv_yz_half (inverse y + z + top-t over half planes)" if nf == 256
                              else "kp_inv_yz_top (inverse y + z + top-t)")
    rows = {}
    for k in passes:
        share = passes[k] / 6.0
        rows[k] = {"kernel": names[k], "ms_per_launch": round(ms[k], 4), "B_rot_share": round(share, 4),
                   "alg_bytes_per_launch": share * b_rot * batch,
                   "GBps": round(share * b_rot * batch / (ms[k] * 1e-3) / 1e9, 1) if share > 0 else None}
    return rows


def rocfft_reference(nf, batch, reps=5):
    """The vendor FFT (hipFFT on rocFFT, megadock_amd/lib/libmdkfftref.so): forward + inverse
    3D C2C of `batch` nf^3 complex64 grids, a reference point next to the whole rotation loop
    (which also voxelizes, multiplies, folds and reduces)."""
    import ctypes as C
    path = os.path.join(ROOT, "megadock_amd", "lib", "libmdkfftref.so")
    if not os.path.exists(path):
        return {"error": "libmdkfftref.so not built"}
    ms = C.c_float(0.0)
    rc = C.CDLL(path).mdkref_fft_time(0, nf, batch, reps, C.byref(ms))
    if rc:
        return {"error": "mdkref_fft_time returned %d" % rc}
    return {"what": "forward + inverse 3D C2C FFT of %d grids of %d^3 complex64 (hipFFT/rocFFT)" % (batch, nf),
            "ms": round(ms.value, 3), "grids_per_s": round(batch / (ms.value * 1e-3), 1),
            "GBps_at_B_rot": round(batch * 32.0 * nf ** 3 / (ms.value * 1e-3) / 1e9, 1)}


class _StdoutToStderr:
    """Sends what native code prints to stdout (the farm's per-job lines) to stderr, so
    that stdout carries the JSON line only."""

    def __enter__(self):
        sys.stdout.flush()
        self.saved = os.dup(1)
        os.dup2(2, 1)

    def __exit__(self, *exc):
        sys.stdout.flush()
        os.dup2(self.saved, 1)
        os.close(self.saved)


def bench_c2(args, M, workloads):
    """c2: a job table of receptor/ligand pairs of uneven sizes (residue counts log-uniform
    in [50, 800], auto grid, 3600 rotations, t = 1) through the in-process farm
    (megadock-hip-dp: host threads prepare jobs while the GPU docks), against the same jobs
    run one after another by the single-pair program."""
    import tempfile
    nr, nl = [int(x) for x in args.pairs.split("x")]
    pairs = [workloads.make_c2_pair(r, l, args.cache) for r in range(nr) for l in range(nl)]
    work = tempfile.mkdtemp(prefix="mdk_c2_")
    lines = ["%s\t%s\t%s" % (a, b, os.path.join(work, "farm_%03d.out" % k)) for k, (a, b) in enumerate(pairs)]
    table = os.path.join(work, "c2.table")
    with open(table, "w") as f:
        f.write("TITLE=c2 sample\nPARAM=-R $1 -L $2 -o $3\n" + "\n".join(lines) + "\n")
    flog = os.path.join(work, "master.log")
    warm = os.path.join(work, "warm.table")  # warm-up: runtime, code objects, file cache
    with open(warm, "w") as f:
        f.write("TITLE=c2 warm-up\nPARAM=-R $1 -L $2 -o $3\n" + "\n".join(lines[:8]) + "\n")
    with _StdoutToStderr():
        M.run_table(["-tb", warm, "-lg", os.path.join(work, "warm.log")])
        t0 = time.time()
        M.run_table(["-tb", table, "-lg", flog])
        farm_s = time.time() - t0
    rows = [l.split("\t") for l in open(flog).read().splitlines() if l[:5].isdigit() and "\t" in l]
    nfs = sorted(2 * int(r[3]) for r in rows)
    rot = 3600 * len(pairs)
    log("farm: %d pairs, %d rotations in %.2f s" % (len(pairs), rot, farm_s))
    seq = None
    if not args.no_sequential:
        t0 = time.time()
        with _StdoutToStderr():
            for k, (a, b) in enumerate(pairs):
                M.run(["-R", a, "-L", b, "-o", os.path.join(work, "single_%03d.out" % k)])
        seq_s = time.time() - t0
        same = all(open(os.path.join(work, "farm_%03d.out" % k), "rb").read()
                   == open(os.path.join(work, "single_%03d.out" % k), "rb").read() for k in range(len(pairs)))
        log("one after another: %.2f s; outputs identical: %s" % (seq_s, same))
        seq = {"seconds": round(seq_s, 3), "rotations_per_s": round(rot / seq_s, 1),
               "farm_speedup": round(seq_s / farm_s, 2), "outputs_identical": same}
    spec = set(M.TUNED_NF + M.FURTHER_NF)
    # the yardstick summed over the pairs: B_rot = 32 nf^3 bytes per rotation of each pair
    alg = sum(3600 * 32.0 * n ** 3 for n in nfs)
    roof = {"bound": "hbm", "achieved": round(alg / farm_s / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(alg / farm_s / 1e9 / HBM_PEAK_GBS, 4), "traffic": None,
            "unit_of_work": "the whole table: sum over its pairs of 3600 rotations x 32 nf^3 bytes, over the farm's "
                            "wall time (preparation, transfers, merges and .out files included)"}
    print(json.dumps({
        "metric": "c2 job table: rotations/s over receptor/ligand pairs of uneven FFT size",
        "value": round(rot / farm_s, 1), "unit": "rotations/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "c2: first %d receptors x %d ligands of the 32 x 32 table, auto grid, 3600 rotations, t = 1"
                   % (nr, nl), "pairs": len(pairs), "nf": {str(n): nfs.count(n) for n in sorted(set(nfs))},
                   "pairs_on_specialised_kernels": sum(1 for n in nfs if n in spec)},
        "roofline": roof,
        "farm_seconds": round(farm_s, 3), "pairs_per_s": round(len(pairs) / farm_s, 2),
        "farm_log": farm_log_summary(open(flog).read()),
        "one_after_another": seq,
    }), flush=True)
    return 0


def farm_log_summary(text):
    """The farm log's last lines: the most prepared jobs waiting at once (and the limit), how
    often the GPU drained between jobs and the contexts made and freed."""
    out = {}
    m = re.search(r"at most (\d+) \(limit (\d+)\)", text)
    if m:
        out["prepared_waiting_max"], out["prepared_waiting_limit"] = int(m.group(1)), int(m.group(2))
    m = re.search(r"drained between jobs: (\d+) with no job prepared, (\d+) for a job of the running job's FFT size; "
                  r"contexts made (\d+), freed (\d+)", text)
    if m:
        out.update({"drained_no_job_prepared": int(m.group(1)), "drained_same_fft_size": int(m.group(2)),
                    "contexts_made": int(m.group(3)), "contexts_freed": int(m.group(4))})
    return out


def bench_c3(args, M, workloads):
    """c3: the best 2000 decoys of c1 refined off the grid on the hard-voxel score
    (megadock_amd.refine.refine_hard: translation along the exact gradient of the
    band-limited hard score, rotation along the trilinear companion's torque, a step kept
    only when the hard score rises).  value = pose steps per second of the refinement.
    --refine-of c4 --refine-top 10000 gives c4's refinement of its best 10000 decoys."""
    from megadock_amd import refine

    rec_pdb, lig_pdb, cfg = workloads.make_pair(args.refine_of, args.cache)
    rec, lig = M.Protein(rec_pdb), M.Protein(lig_pdb)
    nf = cfg["nf"]
    ng, R, t = nf // 2, cfg["rotations"], cfg["t"]
    gr, gi = M.receptor_grids(rec, ng)
    per = 8.0 * (ng * ng * nf + ng * nf * nf + nf * nf * (nf // 2 + 17) + nf * nf * t)
    ctx = M.Context(nf, ng, lig.na, t=t, max_batch=max(1, min(128, int(8 * 2 ** 30 / per))))
    ctx.set_receptor(gr, gi)
    ctx.set_ligand(lig.xyz, lig.center, lig.radius, lig.charge, M.grid_coord(ng), 1.2)
    angles = np.loadtxt(os.path.join(M.DATA, "rotations_%d.txt" % R), dtype=np.float64).astype(np.float32)
    rot = M.rotation_coefs(angles)
    ctx.set_rotations(rot)
    td = time.perf_counter()
    ctx.dock_range(0, R)
    ctx.sync()
    dock_s = time.perf_counter() - td
    top = ctx.topn(R, args.refine_top)
    n = len(top)
    rc = rot[top["index"][:, 0]]
    t0 = top["index"][:, 1:].astype(np.float32)
    log("c3: docked %d rotations in %.2f s; refining the best %d decoys, %d steps" % (R, dock_s, n, args.refine_steps))

    class Counted:
        """The context, counting the poses its two off-grid scores evaluate."""

        def __init__(self, c):
            self.c, self.spectral, self.companion = c, 0, 0

        def score_spectral(self, r, tt):
            self.spectral += len(r)
            return self.c.score_spectral(r, tt)

        def score_companion(self, r, tt):
            self.companion += len(r)
            return self.c.score_companion(r, tt)

    ctx.score_spectral(rc[:8], t0[:8])  # first calls load the kernels
    ctx.score_companion(rc[:8], t0[:8])
    dock = Counted(ctx)
    tr = time.perf_counter()
    _, _, hist = refine.refine_hard(dock, rc, t0, steps=args.refine_steps)
    ref_s = time.perf_counter() - tr
    ctx.close()
    h0, h1 = hist[0], hist[-1]
    scale = np.maximum(np.abs(top["score"].astype(np.float64)), 1e-3)
    gain = (h1 - h0) / np.maximum(np.abs(h0), 1e-3)
    print(json.dumps({
        "metric": "c3: off-grid refinement of the best decoys, pose steps/s",
        "value": round(n * args.refine_steps / ref_s, 1), "unit": "pose steps/s", "n_gpus": 1,
        "higher_is_better": True, "dtype": "f32 (scores), f64 (pose algebra on the host)", "data": "synthetic",
        "config": {"workload": "c3: best %d decoys of %s (%d^3, %d rotations, t = %d), %d refinement steps"
                               % (n, args.refine_of, nf, R, t, args.refine_steps), "decoys": n,
                   "steps": args.refine_steps},
        "refine_seconds": round(ref_s, 3), "dock_seconds": round(dock_s, 3),
        "pose_evaluations": {"spectral": dock.spectral, "companion": dock.companion,
                             "per_second": round((dock.spectral + dock.companion) / ref_s, 1)},
        "hard_score": {
            # at the docked grid translation the band-limited score is the docked score
            "start_vs_docked_max_rel": float(np.max(np.abs(h0 - top["score"]) / scale)),
            "never_dropped": bool(np.all(np.diff(hist, axis=0) >= 0)),
            "improved_fraction": round(float(np.mean(h1 > h0)), 4),
            "mean_rel_gain": round(float(np.mean(gain)), 5), "max_rel_gain": round(float(np.max(gain)), 5)},
    }), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c1", choices=["c0", "c1", "c2", "c3", "c4"])
    ap.add_argument("--refine-steps", type=int, default=10, help="c3: refinement steps")
    ap.add_argument("--refine-of", default="c1", choices=["c1", "c4"], help="c3: the docking run whose decoys are refined")
    ap.add_argument("--refine-top", type=int, default=2000, help="c3: number of decoys refined")
    ap.add_argument("--no-sequential", action="store_true", help="c2: skip the one-after-another comparison")
    ap.add_argument("--pairs", default="4x8", help="c2: receptors x ligands of the 32 x 32 table to run (first ones)")
    ap.add_argument("--batch", type=int, default=0, help="rotations per device batch (0: automatic)")
    ap.add_argument("--plain-rotations", type=int, default=200)
    ap.add_argument("--cpu-rotations", type=int, default=0, help="CPU baseline sample size (0: automatic)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-rocfft", action="store_true", help="skip the vendor FFT (hipFFT) reference point")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default kernel variants")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first copy 1 GiB device to device (calibrates the HBM counters of a profiled run)")
    ap.add_argument("--cache", default=os.path.join(os.environ.get("TMPDIR", "/tmp"), "mdk_workloads"))
    ap.add_argument("--sweep", default="", help="tuning: comma list of correlation-pass variants (v), fused-inverse variants (i), "
                    "stream counts (s), graph replay on/off (g), forward-pass variants (f) and batch sizes (b), e.g. 'v0,v3,i1,s1,b64'; settings accumulate; prints "
                    "one stage-timing line per setting and exits")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the rotation loop runs on a single MI355X here; rotations "
              "would shard as contiguous blocks per GPU with one merge of the per-GPU top-N, which "
              "this build does not run", file=sys.stderr)
        sys.exit(2)

    import megadock_amd as M
    from megadock_amd import workloads

    if args.copy_probe:
        cms = M.copy_probe(1 << 30, reps=2)
        log("copy probe: 1 GiB in %.3f ms" % cms)

    if args.workload == "c2":
        return bench_c2(args, M, workloads)
    if args.workload == "c3":
        return bench_c3(args, M, workloads)

    t_setup = time.time()
    rec_pdb, lig_pdb, cfg = workloads.make_pair(args.workload, args.cache)
    rec, lig = M.Protein(rec_pdb), M.Protein(lig_pdb)
    ng = cfg["nf"] // 2 if cfg["nf"] else M.autogrid(rec, lig)
    nf = 2 * ng
    R, t = cfg["rotations"], cfg["t"]
    log("workload %s: receptor %d atoms, ligand %d atoms, nf=%d, %d rotations, t=%d (setup %.1fs)"
        % (args.workload, rec.na, lig.na, nf, R, t, time.time() - t_setup))
    gr, gi = M.receptor_grids(rec, ng)
    B = args.batch
    if B <= 0:
        H = nf // 2 + 1
        per = 8.0 * (ng * ng * nf + ng * nf * nf + nf * nf * (H + 16) + nf * nf * t)
        cap = 256 if nf <= 128 else 128 if nf <= 192 else 192  # as the host layer's auto_batch
        B = max(1, min(cap, int((36 if nf > 192 else 8) * 2 ** 30 / per)))
    ctx = M.Context(nf, ng, lig.na, t=t, max_batch=B)
    ctx.set_receptor(gr, gi)
    ctx.set_ligand(lig.xyz, lig.center, lig.radius, lig.charge, M.grid_coord(ng), 1.2)
    angles = np.loadtxt(os.path.join(M.DATA, "rotations_%d.txt" % R), dtype=np.float64).astype(np.float32)
    rot = M.rotation_coefs(angles)
    ctx.set_rotations(rot)
    log("context ready: batch %d" % B)
    tuned = None
    if not args.no_autotune and not args.sweep:
        tuned = ctx.autotune(0, min(R, 16 * B))
        log("autotune: correlation variant %d, inverse variant %d, forward variant %d" % tuned)

    if args.sweep:
        nbytes = 1 << 30
        cms = M.copy_probe(nbytes, reps=5)
        print(json.dumps({"copy_probe_bytes": nbytes, "ms": round(cms, 4),
                          "GBps_read_plus_write": round(2 * nbytes / (cms * 1e-3) / 1e9, 1)}), flush=True)
        for item in args.sweep.split(","):
            if item[0] == "v":
                os.environ["MDK_XCORR_VARIANT"] = item[1:]
            if item[0] == "i":
                os.environ["MDK_INV_VARIANT"] = item[1:]
            if item[0] == "s":
                os.environ["MDK_STREAMS"] = item[1:]
            if item[0] == "g":
                os.environ["MDK_GRAPH"] = item[1:]
            if item[0] == "f":
                os.environ["MDK_FWD_VARIANT"] = item[1:]
            bs = int(item[1:]) if item[0] == "b" else B
            sctx = M.Context(nf, ng, lig.na, t=t, max_batch=bs)
            sctx.set_receptor(gr, gi)
            sctx.set_ligand(lig.xyz, lig.center, lig.radius, lig.charge, M.grid_coord(ng), 1.2)
            sctx.set_rotations(rot[:max(4 * bs, 2048)])
            sctx.dock_range(0, 2 * bs)
            sctx.sync()
            ms = sctx.time_range(0, 16 * bs if 16 * bs <= max(4 * bs, 2048) else max(4 * bs, 2048), reps=2)
            n = 16 * bs if 16 * bs <= max(4 * bs, 2048) else max(4 * bs, 2048)
            st = sctx.profile_stages(0, bs)
            print(json.dumps({"setting": item, "batch": bs, "rotations_per_s": round(n / (ms * 1e-3), 1),
                              "stages_ms": {k: round(v, 4) for k, v in st.items()}}), flush=True)
            sctx.close()
        return
    n_out = min(cfg["n_out"], R * t)
    for w in range(args.warmup):
        ctx.dock_range(0, R)
        ctx.topn(R, n_out)
        log("warmup pass %d done" % (w + 1))
    ctx.sync()
    # A step: every rotation docked (batches on the context's streams, HIP events around each
    # stage of each batch), then the merge of the best n_out decoys on the device.
    stage_ms = {k: 0.0 for k in M.STAGES}
    nfull = 0
    t0 = time.perf_counter()
    for s in range(args.steps):
        st, nb = ctx.dock_range_timed(0, R)
        top = ctx.topn(R, n_out)
        for k in st:
            stage_ms[k] += st[k]
        nfull += nb
    wall = time.perf_counter() - t0
    ev_ms = ctx.time_range(0, R, reps=1)  # one pass without the stage events, timed by events on the stream
    log("timed %d steps: %.3f s wall; one pass without stage events %.1f ms" % (args.steps, wall, ev_ms))
    rot_s = args.steps * R / wall
    nb_rot = min(B, R)
    kernels = kernel_table(stage_ms, nfull, nf, nb_rot)
    dominant = max((k for k in kernels if kernels[k]["B_rot_share"] > 0), key=lambda k: kernels[k]["ms_per_launch"])
    dk = kernels[dominant]

    # the same kernels with one batch in flight (one stream): each kernel's duration is then
    # its own, not shared with the other stream's kernels (reported beside, not timed above)
    env = {"MDK_STREAMS": "1"}
    if tuned:  # the variants the timed passes ran
        env.update(MDK_XCORR_VARIANT=str(tuned[0]), MDK_INV_VARIANT=str(tuned[1]), MDK_FWD_VARIANT=str(tuned[2]))
    saved = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    octx = M.Context(nf, ng, lig.na, t=t, max_batch=B)
    for k, v in saved.items():
        if v is None:
            os.environ.pop(k)
        else:
            os.environ[k] = v
    octx.set_receptor(gr, gi)
    octx.set_ligand(lig.xyz, lig.center, lig.radius, lig.charge, M.grid_coord(ng), 1.2)
    octx.set_rotations(rot[:min(R, 16 * B)])
    octx.dock_range(0, min(R, 2 * B))
    octx.sync()
    ost, onb = octx.dock_range_timed(0, min(R, 16 * B))
    octx.close()
    one_stream = {k: v for k, v in kernel_table(ost, onb, nf, nb_rot).items()}

    # plain path: one rotation per call, host round trip each
    P = min(args.plain_rotations, R)
    ids = np.arange(P, dtype=np.int32) * (R // P)
    pctx = M.Context(nf, ng, lig.na, t=t, max_batch=1)
    pctx.set_receptor(gr, gi)
    pctx.set_ligand(lig.xyz, lig.center, lig.radius, lig.charge, M.grid_coord(ng), 1.2)
    pctx.dock(rot[ids[:1]], ids[:1])
    tp = time.perf_counter()
    plain = [pctx.dock(rot[ids[q:q + 1]], ids[q:q + 1]) for q in range(P)]
    plain_rot_s = P / (time.perf_counter() - tp)
    pctx.close()
    log("plain path: %.1f rotations/s over %d rotations" % (plain_rot_s, P))
    res_all = ctx.results(0, R)
    plain_same = all(np.array_equal(plain[q].view(np.uint8), res_all[ids[q] * t:(ids[q] + 1) * t].view(np.uint8))
                     for q in range(P))

    b_rot = 32.0 * nf ** 3
    achieved = dk["alg_bytes_per_launch"] / (dk["ms_per_launch"] * 1e-3) / 1e9
    # measured HBM bytes and VALU instructions per launch (rocprofv3 --pmc runs of this
    # command, summarised by `python -m megadock_amd.profiling`, committed under profiles/)
    prof = None
    pfile = os.path.join(ROOT, "profiles", "pmc_%s_nf%d.json" % (args.workload, nf))
    if os.path.exists(pfile):
        with open(pfile) as fh:
            prof = json.load(fh)

    def pmc(stage, key):
        if not prof:
            return None
        hit = [v for k, v in prof["kernels"].items() if k.startswith(kernels[stage]["kernel"].split(" ")[0] + "<")
               or k == kernels[stage]["kernel"].split(" ")[0]]
        return sum(h[key] for h in hit) / prof["batch"] * nb_rot if hit and all(key in h for h in hit) else None

    traffic = pmc(dominant, "hbm_bytes")
    valu = pmc(dominant, "SQ_INSTS_VALU")
    pipe_traffic = None
    if prof and "pipeline_hbm_bytes" in prof:
        pipe_traffic = prof["pipeline_hbm_bytes"] / prof["batch"]
    out = {
        "metric": METRIC,
        "value": round(rot_s, 1),
        "unit": "rotations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "%s: synthetic receptor %d res (%d atoms) / ligand %d res (%d atoms), %d^3 FP32 complex grid, "
                               "%d rotations, top %d per rotation, %d decoys"
                               % (args.workload, cfg["rec_res"], rec.na, cfg["lig_res"], lig.na, nf, R, t, cfg["n_out"]),
                   "nf": nf, "ng": ng, "rotations": R, "t": t, "decoys": n_out, "batch": B,
                   "step": "dock all rotations + device merge of the best decoys",
                   "variants": {"correlation": tuned[0], "inverse": tuned[1], "forward": tuned[2],
                                "chosen_by": "autotune"} if tuned
                   else "defaults"},
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": round(traffic) if traffic else None,
            "kernel": dk["kernel"], "launches": nfull, "rotations_per_launch": nb_rot,
            "avg_launch_ms": dk["ms_per_launch"],
            "alg_bytes_per_launch": dk["alg_bytes_per_launch"], "B_rot_share": dk["B_rot_share"],
            "measured_hbm_frac": round(traffic / (dk["ms_per_launch"] * 1e-3) / (HBM_PEAK_GBS * 1e9), 4) if traffic else None,
            "limiter": ({"what": "VALU instruction issue", "valu_instructions_per_launch": round(valu),
                         "valu_issue_frac": round(valu / (dk["ms_per_launch"] * 1e-3) / VALU_ISSUE_PEAK, 4)}
                        if valu else None),
            "counters_source": os.path.relpath(pfile, ROOT) if prof else None,
            "pipeline": {"B_rot": b_rot, "achieved": round(rot_s * b_rot / 1e9, 1),
                         "frac": round(rot_s * b_rot / 1e9 / HBM_PEAK_GBS, 4),
                         "measured_hbm_bytes_per_rotation": round(pipe_traffic) if pipe_traffic else None,
                         "measured_hbm_frac": round(rot_s * pipe_traffic / 1e9 / HBM_PEAK_GBS, 4) if pipe_traffic else None},
            "kernels": kernels,
            "kernels_one_stream": {"note": "%d full batches on one stream, outside the timed region: each "
                                           "kernel's own duration" % onb, "kernels": one_stream}},
        "plain": {"value": round(plain_rot_s, 1), "unit": "rotations/s", "rotations": P,
                  "batched_over_plain": round(rot_s / plain_rot_s, 2), "bitwise_equal_to_batched": bool(plain_same)},
        "pass_ms_without_stage_events": round(ev_ms, 2),
        "best_score": float(top["score"][0]),
    }

    if not args.no_rocfft:
        out["rocfft_reference"] = rocfft_reference(nf, min(B, 32))

    if not args.no_cpu:
        from oracle import oracle as O
        # the host cores this job may use: its CPU affinity, capped by OMP_NUM_THREADS when the
        # machine sets it (a one-GPU share of a shared host)
        affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        threads = min(affinity, int(os.environ.get("OMP_NUM_THREADS", "0")) or affinity)
        model = None
        try:
            with open("/proc/cpuinfo") as fh:
                model = next((l.split(":", 1)[1].strip() for l in fh if l.startswith("model name")), None)
        except OSError:
            pass
        n_cpu = args.cpu_rotations or {"c0": 2048, "c1": 768, "c4": 48}[args.workload]
        orec, olig = O.Protein(rec_pdb), O.Protein(lig_pdb)
        ogr, ogi = O.receptor_voxel(ng, 1.2, orec)
        rr, ri = O.receptor_fft(ng, nf, ogr, ogi)
        cids = (np.arange(n_cpu, dtype=np.int32) * (R // n_cpu)).astype(np.int32)
        log("cpu baseline: %d rotations on %d threads" % (n_cpu, threads))
        tc = time.perf_counter()
        cres = O.dock(ng, nf, 1.2, olig, rr, ri, angles, cids, t, threads=threads)
        cpu_s = time.perf_counter() - tc
        g = res_all.reshape(R, t)[cids]
        c = cres.reshape(n_cpu, t)
        rel = np.abs(g["score"].astype(np.float64) - c["score"]) / np.maximum(np.abs(c["score"]), 1e-3)
        out["cpu_baseline"] = {"value": round(n_cpu / cpu_s, 2), "unit": "rotations/s", "cores": threads,
                               "kind": "port",
                               "sample": "%d of the %d rotations (every %d-th), full %d^3 loop per rotation (voxelize, "
                                         "forward FFT, product, inverse FFT, top-t) on OpenMP threads; the oracle's own "
                                         "mixed-radix FFT (FFTW not installed)" % (n_cpu, R, R // n_cpu, nf),
                               "seconds": round(cpu_s, 2),
                               "host": {"nproc": os.cpu_count(), "affinity_cpus": affinity,
                                        "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS"), "model": model}}
        # the global merge over the sample: the n best of its n_cpu * t results, as sets
        n_top = max(1, n_cpu * t // 4)

        def best(r):
            r = r.ravel()
            return {tuple(int(v) for v in r["index"][i])
                    for i in np.argsort(-r["score"].astype(np.float64), kind="stable")[:n_top]}

        out["parity_sample"] = {"rotations": n_cpu, "max_rel_err_top_scores": float(rel.max()),
                                "best_index_identical": float(np.mean(np.all(g["index"][:, 0] == c["index"][:, 0], axis=1))),
                                "top_n": n_top, "top_n_overlap": len(best(g) & best(c)) / n_top}
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
