"""Per-size kernel-variant tuning of the specialised pipeline on the GPU at hand.

    python -m megadock_amd.tune [--sizes 64,80,...] [--rotations N] [--pair RES_R,RES_L]

For every FFT size with specialised kernels (megadock_amd.TUNED_NF + FURTHER_NF, or the
listed ones) it docks a synthetic receptor/ligand pair (megadock_amd.synth, residue counts
scaled to the grid) through a context of that size, runs mdk_autotune over the variants
built for the size (correlation x inverse, then the forward prefetch) and prints one JSON
line per size with the chosen variants and the rotations/s of the defaults and of the
choice.  The choices are the defaults table of megadock_amd/csrc/mdk_pipe.hip
(kDefaults); the library keeps them for every context it makes.
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

from . import synth


def _pair(ng, tmp):
    """A receptor and a ligand that fit a grid of ng voxels of 1.2 A (the receptor about a
    third of the box, the ligand about a sixth): residue counts grow with ng^1.5."""
    r = max(20, int(0.09 * ng ** 1.5))
    l = max(10, r // 3)
    rp, lp = os.path.join(tmp, "r%d.pdb" % ng), os.path.join(tmp, "l%d.pdb" % ng)
    synth.write_pdb(rp, r, seed=700 + ng)
    synth.write_pdb(lp, l, seed=900 + ng)
    return rp, lp


def tune_size(M, nf, rotations, tmp, t=1):
    ng = nf // 2
    rp, lp = _pair(ng, tmp)
    rec, lig = M.Protein(rp), M.Protein(lp)
    gr, gi = M.receptor_grids(rec, ng)
    H = nf // 2 + 1
    per = 8.0 * (ng * ng * nf + ng * nf * nf + nf * nf * (H + 16) + nf * nf * t)
    cap = 256 if nf <= 128 else 128 if nf <= 192 else 192
    B = max(1, min(cap, int((36 if nf > 192 else 8) * 2 ** 30 / per)))
    angles = np.loadtxt(os.path.join(M.DATA, "rotations_3600.txt"), dtype=np.float64).astype(np.float32)
    rot = M.rotation_coefs(angles[:min(rotations, 3600)])
    ctx = M.Context(nf, ng, lig.na, t=t, max_batch=B)
    ctx.set_receptor(gr, gi)
    ctx.set_ligand(lig.xyz, lig.center, lig.radius, lig.charge, M.grid_coord(ng), 1.2)
    ctx.set_rotations(rot)
    n = len(rot)
    ctx.dock_range(0, n)
    ctx.sync()
    before = ctx.time_range(0, n, reps=2)
    choice = ctx.autotune(0, n)
    after = ctx.time_range(0, n, reps=2)
    ctx.close()
    return {"nf": nf, "batch": B, "rotations": n, "correlation": choice[0], "inverse": choice[1],
            "forward": choice[2], "defaults_rot_s": round(n / (before * 1e-3), 1),
            "tuned_rot_s": round(n / (after * 1e-3), 1)}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--sizes", default="", help="comma list of FFT sizes (default: every specialised size)")
    ap.add_argument("--rotations", type=int, default=2048, help="rotations docked per timing")
    a = ap.parse_args(argv)
    import megadock_amd as M
    sizes = [int(x) for x in a.sizes.split(",") if x] or sorted(M.TUNED_NF + M.FURTHER_NF)
    tmp = tempfile.mkdtemp(prefix="mdk_tune_")
    for nf in sizes:
        t0 = time.time()
        row = tune_size(M, nf, a.rotations, tmp)
        row["seconds"] = round(time.time() - t0, 1)
        print(json.dumps(row), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
