"""Write tests/golden/voxel_digests.json from the reference's own CPU voxelizers.

Runs in the build container only (it needs oracle/_ref/libmdkref.so, compiled from the
reference sources by oracle/ref/Makefile).  For each input pair it records SHA-256
digests of the exact float32 bytes of the receptor grids (src/receptor.cpp) and of the
ligand grids (src/ligand.cpp) at a set of rotations, plus a few summary numbers.  The
rotated coordinates come from the reference's ZXZ rotation (decoygen.cpp rotateAtom,
the same float expression as src/docking.cpp:376-393) applied to the centred ligand.

Usage: python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

ROTS = [0, 1, 17, 1000, 2222, 3599]
CASES = [
    # name, receptor, ligand, ng, delta
    ("1gcq", "1gcq_r.pdb", "1gcq_l.pdb", 40, 1.0),
    ("1gcq_delta2", "1gcq_r.pdb", "1gcq_l.pdb", 40, 2.0),
    ("1gcq_small_grid", "1gcq_r.pdb", "1gcq_l.pdb", 24, 1.0),
    ("synth_r120_l40", "synth_r120.pdb", "synth_l40.pdb", 48, 1.0),
]


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, np.float32).tobytes()).hexdigest()


def main():
    from megadock_amd import synth
    for fn, n, seed in (("synth_r120.pdb", 120, 11), ("synth_l40.pdb", 40, 12)):
        p = os.path.join(HERE, fn)
        if not os.path.exists(p):
            synth.write_pdb(p, n, seed)
    ang = O.load_rotations(3600)
    out = {"rotations_set": 3600, "rotations": ROTS, "gw": 1.2, "cases": {}}
    for name, rf, lf, ng, delta in CASES:
        rp, lp = os.path.join(HERE, rf), os.path.join(HERE, lf)
        O.ref_load(rp, lp)
        rg = O.ref_receptor_voxel(ng, 1.2)
        refl = O.ref_protein(1)
        lig_xyz = refl["xyz"] - refl["center"]
        case = {"receptor": rf, "ligand": lf, "ng": ng, "delta": delta,
                "receptor_grid_r": digest(rg[0]), "receptor_grid_i": digest(rg[1]),
                "receptor_core_voxels": int((rg[0] == -45.0).sum()),
                "receptor_grid_r_sum": float(rg[0].astype(np.float64).sum()),
                "ligand_grids": {}}
        for a in ROTS:
            xyz = O.ref_rotate(lig_xyz, ang[a])
            lg = O.ref_ligand_voxel(ng, 1.2, xyz, delta=delta)
            case["ligand_grids"][str(a)] = {"grid_r": digest(lg[0]), "grid_i": digest(lg[1]),
                                      "nonzero_r": int((lg[0] != 0).sum()),
                                      "nonzero_i": int((lg[1] != 0).sum())}
        out["cases"][name] = case
        print(name, ng, case["receptor_core_voxels"])
    with open(os.path.join(HERE, "voxel_digests.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
