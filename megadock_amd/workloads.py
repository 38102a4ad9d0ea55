"""Named synthetic docking workloads (receptor/ligand PDB pairs plus run settings).

  c0  plain, small: 200 / 80 residues, auto grid (about 96^3), 3600 rotations, t=1, N=2000
  c1  plain, medium: 600 / 250 residues, -F 144 (144^3), 54000 rotations, t=3, N=2000
  c2  all-to-all: 32 receptors x 32 ligands, residue counts log-uniform over [50, 800]
      (drawn with seed 2999), auto grids, 3600 rotations per pair, t=1
  c4  large: 2000 / 1000 residues, -F 256 (256^3), 54000 rotations, t=3, N=2000

Proteins come from megadock_amd.synth with seed = 1000*config + 10*index + role
(role 0 receptor, 1 ligand).  For c0 the sphere radius is shrunk in fixed steps until the
receptor's largest box edge lies in [38, 41.6) A, which makes the automatic grid ng=48.
"""
import os

import numpy as np

from . import synth

CONFIGS = {
    "c0": dict(config=0, rec_res=200, lig_res=80, nf=None, rotations=3600, t=1, n_out=2000,
               extent=(38.0, 41.6)),
    "c1": dict(config=1, rec_res=600, lig_res=250, nf=144, rotations=54000, t=3, n_out=2000, extent=None),
    "c4": dict(config=4, rec_res=2000, lig_res=1000, nf=256, rotations=54000, t=3, n_out=2000, extent=None),
}


def _extent(atoms):
    xyz = np.array([a[4] for a in atoms])
    return float((xyz.max(0) - xyz.min(0)).max())


def make_pair(name, outdir):
    """Write <outdir>/<name>_r.pdb and _l.pdb (if absent); return (rec, lig, cfg)."""
    cfg = CONFIGS[name]
    os.makedirs(outdir, exist_ok=True)
    rec = os.path.join(outdir, "%s_r.pdb" % name)
    lig = os.path.join(outdir, "%s_l.pdb" % name)
    c = cfg["config"]
    if not os.path.exists(rec):
        n = cfg["rec_res"]
        base = (3.0 * n * 160.0 / (4.0 * np.pi)) ** (1.0 / 3.0)
        atoms = None
        for step in range(12):
            atoms = synth.build_protein(n, synth.workload_seed(c, 0, 0), radius=base * (1.0 - 0.03 * step))
            if cfg["extent"] is None:
                break
            lo, hi = cfg["extent"]
            if lo <= _extent(atoms) < hi:
                break
        with open(rec, "w") as f:
            f.write(synth.pdb_text(atoms))
    if not os.path.exists(lig):
        atoms = synth.build_protein(cfg["lig_res"], synth.workload_seed(c, 0, 1))
        with open(lig, "w") as f:
            f.write(synth.pdb_text(atoms))
    return rec, lig, cfg


C2_PROTEINS = 32


def c2_sizes():
    """Residue counts of the c2 receptors (first 32) and ligands (next 32)."""
    rng = np.random.default_rng(2999)
    n = np.exp(rng.uniform(np.log(50.0), np.log(800.0), 2 * C2_PROTEINS))
    return [int(round(x)) for x in n[:C2_PROTEINS]], [int(round(x)) for x in n[C2_PROTEINS:]]


def make_c2_pair(receptor, ligand, outdir):
    """Write the c2 receptor `receptor` and ligand `ligand` (0..31) as PDB files (if absent);
    return their paths.  Settings: auto grid, 3600 rotations, t=1."""
    os.makedirs(outdir, exist_ok=True)
    rs, ls = c2_sizes()
    rec = os.path.join(outdir, "c2_r%02d.pdb" % receptor)
    lig = os.path.join(outdir, "c2_l%02d.pdb" % ligand)
    if not os.path.exists(rec):
        with open(rec, "w") as f:
            f.write(synth.pdb_text(synth.build_protein(rs[receptor], synth.workload_seed(2, receptor, 0))))
    if not os.path.exists(lig):
        with open(lig, "w") as f:
            f.write(synth.pdb_text(synth.build_protein(ls[ligand], synth.workload_seed(2, ligand, 1))))
    return rec, lig
