"""GPU: receptor grids built on the device (mdk_set_receptor_atoms, read back with
mdk_get_receptor_grids) are bit-exact with the oracle's restatement of the reference's
Receptor::rpscace/electro (oracle/mdk_oracle.c, itself held bit for bit to the reference's
own code compiled in oracle/_ref by tests/test_oracle_voxel.py and
tests/test_host_parity.py) and with the host producer (mdkh_receptor_voxel), for several
receptors, the c0, c1 and c4 workload receptors, grids and the -a/-d/-e/-v parameters,
including cavities closed by the order-dependent propagation sweeps; the spectrum the
device grids give equals mdk_set_receptor's."""
import os

import numpy as np
import pytest

import megadock_amd as M
from megadock_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _check(prot, ng, gw=1.2, aceratio=1.0, rec_core=-45.0, eratio=1.0, lig=None):
    gr, gi = M.receptor_grids(prot, ng, gw, aceratio, rec_core, eratio)
    na = (lig or prot).na
    ctx = M.Context(2 * ng, ng, na, t=1, max_batch=2)
    xyz = (prot.xyz - prot.center[None, :]).astype(np.float32)  # Protein::shift_center
    ctx.set_receptor_atoms(xyz, prot.radius, prot.charge, prot.ace, M.grid_coord(ng, gw), gw, aceratio, rec_core,
                           eratio)
    dr, di = ctx.receptor_grids()
    orr, ori = O.receptor_voxel(ng, gw, O.Protein(prot.path), aceratio, rec_core, eratio)
    assert np.array_equal(orr.view(np.uint32), dr.view(np.uint32)), "device vs oracle (shape)"
    assert np.array_equal(ori.view(np.uint32), di.view(np.uint32)), "device vs oracle (electrostatics)"
    bad_r = int(np.count_nonzero(dr.view(np.uint32) != gr.view(np.uint32)))
    bad_i = int(np.count_nonzero(di.view(np.uint32) != gi.view(np.uint32)))
    print("ng=%d gw=%g ace=%g core=%g e=%g: %d / %d differing voxels; core %d, cavity-free open %d"
          % (ng, gw, aceratio, rec_core, eratio, bad_r, bad_i, int((gr == rec_core).sum()), int((gr == 0).sum())),
          flush=True)
    assert bad_r == 0 and bad_i == 0
    return ctx, gr, gi


@pytest.mark.parametrize("which,ng,kw", [
    ("1gcq_r", 48, {}), ("1gcq_r", 64, {"aceratio": 0.0}), ("1gcq_l", 40, {"eratio": 0.5, "rec_core": 3.0}),
    ("synth_r120", 72, {"aceratio": 2.0}), ("1gcq_r", 54, {"gw": 1.0}), ("1gcq_r", 20, {}),
])
def test_device_receptor_grids_are_bit_exact(golden_dir, which, ng, kw):
    prot = M.Protein(os.path.join(golden_dir, which + ".pdb"))
    ctx, gr, gi = _check(prot, ng, **kw)
    ctx.close()


@pytest.mark.parametrize("name,ng", [("c0", 48), ("c4", 128)])
def test_device_receptor_grids_workloads(tmp_path, name, ng):
    rec_pdb, lig_pdb, cfg = W.make_pair(name, str(tmp_path))
    ctx, gr, gi = _check(M.Protein(rec_pdb), ng)
    ctx.close()


def test_device_receptor_grids_c1_and_spectrum(tmp_path):
    rec_pdb, lig_pdb, cfg = W.make_pair("c1", str(tmp_path))
    rec = M.Protein(rec_pdb)
    ctx, gr, gi = _check(rec, 72)
    re1, im1 = ctx.receptor_spectrum()
    ctx.set_receptor(gr, gi)
    re2, im2 = ctx.receptor_spectrum()
    assert np.array_equal(re1, re2) and np.array_equal(im1, im2)
    ctx.close()


def _shell_pdb(path, radius=11.0, spacing=1.7, hole=(1.0, 0.0, 0.0), hole_r=5.0):
    """A hollow spherical shell of atoms with one opening: open space reaches the inside
    through the opening only, so how far it spreads depends on the order of the two
    propagation sweeps."""
    n = int(4 * np.pi * radius ** 2 / spacing ** 2)
    g = np.pi * (3 - np.sqrt(5))
    pts = []
    for m in range(n):
        y = 1 - 2 * (m + 0.5) / n
        r = np.sqrt(1 - y * y)
        p = radius * np.array([np.cos(g * m) * r, y, np.sin(g * m) * r])
        if np.linalg.norm(p - radius * np.asarray(hole)) > hole_r:
            pts.append(p)
    with open(path, "w") as f:
        for a, p in enumerate(pts):
            f.write("ATOM  %5d  CA  ALA A%4d    %8.3f%8.3f%8.3f  1.00  0.00           C\n"
                    % (a + 1, a + 1, p[0], p[1], p[2]))
        f.write("END\n")


@pytest.mark.parametrize("hole", [(1.0, 0.0, 0.0), (-1.0, 0.0, 0.0), (0.0, 0.0, -1.0)])
def test_device_receptor_grids_hollow_shell(tmp_path, hole):
    path = str(tmp_path / "shell.pdb")
    _shell_pdb(path, hole=hole)
    prot = M.Protein(path)
    ctx, gr, gi = _check(prot, 32)
    ctx.close()
