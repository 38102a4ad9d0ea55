"""GPU parity on the synthetic workloads other than the benchmark's: c0 (small, auto
grid, t=1), c2 (pairs of uneven size and grid), c4 (256^3, t=3) on samples of their
rotations against the CPU oracle, and c3 (refinement of the best c1 decoys: of a sample,
and the best 2000 of the full 54000-rotation run)."""
import numpy as np
import pytest

import megadock_amd as M
from megadock_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def wl(tmp_path_factory):
    return str(tmp_path_factory.mktemp("workloads"))


def _check_pair(rec_pdb, lig_pdb, nf, t, nrot_set, ids):
    rec, lig = O.Protein(rec_pdb), O.Protein(lig_pdb)
    prec, plig = M.Protein(rec_pdb), M.Protein(lig_pdb)
    if nf is None:
        ng = O.autogrid(rec, lig)
        assert M.autogrid(prec, plig) == ng  # grid sizing: product host layer vs oracle
    else:
        ng = nf // 2
    nf = 2 * ng
    gr, gi = O.receptor_voxel(ng, 1.2, rec)
    rr, ri = O.receptor_fft(ng, nf, gr, gi)
    ang = O.load_rotations(nrot_set)
    ids = np.asarray(ids, np.int32)
    _say("pair %s: ng=%d, %d rotations, t=%d" % (rec_pdb.split("/")[-1], ng, len(ids), t))
    ctx = M.Context(nf, ng, lig.na, t=t, max_batch=len(ids))
    ctx.set_receptor(*M.receptor_grids(prec, ng))
    ctx.set_ligand(lig.xyz, lig.center, lig.radius, lig.charge, O.grid_coord(ng, 1.2), 1.2)
    got = ctx.dock(M.rotation_coefs(ang[ids]), ids)
    ctx.close()
    ref = O.dock(ng, nf, 1.2, lig, rr, ri, ang, ids, t, threads=16)
    remb = O.embed(ng, nf, gr, gi)
    for q in range(len(ids) * t):
        g, r = got[q], ref[q]
        assert g["index"][0] == r["index"][0]
        assert abs(g["score"] - r["score"]) <= 1e-4 * max(abs(r["score"]), 1e-3), (q, g, r)
        if not np.array_equal(g["index"], r["index"]):  # only at a near-tie of the FP64 map
            a = int(ids[q // t])
            xyz = O.rotate_ligand(lig.xyz, lig.center, O.rotation_coef(ang[a]))
            lr, li = O.ligand_voxel(ng, 1.2, xyz, lig.radius, lig.charge)
            s64 = O.correlation_fp64(nf, remb, O.embed(ng, nf, lr, li)).ravel()
            ig = (g["index"][1] * nf + g["index"][2]) * nf + g["index"][3]
            ir = (r["index"][1] * nf + r["index"][2]) * nf + r["index"][3]
            assert abs(s64[ig] - s64[ir]) <= 2e-6 * np.abs(s64).max(), (q, g, r)
    return ng


def test_c0_small_auto_grid(wl):
    rec, lig, cfg = W.make_pair("c0", wl)
    ng = _check_pair(rec, lig, None, cfg["t"], cfg["rotations"], np.arange(0, 3600, 300))
    assert ng == 48  # the c0 receptor extent is chosen to give the 96^3 grid


@pytest.mark.parametrize("r,l", [(1, 0), (4, 5), (3, 7)])
def test_c2_pairs_of_uneven_grids(wl, r, l):
    rec, lig = W.make_c2_pair(r, l, wl)
    _check_pair(rec, lig, None, 1, 3600, [11, 1500, 3333])


def test_c4_large_grid(wl):
    rec, lig, cfg = W.make_pair("c4", wl)
    _check_pair(rec, lig, cfg["nf"], cfg["t"], cfg["rotations"], [7, 40000])


def test_c3_refines_the_best_c1_decoys(wl):
    """The best decoys of a c1 sample refined on the hard score, which never drops."""
    from megadock_amd import refine as F
    rec_pdb, lig_pdb, cfg = W.make_pair("c1", wl)
    prec, plig = M.Protein(rec_pdb), M.Protein(lig_pdb)
    ng = cfg["nf"] // 2
    ctx = M.Context(cfg["nf"], ng, plig.na, t=cfg["t"], max_batch=64)
    ctx.set_receptor(*M.receptor_grids(prec, ng))
    ctx.set_ligand(plig.xyz, plig.center, plig.radius, plig.charge, M.grid_coord(ng), 1.2)
    ang = O.load_rotations(54000)[::108]  # 500 of the 54000
    rc = M.rotation_coefs(ang)
    ctx.set_rotations(rc)
    ctx.dock_range(0, len(rc))
    ctx.sync()
    top = ctx.topn(len(rc), 16)
    rot = rc[top["index"][:, 0]]
    R, t, hist = F.refine_hard(ctx, rot, top["index"][:, 1:].astype(np.float32), steps=4)
    _say("c3 hard scores %s -> %s" % (hist[0][:4], hist[-1][:4]))
    assert np.allclose(hist[0], top["score"], rtol=1e-5, atol=1e-6)
    assert np.all(np.diff(hist, axis=0) >= 0)
    ctx.close()


def test_c3_refines_the_best_2000_of_the_full_c1_run(wl):
    """The config as stated: all 54000 rotations of c1 docked (t = 3), the best 2000 decoys
    refined on the hard score; no decoy's hard score drops, every starting hard score is the
    docked score, and the refinement raises some of them."""
    from megadock_amd import refine as F
    rec_pdb, lig_pdb, cfg = W.make_pair("c1", wl)
    prec, plig = M.Protein(rec_pdb), M.Protein(lig_pdb)
    ng = cfg["nf"] // 2
    ctx = M.Context(cfg["nf"], ng, plig.na, t=cfg["t"], max_batch=128)
    ctx.set_receptor(*M.receptor_grids(prec, ng))
    ctx.set_ligand(plig.xyz, plig.center, plig.radius, plig.charge, M.grid_coord(ng), 1.2)
    rc = M.rotation_coefs(O.load_rotations(54000))
    ctx.set_rotations(rc)
    _say("c3: docking all 54000 c1 rotations")
    ctx.dock_range(0, len(rc))
    ctx.sync()
    top = ctx.topn(len(rc), 2000)
    assert len(top) == 2000 and np.all(np.diff(top["score"]) <= 0)
    rot = rc[top["index"][:, 0]]
    _say("c3: refining the best 2000 decoys")
    R, t, hist = F.refine_hard(ctx, rot, top["index"][:, 1:].astype(np.float32), steps=3)
    _say("c3: improved %d of 2000" % int(np.sum(hist[-1] > hist[0])))
    assert np.allclose(hist[0], top["score"], rtol=1e-5, atol=1e-6)
    assert np.all(np.diff(hist, axis=0) >= 0)
    assert np.sum(hist[-1] > hist[0]) > 0
    ctx.close()
