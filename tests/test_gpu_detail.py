"""GPU: the detailed output (-O, Docking::output_detail, src/docking.cpp:456-898).
mdk_score_detail's real-space float sums bit for bit against the oracle's restatement of
the reference's loop (its order of float additions) and within float rounding of the FP64
sums; the .detail/.csv files of the program byte for byte against the oracle's
restatement of output_detail, every decoy's total equal to its docked score (the FFT path,
an independent computation), the RMSD column as the reference's formula gives it."""
import os

import numpy as np
import pytest

import megadock_amd as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_score_detail_vs_fp64(golden_dir):
    rec = O.Protein(os.path.join(golden_dir, "1gcq_r.pdb"))
    lig = O.Protein(os.path.join(golden_dir, "1gcq_l.pdb"))
    ng = 48
    gr, gi = O.receptor_voxel(ng, 1.2, rec)
    rp, _ = O.receptor_voxel(ng, 1.2, rec, aceratio=0.0)
    ctx = M.Context(2 * ng, ng, lig.na, t=1, max_batch=4)
    ctx.set_receptor(gr, gi)
    ctx.set_ligand(lig.xyz, lig.center, lig.radius, lig.charge, O.grid_coord(ng, 1.2), 1.2)
    ang = O.load_rotations(3600)
    ids = [0, 17, 900, 2222, 3599, 1234]  # more poses than max_batch: two batches
    trans = np.array([[0, 0, 0], [1, 95, 47], [48, 3, 60], [-1, 96, 5], [20, 20, 20], [95, 0, 48]], np.int32)
    print("detail sums for %d poses" % len(ids), flush=True)
    got = ctx.score_detail(rp, M.rotation_coefs(ang[ids]), trans)
    for q, a in enumerate(ids):
        xyz = O.rotate_ligand(lig.xyz, lig.center, O.rotation_coef(ang[a]))
        lr, li = O.ligand_voxel(ng, 1.2, xyz, lig.radius, lig.charge)
        f32, _ = O.detail_sums_f32(ng, gr, gi, rp, lr, li, trans[q])
        assert np.array_equal(got[q].view(np.uint32), f32.view(np.uint32)), (q, got[q], f32)
        ref = O.detail_sums(ng, gr, gi, rp, lr, li, trans[q])
        assert np.allclose(got[q][:6], ref, rtol=1e-5, atol=1e-2), (q, got[q], ref)
    ctx.close()


def test_detail_files(tmp_path, golden_dir):
    r = os.path.join(golden_dir, "1gcq_r.pdb")
    l = os.path.join(golden_dir, "1gcq_l.pdb")
    out = tmp_path / "pair.out"
    M.run(["-R", r, "-L", l, "-o", str(out), "-O", "-N", "30"])
    detail, csv = tmp_path / "pair.detail", tmp_path / "pair.csv"
    assert detail.exists() and csv.exists()
    decoys = [l.split("\t") for l in out.read_text().splitlines()[4:]]
    rows = [l.split(",") for l in csv.read_text().splitlines()[1:]]
    assert len(rows) == len(decoys) == 30
    nf = int(out.read_text().split("\t")[0])
    for k, (d, c) in enumerate(zip(decoys, rows)):
        assert int(c[0]) == k + 1
        assert [int(c[4]), int(c[5]), int(c[6])] == [int(d[3]), int(d[4]), int(d[5])]
        total, rpsc, gain, penalty, elec, ace = [float(x) for x in c[7:13]]
        assert abs(total - float(d[6])) <= 1e-4 * abs(float(d[6])) + 0.01, (k, total, d[6])
        assert abs(rpsc - (gain + penalty)) <= 1e-3 * abs(gain) + 1e-3
        assert abs(total - (rpsc + elec + ace)) <= 1e-4 * abs(total) + 1e-3
    # byte for byte: the oracle's restatement of output_detail on the program's decoys (their
    # rotation numbers found from the .out angles, printed as the reference prints them)
    rec_o, lig_o = O.Protein(r), O.Protein(l)
    ang = O.load_rotations(3600)
    pi = np.float32(4.0 * np.arctan(1.0))
    printed = {}
    for a in range(len(ang)):
        th = [np.float32(float(v) - float(pi) * 2.0) if v >= pi else v for v in ang[a]]
        printed.setdefault("%f\t%f\t%f" % tuple(float(v) for v in th), a)
    dec = [(printed["\t".join(d[:3])], int(d[3]), int(d[4]), int(d[5])) for d in decoys]
    want_detail, want_csv = O.output_detail(rec_o, lig_o, nf // 2, 1.2, r, l, dec, ang)
    assert csv.read_text() == want_csv
    assert detail.read_text() == want_detail
    text = detail.read_text()
    assert "*** score around correct position ***" in text and "*** docking results ***" in text
    near = [l for l in text.split("*** score around correct position ***")[1].split("***")[0].splitlines()
            if l.strip()[:2].rstrip(":").strip().isdigit()]
    assert len(near) == 7
    # RMSD of a translated ligand: the translation of the receptor-ligand centre offset
    rec, lig = O.Protein(r), O.Protein(l)
    gw, ng = 1.2, nf // 2
    diff = (np.asarray(rec.center, np.float32) - np.asarray(lig.center, np.float32)).astype(np.float64)
    for c in rows[:5]:
        t = np.array([int(c[4]), int(c[5]), int(c[6])])
        tem = (3 * ng - t) % (2 * ng)
        shift = gw * (tem - ng) + diff
        assert abs(float(c[13]) - np.linalg.norm(shift)) <= 1e-3 * max(1.0, np.linalg.norm(shift)), (c, shift)
        assert int(c[14]) == (1 if float(c[13]) < 5.0 else 0)
