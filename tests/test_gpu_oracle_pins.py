"""GPU: the product held to the oracle on the workload configurations themselves.

  * ligand voxel grids bit for bit on all 3600 c0 rotations (ng 48, the LDS voxelizer),
    on a stride sample of the 54000 c1 rotations (ng 72, the largest LDS grid) and of the
    54000 c4 rotations (ng 128, the global-memory voxelizer);
  * c1 (144^3, t = 3) on a rotation sample against the oracle's loop;
  * c0 end to end through the program (megadock-hip: PDB reader, grid sizing, receptor
    grids on the device, all 3600 rotations, device merge, .out file) against the oracle's
    loop over all 3600 rotations and its merge (src/fft_process.cpp:399-498): every
    per-rotation best score within 1e-4 relative, best positions identical except at
    near-ties of the FP64 score map, and the 2000 decoys of the .out file at least 99%
    the oracle's top 2000.
"""
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


def _voxels_vs_oracle(lig_pdb, ng, nrot_set, ids, batch):
    lig = O.Protein(lig_pdb)
    ang = O.load_rotations(nrot_set)
    ctx = M.Context(2 * ng, ng, lig.na, t=1, max_batch=batch)
    ctx.set_ligand(lig.xyz, lig.center, lig.radius, lig.charge, O.grid_coord(ng, 1.2), 1.2)
    nz = 0
    for b0 in range(0, len(ids), batch):
        sel = ids[b0:b0 + batch]
        gr, gi = ctx.voxelize(M.rotation_coefs(ang[sel]))
        for q, a in enumerate(sel):
            xyz = O.rotate_ligand(lig.xyz, lig.center, O.rotation_coef(ang[a]))
            er, ei = O.ligand_voxel(ng, 1.2, xyz, lig.radius, lig.charge)
            assert np.array_equal(gr[q].view(np.uint32), er.view(np.uint32)), ("shape", int(a))
            assert np.array_equal(gi[q].view(np.uint32), ei.view(np.uint32)), ("charge", int(a))
            nz += int(np.count_nonzero(er))
    ctx.close()
    return nz


def test_voxels_c0_all_rotations(wl):
    rec, lig, cfg = W.make_pair("c0", wl)
    _say("c0 voxels: 3600 rotations, ng 48")
    assert _voxels_vs_oracle(lig, 48, 3600, np.arange(3600), 128) > 0


@pytest.mark.parametrize("name,ng,stride", [("c1", 72, 540), ("c4", 128, 2700)])
def test_voxels_c1_c4_stride_sample(wl, name, ng, stride):
    rec, lig, cfg = W.make_pair(name, wl)
    ids = np.arange(stride // 3, 54000, stride)
    _say("%s voxels: %d of 54000 rotations, ng %d" % (name, len(ids), ng))
    assert _voxels_vs_oracle(lig, ng, 54000, ids, 16) > 0


def _fp64_map(lig, ng, nf, ang_a, remb):
    xyz = O.rotate_ligand(lig.xyz, lig.center, O.rotation_coef(ang_a))
    lr, li = O.ligand_voxel(ng, 1.2, xyz, lig.radius, lig.charge)
    return O.correlation_fp64(nf, remb, O.embed(ng, nf, lr, li)).ravel()


def _lin(e, nf):
    return (int(e["index"][1]) * nf + int(e["index"][2])) * nf + int(e["index"][3])


def _per_rotation_check(got, ref, t, ids, lig, ng, nf, ang, remb):
    """Per-rotation results: rotation numbers equal, scores within 1e-4 relative, positions
    identical except where the FP64 map calls the two a near-tie (within 2e-6 of the map
    scale, the FP32 error bound of the transforms).  Returns the number of near-ties."""
    ties = 0
    for q in range(len(ids) * t):
        g, r = got[q], ref[q]
        assert g["index"][0] == r["index"][0]
        assert abs(g["score"] - r["score"]) <= 1e-4 * max(abs(r["score"]), 1e-3), (q, g, r)
        if not np.array_equal(g["index"], r["index"]):
            s64 = _fp64_map(lig, ng, nf, ang[int(ids[q // t])], remb)
            assert abs(s64[_lin(g, nf)] - s64[_lin(r, nf)]) <= 2e-6 * np.abs(s64).max(), (q, g, r)
            ties += 1
    return ties


def test_c1_sample_vs_oracle(wl):
    rec_pdb, lig_pdb, cfg = W.make_pair("c1", wl)
    rec, lig = O.Protein(rec_pdb), O.Protein(lig_pdb)
    prec = M.Protein(rec_pdb)
    ng, nf, t = 72, 144, cfg["t"]
    ang = O.load_rotations(54000)
    ids = np.arange(101, 54000, 563, dtype=np.int32)  # 96 rotations
    _say("c1 sample: %d rotations, t=%d, nf=%d" % (len(ids), t, nf))
    ctx = M.Context(nf, ng, lig.na, t=t, max_batch=len(ids))
    ctx.set_receptor_atoms((prec.xyz - prec.center[None, :]).astype(np.float32), prec.radius, prec.charge, prec.ace,
                           M.grid_coord(ng), 1.2)
    ctx.set_ligand(lig.xyz, lig.center, lig.radius, lig.charge, O.grid_coord(ng, 1.2), 1.2)
    got = ctx.dock(M.rotation_coefs(ang[ids]), ids)
    ctx.close()
    gr, gi = O.receptor_voxel(ng, 1.2, rec)
    rr, ri = O.receptor_fft(ng, nf, gr, gi)
    ref = O.dock(ng, nf, 1.2, lig, rr, ri, ang, ids, t, threads=16)
    ties = _per_rotation_check(got, ref, t, ids, lig, ng, nf, ang, O.embed(ng, nf, gr, gi))
    _say("c1 sample: %d near-ties" % ties)


def _out_decoys(path):
    """(angle text x3, tx, ty, tz) -> score of every decoy line of a .out file (the
    columns decoygen reads, src/decoygen.cpp:108-122)."""
    lines = open(path).read().splitlines()[4:]
    out = {}
    for l in lines:
        f = l.split("\t")
        out[(f[0], f[1], f[2], int(f[3]), int(f[4]), int(f[5]))] = float(f[6])
    return out, lines


def test_c0_full_config_through_the_program(wl, tmp_path):
    rec_pdb, lig_pdb, cfg = W.make_pair("c0", wl)
    out = str(tmp_path / "c0.out")
    _say("c0 program: all 3600 rotations, -N 2000")
    M.run(["-R", rec_pdb, "-L", lig_pdb, "-o", out, "-N", "2000"])
    rec, lig = O.Protein(rec_pdb), O.Protein(lig_pdb)
    ng = O.autogrid(rec, lig)
    nf = 2 * ng
    assert ng == 48
    head = open(out).read().splitlines()[0]
    assert head == "%d\t%.2f" % (nf, 1.2)
    dec, lines = _out_decoys(out)
    assert len(lines) == 2000 and len(dec) == 2000

    # per-rotation bests of the same path (device receptor producer, as the program uses it)
    prec = M.Protein(rec_pdb)
    ang = O.load_rotations(3600)
    ctx = M.Context(nf, ng, lig.na, t=1, max_batch=256)
    ctx.set_receptor_atoms((prec.xyz - prec.center[None, :]).astype(np.float32), prec.radius, prec.charge, prec.ace,
                           M.grid_coord(ng), 1.2)
    ctx.set_ligand(lig.xyz, lig.center, lig.radius, lig.charge, O.grid_coord(ng, 1.2), 1.2)
    ctx.set_rotations(M.rotation_coefs(ang))
    ctx.dock_range(0, 3600)
    ctx.sync()
    got = ctx.results(0, 3600)
    top = ctx.topn(3600, 2000)
    ctx.close()

    _say("c0 oracle: 3600 rotations")
    gr, gi = O.receptor_voxel(ng, 1.2, rec)
    rr, ri = O.receptor_fft(ng, nf, gr, gi)
    ids = np.arange(3600, dtype=np.int32)
    ref = O.dock(ng, nf, 1.2, lig, rr, ri, ang, ids, 1, threads=16)
    ties = _per_rotation_check(got, ref, 1, ids, lig, ng, nf, ang, O.embed(ng, nf, gr, gi))
    _say("c0: %d near-ties among 3600 per-rotation bests" % ties)

    # the program's decoys are the library's merged top 2000 ...
    pi_f = float(np.float32(np.pi))

    def key(e):
        th = [float(x) for x in ang[int(e["index"][0])]]
        th = [float(np.float32(x - pi_f * 2.0)) if x >= pi_f else x for x in th]  # src/docking.cpp:437-446
        return tuple("%f" % x for x in th) + tuple(int(v) for v in e["index"][1:])

    assert [key(e) for e in top] == [tuple(l.split("\t")[:3]) + tuple(int(v) for v in l.split("\t")[3:6])
                                     for l in lines]
    # ... and at least 99% of them are the oracle's top 2000
    exp = O.sort_topn(ref, 2000)
    common = set(dec) & {key(e) for e in exp}
    _say("c0: top-2000 overlap %d" % len(common))
    assert len(common) >= 0.99 * 2000
    expd = {key(e): float(e["score"]) for e in exp}
    for k in common:
        assert abs(dec[k] - expd[k]) <= 0.005 + 1e-4 * abs(expd[k]), (k, dec[k], expd[k])  # %.2f column
