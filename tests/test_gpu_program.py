"""GPU: the megadock-hip program end to end (mdkh_run).  Its .out file has the reference's
layout (src/docking.cpp:399-453) and its decoys are the library's merged top-N for the
same pair (mdk_topn over mdk_dock_range), read back the way decoygen reads them
(src/decoygen.cpp:108-122); a forced grid that is too small is refused with the
reference's message (src/control.cpp:382-465)."""
import os

import numpy as np
import pytest

import megadock_amd as M

pytestmark = pytest.mark.gpu


def test_program_out_file_is_the_library_top_n(tmp_path, golden_dir):
    r = os.path.join(golden_dir, "1gcq_r.pdb")
    l = os.path.join(golden_dir, "1gcq_l.pdb")
    out = tmp_path / "x.out"
    M.run(["-R", r, "-L", l, "-o", str(out), "-N", "50", "-t", "2"])
    lines = out.read_text().splitlines()
    rec, lig = M.Protein(r), M.Protein(l)
    ng = M.autogrid(rec, lig)
    nf = 2 * ng
    assert lines[0] == "%d\t%.2f" % (nf, 1.2)
    assert lines[1] == "\t%f\t%f\t%f" % (0.0, 0.0, 0.0)
    assert lines[2] == "%s\t%f\t%f\t%f" % ((r,) + tuple(float(c) for c in rec.center))
    assert lines[3] == "%s\t%f\t%f\t%f" % ((l,) + tuple(float(c) for c in lig.center))
    assert len(lines) == 4 + 50
    print("program nf=%d, 50 decoys" % nf, flush=True)
    ctx = M.Context(nf, ng, lig.na, t=2, max_batch=64)
    ctx.set_receptor(*M.receptor_grids(rec, ng))
    ctx.set_ligand(lig.xyz, lig.center, lig.radius, lig.charge, M.grid_coord(ng), 1.2)
    angles = np.loadtxt(os.path.join(M.DATA, "rotations_3600.txt"), dtype=np.float64).astype(np.float32)
    ctx.set_rotations(M.rotation_coefs(angles))
    ctx.dock_range(0, len(angles))
    ctx.sync()
    top = ctx.topn(len(angles), 50)
    ctx.close()
    pi_f = float(np.float32(np.pi))
    for k, line in enumerate(lines[4:]):
        f = line.split("\t")
        th = [float(x) for x in angles[top[k]["index"][0]]]
        th = [float(np.float32(x - pi_f * 2.0)) if x >= pi_f else x for x in th]  # src/docking.cpp:437-446
        assert f[:3] == ["%f" % x for x in th], (k, f, th)
        assert [int(x) for x in f[3:6]] == list(top[k]["index"][1:]), (k, f)
        assert f[6] == "%.2f" % top[k]["score"], (k, f)
        assert 0 <= int(f[3]) < nf and 0 <= int(f[4]) < nf and 0 <= int(f[5]) < nf


def test_program_refuses_a_forced_grid_that_is_too_small(tmp_path, golden_dir):
    r = os.path.join(golden_dir, "1gcq_r.pdb")
    l = os.path.join(golden_dir, "1gcq_l.pdb")
    with pytest.raises(M.MdkError, match="too big"):
        M.run(["-R", r, "-L", l, "-o", str(tmp_path / "y.out"), "-F", "24"])


def test_entry_points_refuse_invalid_arguments(golden_dir):
    """Errors come back as a status and a message, with nothing launched: a ligand larger
    than the context holds, detail translations outside -ng..3ng-1, a non-positive grid
    width for the device receptor producer, an empty detail call."""
    lig = M.Protein(os.path.join(golden_dir, "1gcq_l.pdb"))
    rec = M.Protein(os.path.join(golden_dir, "1gcq_r.pdb"))
    ng = 48
    ctx = M.Context(2 * ng, ng, 100, t=1, max_batch=2)
    with pytest.raises(AssertionError):  # the binding checks the count first
        ctx.set_ligand(lig.xyz, lig.center, lig.radius, lig.charge, M.grid_coord(ng), 1.2)
    ctx.set_ligand(lig.xyz[:100], lig.center, lig.radius[:100], lig.charge[:100], M.grid_coord(ng), 1.2)
    xyz = (rec.xyz - rec.center[None, :]).astype(np.float32)
    with pytest.raises(M.MdkError, match="invalid grid"):
        ctx.set_receptor_atoms(xyz, rec.radius, rec.charge, rec.ace, M.grid_coord(ng), gw=0.0)
    ctx.set_receptor_atoms(xyz, rec.radius, rec.charge, rec.ace, M.grid_coord(ng))
    rp = np.zeros(ng ** 3, np.float32)
    rc = M.rotation_coefs(np.zeros((1, 3), np.float32))
    with pytest.raises(M.MdkError, match="translation out of range"):
        ctx.score_detail(rp, rc, np.array([[3 * ng, 0, 0]], np.int32))
    assert ctx.score_detail(rp, rc[:0], np.zeros((0, 3), np.int32)).shape == (0, 9)
    ctx.close()
