"""The oracle's voxelizers and rotation against the reference's own CPU code.

Pinned three ways: the committed digests written by tests/golden/make_golden.py from
the reference build; the reference build itself (oracle/_ref, when present); and
closed forms (identity rotation, orthonormality, one-atom grids).
"""
import hashlib
import json
import os

import numpy as np
import pytest

from oracle import oracle as O


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, np.float32).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def golden(golden_dir):
    with open(os.path.join(golden_dir, "voxel_digests.json")) as f:
        return json.load(f)


def test_golden_receptor_and_ligand_grids(golden, golden_dir):
    ang = O.load_rotations(golden["rotations_set"])
    for name, case in golden["cases"].items():
        rec = O.Protein(os.path.join(golden_dir, case["receptor"]))
        lig = O.Protein(os.path.join(golden_dir, case["ligand"]))
        ng = case["ng"]
        gr, gi = O.receptor_voxel(ng, golden["gw"], rec)
        assert digest(gr) == case["receptor_grid_r"], name
        assert digest(gi) == case["receptor_grid_i"], name
        for a, d in case["ligand_grids"].items():
            xyz = O.rotate_ligand(lig.xyz, lig.center, O.rotation_coef(ang[int(a)]))
            lr, li = O.ligand_voxel(ng, golden["gw"], xyz, lig.radius, lig.charge, case["delta"])
            assert digest(lr) == d["grid_r"], (name, a)
            assert digest(li) == d["grid_i"], (name, a)


def test_autogrid_product_equals_oracle(golden_dir):
    """Grid sizing of the golden pairs: the product's host layer and the oracle agree (both
    are pinned to values derived by hand from src/control.cpp:287-379 in
    tests/test_host_parity.py::test_autogrid_hand_derived)."""
    import megadock_amd as M
    for r, l in (("synth_r120.pdb", "synth_l40.pdb"), ("1gcq_r.pdb", "1gcq_l.pdb")):
        rp, lp = os.path.join(golden_dir, r), os.path.join(golden_dir, l)
        assert M.autogrid(M.Protein(rp), M.Protein(lp)) == O.autogrid(O.Protein(rp), O.Protein(lp))


needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) not present")


@needs_ref
def test_pdb_and_tables_match_reference(golden_dir):
    for r, l in (("1gcq_r.pdb", "1gcq_l.pdb"), ("synth_r120.pdb", "synth_l40.pdb")):
        rp, lp = os.path.join(golden_dir, r), os.path.join(golden_dir, l)
        O.ref_load(rp, lp)
        for which, path in ((0, rp), (1, lp)):
            mine = O.Protein(path)
            ref = O.ref_protein(which)
            assert ref["keys"] == mine.keys
            for k in ("xyz", "radius", "charge", "ace", "center"):
                assert np.array_equal(ref[k], getattr(mine, k)), (path, k)


@needs_ref
@pytest.mark.parametrize("ng,delta", [(40, 1.0), (40, 0.5), (32, 1.0), (44, -2.0)])
def test_ligand_voxel_matches_reference_many_rotations(golden_dir, ng, delta):
    rp, lp = os.path.join(golden_dir, "1gcq_r.pdb"), os.path.join(golden_dir, "1gcq_l.pdb")
    O.ref_load(rp, lp)
    lig = O.Protein(lp)
    ang = O.load_rotations(3600)
    for a in range(0, 3600, 97):
        r = O.rotation_coef(ang[a])
        xyz = O.rotate_ligand(lig.xyz, lig.center, r)
        assert np.array_equal(xyz, O.ref_rotate(lig.xyz - lig.center, ang[a])), a
        e = O.ref_ligand_voxel(ng, 1.2, xyz, delta=delta)
        g = O.ligand_voxel(ng, 1.2, xyz, lig.radius, lig.charge, delta)
        assert np.array_equal(e[0], g[0]) and np.array_equal(e[1], g[1]), a


@needs_ref
@pytest.mark.parametrize("ng,aceratio,core,eratio", [(40, 1.0, -45.0, 1.0), (48, 0.5, -30.0, 2.0), (36, 1.0, -45.0, 1.0)])
def test_receptor_voxel_matches_reference(golden_dir, ng, aceratio, core, eratio):
    rp, lp = os.path.join(golden_dir, "synth_r120.pdb"), os.path.join(golden_dir, "synth_l40.pdb")
    O.ref_load(rp, lp)
    e = O.ref_receptor_voxel(ng, 1.2, aceratio=aceratio, rec_core=core, eratio=eratio)
    g = O.receptor_voxel(ng, 1.2, O.Protein(rp), aceratio=aceratio, rec_core=core, eratio=eratio)
    assert np.array_equal(e[0], g[0]) and np.array_equal(e[1], g[1])


def test_rotation_closed_forms():
    assert np.array_equal(O.rotation_coef(np.zeros(3, np.float32)), np.eye(3, dtype=np.float32).ravel())
    ang = O.load_rotations(3600)
    for a in range(0, 3600, 61):
        r = O.rotation_coef(ang[a]).astype(np.float64).reshape(3, 3)
        assert np.allclose(r @ r.T, np.eye(3), atol=1e-6)
        assert abs(np.linalg.det(r) - 1.0) < 1e-6
        # ZXZ composition: Rz(phi) Rx(theta) Rz(psi)
        p, t, s = [float(v) for v in ang[a]]
        rz = lambda x: np.array([[np.cos(x), -np.sin(x), 0], [np.sin(x), np.cos(x), 0], [0, 0, 1]])
        rx = lambda x: np.array([[1, 0, 0], [0, np.cos(x), -np.sin(x)], [0, np.sin(x), np.cos(x)]])
        assert np.allclose(r, rz(p) @ rx(t) @ rz(s), atol=1e-6)


def test_rotation_sets():
    for n in (1, 3, 24, 360, 3600, 54000):
        assert O.load_rotations(n).shape == (n, 3)


def test_one_atom_ligand_grid():
    """One carbon at the grid centre: core sphere of 1.5 r^2, scraped to its interior,
    surface sphere of r^2, charge in the nearest voxel."""
    ng, gw = 12, 1.2
    xyz = np.zeros((1, 3), np.float32)
    gr, gi = O.ligand_voxel(ng, gw, xyz, np.array([1.8], np.float32), np.array([0.25], np.float32))
    gc = O.grid_coord(ng, gw).astype(np.float64)
    d2 = gc[:, None, None] ** 2 + gc[None, :, None] ** 2 + gc[None, None, :] ** 2
    surf = d2 < 1.8 ** 2
    # the one-atom grid is {0, 1}: every voxel within r is marked 1 (core or surface)
    assert set(np.unique(gr)) <= {0.0, 1.0}
    assert np.array_equal(gr.reshape(ng, ng, ng) == 1.0, surf | ((d2 < 1.5 * 1.8 ** 2) & ~_scraped(d2 < 1.5 * 1.8 ** 2)))
    assert gi.sum() == np.float32(0.25) and (gi != 0).sum() == 1
    assert gi.reshape(ng, ng, ng)[ng // 2, ng // 2, ng // 2] == np.float32(0.25)


def _scraped(core):
    ng = core.shape[0]
    out = np.zeros_like(core)
    for i in range(1, ng - 1):
        for j in range(1, ng - 1):
            for k in range(1, ng - 1):
                if core[i, j, k] and not (core[i - 1, j, k] and core[i + 1, j, k] and core[i, j - 1, k]
                                          and core[i, j + 1, k] and core[i, j, k - 1] and core[i, j, k + 1]):
                    out[i, j, k] = True
    return out


def test_charges_sum_in_atom_order():
    """Three charges in one voxel add left to right from +0.0f (float, not exact)."""
    ng, gw = 8, 1.2
    xyz = np.array([[0.1, 0.1, 0.1], [0.2, 0.2, 0.2], [0.3, 0.3, 0.3]], np.float32)
    q = np.array([1e8, 1.0, -1e8], np.float32)
    _, gi = O.ligand_voxel(ng, gw, xyz, np.full(3, 1.0, np.float32), q)
    expect = np.float32(np.float32(np.float32(0.0) + q[0]) + q[1]) + q[2]
    assert gi.reshape(ng, ng, ng)[4, 4, 4] == expect == np.float32(0.0)
