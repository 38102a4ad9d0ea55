"""The product's host layer (libmdk's C++ input side and receptor producer, called through
include/mdk_host.h) held to the reference's own CPU code compiled in oracle/_ref and to
the oracle, on the golden PDB files and on the synthetic workloads:

  * PDB reader and per-atom table lookups (mdkh_protein_load, restating
    src/pdb_entry.cpp:18-105 and src/protein.cpp:19-64) against the reference's
    PDBEntry/Protein/Parameter;
  * receptor grids (mdkh_receptor_voxel, restating src/receptor.cpp:20-372) bit for bit
    against the reference's Receptor::rpscace + electro and against the oracle;
  * built-in rotation sets (the data files read by mdkh_rotation_set and by the oracle)
    against the reference's _Zangle after Parameter::initialize with -r n;
  * grid sizing (mdkh_autogrid, src/control.cpp:287-379) against values derived by hand
    from those lines for extents on, and one float ulp either side of, a table entry.

CPU only: libmdk.so loads without a GPU and these entry points touch none.
"""
import os

import numpy as np
import pytest

import megadock_amd as M
from megadock_amd import workloads as W
from oracle import oracle as O

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) not present")


@pytest.fixture(scope="module")
def wl(tmp_path_factory):
    """Workload PDB files: c0, c1, c4 pairs and two c2 proteins."""
    d = str(tmp_path_factory.mktemp("wl"))
    out = {}
    for name in ("c0", "c1", "c4"):
        out[name] = W.make_pair(name, d)[:2]
    out["c2_4_5"] = W.make_c2_pair(4, 5, d)
    return out


def _bits_equal(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))


def _pairs(golden_dir, wl):
    g = [(os.path.join(golden_dir, a), os.path.join(golden_dir, b))
         for a, b in (("1gcq_r.pdb", "1gcq_l.pdb"), ("synth_r120.pdb", "synth_l40.pdb"))]
    return g + [wl[k] for k in ("c0", "c1", "c4", "c2_4_5")]


@needs_ref
def test_product_pdb_reader_matches_reference(golden_dir, wl):
    """Coordinates, radius, charge, ACE and box centre of every atom, bit for bit."""
    for rp, lp in _pairs(golden_dir, wl):
        O.ref_load(rp, lp)
        for which, path in ((0, rp), (1, lp)):
            ref = O.ref_protein(which)
            mine = M.Protein(path)
            assert mine.na == len(ref["xyz"]) > 0, path
            for k in ("xyz", "radius", "charge", "ace", "center"):
                assert _bits_equal(getattr(mine, k), ref[k]), (path, k)


@needs_ref
@pytest.mark.parametrize("case", [
    ("1gcq", 40, {}), ("1gcq", 48, {"aceratio": 0.5, "rec_core": -30.0, "eratio": 2.0}),
    ("synth", 48, {}), ("synth", 36, {"gw": 1.0}), ("c0", 48, {}), ("c1", 72, {}), ("c4", 128, {}),
    ("c2_4_5", 55, {}),
])
def test_product_host_receptor_grids_match_reference_and_oracle(golden_dir, wl, case):
    name, ng, kw = case
    files = {"1gcq": (os.path.join(golden_dir, "1gcq_r.pdb"), os.path.join(golden_dir, "1gcq_l.pdb")),
             "synth": (os.path.join(golden_dir, "synth_r120.pdb"), os.path.join(golden_dir, "synth_l40.pdb"))}
    rp, lp = files.get(name) or wl[name]
    gw = kw.get("gw", 1.2)
    prm = dict(aceratio=kw.get("aceratio", 1.0), rec_core=kw.get("rec_core", -45.0), eratio=kw.get("eratio", 1.0))
    mine = M.receptor_grids(M.Protein(rp), ng, gw, **prm)
    orc = O.receptor_voxel(ng, gw, O.Protein(rp), **prm)
    O.ref_load(rp, lp, ["-v", repr(gw)])  # the reference's Protein takes its pitch from -v
    ref = O.ref_receptor_voxel(ng, gw, **prm)
    for q in range(2):
        assert _bits_equal(mine[q], ref[q]), (name, ng, q, "product vs reference")
        assert _bits_equal(orc[q], ref[q]), (name, ng, q, "oracle vs reference")
    assert int((ref[0] == prm["rec_core"]).sum()) > 0  # the core fill ran


@needs_ref
@pytest.mark.parametrize("n", [1, 3, 24, 360, 3600, 54000])
def test_rotation_sets_match_reference_table(golden_dir, n):
    """The data files hold the reference's own tables: the reference's _Zangle after
    Parameter::initialize(-r n) (src/parameter.cpp:376-395) equals, bit for bit, what the
    product's host layer and the oracle load."""
    ref = O.ref_rotations(n, os.path.join(golden_dir, "1gcq_r.pdb"), os.path.join(golden_dir, "1gcq_l.pdb"))
    assert ref.shape == (n, 3)
    assert _bits_equal(M.rotation_set(n), ref)
    assert _bits_equal(O.load_rotations(n), ref)


def test_rotation_set_unknown_is_refused():
    with pytest.raises(M.MdkError):
        M.rotation_set(7)


def _pdb(path, coords):
    with open(path, "w") as f:
        for a, (x, y, z) in enumerate(coords):
            f.write("ATOM  %5d  CA  ALA A%4d    %8s%8s%8s  1.00  0.00           C\n" % (a + 1, a + 1, x, y, z))
        f.write("END\n")
    return path


# Expected ng worked out from src/control.cpp:287-379 for the 11-base table
# (src/control.cpp:91-125).  Receptor: size = max over axes of (max - min) of the float
# coordinates, in float; size = float(size + 2 * 8.0); n = 1 + int(size / gw) with float
# division and truncation; ng = smallest table entry >= n.  Ligand: the same with
# size = sqrt of the largest float squared pair distance and a margin of 2 * 3.0.
_TINY = [("0.000", "0.000", "0.000"), ("1.000", "0.000", "0.000")]  # ligand or receptor that never binds
AUTOGRID_CASES = [
    # extent exactly 32: size 48.0, n = 1 + int(48.0) = 49, in the table (a ceiling would give 48)
    ("rec_on_entry", 1.0, [("0.000",) * 3, ("32.000", "0.000", "0.000")], _TINY, 49),
    # float(-7.996) - float(-39.996) is one ulp below 32 (31.999998); + 16.0 rounds, in float,
    # to 48.0 exactly, so n = 49 (summed in double it would stay below 48 and give 48)
    ("rec_one_ulp_below", 1.0, [("-39.996", "0.000", "0.000"), ("-7.996", "0.000", "0.000")], _TINY, 49),
    # two ulps below 32 (31.999996): size 47.999996 is a float, n = 1 + 47 = 48
    ("rec_two_ulps_below", 1.0, [("32.004", "0.000", "0.000"), ("64.004", "0.000", "0.000")], _TINY, 48),
    # one ulp above 32 (32.000004): size 48.000004, n = 49
    ("rec_one_ulp_above", 1.0, [("32.007", "0.000", "0.000"), ("64.007", "0.000", "0.000")], _TINY, 49),
    # extent 31.999: size 47.999, n = 48
    ("rec_below_entry", 1.0, [("0.000",) * 3, ("31.999", "0.000", "0.000")], _TINY, 48),
    # extent 34: n = 51, no table entry 51..53, so 54
    ("rec_table_gap", 1.0, [("0.000",) * 3, ("0.000", "34.000", "0.000")], _TINY, 54),
    # gw 1.2: float(41.6) + 16 = 57.599998; / float(1.2) = 47.999996 in float, n = 48
    # (exact decimals, 57.6 / 1.2 = 48, would give 49)
    ("rec_float_quotient", 1.2, [("0.000",) * 3, ("0.000", "0.000", "41.600")], _TINY, 48),
    # extent 250: n = 267, above 256: the table continues, 270
    ("rec_above_256", 1.0, [("0.000",) * 3, ("250.000", "0.000", "0.000")], _TINY, 270),
    # ligand (0,0,0)-(20,20,0): sqrt(800) = 28.284271, + 6 = 34.28427, n = 35, in the table;
    # its box edge (20) would give 27
    ("lig_pair_distance", 1.0, _TINY, [("0.000",) * 3, ("20.000", "20.000", "0.000")], 35),
    # three ligand atoms, largest distance between the last two (20): 26, n = 27
    ("lig_inner_pair", 1.0, _TINY, [("0.000",) * 3, ("10.000", "0.000", "0.000"), ("-10.000", "0.000", "0.000")], 27),
    # ligand distance 42: size 48.0, n = 49
    ("lig_on_entry", 1.0, _TINY, [("0.000",) * 3, ("0.000", "42.000", "0.000")], 49),
]


@pytest.mark.parametrize("name,gw,rec,lig,expect", AUTOGRID_CASES, ids=[c[0] for c in AUTOGRID_CASES])
def test_autogrid_hand_derived(tmp_path, name, gw, rec, lig, expect):
    rp = _pdb(str(tmp_path / "r.pdb"), rec)
    lp = _pdb(str(tmp_path / "l.pdb"), lig)
    assert M.autogrid(M.Protein(rp), M.Protein(lp), gw) == expect
    assert O.autogrid(O.Protein(rp), O.Protein(lp), gw) == expect
