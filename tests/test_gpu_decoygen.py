"""GPU: megadock-hip-decoygen writes the decoy PDB files of the reference's decoygen
(src/decoygen.cpp, compiled into oracle/_ref and run as the checker) byte for byte, one
decoy at a time and for a range of decoys at once (coordinates on the device,
mdk_decoy_coords)."""
import os
import subprocess

import pytest

import megadock_amd as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOOL = os.path.join(os.path.dirname(M.LIB_PATH), "..", "bin", "megadock-hip-decoygen")


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref (the reference's decoygen) is not built")
def test_decoygen_matches_reference(tmp_path, golden_dir):
    r = os.path.join(golden_dir, "1gcq_r.pdb")
    l = os.path.join(golden_dir, "1gcq_l.pdb")
    out = tmp_path / "pair.out"
    M.run(["-R", r, "-L", l, "-o", str(out), "-N", "30"])
    print("decoys of %s" % out, flush=True)
    for n in (1, 7, 30):
        mine, ref = tmp_path / ("mine_%d.pdb" % n), tmp_path / ("ref_%d.pdb" % n)
        subprocess.run([TOOL, str(mine), l, str(out), str(n)], check=True, timeout=120)
        assert O.ref_decoygen(ref, l, out, n) == 0
        assert mine.read_bytes() == ref.read_bytes(), n
    subprocess.run([TOOL, str(tmp_path / "all.pdb"), l, str(out), "1-30"], check=True, timeout=120)
    for n in (1, 2, 15, 30):
        ref = tmp_path / ("ref_%d.pdb" % n)
        if not ref.exists():
            assert O.ref_decoygen(ref, l, out, n) == 0
        assert (tmp_path / ("all.pdb.%d" % n)).read_bytes() == ref.read_bytes(), n
