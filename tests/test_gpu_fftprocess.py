"""GPU: FFTProcess with the reference's class interface (include/mdk_fft_process.h,
src/fft_process.h:45-196) driven as the reference's Docking::dockz drives it
(src/docking.cpp:194-346), by the C++ driver tests/native/fftprocess_dockz.cpp: OpenMP
threads with dynamic scheduling, the GPU threads through rotation_index + cuda_fft, the
others through ligand_preparation + fft3d + convolution + fft3d + score_sort, then
sort_index.  _Top is checked against the oracle's loop (src/fft_process.cpp:399-458) and
the reference's partial sort (:462-479)."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "native", "_build", "fftprocess_dockz")


def _run(rec, lig, nrot, t, n, nproc2, num_gpu):
    assert os.path.exists(DRIVER), "tests/native/_build/fftprocess_dockz is built by __graft_entry__.build()"
    out = subprocess.run([DRIVER, rec, lig, str(nrot), str(t), str(n), str(nproc2), str(num_gpu)],
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    rows = [l.split() for l in out.stdout.splitlines() if l.startswith("E ")]
    score = np.array([float(r[2]) for r in rows])
    index = np.array([[int(x) for x in r[3:7]] for r in rows])
    return score, index


@pytest.mark.parametrize("t,nproc2,num_gpu", [(1, 1, 1), (2, 3, 1), (3, 4, 2)])
def test_fftprocess_like_dockz_vs_oracle(golden_dir, t, nproc2, num_gpu):
    """nproc2 host threads of which num_gpu take the GPU branch (with one GPU, the second
    GPU thread's context sits on the same device)."""
    rp, lp = os.path.join(golden_dir, "1gcq_r.pdb"), os.path.join(golden_dir, "1gcq_l.pdb")
    nrot, n = 360, 200
    print("FFTProcess driver: t %d, %d threads, %d GPU threads" % (t, nproc2, num_gpu), flush=True)
    score, index = _run(rp, lp, nrot, t, n, nproc2, num_gpu)
    assert len(score) == nrot * t
    rec, lig = O.Protein(rp), O.Protein(lp)
    ng = O.autogrid(rec, lig)
    nf = 2 * ng
    gr, gi = O.receptor_voxel(ng, 1.2, rec)
    rr, ri = O.receptor_fft(ng, nf, gr, gi)
    ang = O.load_rotations(nrot)
    ref = O.dock(ng, nf, 1.2, lig, rr, ri, ang, np.arange(nrot, dtype=np.int32), t, threads=16)
    # every rotation's t results are in _Top (each rotation exactly t times)
    assert np.array_equal(np.bincount(index[:, 0], minlength=nrot), np.full(nrot, t))
    remb = None
    for a in range(nrot):
        g = np.sort(score[index[:, 0] == a])[::-1]
        r = np.sort(ref["score"][a * t:(a + 1) * t].astype(np.float64))[::-1]
        assert np.all(np.abs(g - r) <= 1e-4 * np.maximum(np.abs(r), 1e-3)), (a, g, r)
        gset = {tuple(x) for x in index[index[:, 0] == a]}
        rset = {tuple(int(v) for v in x) for x in ref["index"][a * t:(a + 1) * t]}
        if gset != rset:  # only at a near-tie of the FP64 map
            if remb is None:
                remb = O.embed(ng, nf, gr, gi)
            xyz = O.rotate_ligand(lig.xyz, lig.center, O.rotation_coef(ang[a]))
            lr, li = O.ligand_voxel(ng, 1.2, xyz, lig.radius, lig.charge)
            s64 = O.correlation_fp64(nf, remb, O.embed(ng, nf, lr, li)).ravel()
            for gi_, ri_ in zip(sorted(gset - rset), sorted(rset - gset)):
                ig = (gi_[1] * nf + gi_[2]) * nf + gi_[3]
                ir = (ri_[1] * nf + ri_[2]) * nf + ri_[3]
                assert abs(s64[ig] - s64[ir]) <= 2e-6 * np.abs(s64).max(), (a, gi_, ri_)
    # sort_index: the first n entries are the n best, by score descending
    assert np.all(np.diff(score[:n]) <= 0)
    assert np.allclose(score[:n], np.sort(score)[::-1][:n], rtol=0, atol=0)
