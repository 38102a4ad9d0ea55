"""GPU: mdk_topn_noalloc, the merge that allocates nothing (its state, keys and results
live in a fixed per-GPU device array), against mdk_topn on the same device results: two
calls in a row, the second with fewer results, bitwise equal both times; sizes beyond the
array's maximum are refused with an error."""
import ctypes as C
import os

import numpy as np
import pytest

import megadock_amd as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu

NOALLOC_MAX = 32768


def _noalloc(ctx, nrot, n):
    out = np.zeros(n, M.SORTSCORE_DTYPE)
    rc = M.lib().mdk_topn_noalloc(ctx.h, C.c_int(nrot), C.c_int(n), C.c_void_p(out.ctypes.data))
    return rc, out


def test_topn_noalloc_matches_topn(golden_dir):
    print("no-alloc merge: 1gcq, nf 80, 3600 rotations, t 10", flush=True)
    rec = O.Protein(os.path.join(golden_dir, "1gcq_r.pdb"))
    lig = O.Protein(os.path.join(golden_dir, "1gcq_l.pdb"))
    ng, nf, t = 40, 80, 10
    gr, gi = O.receptor_voxel(ng, 1.2, rec)
    ang = O.load_rotations(3600)
    ctx = M.Context(nf, ng, lig.na, t=t, max_batch=128)
    try:
        ctx.set_receptor(gr, gi)
        ctx.set_ligand(lig.xyz, lig.center, lig.radius, lig.charge, O.grid_coord(ng, 1.2), 1.2)
        ctx.set_rotations(M.rotation_coefs(ang))
        ctx.dock_range(0, 3600)
        ctx.sync()
        for nrot, n in [(3600, NOALLOC_MAX), (1000, 2000)]:  # the second call merges fewer results
            print("  nrot %d, n %d" % (nrot, n), flush=True)
            ref = ctx.topn(nrot, n)
            rc, got = _noalloc(ctx, nrot, n)
            assert rc == 0, M.lib().mdk_last_error()
            assert np.array_equal(got.view(np.uint8), ref.view(np.uint8)), (nrot, n)
            s = got["score"]
            assert np.all(s[:-1] >= s[1:])
        # every rotation's results are in the merge of all of them at its maximum size
        rc, _ = _noalloc(ctx, 3600, NOALLOC_MAX + 1)
        assert rc != 0
        assert b"32768" in M.lib().mdk_last_error()
    finally:
        ctx.close()


def test_topn_noalloc_second_context_smaller(golden_dir):
    """A second context on the same GPU shares the array; its results are its own."""
    print("no-alloc merge: second context, nf 80, 360 rotations, t 1", flush=True)
    rec = O.Protein(os.path.join(golden_dir, "1gcq_r.pdb"))
    lig = O.Protein(os.path.join(golden_dir, "1gcq_l.pdb"))
    ng, nf = 40, 80
    gr, gi = O.receptor_voxel(ng, 1.2, rec)
    ang = O.load_rotations(360)
    ctx = M.Context(nf, ng, lig.na, t=1, max_batch=64)
    try:
        ctx.set_receptor(gr, gi)
        ctx.set_ligand(lig.xyz, lig.center, lig.radius, lig.charge, O.grid_coord(ng, 1.2), 1.2)
        ctx.set_rotations(M.rotation_coefs(ang))
        ctx.dock_range(0, 360)
        ctx.sync()
        for n in (360, 100):
            ref = ctx.topn(360, n)
            rc, got = _noalloc(ctx, 360, n)
            assert rc == 0, M.lib().mdk_last_error()
            assert np.array_equal(got.view(np.uint8), ref.view(np.uint8)), n
        res = ctx.results(0, 360)
        order = np.lexsort((np.arange(360), -res["score"].astype(np.float64)))
        assert np.array_equal(got.view(np.uint8), res[order[:100]].view(np.uint8))
    finally:
        ctx.close()
