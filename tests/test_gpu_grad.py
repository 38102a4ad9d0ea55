"""GPU: off-grid scores and their gradients (mdk_score_spectral, mdk_score_companion)
against the FP64 references of the oracle, and against central finite differences of
those references (h = 1e-3 A in translation; see the companion test for rotation)."""
import os

import numpy as np
import pytest

import megadock_amd as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def pair(golden_dir):
    rec = O.Protein(os.path.join(golden_dir, "1gcq_r.pdb"))
    lig = O.Protein(os.path.join(golden_dir, "1gcq_l.pdb"))
    ng = 48  # nf = 96: specialised kernels
    gr, gi = O.receptor_voxel(ng, 1.2, rec)
    return rec, lig, ng, gr, gi


def _ctx(lig, ng, gr, gi, t=1, max_batch=8):
    ctx = M.Context(2 * ng, ng, lig.na, t=t, max_batch=max_batch)
    ctx.set_receptor(gr, gi)
    ctx.set_ligand(lig.xyz, lig.center, lig.radius, lig.charge, O.grid_coord(ng, 1.2), 1.2)
    return ctx


@pytest.mark.parametrize("ngs", [None, 50, 63])
def test_spectral_score_and_gradient_vs_fp64(pair, ngs):
    """nf = 96 (tuned kernels) and 100, 126 (further sizes, default variants)."""
    rec, lig, ng, gr, gi = pair
    if ngs is not None:
        ng = ngs
        gr, gi = O.receptor_voxel(ng, 1.2, rec)
    nf = 2 * ng
    ctx = _ctx(lig, ng, gr, gi)
    ang = O.load_rotations(3600)
    ids = [17, 1234, 2999]
    rc = M.rotation_coefs(ang[ids])
    rng = np.random.default_rng(11)
    t = rng.uniform(0, nf, (len(ids), 3)).astype(np.float32)
    _say("spectral score nf=%d, %d poses" % (nf, len(ids)))
    got = ctx.score_spectral(rc, t)
    remb = O.embed(ng, nf, gr, gi)
    for q, a in enumerate(ids):
        xyz = O.rotate_ligand(lig.xyz, lig.center, O.rotation_coef(ang[a]))
        lr, li = O.ligand_voxel(ng, 1.2, xyz, lig.radius, lig.charge)
        Q = O.folded_spectrum_fp64(nf, remb, O.embed(ng, nf, lr, li))
        s, g = O.spectral_score_fp64(nf, Q, t[q].astype(np.float64))
        scale = np.abs(np.fft.ifftn(Q)).max()  # the score map's largest magnitude / nf^3
        _say("  pose %d: s %.6g vs %.6g, |g| %.4g, grad err %.3e" % (a, got[q, 0], s, np.linalg.norm(g),
                                                                   np.linalg.norm(got[q, 1:] - g) / np.linalg.norm(g)))
        assert abs(got[q, 0] - s) <= 2e-6 * scale, (q, got[q], s)
        assert np.linalg.norm(got[q, 1:] - g) <= 1e-3 * np.linalg.norm(g), (q, got[q], g)
    ctx.close()


@pytest.mark.parametrize("ngs", [None, 128])
def test_spectral_score_at_grid_points_is_the_docked_score(pair, ngs):
    """nf = 96 and 256 (whose kz = 128 Nyquist plane enters the sum with weight 1): the
    band-limited score at each docked translation equals the docked score."""
    rec, lig, ng, gr, gi = pair
    if ngs is not None:
        ng = ngs
        gr, gi = O.receptor_voxel(ng, 1.2, rec)
    ctx = _ctx(lig, ng, gr, gi, t=3)
    ang = O.load_rotations(3600)
    ids = np.array([5, 777, 3000], np.int32)
    rc = M.rotation_coefs(ang[ids])
    _say("spectral at docked grid points nf=%d" % (2 * ng))
    res = ctx.dock(rc, ids)
    t = res["index"][:, 1:].astype(np.float32)
    got = ctx.score_spectral(np.repeat(rc, 3, axis=0), t)
    for q in range(len(res)):
        assert abs(got[q, 0] - res["score"][q]) <= 1e-5 * max(abs(res["score"][q]), 1e-3), (q, got[q], res[q])
    ctx.close()


def kink_free_shift(u):
    """Shift (voxels) that moves the cell boundaries of the trilinear interpolation to the
    middle of the widest gap between the atoms' fractional coordinates u (n,), so that a
    central difference along this axis crosses no kink of the piecewise-linear score."""
    f = np.sort(np.mod(u, 1.0))
    gaps = np.diff(np.concatenate([f, [f[0] + 1.0]]))
    i = int(np.argmax(gaps))
    return float(np.mod(f[i] + 0.5 * gaps[i], 1.0)), float(gaps[i])


def place_for_fd(lig, R, t, gw, ng, nf):
    """Translation near t whose finite differences along x, y and z cross no cell boundary."""
    X = (np.asarray(lig.xyz, np.float64) - np.asarray(lig.center, np.float64)) @ R.T
    t = np.array(t, np.float64)
    for a in range(3):
        tp = t[a] if t[a] < nf / 2 else t[a] - nf
        u = X[:, a] / gw + 0.5 * ng - 0.5 - tp
        d, gap = kink_free_shift(u)
        assert gap > 4e-3 / gw, gap  # room for +-1e-3 A
        t[a] = (t[a] + d) % nf  # u - d: the boundary sits mid-gap
    return t


def test_companion_score_and_gradient_vs_fp64_and_fd(pair):
    """Translation: central differences at h = 1e-3 A from a translation whose
    differences cross no cell boundary.  Rotation: a rotation by 1e-3 rad moves atoms
    10 A from the centre by 0.01 A and crosses boundaries for many of them, where the
    trilinear score is not differentiable, so the rotation check uses h = 1e-5 rad."""
    rec, lig, ng, gr, gi = pair
    nf = 2 * ng
    ctx = _ctx(lig, ng, gr, gi)
    ang = O.load_rotations(3600)
    ids = [3, 1800]
    rc = M.rotation_coefs(ang[ids])
    t = np.array([[1.25, -2.5, 0.75], [nf - 3.4, 2.2, nf - 0.6]], np.float64) % nf
    t = np.array([place_for_fd(lig, rc[q].reshape(3, 3).astype(np.float64), t[q], 1.2, ng, nf)
                  for q in range(len(ids))]).astype(np.float32)
    got = ctx.score_companion(rc, t)
    cen = np.asarray(lig.center, np.float64)
    for q in range(len(ids)):
        R = rc[q].reshape(3, 3).astype(np.float64)
        tq = t[q].astype(np.float64)
        args = (gr, gi, ng, 1.2, lig.xyz, cen, lig.radius, lig.charge)
        s, dt, dw = O.companion_fp64(*args, R, tq, nf)
        _say("companion pose %d: s %.6g vs %.6g" % (ids[q], got[q, 0], s))
        assert abs(got[q, 0] - s) <= 1e-4 * abs(s), (got[q], s)
        assert np.linalg.norm(got[q, 1:4] - dt) <= 1e-3 * np.linalg.norm(dt), (got[q, 1:4], dt)
        assert np.linalg.norm(got[q, 4:7] - dw) <= 1e-3 * np.linalg.norm(dw), (got[q, 4:7], dw)
        ht, hw = 1e-3 / 1.2, 1e-5
        fdt = np.array([(O.companion_fp64(*args, R, tq + ht * e, nf)[0] - O.companion_fp64(*args, R, tq - ht * e, nf)[0])
                        / (2 * ht) for e in np.eye(3)])
        fdw = np.array([(O.companion_fp64(*args, O.rotvec_matrix(hw * e) @ R, tq, nf)[0]
                         - O.companion_fp64(*args, O.rotvec_matrix(-hw * e) @ R, tq, nf)[0]) / (2 * hw) for e in np.eye(3)])
        assert np.linalg.norm(got[q, 1:4] - fdt) <= 1e-3 * np.linalg.norm(fdt), (got[q, 1:4], fdt)
        assert np.linalg.norm(got[q, 4:7] - fdw) <= 1e-3 * np.linalg.norm(fdw), (got[q, 4:7], fdw)
    ctx.close()


def test_scores_over_more_poses_than_the_batch(pair):
    """Both off-grid scores take any number of poses: 21 poses through a context of batch 8
    run as batches of 8, 8 and 5 and equal the results of those batches called one by one
    (same kernels, fixed reduction order: bitwise)."""
    rec, lig, ng, gr, gi = pair
    nf = 2 * ng
    ctx = _ctx(lig, ng, gr, gi, max_batch=8)
    ang = O.load_rotations(3600)
    ids = np.arange(5, 3600, 171)[:21]
    rc = M.rotation_coefs(ang[ids])
    t = np.random.default_rng(4).uniform(0, nf, (len(ids), 3)).astype(np.float32)
    _say("off-grid scores of %d poses, batch 8" % len(ids))
    for fn in (ctx.score_spectral, ctx.score_companion):
        whole = fn(rc, t)
        parts = np.concatenate([fn(rc[q:q + 8], t[q:q + 8]) for q in range(0, len(ids), 8)])
        assert np.array_equal(whole.view(np.uint32), parts.view(np.uint32)), fn.__name__
        assert np.all(np.isfinite(whole))
    assert ctx.score_spectral(rc[:0], t[:0]).shape == (0, 4)
    ctx.close()
