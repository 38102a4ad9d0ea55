"""GPU: the torch.autograd.Function around the companion score and the refinement loop.

The autograd gradient (device gradient mapped through the left Jacobian of SO(3)) is
checked against central finite differences of the FP64 companion away from the identity
(1e-3 A in translation from a kink-free placement, 1e-5 rad in rotation: see
test_gpu_grad.py); refinement must not lower any decoy's companion score,
and the refined poses are re-judged with the hard-voxel score (mdk_score_spectral)."""
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import megadock_amd as M
from megadock_amd import refine as F
from oracle import oracle as O
from test_gpu_grad import place_for_fd

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(golden_dir):
    rec = O.Protein(os.path.join(golden_dir, "1gcq_r.pdb"))
    lig = O.Protein(os.path.join(golden_dir, "1gcq_l.pdb"))
    ng = 48
    gr, gi = O.receptor_voxel(ng, 1.2, rec)
    ctx = M.Context(2 * ng, ng, lig.na, t=1, max_batch=16)
    ctx.set_receptor(gr, gi)
    ctx.set_ligand(lig.xyz, lig.center, lig.radius, lig.charge, O.grid_coord(ng, 1.2), 1.2)
    yield lig, ng, gr, gi, ctx
    ctx.close()


def test_autograd_gradient_matches_fp64_finite_differences(setup):
    lig, ng, gr, gi, ctx = setup
    nf = 2 * ng
    ang = O.load_rotations(3600)
    base = torch.from_numpy(M.rotation_coefs(ang[[41, 2500]]).reshape(-1, 3, 3).astype(np.float64))
    w0 = np.array([[0.05, -0.03, 0.02], [-0.1, 0.04, 0.07]])
    t0 = np.array([place_for_fd(lig, O.rotvec_matrix(w0[q]) @ base[q].numpy(), np.array(tq) % nf, 1.2, ng, nf)
                   for q, tq in enumerate([[1.5, -0.5, 2.25], [0.75, 1.0, -1.5]])])
    w = torch.tensor(w0, dtype=torch.float64, requires_grad=True)
    t = torch.tensor(t0, dtype=torch.float64, requires_grad=True)
    s = F.CompanionScore.apply(w, t, base, ctx)
    s.sum().backward()
    cen = np.asarray(lig.center, np.float64)
    args = (gr, gi, ng, 1.2, lig.xyz, cen, lig.radius, lig.charge)

    def f(wq, tq, q):
        R = O.rotvec_matrix(wq) @ base[q].numpy()
        return O.companion_fp64(*args, R, tq % nf, nf)[0]

    for q in range(2):
        wq, tq = w[q].detach().numpy(), t[q].detach().numpy()
        hw, ht = 1e-5, 1e-3 / 1.2
        fdw = np.array([(f(wq + hw * e, tq, q) - f(wq - hw * e, tq, q)) / (2 * hw) for e in np.eye(3)])
        fdt = np.array([(f(wq, tq + ht * e, q) - f(wq, tq - ht * e, q)) / (2 * ht) for e in np.eye(3)])
        gw, gt = w.grad[q].numpy(), t.grad[q].numpy()
        print("pose %d: dS/dw %s fd %s; dS/dt %s fd %s" % (q, gw, fdw, gt, fdt), flush=True)
        assert np.linalg.norm(gw - fdw) <= 1e-3 * np.linalg.norm(fdw)
        assert np.linalg.norm(gt - fdt) <= 1e-3 * np.linalg.norm(fdt)


def test_refinement_is_monotone_and_rejudged(setup):
    lig, ng, gr, gi, ctx = setup
    ang = O.load_rotations(360)
    rc = M.rotation_coefs(ang)
    ctx.set_rotations(rc)
    ctx.dock_range(0, 360)
    ctx.sync()
    top = ctx.topn(360, 8)
    rot = rc[top["index"][:, 0]]
    t0 = top["index"][:, 1:].astype(np.float32)
    hard0 = ctx.score_spectral(rot, t0)[:, 0]
    assert np.allclose(hard0, top["score"], rtol=1e-5, atol=1e-6)
    R, t, hist = F.refine(ctx, rot, t0, steps=12)
    print("companion first/last:", hist[0], hist[-1], flush=True)
    assert np.all(np.diff(hist, axis=0) >= 0)
    assert np.all(hist[-1] >= hist[0])
    hard1 = ctx.score_spectral(R, t)[:, 0]
    print("hard score before/after:", hard0, hard1, flush=True)
    assert np.all(np.isfinite(hard1))
    assert np.allclose(np.einsum("bij,bkj->bik", R.reshape(-1, 3, 3), R.reshape(-1, 3, 3)),
                       np.eye(3)[None], atol=1e-5)


def test_hard_refinement_never_lowers_the_hard_score(setup):
    lig, ng, gr, gi, ctx = setup
    ang = O.load_rotations(360)
    rc = M.rotation_coefs(ang)
    ctx.set_rotations(rc)
    ctx.dock_range(0, 360)
    ctx.sync()
    top = ctx.topn(360, 8)
    rot = rc[top["index"][:, 0]]
    t0 = top["index"][:, 1:].astype(np.float32)
    R, t, hist = F.refine_hard(ctx, rot, t0, steps=8)
    print("hard score first/last:", hist[0], hist[-1], flush=True)
    assert np.allclose(hist[0], top["score"], rtol=1e-5, atol=1e-6)
    assert np.all(np.diff(hist, axis=0) >= 0)
    # the re-judged score of the returned poses is the last history entry
    assert np.allclose(ctx.score_spectral(R, t)[:, 0], hist[-1], rtol=1e-6)
