"""GPU parity: the HIP path (through the libmdk C ABI) against the CPU oracle.

Bit-exact for voxel grids; FP32 transforms against float64 numpy within the stated
tolerances; per-rotation best scores within 1e-4 relative of the oracle, best
positions identical except at near-ties judged on the FP64 score map.
"""
import os
import sys

import numpy as np
import pytest

import megadock_amd as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FFT_SIZES = [8, 12, 16, 18, 20, 22, 26, 28, 36, 40, 42, 44, 48, 64, 66, 70, 72, 80, 88, 90,
             96, 98, 100, 110, 128, 144, 154, 160, 192, 256]


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def gpu():
    n = M.device_count()
    assert n >= 1, "no GPU visible"
    return 0


@pytest.fixture(scope="module")
def gcq(golden_dir):
    rec = O.Protein(os.path.join(golden_dir, "1gcq_r.pdb"))
    lig = O.Protein(os.path.join(golden_dir, "1gcq_l.pdb"))
    ng = O.autogrid(rec, lig)
    return rec, lig, ng


@pytest.mark.parametrize("n", FFT_SIZES)
def test_fft1d_matches_numpy(gpu, n):
    _say("fft1d n=%d" % n)
    rng = np.random.default_rng(n)
    x = (rng.standard_normal((37, n)) + 1j * rng.standard_normal((37, n))).astype(np.complex64)
    for sign in (-1, 1):
        y = M.fft1d(x, sign)
        ref = np.fft.fft(x.astype(np.complex128), axis=-1) if sign < 0 else np.fft.ifft(x.astype(np.complex128), axis=-1) * n
        err = np.linalg.norm(y - ref) / np.linalg.norm(ref)
        assert err < 5e-7 * max(1.0, np.log2(n)), (n, sign, err)


def _lig_ctx(lig, ng, t=1, max_batch=16):
    nf = 2 * ng
    ctx = M.Context(nf, ng, lig.na, t=t, max_batch=max_batch)
    ctx.set_ligand(lig.xyz, lig.center, lig.radius, lig.charge, O.grid_coord(ng, 1.2), 1.2)
    return ctx


@pytest.fixture(params=["specialised", "generic"])
def kernels(request):
    """Run with the compile-time specialised pipeline (when built for the size) and with
    the generic mixed-radix kernels (MDK_FORCE_GENERIC=1, read when a context is made)."""
    old = os.environ.get("MDK_FORCE_GENERIC")
    os.environ["MDK_FORCE_GENERIC"] = "1" if request.param == "generic" else "0"
    yield request.param
    if old is None:
        os.environ.pop("MDK_FORCE_GENERIC", None)
    else:
        os.environ["MDK_FORCE_GENERIC"] = old


@pytest.fixture(scope="module")
def synth(golden_dir):
    rec = O.Protein(os.path.join(golden_dir, "synth_r120.pdb"))
    lig = O.Protein(os.path.join(golden_dir, "synth_l40.pdb"))
    return rec, lig


def test_voxelize_bit_exact_1gcq(gpu, gcq):
    rec, lig, ng = gcq
    _say("voxelize 1gcq ng=%d" % ng)
    ctx = _lig_ctx(lig, ng)
    ang = O.load_rotations(3600)
    rots = list(range(0, 3600, 50))
    for b0 in range(0, len(rots), 16):
        ids = rots[b0:b0 + 16]
        rc = M.rotation_coefs(ang[ids])
        gr, gi = ctx.voxelize(rc)
        for q, a in enumerate(ids):
            xyz = O.rotate_ligand(lig.xyz, lig.center, O.rotation_coef(ang[a]))
            er, ei = O.ligand_voxel(ng, 1.2, xyz, lig.radius, lig.charge)
            assert np.array_equal(gr[q], er), ("grid_r", a)
            assert np.array_equal(gi[q], ei), ("grid_i", a)
    ctx.close()


@pytest.mark.parametrize("delta,ng", [(2.0, None), (1.0, 24), (-3.5, 30), (2.0, 80), (1.0, 88)])
def test_voxelize_bit_exact_edge(gpu, gcq, delta, ng):
    """core value != 1 (the extra surface sweep), grids too small for the ligand (atoms
    outside the box: clamped charges, truncated search boxes) and grids whose bit masks
    exceed one workgroup's LDS (ng 80, 88: the global-memory voxelizer)."""
    rec, lig, ng0 = gcq
    ng = ng or ng0
    _say("voxelize edge delta=%g ng=%d" % (delta, ng))
    nf = 2 * ng
    ctx = M.Context(nf, ng, lig.na, t=1, max_batch=8)
    ctx.set_ligand(lig.xyz, lig.center, lig.radius, lig.charge, O.grid_coord(ng, 1.2), 1.2, delta)
    ang = O.load_rotations(360)
    ids = [0, 7, 99, 200, 311, 359]
    gr, gi = ctx.voxelize(M.rotation_coefs(ang[ids]))
    for q, a in enumerate(ids):
        xyz = O.rotate_ligand(lig.xyz, lig.center, O.rotation_coef(ang[a]))
        er, ei = O.ligand_voxel(ng, 1.2, xyz, lig.radius, lig.charge, delta)
        assert np.array_equal(gr[q], er), ("grid_r", a)
        assert np.array_equal(gi[q], ei), ("grid_i", a)
    ctx.close()


def test_receptor_spectrum_vs_fp64(gpu, gcq):
    rec, lig, ng = gcq
    nf = 2 * ng
    _say("receptor spectrum nf=%d" % nf)
    gr, gi = O.receptor_voxel(ng, 1.2, rec)
    ctx = _lig_ctx(lig, ng)
    ctx.set_receptor(gr, gi)
    re, im = ctx.receptor_spectrum()
    emb = O.embed(ng, nf, gr, gi).astype(np.complex128).reshape(nf, nf, nf)
    ref = np.fft.fftn(emb).ravel()
    got = re.astype(np.float64) + 1j * im.astype(np.float64)
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    _say("  rel L2 err %.3e" % rel)
    assert rel < 2e-6
    ctx.close()


@pytest.mark.parametrize("which,ng", [("gcq", None), ("synth", 48), ("synth", 72), ("synth", 128), ("synth", 50)])
def test_score_map_vs_fp64(gpu, gcq, synth, kernels, which, ng):
    """nf = 80, 96, 144, 256 (tuned specialised kernels) and 100 (specialised with the
    default variants; the generic kernels through the `kernels` fixture)."""
    if which == "gcq":
        rec, lig, ng = gcq
    else:
        rec, lig = synth
    nf = 2 * ng
    gr, gi = O.receptor_voxel(ng, 1.2, rec)
    ctx = _lig_ctx(lig, ng)
    ctx.set_receptor(gr, gi)
    ang = O.load_rotations(3600)
    remb = O.embed(ng, nf, gr, gi)
    for a in ((0, 1234, 3599) if nf < 200 else (1234,)):
        _say("score map nf=%d rotation %d (%s)" % (nf, a, kernels))
        smap = ctx.score_map(M.rotation_coefs(ang[a:a + 1])[0]).astype(np.float64)
        xyz = O.rotate_ligand(lig.xyz, lig.center, O.rotation_coef(ang[a]))
        lr, li = O.ligand_voxel(ng, 1.2, xyz, lig.radius, lig.charge)
        ref = O.correlation_fp64(nf, remb, O.embed(ng, nf, lr, li)).ravel()
        err = np.abs(smap - ref).max() / np.abs(ref).max()
        _say("  max err / max|score| = %.3e" % err)
        assert err < 2e-6
    ctx.close()


def _near_tie_ok(smap64, i_gpu, i_ref, tol):
    return abs(smap64[i_gpu] - smap64[i_ref]) <= tol * abs(smap64).max()


@pytest.mark.parametrize("t,which,ng", [(1, "gcq", None), (3, "gcq", None), (3, "synth", 72), (2, "synth", 48)])
def test_dock_vs_oracle(gpu, gcq, synth, kernels, t, which, ng):
    if which == "gcq":
        rec, lig, ng = gcq
    else:
        rec, lig = synth
    nf = 2 * ng
    gr, gi = O.receptor_voxel(ng, 1.2, rec)
    rr, ri = O.receptor_fft(ng, nf, gr, gi)
    ctx = _lig_ctx(lig, ng, t=t)
    ctx.set_receptor(gr, gi)
    ang = O.load_rotations(3600)
    ids = np.arange(0, 3600, 150 if nf < 120 else 450, dtype=np.int32)
    _say("dock %d rotations t=%d nf=%d (%s)" % (len(ids), t, nf, kernels))
    got = np.concatenate([ctx.dock(M.rotation_coefs(ang[ids[q:q + 16]]), ids[q:q + 16])
                          for q in range(0, len(ids), 16)])
    ref = O.dock(ng, nf, 1.2, lig, rr, ri, ang, ids, t, threads=8)
    remb = O.embed(ng, nf, gr, gi)
    nf3 = nf ** 3
    for q in range(len(ids) * t):
        g, r = got[q], ref[q]
        assert g["index"][0] == r["index"][0]
        assert abs(g["score"] - r["score"]) <= 1e-4 * max(abs(r["score"]), 1e-3), (q, g, r)
        if not np.array_equal(g["index"], r["index"]):
            a = int(ids[q // t])
            xyz = O.rotate_ligand(lig.xyz, lig.center, O.rotation_coef(ang[a]))
            lr, li = O.ligand_voxel(ng, 1.2, xyz, lig.radius, lig.charge)
            s64 = O.correlation_fp64(nf, remb, O.embed(ng, nf, lr, li)).ravel()
            ig = (g["index"][1] * nf + g["index"][2]) * nf + g["index"][3]
            ir = (r["index"][1] * nf + r["index"][2]) * nf + r["index"][3]
            assert _near_tie_ok(s64, ig, ir, 2e-6), (q, g, r, s64[ig] / nf3, s64[ir] / nf3)
    ctx.close()


@pytest.mark.parametrize("t", [1, 2])
@pytest.mark.parametrize("nf", [n for n in M.FURTHER_NF if n <= 192])
def test_further_sizes_vs_fp64_and_oracle(gpu, synth, nf, t):
    """The sizes specialised with the default kernel variants only (a table farm meets
    them: nf = 2 ng from the reference's FFT size table, including ng odd, where the
    ligand block sits at floor(ng/2)): score map vs FP64 and per-rotation top-t vs the
    oracle, as for the tuned sizes, for t = 1 (the reference's default: raw keys, a 0.0
    starting value) and t = 2 (keys normalised by nf^3, a -99999 starting value)."""
    rec, lig = synth
    ng = nf // 2
    gr, gi = O.receptor_voxel(ng, 1.2, rec)
    rr, ri = O.receptor_fft(ng, nf, gr, gi)
    ctx = _lig_ctx(lig, ng, t=t, max_batch=8)
    ctx.set_receptor(gr, gi)
    ang = O.load_rotations(3600)
    remb = O.embed(ng, nf, gr, gi)
    if t == 2:
        _say("further size nf=%d: score map" % nf)
        a = 1234
        smap = ctx.score_map(M.rotation_coefs(ang[a:a + 1])[0]).astype(np.float64)
        xyz = O.rotate_ligand(lig.xyz, lig.center, O.rotation_coef(ang[a]))
        lr, li = O.ligand_voxel(ng, 1.2, xyz, lig.radius, lig.charge)
        ref = O.correlation_fp64(nf, remb, O.embed(ng, nf, lr, li)).ravel()
        assert np.abs(smap - ref).max() <= 2e-6 * np.abs(ref).max()
    ids = np.arange(7 + t, 3600, 900 if nf >= 160 else 450, dtype=np.int32)
    _say("further size nf=%d: dock %d rotations, t=%d" % (nf, len(ids), t))
    got = ctx.dock(M.rotation_coefs(ang[ids]), ids)
    exp = O.dock(ng, nf, 1.2, lig, rr, ri, ang, ids, t, threads=16)
    for q in range(len(ids) * t):
        g, r = got[q], exp[q]
        assert g["index"][0] == r["index"][0]
        assert abs(g["score"] - r["score"]) <= 1e-4 * max(abs(r["score"]), 1e-3), (q, g, r)
        if not np.array_equal(g["index"], r["index"]):
            b = int(ids[q // t])
            xyz = O.rotate_ligand(lig.xyz, lig.center, O.rotation_coef(ang[b]))
            lr, li = O.ligand_voxel(ng, 1.2, xyz, lig.radius, lig.charge)
            s64 = O.correlation_fp64(nf, remb, O.embed(ng, nf, lr, li)).ravel()
            ig = (g["index"][1] * nf + g["index"][2]) * nf + g["index"][3]
            ir = (r["index"][1] * nf + r["index"][2]) * nf + r["index"][3]
            assert _near_tie_ok(s64, ig, ir, 2e-6), (q, g, r)
    ctx.close()


@pytest.mark.parametrize("nf,t", [(n, 2) for n in M.FURTHER_NF if n > 192] + [(n, 1) for n in (196, 250, 320)])
def test_sizes_above_192_vs_oracle(gpu, synth, nf, t):
    """The table sizes above 192 (specialised with the passes that do not need a plane in
    LDS: forward z then y through HBM, the correlation pass, inverse y in place, then the
    z step over row groups, the last one partial where 16 does not divide nf): per-rotation
    top-2 against the oracle on two rotations, and at three sizes the reference's default
    t = 1 (raw score keys); at 200, 252 and 320 also the score map against FP64."""
    rec, lig = synth
    ng = nf // 2
    gr, gi = O.receptor_voxel(ng, 1.2, rec)
    rr, ri = O.receptor_fft(ng, nf, gr, gi)
    ctx = _lig_ctx(lig, ng, t=t, max_batch=4)
    ctx.set_receptor(gr, gi)
    ang = O.load_rotations(3600)
    remb = O.embed(ng, nf, gr, gi)
    if nf in (200, 252, 320) and t == 2:
        _say("size nf=%d: score map" % nf)
        a = 2345
        smap = ctx.score_map(M.rotation_coefs(ang[a:a + 1])[0]).astype(np.float64)
        xyz = O.rotate_ligand(lig.xyz, lig.center, O.rotation_coef(ang[a]))
        lr, li = O.ligand_voxel(ng, 1.2, xyz, lig.radius, lig.charge)
        ref = O.correlation_fp64(nf, remb, O.embed(ng, nf, lr, li)).ravel()
        assert np.abs(smap - ref).max() <= 2e-6 * np.abs(ref).max()
        del smap, ref
    ids = np.array([11, 2000], dtype=np.int32)
    _say("size nf=%d: dock %d rotations, t = %d" % (nf, len(ids), t))
    got = ctx.dock(M.rotation_coefs(ang[ids]), ids)
    exp = O.dock(ng, nf, 1.2, lig, rr, ri, ang, ids, t, threads=16)
    for q in range(len(ids) * t):
        g, r = got[q], exp[q]
        assert g["index"][0] == r["index"][0]
        assert abs(g["score"] - r["score"]) <= 1e-4 * max(abs(r["score"]), 1e-3), (q, g, r)
        if not np.array_equal(g["index"], r["index"]):
            b = int(ids[q // t])
            xyz = O.rotate_ligand(lig.xyz, lig.center, O.rotation_coef(ang[b]))
            lr, li = O.ligand_voxel(ng, 1.2, xyz, lig.radius, lig.charge)
            s64 = O.correlation_fp64(nf, remb, O.embed(ng, nf, lr, li)).ravel()
            ig = (g["index"][1] * nf + g["index"][2]) * nf + g["index"][3]
            ir = (r["index"][1] * nf + r["index"][2]) * nf + r["index"][3]
            assert _near_tie_ok(s64, ig, ir, 2e-6), (q, g, r)
    ctx.close()


def test_dock_range_matches_batches_and_topn(gpu, gcq, kernels):
    rec, lig, ng = gcq
    nf = 2 * ng
    gr, gi = O.receptor_voxel(ng, 1.2, rec)
    ctx = _lig_ctx(lig, ng, t=2, max_batch=32)
    ctx.set_receptor(gr, gi)
    ang = O.load_rotations(360)
    rc = M.rotation_coefs(ang)
    ctx.set_rotations(rc)
    ctx.dock_range(0, 360)
    ctx.sync()
    res = ctx.results(0, 360)
    _say("dock_range 360 rotations")
    # same kernels, different batch composition: bitwise identical per rotation
    ids = np.arange(100, 117, dtype=np.int32)
    b = ctx.dock(rc[ids], ids)
    assert np.array_equal(b.view(np.uint8), res[100 * 2:117 * 2].view(np.uint8))
    top = ctx.topn(360, 50)
    order = sorted(range(len(res)), key=lambda i: (-res[i]["score"], i))[:50]
    assert np.array_equal(top.view(np.uint8), res[order].view(np.uint8))
    ctx.close()


@pytest.mark.parametrize("nf,nrot,t,n", [(96, 360, 1, 100), (144, 240, 3, 200)])
def test_topn_overlap_with_oracle(gpu, synth, nf, nrot, t, n):
    """The decoys the program writes: the device's top n over a rotation sample
    (mdk_dock_range + mdk_topn) against the oracle's loop and merge over the same
    rotations (src/fft_process.cpp:407-498).  At least 99% of the n (rotation, translation)
    entries agree as a set, and an entry both keep has the same score within 1e-4."""
    rec, lig = synth
    ng = nf // 2
    gr, gi = O.receptor_voxel(ng, 1.2, rec)
    rr, ri = O.receptor_fft(ng, nf, gr, gi)
    ang = O.load_rotations(54000)[::54000 // nrot][:nrot]
    ctx = _lig_ctx(lig, ng, t=t, max_batch=64)
    ctx.set_receptor(gr, gi)
    ctx.set_rotations(M.rotation_coefs(ang))
    _say("top-%d overlap nf=%d: dock %d rotations" % (n, nf, nrot))
    ctx.dock_range(0, nrot)
    ctx.sync()
    top = ctx.topn(nrot, n)
    ctx.close()
    _say("top-%d overlap nf=%d: oracle" % (n, nf))
    ref = O.sort_topn(O.dock(ng, nf, 1.2, lig, rr, ri, ang, np.arange(nrot, dtype=np.int32), t, threads=16), n)
    got = {tuple(int(v) for v in e["index"]): float(e["score"]) for e in top}
    exp = {tuple(int(v) for v in e["index"]): float(e["score"]) for e in ref}
    common = set(got) & set(exp)
    _say("top-%d overlap nf=%d: %d common" % (n, nf, len(common)))
    assert len(common) >= 0.99 * n
    for k in common:
        assert abs(got[k] - exp[k]) <= 1e-4 * max(abs(exp[k]), 1e-3), (k, got[k], exp[k])
    # the scores ranked r-th agree as well, entry sets aside
    assert np.allclose(top["score"], ref["score"], rtol=1e-4, atol=0)


@pytest.mark.parametrize("t", [1, 3])
@pytest.mark.parametrize("ng", [40, 72, 128])
def test_dock_all_tied_scores(gpu, gcq, kernels, t, ng):
    """A zero receptor spectrum makes every score exactly 0: the reference's rules give
    (0, index 0) for t == 1 (nothing beats the starting 0.0) and the first t grid indices
    for t > 1 (0 > -99999, ties kept in index order), src/fft_process.cpp:407-455."""
    rec, lig, _ = gcq
    nf = 2 * ng
    ctx = _lig_ctx(lig, ng, t=t, max_batch=4)
    z = np.zeros(nf ** 3, np.float32)
    ctx.set_receptor_spectrum(z, z)
    ang = O.load_rotations(24)
    ids = np.array([3, 17, 5], np.int32)
    _say("all-tied dock t=%d nf=%d (%s)" % (t, nf, kernels))
    got = ctx.dock(M.rotation_coefs(ang[ids]), ids)
    for q in range(len(ids)):
        for k in range(t):
            g = got[q * t + k]
            assert g["score"] == 0.0, (q, k, g)
            assert list(g["index"]) == [int(ids[q]), 0, 0, k], (q, k, g)
    ctx.close()


@pytest.mark.parametrize("nrot,t,n", [(360, 2, 720), (4096, 3, 10000), (4096, 3, 1)])
def test_topn_device_merge(gpu, gcq, nrot, t, n):
    """mdk_topn merges on the device: (score desc, rotation asc, rank asc), as a stable
    sort of the per-rotation results by descending score.  n = 10000 takes the
    global-memory sorting network (more than 8192 keys)."""
    rec, lig, ng = gcq
    gr, gi = O.receptor_voxel(ng, 1.2, rec)
    ctx = _lig_ctx(lig, ng, t=t, max_batch=64)
    ctx.set_receptor(gr, gi)
    ang = O.load_rotations(54000)[:nrot]
    ctx.set_rotations(M.rotation_coefs(ang))
    ctx.dock_range(0, nrot)
    ctx.sync()
    res = ctx.results(0, nrot)
    _say("topn %d of %d results" % (n, len(res)))
    top = ctx.topn(nrot, n)
    order = np.argsort(-res["score"].astype(np.float64), kind="stable")[:n]
    assert np.array_equal(top.view(np.uint8), res[order].view(np.uint8))
    ctx.close()


def test_topn_started_then_finished_equals_topn(gpu, gcq):
    """mdk_topn_start queues the merge behind the queued batches and returns; work of
    another context queued meanwhile does not change it; mdk_topn_finish returns what
    mdk_topn returns.  A second start before the finish, and a finish without a start,
    are errors."""
    rec, lig, ng = gcq
    gr, gi = O.receptor_voxel(ng, 1.2, rec)
    rc = M.rotation_coefs(O.load_rotations(3600)[:300])
    a, b = _lig_ctx(lig, ng, t=2, max_batch=64), _lig_ctx(lig, ng, t=1, max_batch=32)
    for c in (a, b):
        c.set_receptor(gr, gi)
        c.set_rotations(rc)
    _say("started merge over 300 rotations, another context queued behind it")
    a.dock_range(0, 300)
    a.topn_start(300, 200)
    with pytest.raises(M.MdkError):
        a.topn_start(300, 200)
    b.dock_range(0, 300)
    started = a.topn_finish()
    with pytest.raises(M.MdkError):
        a.topn_finish()
    assert len(started) == 200
    assert np.array_equal(started.view(np.uint8), a.topn(300, 200).view(np.uint8))
    b.sync()
    a.close()
    b.close()


def test_setup_calls_queue_behind_docking(gpu, gcq, synth):
    """mdk_set_receptor and mdk_set_ligand_atoms return without waiting for the batches
    queued before them: those batches see the previous receptor and ligand, the batches
    queued after see the new ones, bit for bit as contexts set up once with each pair."""
    rec, lig, ng = gcq
    _, lig2 = synth
    gr, gi = O.receptor_voxel(ng, 1.2, rec)
    gr2 = np.roll(gr.reshape(ng, ng, ng), 5, axis=0).ravel().copy()
    gi2 = np.roll(gi.reshape(ng, ng, ng), 5, axis=0).ravel().copy()
    rc = M.rotation_coefs(O.load_rotations(3600)[:300])
    gc = O.grid_coord(ng, 1.2)

    def ctx_with(r, l):
        ctx = M.Context(2 * ng, ng, lig.na, t=2, max_batch=64)
        ctx.set_receptor(*r)
        ctx.set_ligand(l.xyz, l.center, l.radius, l.charge, gc, 1.2)
        ctx.set_rotations(rc)
        return ctx

    refs = []
    for r, l in (((gr, gi), lig), ((gr2, gi2), lig), ((gr2, gi2), lig2)):
        ctx = ctx_with(r, l)
        ctx.dock_range(0, 300)
        ctx.sync()
        refs.append(ctx.results(0, 300))
        ctx.close()
    _say("set-up calls between queued ranges")
    ctx = ctx_with((gr, gi), lig)
    ctx.dock_range(0, 100)
    ctx.set_receptor(gr2, gi2)  # queued behind rotations 0..99
    ctx.dock_range(100, 100)
    ctx.set_ligand(lig2.xyz, lig2.center, lig2.radius, lig2.charge, gc, 1.2)  # behind 100..199
    ctx.dock_range(200, 100)
    ctx.sync()
    got = ctx.results(0, 300)
    ctx.close()
    for k in range(3):
        sl = slice(200 * k, 200 * (k + 1))  # t = 2 results per rotation
        assert np.array_equal(got[sl].view(np.uint8), refs[k][sl].view(np.uint8)), k
    assert not np.array_equal(refs[0][:200].view(np.uint8), refs[1][:200].view(np.uint8))


def test_topn_device_merge_all_tied(gpu, gcq):
    """Every score 0 (zero receptor spectrum): the merge keeps result order."""
    rec, lig, ng = gcq
    nf = 2 * ng
    ctx = _lig_ctx(lig, ng, t=3, max_batch=16)
    z = np.zeros(nf ** 3, np.float32)
    ctx.set_receptor_spectrum(z, z)
    ang = O.load_rotations(360)[:40]
    ctx.set_rotations(M.rotation_coefs(ang))
    ctx.dock_range(0, 40)
    ctx.sync()
    res = ctx.results(0, 40)
    top = ctx.topn(40, 100)
    assert np.array_equal(top.view(np.uint8), res[:100].view(np.uint8))
    ctx.close()


@pytest.mark.parametrize("ng,env", [(48, {"MDK_XCORR_VARIANT": v}) for v in "01234567"] +
                         [(50, {"MDK_XCORR_VARIANT": v}) for v in "57"] + [(50, {"MDK_INV_VARIANT": "2"}),
                                                                           (50, {"MDK_FWD_VARIANT": "2"})] +
                         [(48, {"MDK_INV_VARIANT": v}) for v in "123"] +
                         [(48, {"MDK_STREAMS": "1"}), (48, {"MDK_FWD_VARIANT": "1"}), (48, {"MDK_FWD_VARIANT": "2"})] +
                         [(72, {"MDK_XCORR_VARIANT": v}) for v in "4567"] + [(72, {"MDK_FWD_VARIANT": v}) for v in "12"] +
                         [(40, {"MDK_XCORR_VARIANT": "7"}), (128, {"MDK_XCORR_VARIANT": "7"})])
def test_tuning_variants_give_identical_results(gpu, gcq, ng, env):
    """Every kernel variant computes the same operations in the same order: results are
    bitwise equal to the default's (nf = 80, 96, 144 and 256)."""
    rec, lig, _ = gcq
    gr, gi = O.receptor_voxel(ng, 1.2, rec)
    ang = O.load_rotations(3600)[:300]
    rc = M.rotation_coefs(ang)

    def run():
        ctx = _lig_ctx(lig, ng, t=3, max_batch=64)
        ctx.set_receptor(gr, gi)
        ctx.set_rotations(rc)
        ctx.dock_range(0, len(rc))
        ctx.sync()
        out = ctx.results(0, len(rc))
        ctx.close()
        return out

    base = run()
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        _say("variant %s" % env)
        other = run()
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    assert np.array_equal(base.view(np.uint8), other.view(np.uint8))


def _with_env(env, fn):
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return fn()
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


@pytest.mark.parametrize("ng,t", [(48, 1), (48, 2), (48, 4), (40, 7), (72, 3), (48, 16)])
@pytest.mark.parametrize("tied", [False, True])
def test_segmented_top_t_lists(gpu, gcq, ng, t, tied):
    """Fused-inverse variants 2 and 3 keep running top-t lists over the planes of a rotation
    that one workgroup walks (segments, some cut short where a workgroup's range of planes
    ends inside a rotation); per-rotation results must be bitwise those of the per-plane
    lists (variant 0), for t over every compile-time list length, for a batch whose plane
    count does not split evenly, and with every score tied."""
    rec, lig, _ = gcq
    nf = 2 * ng
    ang = O.load_rotations(3600)[:150]
    rc = M.rotation_coefs(ang)
    if tied:
        z = np.zeros(nf ** 3, np.float32)

    def run():
        ctx = _lig_ctx(lig, ng, t=t, max_batch=53)  # batches of 53, 53 and 44 rotations
        if tied:
            ctx.set_receptor_spectrum(z, z)
        else:
            ctx.set_receptor(*O.receptor_voxel(ng, 1.2, rec))
        ctx.set_rotations(rc)
        ctx.dock_range(0, len(rc))
        ctx.sync()
        out = ctx.results(0, len(rc))
        ctx.close()
        return out

    _say("segmented top-t nf=%d t=%d tied=%s" % (nf, t, tied))
    base = _with_env({"MDK_INV_VARIANT": "0"}, run)
    for v in "23":
        other = _with_env({"MDK_INV_VARIANT": v}, run)
        assert np.array_equal(base.view(np.uint8), other.view(np.uint8)), v
    if tied:
        for q in range(len(rc)):
            assert list(base[q * t]["index"][1:]) == [0, 0, 0]


def test_autotune_keeps_results(gpu, gcq):
    rec, lig, _ = gcq
    ng = 48
    gr, gi = O.receptor_voxel(ng, 1.2, rec)
    rc = M.rotation_coefs(O.load_rotations(3600)[:256])
    ctx = _lig_ctx(lig, ng, t=2, max_batch=64)
    ctx.set_receptor(gr, gi)
    ctx.set_rotations(rc)
    ctx.dock_range(0, 256)
    ctx.sync()
    before = ctx.results(0, 256)
    xv, iv, fv = ctx.autotune(0, 128)
    _say("autotune chose correlation variant %d, inverse variant %d, forward variant %d" % (xv, iv, fv))
    assert 3 <= xv <= 7 and 0 <= iv <= 3 and fv in (0, 1, 2)
    ctx.dock_range(0, 256)
    ctx.sync()
    assert np.array_equal(before.view(np.uint8), ctx.results(0, 256).view(np.uint8))
    ctx.close()


def test_graph_replays_match_eager(gpu, gcq):
    """MDK_GRAPH=1: the full batches of mdk_dock_range replay a captured graph.  Three
    ranges in a row, with eager batches (the tails) and synchronisations between the
    replays, each checked against the eager path; then a new rotation table (the graph is
    captured again)."""
    rec, lig, _ = gcq
    ng = 48
    gr, gi = O.receptor_voxel(ng, 1.2, rec)
    ang = O.load_rotations(3600)
    rc1, rc2 = M.rotation_coefs(ang[:400]), M.rotation_coefs(ang[1000:1300])

    def make():
        ctx = _lig_ctx(lig, ng, t=3, max_batch=64)
        ctx.set_receptor(gr, gi)
        return ctx

    eager = make()
    eager.set_rotations(rc1)
    eager.dock_range(0, 400)
    eager.sync()
    ref1 = eager.results(0, 400)
    eager.set_rotations(rc2)
    eager.dock_range(0, 300)
    eager.sync()
    ref2 = eager.results(0, 300)
    eager.close()
    old = os.environ.get("MDK_GRAPH")
    os.environ["MDK_GRAPH"] = "1"
    try:
        g = make()
    finally:
        if old is None:
            os.environ.pop("MDK_GRAPH", None)
        else:
            os.environ["MDK_GRAPH"] = old
    g.set_rotations(rc1)
    for first, count in [(0, 400), (64, 200), (7, 333)]:
        _say("graph replay range %d+%d" % (first, count))
        g.dock_range(first, count)
        g.sync()
        got = g.results(first, count)
        assert np.array_equal(got.view(np.uint8), ref1[first * 3:(first + count) * 3].view(np.uint8))
    g.set_rotations(rc2)
    g.dock_range(0, 300)
    g.sync()
    assert np.array_equal(g.results(0, 300).view(np.uint8), ref2.view(np.uint8))
    g.close()


_NF256_REF = {}


@pytest.mark.parametrize("inv,fwd", [("0", "1"), ("2", "0"), ("2", "1"), ("0", "0"), ("1", "0")])
def test_nf256_inverse_forms_vs_oracle(gpu, synth, inv, fwd):
    """At nf = 256 a plane does not fit one workgroup's LDS.  The default fused forward pass
    takes each z-transformed plane as its even-kz and odd-kz halves and splits the y
    transform by ky parity (MDK_FWD_VARIANT 0), variant 1 runs the z and y passes apart; the
    default fused inverse takes each plane as its even and odd rows (a radix-2 split of the
    y transform, MDK_INV_VARIANT 2-3; variant 1 with 512 threads reading the plane once),
    variant 0 runs the in-place y pass and the z pass apart.  Each pairing against the
    oracle, per rotation, for t = 1 and t = 3 lists."""
    rec, lig = synth
    ng, nf = 128, 256
    gr, gi = O.receptor_voxel(ng, 1.2, rec)
    remb = O.embed(ng, nf, gr, gi)
    ang = O.load_rotations(3600)
    ids = np.array([5, 1717, 3333], dtype=np.int32)
    for t in (1, 3):
        env = {"MDK_INV_VARIANT": inv, "MDK_FWD_VARIANT": fwd}
        got = _with_env(env, lambda: _dock_ids(lig, ng, t, gr, gi, ang, ids))
        _say("nf 256 inverse variant %s, forward variant %s, t %d: %d rotations" % (inv, fwd, t, len(ids)))
        if t not in _NF256_REF:
            rr, ri = O.receptor_fft(ng, nf, gr, gi)
            _NF256_REF[t] = O.dock(ng, nf, 1.2, lig, rr, ri, ang, ids, t, threads=16)
        ref = _NF256_REF[t]
        for q in range(len(ids) * t):
            g, r = got[q], ref[q]
            assert g["index"][0] == r["index"][0]
            assert abs(g["score"] - r["score"]) <= 1e-4 * max(abs(r["score"]), 1e-3), (q, g, r)
            if not np.array_equal(g["index"], r["index"]):
                a = int(ids[q // t])
                xyz = O.rotate_ligand(lig.xyz, lig.center, O.rotation_coef(ang[a]))
                lr, li = O.ligand_voxel(ng, 1.2, xyz, lig.radius, lig.charge)
                s64 = O.correlation_fp64(nf, remb, O.embed(ng, nf, lr, li)).ravel()
                ig = (g["index"][1] * nf + g["index"][2]) * nf + g["index"][3]
                ir = (r["index"][1] * nf + r["index"][2]) * nf + r["index"][3]
                assert _near_tie_ok(s64, ig, ir, 2e-6), (q, g, r)


def _dock_ids(lig, ng, t, gr, gi, ang, ids):
    ctx = _lig_ctx(lig, ng, t=t, max_batch=4)
    ctx.set_receptor(gr, gi)
    out = ctx.dock(M.rotation_coefs(ang[ids]), ids)
    ctx.close()
    return out
