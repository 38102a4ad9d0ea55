"""The oracle's FFT, product, top-t and merge pinned by mathematics.

FFTW is not in this image, so the FFT stage is pinned by numpy's float64 FFT, a
brute-force DFT, and closed forms (delta -> constant spectrum, plane wave -> delta,
shift theorem, Parseval).  The composite embed+FFT+product+IFFT is pinned by a direct
real-space correlation using the reference's own FFT-index -> translation map
(src/docking.cpp:716, tem = (3 ng - t) mod 2 ng).  Top-t follows the reference's
initial values and strict comparisons (src/fft_process.cpp:407-446) and is pinned by
brute force on maps built with ties.
"""
import numpy as np
import pytest

from oracle import oracle as O

SIZES = [2, 3, 4, 5, 6, 7, 8, 10, 11, 12, 13, 14, 16, 18, 20, 22, 26, 28, 33, 35, 40, 44, 48,
         49, 55, 64, 66, 70, 72, 77, 80, 96, 98, 121, 128, 144, 154, 160, 192]


def brute_dft(x, sign):
    n = x.shape[-1]
    k = np.arange(n)
    w = np.exp(sign * 2j * np.pi * np.outer(k, k) / n)
    return x @ w.T


@pytest.mark.parametrize("n", SIZES)
def test_fft1d_vs_numpy(n):
    rng = np.random.default_rng(n)
    x = (rng.standard_normal((6, n)) + 1j * rng.standard_normal((6, n))).astype(np.complex64)
    for sign in (-1, 1):
        y = O.fft1d(n, x, sign)
        ref = np.fft.fft(x.astype(np.complex128), axis=-1) if sign < 0 else np.fft.ifft(x.astype(np.complex128), axis=-1) * n
        assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 4e-7 * max(1, np.log2(n))
        if n <= 16:
            assert np.allclose(y, brute_dft(x.astype(np.complex128), sign), rtol=0, atol=1e-5 * np.abs(ref).max())


def test_fft3d_closed_forms():
    n = 12
    d = np.zeros((n, n, n), np.complex64)
    d[0, 0, 0] = 1.0
    assert np.allclose(O.fft3d(n, d, -1), 1.0, atol=1e-6)  # delta -> constant
    g = np.indices((n, n, n))
    k = (2, 5, 7)
    wave = np.exp(2j * np.pi * (k[0] * g[0] + k[1] * g[1] + k[2] * g[2]) / n).astype(np.complex64)
    spec = O.fft3d(n, wave, -1).reshape(n, n, n)
    expect = np.zeros((n, n, n))
    expect[k] = n ** 3
    assert np.abs(spec - expect).max() < 1e-3  # plane wave -> delta at k
    rng = np.random.default_rng(3)
    x = (rng.standard_normal((n, n, n)) + 1j * rng.standard_normal((n, n, n))).astype(np.complex64)
    X = O.fft3d(n, x, -1).reshape(n, n, n)
    xs = np.roll(x, (1, 2, 3), axis=(0, 1, 2))
    Xs = O.fft3d(n, xs, -1).reshape(n, n, n)
    phase = np.exp(-2j * np.pi * (1 * g[0] + 2 * g[1] + 3 * g[2]) / n)
    assert np.abs(Xs - X * phase).max() < 2e-4 * np.abs(X).max()  # shift theorem
    assert abs(np.sum(np.abs(X) ** 2) / n ** 3 - np.sum(np.abs(x) ** 2)) < 1e-4 * np.sum(np.abs(x) ** 2)  # Parseval
    inv = O.fft3d(n, X, +1).reshape(n, n, n) / n ** 3
    assert np.abs(inv - x).max() < 1e-5


def test_correlation_composite_matches_real_space():
    """embed + forward FFT + conj(R) L + inverse FFT == direct real-space score at the
    translation given by the reference's index map (src/docking.cpp:712-819)."""
    ng = 6
    nf = 2 * ng
    rng = np.random.default_rng(7)
    rec_r = rng.integers(-2, 3, ng ** 3).astype(np.float32)
    rec_i = rng.standard_normal(ng ** 3).astype(np.float32)
    lig_r = (rng.random(ng ** 3) < 0.3).astype(np.float32)
    lig_i = np.where(rng.random(ng ** 3) < 0.2, rng.standard_normal(ng ** 3), 0).astype(np.float32)
    rr, ri = O.receptor_fft(ng, nf, rec_r, rec_i)
    L = O.fft3d(nf, O.embed(ng, nf, lig_r, lig_i), -1)
    out = O.fft3d(nf, O.convolution(rr, ri, L), +1).reshape(nf, nf, nf)
    R3 = rec_r.reshape(ng, ng, ng).astype(np.float64)
    I3 = rec_i.reshape(ng, ng, ng).astype(np.float64)
    l3 = lig_r.reshape(ng, ng, ng).astype(np.float64)
    li3 = lig_i.reshape(ng, ng, ng).astype(np.float64)
    for t in [(0, 0, 0), (1, 2, 3), (11, 0, 5), (6, 6, 6), (3, 10, 1)]:
        tem = [(3 * ng - c) % (2 * ng) for c in t]
        s = 0.0
        for l in range(ng):
            for m in range(ng):
                for n in range(ng):
                    rx, ry, rz = l + tem[0] - ng, m + tem[1] - ng, n + tem[2] - ng
                    if 0 <= rx < ng and 0 <= ry < ng and 0 <= rz < ng:
                        s += R3[rx, ry, rz] * l3[l, m, n] + I3[rx, ry, rz] * li3[l, m, n]
        assert abs(out[t].real / nf ** 3 - s) < 1e-4 * max(1.0, abs(s)), (t, out[t].real, s)  # unnormalised transforms


def _brute_top(vals, t):
    nf3 = vals.size
    if t == 1:
        best, bi = np.float32(0.0), 0
        for i, v in enumerate(vals):
            if best < v:
                best, bi = v, i
        return [(np.float32(best / np.float32(nf3)), bi)]
    raw = (vals / np.float32(nf3)).astype(np.float32)
    order = sorted(range(nf3), key=lambda i: (-raw[i], i))
    sel = [(raw[i], i) for i in order if raw[i] > np.float32(-99999.0)][:t]
    while len(sel) < t:
        sel.append((np.float32(-99999.0), 0))
    return sel


@pytest.mark.parametrize("t", [1, 2, 3, 5])
def test_score_sort_tie_rules(t):
    nf = 6
    rng = np.random.default_rng(t)
    for trial in range(6):
        vals = rng.integers(-3, 4, nf ** 3).astype(np.float32) * np.float32(2.5)
        if trial == 3:
            vals = -np.abs(vals)  # all non-positive: t == 1 reports (0, index 0)
        fo = np.zeros(nf ** 3, np.complex64)
        fo.real = vals
        got = O.score_sort(nf, t, fo, 42)
        exp = _brute_top(vals, t)
        for q in range(t):
            s, i = exp[q]
            assert got[q]["score"] == s
            assert tuple(got[q]["index"]) == (42, i // nf ** 2, (i // nf) % nf, i % nf)


def test_sort_topn_order():
    rng = np.random.default_rng(0)
    top = np.zeros(500, O.SORTSCORE_DTYPE)
    top["score"] = rng.integers(0, 20, 500).astype(np.float32)
    top["index"][:, 0] = np.arange(500) // 2
    top["index"][:, 1] = np.arange(500) % 2
    got = O.sort_topn(top, 100)
    key = sorted(range(500), key=lambda i: (-top[i]["score"], top[i]["index"][0], top[i]["index"][1]))[:100]
    assert np.array_equal(got.view(np.uint8), top[key].view(np.uint8))


def test_dock_loop_equals_stepwise(golden_dir):
    """or_dock (the OpenMP loop) equals the stepwise calls for the same rotations."""
    import os
    rec = O.Protein(os.path.join(golden_dir, "1gcq_r.pdb"))
    lig = O.Protein(os.path.join(golden_dir, "1gcq_l.pdb"))
    ng, nf = 20, 40
    gr, gi = O.receptor_voxel(ng, 1.2, rec)
    rr, ri = O.receptor_fft(ng, nf, gr, gi)
    ang = O.load_rotations(360)
    ids = np.array([0, 5, 200], np.int32)
    got = O.dock(ng, nf, 1.2, lig, rr, ri, ang, ids, 2, threads=3)
    for q, a in enumerate(ids):
        xyz = O.rotate_ligand(lig.xyz, lig.center, O.rotation_coef(ang[a]))
        lr, li = O.ligand_voxel(ng, 1.2, xyz, lig.radius, lig.charge)
        L = O.fft3d(nf, O.embed(ng, nf, lr, li), -1)
        out = O.fft3d(nf, O.convolution(rr, ri, L), +1)
        exp = O.score_sort(nf, 2, out, int(a))
        assert np.array_equal(got[2 * q:2 * q + 2].view(np.uint8), exp.view(np.uint8))


@pytest.mark.parametrize("ng", [6, 7])
def test_detail_sums_are_the_correlation_score(ng):
    """The real-space sums of the detailed output add up to the docking score at the same
    translation (rPSC + ACE = Rr Lr and ELEC = Ri Li over the overlap), which the FFT
    correlation of the padded grids gives independently, for translations that wrap and
    for the edge values -1 and nf a neighbour of the correct position can take."""
    rng = np.random.default_rng(ng)
    nf = 2 * ng
    rr, ri, rp, lr, li = [rng.standard_normal(ng ** 3).astype(np.float32) for _ in range(5)]
    # the unnormalised inverse transform: nf^3 times the real-space correlation
    smap = O.correlation_fp64(nf, O.embed(ng, nf, rr, ri), O.embed(ng, nf, lr, li)).reshape(nf, nf, nf)
    for t in [(0, 0, 0), (1, 2, 3), (nf - 1, ng, 5), (ng // 2, nf - 2, ng + 1), (3, 4, nf - 1)]:
        d = O.detail_sums(ng, rr, ri, rp, lr, li, t)
        total = d.sum() * nf ** 3
        assert abs(total - smap[t]) <= 1e-5 * np.abs(smap).max(), (t, total, smap[t])
        assert d[0] >= 0 and d[2] >= 0 and d[4] >= 0 and d[1] <= 0 and d[3] <= 0 and d[5] <= 0
    for t, same in [((-1, 0, 0), (nf - 1, 0, 0)), ((nf, 3, 1), (0, 3, 1))]:
        assert np.array_equal(O.detail_sums(ng, rr, ri, rp, lr, li, t), O.detail_sums(ng, rr, ri, rp, lr, li, same))


def test_detail_sums_split_and_ace_part():
    """Constant grids: every overlap voxel contributes the same products, so the sums are
    the overlap count times them, with the signs going to the right halves."""
    ng = 8
    one = np.ones(ng ** 3, np.float32)
    d = O.detail_sums(ng, 3 * one, -2 * one, one, 2 * one, 0.5 * one, (1, 0, 3))
    cnt = (ng - 1) * ng * (ng - 3)  # a shift by t voxels overlaps ng - t planes
    assert d[0] == 2.0 * cnt and d[1] == 0.0        # rPSC = 1 * 2
    assert d[2] == 0.0 and d[3] == -1.0 * cnt       # ELEC = -2 * 0.5
    assert d[4] == 4.0 * cnt and d[5] == 0.0        # ACE = (3 - 1) * 2


@pytest.mark.parametrize("t", [(0, 0, 0), (1, 2, 3), (13, 7, 9), (3, 15, 1)])
def test_detail_float_sums_exact_on_integer_grids(t):
    """Integer-valued grids of small magnitude: every product and every partial sum is an
    integer below 2^24, so the reference's float sums are exact in any order and equal the
    float64 sums; the unsplit sums equal the split halves added."""
    ng = 8
    rng = np.random.default_rng(sum(t))
    rr, ri, rp, lr, li = [rng.integers(-3, 4, ng ** 3).astype(np.float32) for _ in range(5)]
    f, cnt = O.detail_sums_f32(ng, rr, ri, rp, lr, li, t)
    d = O.detail_sums(ng, rr, ri, rp, lr, li, t)
    assert np.array_equal(f[:6].astype(np.float64), d), (f, d)
    assert np.array_equal(f[6:], f[0:6:2] + f[1:6:2])
    # the counts are the non-zero products of each sign
    assert cnt.sum() > 0 and np.all(cnt >= 0)


def test_detail_float_sums_follow_the_raster_order():
    """Products 2^24 then 1, 1, 1 along the reference's order (l, m, n ascending): added one
    by one in float each 1 is lost (2^24 + 1 rounds back to 2^24), so the sum is 2^24; any
    order that adds the ones first gives 2^24 + 4.  One receptor voxel per product, the
    ligand all ones, translation 0 (s = ng: voxel i meets voxel i)."""
    ng = 4
    rp = np.zeros(ng ** 3, np.float32)
    cells = [(1, 2, 3), (2, 0, 1), (2, 3, 0), (3, 1, 2)]  # raster order: the big one first
    for k, (l, m, n) in enumerate(cells):
        rp[(l * ng + m) * ng + n] = 2.0 ** 24 if k == 0 else 1.0
    one = np.ones(ng ** 3, np.float32)
    zero = np.zeros(ng ** 3, np.float32)
    f, cnt = O.detail_sums_f32(ng, rp, zero, rp, one, zero, (0, 0, 0))
    assert f[0] == np.float32(2.0 ** 24) and f[6] == np.float32(2.0 ** 24)
    assert cnt[0] == 4
    # the big product last in raster order: the ones add up first
    rp2 = np.zeros(ng ** 3, np.float32)
    for k, (l, m, n) in enumerate(cells):
        rp2[(l * ng + m) * ng + n] = 2.0 ** 24 if k == 3 else 1.0
    f2, _ = O.detail_sums_f32(ng, rp2, zero, rp2, one, zero, (0, 0, 0))
    assert f2[0] == np.float32(2.0 ** 24 + 4)
