"""Pins of the FP64 off-grid score references (CPU): the band-limited score equals the
hard score map at grid points, gradients match central finite differences, and the
trilinear companion reproduces linear receptor fields exactly."""
import numpy as np
import pytest

from oracle import oracle as O


def _grids(nf, seed):
    rng = np.random.default_rng(seed)
    r = (rng.standard_normal(nf ** 3) + 1j * rng.standard_normal(nf ** 3)).astype(np.complex128)
    l = (rng.standard_normal(nf ** 3) + 1j * rng.standard_normal(nf ** 3)).astype(np.complex128)
    return r, l


@pytest.mark.parametrize("nf", [16, 20])
def test_spectral_score_equals_score_map_at_grid_points(nf):
    r, l = _grids(nf, nf)
    smap = O.correlation_fp64(nf, r, l).reshape(nf, nf, nf) / nf ** 3
    Q = O.folded_spectrum_fp64(nf, r, l)
    for t in [(0, 0, 0), (1, 2, 3), (nf - 1, nf // 2, 5), (nf // 2, nf - 3, nf // 2 + 1)]:
        s, _ = O.spectral_score_fp64(nf, Q, np.array(t, np.float64))
        assert abs(s - smap[t]) <= 1e-9 * np.abs(smap).max(), (t, s, smap[t])


def test_spectral_gradient_matches_finite_differences():
    nf = 16
    r, l = _grids(nf, 7)
    Q = O.folded_spectrum_fp64(nf, r, l)
    rng = np.random.default_rng(3)
    for _ in range(4):
        t = rng.uniform(0, nf, 3)
        _, g = O.spectral_score_fp64(nf, Q, t)
        h = 1e-3
        fd = np.array([(O.spectral_score_fp64(nf, Q, t + h * e)[0] - O.spectral_score_fp64(nf, Q, t - h * e)[0]) / (2 * h)
                       for e in np.eye(3)])
        # white-noise grids reach the Nyquist frequency: the O(h^2) truncation of the
        # central difference is then about (h pi)^2 / 6 ~ 2e-6 relative
        assert np.linalg.norm(fd - g) <= 1e-5 * np.linalg.norm(g), (t, g, fd)


def test_spectral_single_frequency_closed_form():
    """Q = delta at k0 (kz = 0 plane, partner -k0 included): s(t) = 2 cos(2 pi k0.t / nf) / nf^3."""
    nf = 16
    Q = np.zeros((nf, nf, nf), np.complex128)
    k0 = (3, nf - 2, 0)
    Q[k0] = 1.0
    Q[(nf - 3) % nf, 2, 0] = 1.0
    t = np.array([0.3, 1.7, 4.2])
    s, g = O.spectral_score_fp64(nf, Q, t)
    kv = np.array([3, -2, 0], np.float64)
    ph = 2 * np.pi * kv.dot(t) / nf
    assert abs(s - 2 * np.cos(ph) / nf ** 3) < 1e-15
    assert np.allclose(g, -2 * np.sin(ph) * 2 * np.pi * kv / nf / nf ** 3, atol=1e-16)


def _ligand(na, seed):
    rng = np.random.default_rng(seed)
    xyz = rng.uniform(-6, 6, (na, 3))
    radius = rng.uniform(1.5, 2.0, na)
    charge = rng.uniform(-0.5, 0.5, na)
    return xyz, xyz.mean(axis=0), radius, charge


def test_companion_linear_field_closed_form():
    ng, gw, nf = 24, 1.2, 48
    i, j, k = np.meshgrid(np.arange(ng), np.arange(ng), np.arange(ng), indexing="ij")
    a, b = np.array([0.7, -1.1, 0.4]), np.array([-0.2, 0.3, 0.9])
    gr = (a[0] * i + a[1] * j + a[2] * k + 2.0).ravel()
    gi = (b[0] * i + b[1] * j + b[2] * k - 1.0).ravel()
    xyz, c, radius, charge = _ligand(20, 1)
    R = O.rotvec_matrix([0.3, -0.2, 0.5])
    t = np.array([1.3, -0.7, 0.4]) % nf
    s, dt, dw = O.companion_fp64(gr, gi, ng, gw, xyz, c, radius, charge, R, t, nf)
    X = (xyz - c) @ R.T
    tp = np.where(t < nf / 2, t, t - nf)
    u = (X - gw * tp) / gw + ng / 2 - 0.5
    assert np.all((u > 0) & (u < ng - 1))
    ws = 4 / 3 * np.pi * radius ** 3 / gw ** 3
    s_ref = (ws * (u @ a + 2.0) + charge * (u @ b - 1.0)).sum()
    g = (ws[:, None] * a + charge[:, None] * b) / gw
    assert abs(s - s_ref) < 1e-9 * abs(s_ref)
    assert np.allclose(dt, -gw * g.sum(axis=0), rtol=1e-12)
    assert np.allclose(dw, np.cross(X, g).sum(axis=0), rtol=1e-10, atol=1e-12)


def test_companion_gradient_matches_finite_differences():
    """h = 1e-3 A in translation, 1e-3 rad in rotation (FP64)."""
    ng, gw, nf = 24, 1.2, 48
    rng = np.random.default_rng(5)
    gr = rng.standard_normal(ng ** 3) * 3
    gi = rng.standard_normal(ng ** 3)
    xyz, c, radius, charge = _ligand(25, 2)
    R = O.rotvec_matrix([0.1, 0.7, -0.4])
    t = np.array([0.37, -1.21, 2.05]) % nf
    s, dt, dw = O.companion_fp64(gr, gi, ng, gw, xyz, c, radius, charge, R, t, nf)
    ht = 1e-3 / gw
    fdt = np.array([(O.companion_fp64(gr, gi, ng, gw, xyz, c, radius, charge, R, t + ht * e, nf)[0]
                     - O.companion_fp64(gr, gi, ng, gw, xyz, c, radius, charge, R, t - ht * e, nf)[0]) / (2 * ht)
                    for e in np.eye(3)])
    hw = 1e-3
    fdw = np.array([(O.companion_fp64(gr, gi, ng, gw, xyz, c, radius, charge, O.rotvec_matrix(hw * e) @ R, t, nf)[0]
                     - O.companion_fp64(gr, gi, ng, gw, xyz, c, radius, charge, O.rotvec_matrix(-hw * e) @ R, t, nf)[0])
                    / (2 * hw) for e in np.eye(3)])
    assert np.linalg.norm(fdt - dt) <= 1e-3 * np.linalg.norm(dt), (dt, fdt)
    assert np.linalg.norm(fdw - dw) <= 1e-3 * np.linalg.norm(dw), (dw, fdw)
