"""Python face of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Loads ``oracle/liboracle.so`` (the C restatement of the reference's CPU path) and,
when it has been built, ``oracle/_ref/libmdkref.so`` (the reference's own CPU sources
compiled with a small driver).  Only tests, the smoke check and the benchmark's CPU
baseline import this module, and only as the checker.

Also holds the oracle's own PDB reader and table lookups (src/pdb_entry.cpp:18-105,
src/parameter.cpp:425-470) and an FP64 numpy path for the FFT stages.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DATA = os.path.join(os.path.dirname(HERE), "megadock_amd", "data")


class SortScore(C.Structure):
    _fields_ = [("score", C.c_float), ("index", C.c_int * 4)]


SORTSCORE_DTYPE = np.dtype([("score", np.float32), ("index", np.int32, (4,))])

_lib = None
_ref = None


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        _lib = C.CDLL(path)
    return _lib


def ref_available():
    return os.path.exists(os.path.join(HERE, "_ref", "libmdkref.so"))


def ref():
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libmdkref.so")
        if not os.path.exists(path):
            raise RuntimeError("reference build missing: oracle/_ref/libmdkref.so")
        _ref = C.CDLL(path)
    return _ref


# ------------------------------------------------------------------ tables / PDB
def _load_table(name):
    d = {}
    with open(os.path.join(DATA, name)) as f:
        for line in f:
            if line.startswith("#"):
                continue
            k, v = line.rstrip("\n").split("\t")
            if k not in d:  # std::map::insert keeps the first value
                d[k] = np.float32(float(v))
    return d


_TABLES = {}


def tables():
    if not _TABLES:
        _TABLES["radius"] = _load_table("charmm_radius.tsv")
        _TABLES["charge"] = _load_table("charmm_charge.tsv")
        _TABLES["ace"] = _load_table("ace.tsv")
    return _TABLES


def load_rotations(n):
    """float32[n,3] rotation set (src/parameter.cpp:3793-61886)."""
    rows = []
    with open(os.path.join(DATA, "rotations_%d.txt" % n)) as f:
        for line in f:
            if not line.startswith("#"):
                rows.append([float(v) for v in line.split()])
    return np.array(rows, dtype=np.float64).astype(np.float32)


def erase_space(s):
    """src/pdb_entry.cpp:78-105"""
    while "  " in s:
        n = s.find("  ")
        s = s[:n] + s[n + 1:]
    if s[:1] == " ":
        s = s[1:]
    if s[-1:] == " ":
        s = s[:-1]
    return s


def read_pdb(path):
    """src/pdb_entry.cpp:18-75 -> (coords float32[n,3], keys list)."""
    xyz, keys = [], []
    with open(path) as f:
        for line in f:
            line = line.rstrip("\n")
            if line[0:6] in ("ATOM  ", "HETATM"):
                if line[76:78] == " H":
                    continue
                xyz.append([float(line[30:38]), float(line[38:46]), float(line[46:54])])
                keys.append(erase_space(line[11:21]))
    return np.array(xyz, dtype=np.float64).astype(np.float32).reshape(-1, 3), keys


class Protein:
    """Atoms plus per-atom radius/charge/ACE as src/protein.cpp:19-64 looks them up."""

    def __init__(self, path):
        self.path = path
        self.xyz, self.keys = read_pdb(path)
        t = tables()
        self.radius = np.array([t["radius"].get(k, np.float32(1.8)) for k in self.keys], np.float32)
        self.charge = np.array([t["charge"].get(k, np.float32(0.0)) for k in self.keys], np.float32)
        self.ace = np.array([t["ace"].get(k, np.float32(0.0)) for k in self.keys], np.float32)
        self.center = center(self.xyz)

    @property
    def na(self):
        return len(self.xyz)


# ------------------------------------------------------------------ C oracle wrappers
def rotation_coef(theta):
    r = np.zeros(9, np.float32)
    th = np.ascontiguousarray(theta, np.float32)
    lib().or_rotation_coef(_fp(th), _fp(r))
    return r


def rotate_ligand(xyz, center_, r):
    xyz = np.ascontiguousarray(xyz, np.float32)
    out = np.zeros_like(xyz)
    c = np.ascontiguousarray(center_, np.float32)
    lib().or_rotate_ligand(C.c_int(len(xyz)), _fp(xyz), _fp(c), _fp(np.ascontiguousarray(r, np.float32)), _fp(out))
    return out


def grid_coord(ng, gw):
    out = np.zeros(ng, np.float32)
    lib().or_grid_coord(C.c_int(ng), C.c_float(gw), _fp(out))
    return out


def center(xyz):
    xyz = np.ascontiguousarray(xyz, np.float32)
    c = np.zeros(3, np.float32)
    lib().or_center(C.c_int(len(xyz)), _fp(xyz), _fp(c))
    return c


def shift_center(xyz):
    xyz = np.ascontiguousarray(xyz, np.float32)
    out = np.zeros_like(xyz)
    lib().or_shift_center(C.c_int(len(xyz)), _fp(xyz), _fp(out))
    return out


def ligand_voxel(ng, gw, xyz_rot, radius, charge, delta=1.0):
    gr = np.zeros(ng ** 3, np.float32)
    gi = np.zeros(ng ** 3, np.float32)
    gc = grid_coord(ng, gw)
    xyz_rot = np.ascontiguousarray(xyz_rot, np.float32)
    lib().or_ligand_voxel(C.c_int(ng), C.c_float(gw), C.c_int(len(xyz_rot)), _fp(xyz_rot),
                          _fp(np.ascontiguousarray(radius, np.float32)),
                          _fp(np.ascontiguousarray(charge, np.float32)), _fp(gc),
                          C.c_float(delta), _fp(gr), _fp(gi))
    return gr, gi


def receptor_voxel(ng, gw, prot, aceratio=1.0, rec_core=-45.0, eratio=1.0):
    gr = np.zeros(ng ** 3, np.float32)
    gi = np.zeros(ng ** 3, np.float32)
    gc = grid_coord(ng, gw)
    xyz = shift_center(prot.xyz)
    lib().or_receptor_voxel(C.c_int(ng), C.c_float(gw), C.c_int(prot.na), _fp(xyz),
                            _fp(prot.radius), _fp(prot.charge), _fp(prot.ace), _fp(gc),
                            C.c_float(aceratio), C.c_float(rec_core), C.c_float(eratio),
                            _fp(gr), _fp(gi))
    return gr, gi


def embed(ng, nf, gr, gi):
    out = np.zeros(2 * nf ** 3, np.float32)
    lib().or_embed(C.c_int(ng), C.c_int(nf), _fp(np.ascontiguousarray(gr, np.float32)),
                   _fp(np.ascontiguousarray(gi, np.float32)), _fp(out))
    return out.view(np.complex64)


def fft3d(nf, data, sign):
    a = np.ascontiguousarray(data, np.complex64).ravel().copy()
    lib().or_fft3d(C.c_int(nf), _fp(a.view(np.float32)), _fp(a.view(np.float32)), C.c_int(sign))
    return a


def fft1d(n, lines, sign):
    a = np.ascontiguousarray(lines, np.complex64).copy()
    count = a.size // n
    lib().or_fft1d(C.c_int(n), C.c_int(count), _fp(a.view(np.float32)), _fp(a.view(np.float32)), C.c_int(sign))
    return a


def receptor_fft(ng, nf, gr, gi):
    nf3 = nf ** 3
    rr = np.zeros(nf3, np.float32)
    ri = np.zeros(nf3, np.float32)
    lib().or_receptor_fft(C.c_int(ng), C.c_int(nf), _fp(gr), _fp(gi), _fp(rr), _fp(ri))
    return rr, ri


def convolution(rr, ri, lig):
    lig = np.ascontiguousarray(lig, np.complex64)
    out = np.zeros_like(lig)
    lib().or_convolution(C.c_size_t(len(rr)), _fp(rr), _fp(ri), _fp(lig.view(np.float32)),
                         _fp(out.view(np.float32)))
    return out


def score_sort(nf, t, fftout, rot):
    out = np.zeros(t, SORTSCORE_DTYPE)
    fo = np.ascontiguousarray(fftout, np.complex64)
    lib().or_score_sort(C.c_int(nf), C.c_int(t), _fp(fo.view(np.float32)), C.c_int(rot),
                        C.c_void_p(out.ctypes.data))
    return out


def sort_topn(top, n):
    top = np.ascontiguousarray(top, SORTSCORE_DTYPE).copy()
    lib().or_sort_topn(C.c_void_p(top.ctypes.data), C.c_int(len(top)), C.c_int(n))
    return top[:n]


def dock(ng, nf, gw, lig, rr, ri, angles, rot_ids, t, threads=1, delta=1.0):
    rot_ids = np.ascontiguousarray(rot_ids, np.int32)
    out = np.zeros(len(rot_ids) * t, SORTSCORE_DTYPE)
    ang = np.ascontiguousarray(angles, np.float32).ravel()
    rc = lib().or_dock(C.c_int(ng), C.c_int(nf), C.c_float(gw), C.c_int(lig.na), _fp(lig.xyz),
                       _fp(lig.center), _fp(lig.radius), _fp(lig.charge), C.c_float(delta),
                       _fp(rr), _fp(ri), _fp(ang), _ip(rot_ids), C.c_int(len(rot_ids)), C.c_int(t),
                       C.c_int(threads), C.c_void_p(out.ctypes.data))
    assert rc == 0
    return out


# ------------------------------------------------------------------ FP64 path (numpy)
def correlation_fp64(nf, rec_grid_c, lig_grid_c):
    """Re(IFFT(conj(FFT r) * FFT l)) in float64 for complex nf^3 grids (unnormalised)."""
    R = np.fft.fftn(np.asarray(rec_grid_c, np.complex128).reshape(nf, nf, nf))
    L = np.fft.fftn(np.asarray(lig_grid_c, np.complex128).reshape(nf, nf, nf))
    return np.real(np.fft.ifftn(np.conj(R) * L)) * nf ** 3


def detail_sums(ng, rec_r, rec_i, rec_rpsc, lig_r, lig_i, trans):
    """Real-space sums of Docking::output_detail (src/docking.cpp:596-657, :710-801) at
    the grid translation `trans` (3 ints, a decoy's index[1..3]): per axis
    s = (3 ng - t) mod 2 ng, ligand voxel i meets receptor voxel i + s - ng inside the
    grid; products in float32 as the reference forms them, sums in float64.  Returns
    (rPSC+, rPSC-, ELEC+, ELEC-, ACE+, ACE-) with rPSC = Rp Lr, ELEC = Ri Li and
    ACE = (Rr - Rp) Lr (Rp: receptor shape grid built without the ACE term)."""
    g = [np.asarray(a, np.float32).reshape(ng, ng, ng) for a in (rec_r, rec_i, rec_rpsc, lig_r, lig_i)]
    rr, ri, rp, lr, li = g
    lsl, rsl = [], []
    for a in range(3):
        s = (3 * ng - int(trans[a])) % (2 * ng)
        lo, hi = (ng - s, ng) if s < ng else (0, 2 * ng - s)
        lsl.append(slice(lo, hi))
        rsl.append(slice(lo + s - ng, hi + s - ng))
    lsl, rsl = tuple(lsl), tuple(rsl)
    p = rp[rsl] * lr[lsl]
    e = ri[rsl] * li[lsl]
    c = (rr[rsl] - rp[rsl]) * lr[lsl]
    out = []
    for v in (p, e, c):
        v = v.astype(np.float64)
        out += [v[v > 0].sum(), v[v < 0].sum()]
    return np.array(out)


# ------------------------------------------------------------------ reference build
def ref_load(rec_pdb, lig_pdb, extra=()):
    r = ref()
    arr = (C.c_char_p * max(1, len(extra)))(*[e.encode() for e in extra])
    return r.ref_load(rec_pdb.encode(), lig_pdb.encode(), C.c_int(len(extra)), arr)


def ref_protein(which):
    r = ref()
    na = r.ref_num_atoms(C.c_int(which))
    xyz = np.zeros((na, 3), np.float32)
    r.ref_coords(C.c_int(which), _fp(xyz))
    rad = np.zeros(na, np.float32)
    chg = np.zeros(na, np.float32)
    ace = np.zeros(na, np.float32)
    r.ref_atom_params(C.c_int(which), _fp(rad), _fp(chg), _fp(ace))
    c = np.zeros(3, np.float32)
    r.ref_center(C.c_int(which), _fp(c))
    keys = []
    buf = C.create_string_buffer(64)
    for i in range(na):
        r.ref_atom_type(C.c_int(which), C.c_int(i), buf, C.c_int(64))
        keys.append(buf.value.decode())
    return dict(xyz=xyz, radius=rad, charge=chg, ace=ace, center=c, keys=keys)


def ref_receptor_voxel(ng, gw, aceratio=1.0, rec_core=-45.0, lig_core=1.0, eratio=1.0):
    gr = np.zeros(ng ** 3, np.float32)
    gi = np.zeros(ng ** 3, np.float32)
    ref().ref_receptor_voxel(C.c_int(ng), C.c_float(gw), C.c_float(aceratio), C.c_float(rec_core),
                             C.c_float(lig_core), C.c_float(eratio), _fp(gr), _fp(gi))
    return gr, gi


def ref_ligand_voxel(ng, gw, xyz_rot, delta=1.0, eratio=1.0):
    gr = np.zeros(ng ** 3, np.float32)
    gi = np.zeros(ng ** 3, np.float32)
    xyz_rot = np.ascontiguousarray(xyz_rot, np.float32)
    ref().ref_ligand_voxel(C.c_int(ng), C.c_float(gw), _fp(xyz_rot), C.c_float(1.0), C.c_float(-45.0),
                           C.c_float(delta), C.c_float(eratio), _fp(gr), _fp(gi))
    return gr, gi


def ref_rotations(n, rec_pdb, lig_pdb):
    """float32[n, 3]: the reference's own rotation table for -r n (Parameter::_Zangle after
    Parameter::initialize, read through the HeaderDB friend in ref_driver.cpp)."""
    ref_load(rec_pdb, lig_pdb, ["-r", str(n)])
    r = ref()
    cnt = r.ref_rotation_count()
    out = np.zeros((cnt, 3), np.float32)
    r.ref_rotation_angles(_fp(out))
    return out


def ref_rotate(xyz, theta):
    xyz = np.ascontiguousarray(xyz, np.float32)
    out = np.zeros_like(xyz)
    th = np.ascontiguousarray(theta, np.float32)
    ref().ref_rotate(_fp(xyz), C.c_int(len(xyz)), _fp(th), _fp(out))
    return out


def ref_decoygen(new_pdb, ligand_pdb, out_file, n):
    """The reference's decoygen program (the main of src/decoygen.cpp, compiled into
    oracle/_ref): writes decoy n of out_file as new_pdb."""
    argv = [b"decoygen", str(new_pdb).encode(), str(ligand_pdb).encode(), str(out_file).encode(), str(n).encode()]
    arr = (C.c_char_p * len(argv))(*argv)
    return ref().main(C.c_int(len(argv)), arr)


# ------------------------------------------------------------------ A16 grid sizing
GRID_TABLE_11 = [2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 14, 15, 16, 18, 20, 21, 22, 24, 25, 27,
                 28, 30, 32, 33, 35, 36, 40, 42, 44, 45, 48, 49, 50, 54, 55, 56, 60, 63,
                 64, 66, 70, 72, 75, 77, 80, 81, 84, 88, 90, 96, 98, 99, 100, 105, 108,
                 110, 112, 120, 121, 125, 126, 128, 132, 135, 140, 144, 147, 150,
                 154, 160, 162, 165, 168, 175, 176, 180, 189, 192, 196, 198, 200,
                 210, 216, 220, 224, 225, 231, 240, 242, 243, 245, 250, 252, 256,
                 264, 270, 275, 280, 288, 294, 297, 300, 308, 315, 320, 324, 330,
                 336, 343, 350, 352, 360, 363, 375, 378, 384, 385, 392, 396, 400,
                 405, 420, 432, 440, 441, 448, 450, 462, 480, 484, 486, 490, 495,
                 500, 504, 512, 525, 528, 539, 540, 550, 560, 567, 576, 588, 594,
                 600, 605, 616, 625, 630, 640, 648, 660, 672, 675, 686, 693, 700,
                 704, 720, 726, 729, 735, 750, 756, 768, 770, 784, 792, 800, 810,
                 825, 840, 847, 864, 875, 880, 882, 891, 896, 900, 924, 945, 960,
                 968, 972, 980, 990, 1000, 1008, 1024]


def autogrid(rec, lig, gw=1.2, table=GRID_TABLE_11):
    """src/control.cpp:287-379 (11-base table, src/control.cpp:91-125):
    returns ng = max(receptor ng, ligand ng)."""
    gw = np.float32(gw)
    edge_lo = rec.xyz.min(axis=0)
    edge_hi = rec.xyz.max(axis=0)
    size_rec = np.float32(0.0)
    for i in range(3):
        s = np.float32(edge_hi[i] - edge_lo[i])
        if s > size_rec:
            size_rec = s
    size_rec = np.float32(np.float64(size_rec) + 2.0 * 8.0)
    ngr = 1 + int(np.float32(size_rec / gw))
    x = lig.xyz
    d2max = np.float32(0.0)
    for i in range(len(x) - 1):
        d = x[i + 1:] - x[i]
        d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        m = d2.max()
        if m > d2max:
            d2max = m
    size_lig = np.float32(np.sqrt(d2max))
    size_lig = np.float32(np.float64(size_lig) + 2.0 * 3.0)
    ngl = 1 + int(np.float32(size_lig / gw))

    def pick(n):
        for v in table:
            if v >= n:
                return v
        return n
    return max(pick(ngr), pick(ngl))


# ---------------------------------------------------------------- off-grid scores (FP64)
# The reference has no off-grid score; these FP64 restatements check the device kernels
# for scores between grid points and their gradients.  They are pinned by the hard score
# map at grid points, by central finite differences and by closed forms (tests).

def signed_freq(nf):
    k = np.arange(nf)
    return np.where(k < nf // 2, k, k - nf)


def folded_spectrum_fp64(nf, rec_grid_c, lig_grid_c):
    """Q = (P(k) + conj(P(-k))) / 2 with P = conj(FFT r) FFT l, FP64, full nf^3."""
    R = np.fft.fftn(np.asarray(rec_grid_c, np.complex128).reshape(nf, nf, nf))
    L = np.fft.fftn(np.asarray(lig_grid_c, np.complex128).reshape(nf, nf, nf))
    P = np.conj(R) * L
    Pm = np.roll(np.flip(P, axis=(0, 1, 2)), 1, axis=(0, 1, 2))  # P(-k)
    return 0.5 * (P + np.conj(Pm))


def spectral_score_fp64(nf, Q, t):
    """Band-limited score s(t) = 1/nf^3 sum_{kz <= nf/2} w(kz) Re[Q e^{i phi}] at a
    continuous translation t (voxels) and its gradient, phi = 2 pi k'.t / nf with signed
    frequencies k', w = 1 at kz = 0 and nf/2, else 2."""
    H = nf // 2 + 1
    k = signed_freq(nf).astype(np.float64)
    ex = np.exp(2j * np.pi * k * t[0] / nf)
    ey = np.exp(2j * np.pi * k * t[1] / nf)
    ez = np.exp(2j * np.pi * k[:H] * t[2] / nf)
    w = np.full(H, 2.0)
    w[0] = 1.0
    w[nf // 2] = 1.0
    z = Q[:, :, :H] * ex[:, None, None] * ey[None, :, None] * (ez * w)[None, None, :]
    n3 = float(nf) ** 3
    s = np.real(z).sum() / n3
    im = np.imag(z)
    g = -2.0 * np.pi / nf / n3
    grad = np.array([(im.sum(axis=(1, 2)) * k).sum(), (im.sum(axis=(0, 2)) * k).sum(),
                     (im.sum(axis=(0, 1)) * k[:H]).sum()]) * g
    return s, grad


def trilinear_fp64(grid, ng, u):
    """Value and gradient (per voxel) of the trilinear interpolation of an ng^3 grid at
    voxel coordinate u (3,); corners outside the grid count as 0."""
    i0 = np.floor(u).astype(np.int64)
    f = u - i0
    val = 0.0
    grad = np.zeros(3)
    g3 = grid.reshape(ng, ng, ng)
    for c in range(8):
        d = np.array([(c >> 2) & 1, (c >> 1) & 1, c & 1])
        i = i0 + d
        if np.any(i < 0) or np.any(i >= ng):
            continue
        g = float(g3[i[0], i[1], i[2]])
        wv = np.where(d == 1, f, 1.0 - f)
        sv = np.where(d == 1, 1.0, -1.0)
        val += wv.prod() * g
        grad += np.array([sv[0] * wv[1] * wv[2], wv[0] * sv[1] * wv[2], wv[0] * wv[1] * sv[2]]) * g
    return val, grad


def companion_fp64(gr, gi, ng, gw, xyz, center_, radius, charge, R, t, nf):
    """Companion score s = sum_a [w_a Rr(P_a) + q_a Ri(P_a)], P_a = R (a - c) - gw t',
    w_a = (4/3) pi r_a^3 / gw^3, with (s, ds/dt' (3), ds/dw (3)) for a rotation exp([w]x)
    applied after R."""
    R = np.asarray(R, np.float64).reshape(3, 3)
    t = np.asarray(t, np.float64)
    tp = np.where(t < 0.5 * nf, t, t - nf)
    X = (np.asarray(xyz, np.float64) - np.asarray(center_, np.float64)) @ R.T
    P = X - gw * tp
    u = P / gw + 0.5 * ng - 0.5
    ws = 4.0 / 3.0 * np.pi * np.asarray(radius, np.float64) ** 3 / gw ** 3
    q = np.asarray(charge, np.float64)
    s = 0.0
    dt = np.zeros(3)
    dw = np.zeros(3)
    for a in range(len(X)):
        vr, gr_ = trilinear_fp64(gr, ng, u[a])
        vi, gi_ = trilinear_fp64(gi, ng, u[a])
        s += ws[a] * vr + q[a] * vi
        g = (ws[a] * gr_ + q[a] * gi_) / gw
        dt += -gw * g
        dw += np.cross(X[a], g)
    return s, dt, dw


def rotvec_matrix(w):
    """exp([w]x) (Rodrigues)."""
    w = np.asarray(w, np.float64)
    th = np.linalg.norm(w)
    K = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    if th < 1e-12:
        return np.eye(3) + K
    return np.eye(3) + np.sin(th) / th * K + (1 - np.cos(th)) / th ** 2 * (K @ K)


# ------------------------------------------------------------------ detailed output (-O)
def detail_sums_f32(ng, rec_r, rec_i, rec_rpsc, lig_r, lig_i, trans):
    """The float sums of Docking::output_detail in the reference's order
    (src/docking.cpp:646-654, :720-807): (rPSC+, rPSC-, ELEC+, ELEC-, ACE+, ACE-, rPSC,
    ELEC, ACE unsplit) as float32, and the counts of the six split sums."""
    a = [np.ascontiguousarray(x, np.float32).ravel() for x in (rec_r, rec_i, rec_rpsc, lig_r, lig_i)]
    tr = np.ascontiguousarray(trans, np.int32)
    out = np.zeros(9, np.float32)
    cnt = np.zeros(6, np.int32)
    lib().or_detail_sums_f32(C.c_int(ng), *[_fp(x) for x in a], _ip(tr), _fp(out), _ip(cnt))
    return out, cnt


def _atom_class(key):
    """src/protein.cpp:50-60: 1 for CA, 2 for C and N, else 0."""
    if key[:3] == "CA ":
        return 1
    if key[:2] in ("C ", "N "):
        return 2
    return 0


def output_detail(rec, lig, ng, gw, rec_file, lig_file, decoys, angles, elec_ratio=1.0, ace_ratio=1.0):
    """The .detail and .csv texts of Docking::output_detail (src/docking.cpp:456-898) for
    the decoys (rotation number, tx, ty, tz) of angle table `angles`: header, the seven
    poses around the input position with the unrotated ligand, then the decoys; float
    arithmetic as the reference's (float sums in its order, RMSD terms squared in double
    and accumulated in float), printed with its formats."""
    f32 = np.float32
    nf1 = 2 * ng
    gwf = f32(gw)
    pi = f32(4.0 * np.arctan(1.0))
    gr, gi = receptor_voxel(ng, gw, rec, aceratio=ace_ratio, eratio=elec_ratio)
    rp, _ = receptor_voxel(ng, gw, rec, aceratio=0.0, eratio=elec_ratio)
    coordr = np.asarray(rec.center, np.float32)
    coordl = np.asarray(lig.center, np.float32)
    diff = (coordr - coordl).astype(np.float32)
    cor = []
    for i in range(3):
        q = f32(diff[i] / gwf)
        c = int(np.floor(abs(float(q)) + 0.5)) * (1 if q >= 0 else -1)  # round(), halves away from zero
        cor.append(c + nf1 if c < 0 else c)
    cls = [_atom_class(k) for k in lig.keys]
    n_lca = sum(1 for c in cls if c == 1)
    xyz = np.asarray(lig.xyz, np.float32)

    def rmsd(tem):
        lca = f32(0.0)
        for i in range(lig.na):
            if cls[i] > 0:
                for k in range(3):
                    moved = f32(f32(xyz[i, k] + f32(gwf * f32(tem[k] - ng))) + diff[k])
                    d = float(f32(moved - xyz[i, k]))
                    if cls[i] == 1:
                        lca = f32(float(lca) + d * d)
        r = f32(np.sqrt(f32(lca / f32(n_lca))))
        return r, 1 if r < f32(5.0) else 0

    def lig_grids(theta):
        xyz_r = rotate_ligand(lig.xyz, lig.center, rotation_coef(np.asarray(theta, np.float32)))
        return ligand_voxel(ng, gw, xyz_r, lig.radius, lig.charge)

    fmt = "%5d: (%5.2f,%5.2f,%5.2f) (%3d,%3d,%3d) %11.2f %8.0f %8.0f %8.0f %10.2f %10.2f %7.3f %7d"
    columns = ("Rank   Angle               Trans         total       rPSC     rPSC(+)  rPSC(-)  ELEC       RecACE"
               "     LCARMSD NearNat")
    d = ["%d\t%.2f" % (nf1, gw), "\t%f\t%f\t%f" % (0.0, 0.0, 0.0),
         rec_file + "\t%f\t%f\t%f" % tuple(float(v) for v in coordr),
         lig_file + "\t%f\t%f\t%f" % tuple(float(v) for v in coordl),
         "\nElec score ratio: %6.2f\nACE  score ratio: %6.2f" % (elec_ratio, ace_ratio),
         "\ncorrect position: angle = (%5.2f,%5.2f,%5.2f), trans = (%3d,%3d,%3d) \n" % (0.0, 0.0, 0.0, *cor),
         " *** score around correct position *** \n", columns]
    lr0, li0 = lig_grids((0.0, 0.0, 0.0))
    c = list(cor)
    for rank in range(7):  # src/docking.cpp:596-700
        if rank == 1:
            c[0] -= 1
        if rank == 2:
            c[0] += 2
        if rank == 3:
            c[0] -= 1
            c[1] -= 1
        if rank == 4:
            c[1] += 2
        if rank == 5:
            c[1] -= 1
            c[2] -= 1
        if rank == 6:
            c[2] += 2
        s, _ = detail_sums_f32(ng, gr, gi, rp, lr0, li0, c)
        total = f32(f32(s[6] + s[7]) + s[8])
        tem = [(3 * ng - v) % (2 * ng) for v in c]
        r, nn = rmsd(tem)
        d.append(fmt % (rank + 1, 0.0, 0.0, 0.0, c[0], c[1], c[2], total, s[6], s[0], s[1], s[7], s[8], r, nn))
    d += ["\n *** docking results *** \n", columns]
    csv = ["Rank,Anglex,Angley,Anglez,Transi,Transj,Transk,total,rPSC,rPSC_gain,rPSC_penalty,ELEC ,RecACE,LCARMSD,NearNat"]
    for rank, (a, tx, ty, tz) in enumerate(decoys):  # src/docking.cpp:711-857
        theta = np.asarray(angles[a], np.float32)
        lr, li = lig_grids(theta)
        s, _ = detail_sums_f32(ng, gr, gi, rp, lr, li, (tx, ty, tz))
        rpsc, elec, ace = f32(s[0] + s[1]), f32(s[2] + s[3]), f32(s[4] + s[5])
        total = f32(f32(rpsc + elec) + ace)
        th = [f32(float(v) - float(pi) * 2.0) if v >= pi else v for v in theta]
        tem = [(3 * ng - v) % (2 * ng) for v in (tx, ty, tz)]
        r, nn = rmsd(tem)
        vals = (total, rpsc, s[0], s[1], elec, ace, r)
        d.append(fmt % (rank + 1, *th, tx, ty, tz, *vals, nn))
        csv.append("%d, %f,%f,%f,%d,%d,%d,%f,%f,%f,%f,%f,%f,%f,%d" % (rank + 1, *th, tx, ty, tz, *vals, nn))
    return "\n".join(d) + "\n", "\n".join(csv) + "\n"
