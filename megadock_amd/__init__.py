"""megadock_amd: MI355X-native FFT docking rotation loop.

Python face of ``libmdk.so`` (the HIP kernels behind the C ABI declared in
``include/mdk_hip.h``).  Loading fails loudly when the library has not been built;
there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmdk.so")
DATA = os.path.join(HERE, "data")

SORTSCORE_DTYPE = np.dtype([("score", np.float32), ("index", np.int32, (4,))])

# every symbol declared in include/mdk_hip.h
ABI_SYMBOLS = [
    "mdk_version", "mdk_last_error", "mdk_rotation_coef", "mdk_device_count", "mdk_create", "mdk_destroy",
    "mdk_set_receptor", "mdk_set_receptor_spectrum", "mdk_get_receptor_spectrum",
    "mdk_set_ligand", "mdk_dock", "mdk_set_rotations", "mdk_dock_range", "mdk_sync",
    "mdk_get_results", "mdk_topn", "mdk_time_range", "mdk_profile_stages", "mdk_voxelize",
    "mdk_score_map", "mdk_fft1d", "mdk_copy_probe", "mdk_score_spectral", "mdk_score_companion",
    "mdk_autotune", "mdk_set_ligand_atoms", "mdk_score_detail", "mdk_set_receptor_atoms", "mdk_get_receptor_grids",
    "mdk_decoy_coords", "mdk_topn_noalloc", "mdk_dock_range_timed", "mdk_dock_grid",
    "mdk_topn_start", "mdk_topn_finish",
]
# every symbol declared in include/mdk_host.h
HOST_SYMBOLS = [
    "mdkh_protein_load", "mdkh_protein_free", "mdkh_protein_size", "mdkh_protein_atoms",
    "mdkh_autogrid", "mdkh_rotation_set", "mdkh_grid_coord", "mdkh_receptor_voxel", "mdkh_run", "mdkh_run_table", "mdkh_last_error",
]
STAGES = ["voxelize", "fwd_z", "fwd_y", "x_corr", "inv_y", "inv_z_top", "reduce"]
# FFT sizes with compile-time specialised kernels (megadock_amd/csrc/mdk_pipe.hip): the
# tuned sizes, with every kernel variant, and the further sizes, default variants only;
# other sizes run the generic mixed-radix kernels
TUNED_NF = [80, 96, 144, 256]
FURTHER_NF = [64, 66, 70, 72, 84, 88, 90, 98, 100, 108, 110, 112, 120, 126, 128, 132, 140, 150, 154, 160, 162, 168,
              176, 180, 192, 196, 198, 200, 210, 216, 220, 224, 240, 242, 250, 252, 264, 270, 280, 288, 294, 300, 308,
              320]

_lib = None


class MdkError(RuntimeError):
    pass


def lib():
    """Load libmdk.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MdkError("libmdk.so not built (run __graft_entry__.build() or make -C megadock_amd/csrc)")
        L = C.CDLL(LIB_PATH)
        L.mdk_last_error.restype = C.c_char_p
        L.mdkh_last_error.restype = C.c_char_p
        L.mdk_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_void_p)]
        for name in ABI_SYMBOLS + HOST_SYMBOLS:
            getattr(L, name)  # raises AttributeError if an export is missing
        _lib = L
    return _lib


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _check(rc, what):
    if rc != 0:
        raise MdkError("%s failed: %s" % (what, lib().mdk_last_error().decode()))


def device_count():
    n = C.c_int(0)
    _check(lib().mdk_device_count(C.byref(n)), "mdk_device_count")
    return n.value


def rotation_coefs(angles):
    """float32[R,9] ZXZ coefficients computed on the host by mdk_rotation_coef
    (float cosf/sinf, src/docking.cpp:376-384)."""
    a = np.ascontiguousarray(angles, np.float32).reshape(-1, 3)
    r = np.empty((len(a), 9), np.float32)
    _check(lib().mdk_rotation_coef(_fp(a), C.c_int(len(a)), _fp(r)), "mdk_rotation_coef")
    return r


class Context:
    """One receptor-ligand docking context on one GPU (wraps mdk_ctx)."""

    def __init__(self, nf, ng, na, t=1, max_batch=16, device=0):
        self.nf, self.ng, self.na, self.t, self.max_batch = nf, ng, na, t, max_batch
        h = C.c_void_p()
        _check(lib().mdk_create(device, nf, ng, na, t, max_batch, C.byref(h)), "mdk_create")
        self.h = h

    def close(self):
        if self.h:
            lib().mdk_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_receptor(self, grid_r, grid_i):
        gr = np.ascontiguousarray(grid_r, np.float32)
        gi = np.ascontiguousarray(grid_i, np.float32)
        assert gr.size == self.ng ** 3 and gi.size == self.ng ** 3
        _check(lib().mdk_set_receptor(self.h, _fp(gr), _fp(gi)), "mdk_set_receptor")

    def set_receptor_atoms(self, xyz_centred, radius, charge, ace, grid_coord, gw=1.2, aceratio=1.0,
                           rec_core=-45.0, eratio=1.0):
        """Receptor grids built on the device from the atoms (centred coordinates), then
        transformed; bit-exact with the host producer (receptor_grids)."""
        xyz = np.ascontiguousarray(xyz_centred, np.float32).ravel()
        na = xyz.size // 3
        arrs = [np.ascontiguousarray(a, np.float32) for a in (radius, charge, ace, grid_coord)]
        _check(lib().mdk_set_receptor_atoms(self.h, C.c_int(na), _fp(xyz), *[_fp(a) for a in arrs], C.c_float(gw),
                                            C.c_float(aceratio), C.c_float(rec_core), C.c_float(eratio)),
               "mdk_set_receptor_atoms")

    def receptor_grids(self):
        n3 = self.ng ** 3
        gr = np.zeros(n3, np.float32)
        gi = np.zeros(n3, np.float32)
        _check(lib().mdk_get_receptor_grids(self.h, _fp(gr), _fp(gi)), "mdk_get_receptor_grids")
        return gr, gi

    def set_receptor_spectrum(self, re, im):
        re = np.ascontiguousarray(re, np.float32)
        im = np.ascontiguousarray(im, np.float32)
        assert re.size == self.nf ** 3 and im.size == self.nf ** 3
        _check(lib().mdk_set_receptor_spectrum(self.h, _fp(re), _fp(im)), "mdk_set_receptor_spectrum")

    def receptor_spectrum(self):
        n3 = self.nf ** 3
        re = np.zeros(n3, np.float32)
        im = np.zeros(n3, np.float32)
        _check(lib().mdk_get_receptor_spectrum(self.h, _fp(re), _fp(im)), "mdk_get_receptor_spectrum")
        return re, im

    def set_ligand(self, xyz, center, radius, charge, grid_coord, gw, delta=1.0):
        """Ligand atoms; fewer than the context's na go through mdk_set_ligand_atoms."""
        xyz = np.ascontiguousarray(xyz, np.float32)
        na = xyz.size // 3
        assert xyz.size == 3 * na and 1 <= na <= self.na
        args = [np.ascontiguousarray(a, np.float32) for a in (center, radius, charge, grid_coord)]
        if na == self.na:
            _check(lib().mdk_set_ligand(self.h, _fp(xyz), *[_fp(a) for a in args], C.c_float(gw),
                                        C.c_float(delta)), "mdk_set_ligand")
        else:
            _check(lib().mdk_set_ligand_atoms(self.h, C.c_int(na), _fp(xyz), *[_fp(a) for a in args],
                                              C.c_float(gw), C.c_float(delta)), "mdk_set_ligand_atoms")

    def dock(self, rotcoef, rot_id):
        rc = np.ascontiguousarray(rotcoef, np.float32).reshape(-1, 9)
        ids = np.ascontiguousarray(rot_id, np.int32)
        B = len(rc)
        out = np.zeros(B * self.t, SORTSCORE_DTYPE)
        _check(lib().mdk_dock(self.h, _fp(rc), _ip(ids), C.c_int(B), C.c_void_p(out.ctypes.data)), "mdk_dock")
        return out

    def set_rotations(self, rotcoef):
        rc = np.ascontiguousarray(rotcoef, np.float32).reshape(-1, 9)
        self.nrot = len(rc)
        _check(lib().mdk_set_rotations(self.h, _fp(rc), C.c_int(len(rc))), "mdk_set_rotations")

    def dock_range(self, first, count):
        _check(lib().mdk_dock_range(self.h, C.c_int(first), C.c_int(count)), "mdk_dock_range")

    def sync(self):
        _check(lib().mdk_sync(self.h), "mdk_sync")

    def results(self, first, count):
        out = np.zeros(count * self.t, SORTSCORE_DTYPE)
        _check(lib().mdk_get_results(self.h, C.c_int(first), C.c_int(count), C.c_void_p(out.ctypes.data)),
               "mdk_get_results")
        return out

    def topn(self, nrot, n):
        out = np.zeros(n, SORTSCORE_DTYPE)
        _check(lib().mdk_topn(self.h, C.c_int(nrot), C.c_int(n), C.c_void_p(out.ctypes.data)), "mdk_topn")
        return out

    def topn_start(self, nrot, n):
        """Queue the merge of rotations [0, nrot) behind the batches queued so far."""
        _check(lib().mdk_topn_start(self.h, C.c_int(nrot), C.c_int(n)), "mdk_topn_start")
        self._topn_n = min(n, nrot * self.t)

    def topn_finish(self):
        """Wait for the started merge; its best min(n, nrot * t) results."""
        out = np.zeros(max(1, getattr(self, "_topn_n", 0)), SORTSCORE_DTYPE)
        _check(lib().mdk_topn_finish(self.h, C.c_void_p(out.ctypes.data)), "mdk_topn_finish")
        return out[:getattr(self, "_topn_n", 0)]

    def dock_range_timed(self, first, count):
        """dock_range + sync with HIP events around every stage of every batch: returns
        ({stage: summed ms over the full batches}, number of full batches)."""
        ms = (C.c_float * 16)()
        ns, nfull = C.c_int(0), C.c_int(0)
        _check(lib().mdk_dock_range_timed(self.h, C.c_int(first), C.c_int(count), ms, C.byref(ns), C.byref(nfull)),
               "mdk_dock_range_timed")
        return {STAGES[q]: ms[q] for q in range(ns.value)}, nfull.value

    def time_range(self, first, count, reps=1):
        ms = C.c_float(0.0)
        _check(lib().mdk_time_range(self.h, C.c_int(first), C.c_int(count), C.c_int(reps), C.byref(ms)),
               "mdk_time_range")
        return ms.value

    def profile_stages(self, first, count):
        ms = np.zeros(16, np.float32)
        n = C.c_int(0)
        _check(lib().mdk_profile_stages(self.h, C.c_int(first), C.c_int(count), _fp(ms), C.byref(n)),
               "mdk_profile_stages")
        return dict(zip(STAGES, [float(v) for v in ms[:n.value]]))

    def voxelize(self, rotcoef):
        rc = np.ascontiguousarray(rotcoef, np.float32).reshape(-1, 9)
        B = len(rc)
        gr = np.zeros(B * self.ng ** 3, np.float32)
        gi = np.zeros(B * self.ng ** 3, np.float32)
        _check(lib().mdk_voxelize(self.h, _fp(rc), C.c_int(B), _fp(gr), _fp(gi)), "mdk_voxelize")
        return gr.reshape(B, -1), gi.reshape(B, -1)

    def autotune(self, first, count):
        """Time the kernel variants on rotations [first, first+count) of the table and keep
        the fastest; returns the chosen (correlation, inverse, forward) variant numbers."""
        ch = (C.c_int * 3)()
        _check(lib().mdk_autotune(self.h, C.c_int(first), C.c_int(count), ch), "mdk_autotune")
        return ch[0], ch[1], ch[2]

    def score_spectral(self, rotcoef, t):
        """(B, 4): band-limited score at continuous translations t (B, 3) in voxels and
        its gradient (s, ds/dtx, ds/dty, ds/dtz)."""
        rc = np.ascontiguousarray(rotcoef, np.float32).reshape(-1, 9)
        tt = np.ascontiguousarray(t, np.float32).reshape(-1, 3)
        assert len(rc) == len(tt)
        out = np.zeros((len(rc), 4), np.float32)
        _check(lib().mdk_score_spectral(self.h, _fp(rc), _fp(tt), C.c_int(len(rc)), _fp(out)), "mdk_score_spectral")
        return out

    def score_detail(self, rec_rpsc, rotcoef, trans):
        """(B, 9) float32: real-space rPSC(+), rPSC(-), ELEC(+), ELEC(-), ACE(+), ACE(-), then
        rPSC, ELEC, ACE unsplit, of poses at grid translations trans (B, 3) int, the float
        sums of the reference's detailed output in its order; rec_rpsc is the receptor shape
        grid built without the ACE term."""
        rp = np.ascontiguousarray(rec_rpsc, np.float32).ravel()
        assert rp.size == self.ng ** 3
        rc = np.ascontiguousarray(rotcoef, np.float32).reshape(-1, 9)
        tr = np.ascontiguousarray(trans, np.int32).reshape(-1, 3)
        out = np.zeros((len(rc), 9), np.float32)
        _check(lib().mdk_score_detail(self.h, _fp(rp), _fp(rc), tr.ctypes.data_as(C.POINTER(C.c_int)),
                                      C.c_int(len(rc)), _fp(out)),
               "mdk_score_detail")
        return out

    def score_companion(self, rotcoef, t):
        """(B, 7): trilinear companion score and its gradient (s, ds/dt' (3), ds/dw (3))."""
        rc = np.ascontiguousarray(rotcoef, np.float32).reshape(-1, 9)
        tt = np.ascontiguousarray(t, np.float32).reshape(-1, 3)
        assert len(rc) == len(tt)
        out = np.zeros((len(rc), 7), np.float32)
        _check(lib().mdk_score_companion(self.h, _fp(rc), _fp(tt), C.c_int(len(rc)), _fp(out)), "mdk_score_companion")
        return out

    def score_map(self, rotcoef):
        rc = np.ascontiguousarray(rotcoef, np.float32).reshape(9)
        out = np.zeros(self.nf ** 3, np.float32)
        _check(lib().mdk_score_map(self.h, _fp(rc), _fp(out)), "mdk_score_map")
        return out


def copy_probe(nbytes, reps=10, device=0):
    """Milliseconds per device-to-device copy of nbytes (hipMemcpy blit kernel)."""
    ms = C.c_float(0.0)
    _check(lib().mdk_copy_probe(C.c_int(device), C.c_size_t(nbytes), C.c_int(reps), C.byref(ms)), "mdk_copy_probe")
    return ms.value


def fft1d(lines, sign=-1, device=0):
    a = np.ascontiguousarray(lines, np.complex64)
    n = a.shape[-1]
    count = a.size // n
    out = np.zeros_like(a)
    _check(lib().mdk_fft1d(C.c_int(device), C.c_int(n), C.c_int(count), _fp(a.view(np.float32)),
                           _fp(out.view(np.float32)), C.c_int(sign)), "mdk_fft1d")
    return out


# ---------------------------------------------------------------- host layer
def _hcheck(rc, what):
    if rc != 0:
        raise MdkError("%s failed: %s" % (what, lib().mdkh_last_error().decode()))


class Protein:
    """PDB atoms with CHARMM radius/charge and ACE per atom (mdkh_protein_*)."""

    def __init__(self, path):
        self.path = path
        h = C.c_void_p()
        _hcheck(lib().mdkh_protein_load(path.encode(), C.byref(h)), "mdkh_protein_load")
        self.h = h
        n = C.c_int(0)
        _hcheck(lib().mdkh_protein_size(h, C.byref(n)), "mdkh_protein_size")
        na = n.value
        self.xyz = np.zeros((na, 3), np.float32)
        self.radius = np.zeros(na, np.float32)
        self.charge = np.zeros(na, np.float32)
        self.ace = np.zeros(na, np.float32)
        self.center = np.zeros(3, np.float32)
        _hcheck(lib().mdkh_protein_atoms(h, _fp(self.xyz), _fp(self.radius), _fp(self.charge), _fp(self.ace),
                                         _fp(self.center)), "mdkh_protein_atoms")

    @property
    def na(self):
        return len(self.xyz)

    def __del__(self):
        try:
            lib().mdkh_protein_free(self.h)
        except Exception:
            pass


def autogrid(rec, lig, gw=1.2, table=11):
    ng = C.c_int(0)
    _hcheck(lib().mdkh_autogrid(rec.h, lig.h, C.c_float(gw), C.c_int(table), C.byref(ng)), "mdkh_autogrid")
    return ng.value


def rotation_set(n):
    """float32[n, 3]: built-in rotation set n as the host layer loads it (mdkh_rotation_set)."""
    cnt = C.c_int(0)
    _hcheck(lib().mdkh_rotation_set(C.c_int(n), None, C.byref(cnt)), "mdkh_rotation_set")
    out = np.zeros((cnt.value, 3), np.float32)
    _hcheck(lib().mdkh_rotation_set(C.c_int(n), _fp(out), C.byref(cnt)), "mdkh_rotation_set")
    return out


def grid_coord(ng, gw=1.2):
    out = np.zeros(ng, np.float32)
    _hcheck(lib().mdkh_grid_coord(C.c_int(ng), C.c_float(gw), _fp(out)), "mdkh_grid_coord")
    return out


def receptor_grids(rec, ng, gw=1.2, aceratio=1.0, rec_core=-45.0, eratio=1.0):
    gr = np.zeros(ng ** 3, np.float32)
    gi = np.zeros(ng ** 3, np.float32)
    _hcheck(lib().mdkh_receptor_voxel(rec.h, C.c_int(ng), C.c_float(gw), C.c_float(aceratio),
                                      C.c_float(rec_core), C.c_float(eratio), _fp(gr), _fp(gi)),
            "mdkh_receptor_voxel")
    return gr, gi


def run(args):
    """The megadock command line (e.g. ["-R", r, "-L", l, "-o", out])."""
    argv = ["megadock-hip"] + list(args)
    arr = (C.c_char_p * len(argv))(*[a.encode() for a in argv])
    _hcheck(lib().mdkh_run(C.c_int(len(argv)), arr), "mdkh_run")


def run_table(args):
    """The job-table farm (megadock-hip-dp), e.g. ["-tb", table, "-lg", log]: every job
    line of the table expands the table's PARAM line into a megadock command line."""
    argv = ["megadock-hip-dp"] + list(args)
    arr = (C.c_char_p * len(argv))(*[a.encode() for a in argv])
    _hcheck(lib().mdkh_run_table(C.c_int(len(argv)), arr), "mdkh_run_table")
