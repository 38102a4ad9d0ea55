"""The C-ABI library loads and exports every entry point the headers declare; host-only
entry points work without a GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import megadock_amd as M
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(mdk_\w+)\s*\(", text, re.M)))


def test_every_declared_symbol_is_exported():
    names = declared("mdk_hip.h")
    assert len(names) >= 15
    lib = C.CDLL(M.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(M.ABI_SYMBOLS) == names


def test_rotation_coefficients_match_oracle():
    ang = O.load_rotations(54000)
    got = M.rotation_coefs(ang)
    for a in range(0, 54000, 13):
        assert np.array_equal(got[a], O.rotation_coef(ang[a])), a


def test_create_rejects_bad_sizes_without_touching_the_gpu():
    h = C.c_void_p()
    L = M.lib()
    assert L.mdk_create(0, 96, 40, 100, 1, 8, C.byref(h)) != 0  # nf != 2 ng
    assert b"2*ng" in L.mdk_last_error()
    assert L.mdk_create(0, 96, 48, 100, 0, 8, C.byref(h)) != 0  # t < 1
    assert L.mdk_create(0, 96, 48, 100, 17, 8, C.byref(h)) != 0  # t > 16
    assert L.mdk_create(0, 2 * 17, 17, 100, 1, 8, C.byref(h)) != 0  # prime 17
    assert L.mdk_version() == 1
    with pytest.raises(M.MdkError):
        M.Context(96, 40, 100)
