"""The C-ABI library loads and exports every entry point declared in
include/gaspacho2.h; argument validation works without a GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "gaspacho2.h")


def _declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(gp_\w+)\s*\(", txt, flags=re.M)
    return sorted(set(n for n in names if n != "gp_outer_cb"))


def test_header_declares_entry_points():
    names = _declared()
    for must in ("gp_create", "gp_elbo", "gp_grad", "gp_irls_step", "gp_lbfgs_step", "gp_fit"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from gaspacho2_amd import lib
    L = lib.load()
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing


def test_create_rejects_bad_arguments_without_device():
    from gaspacho2_amd import lib
    L = lib.load()
    h = C.c_void_p()
    nh = np.array([0, 1, 5], np.int32)
    Yp = np.array([0, 1], np.int32)
    Yi = np.array([7], np.int32)        # row index out of range for N = 3
    Yx = np.array([1.0])
    z = np.ones(3)
    st = L.gp_create(C.byref(h), 3, 1, 4, 2, 2, nh.ctypes.data_as(C.POINTER(C.c_int)),
                     Yp.ctypes.data_as(C.POINTER(C.c_int)), Yi.ctypes.data_as(C.POINTER(C.c_int)),
                     Yx.ctypes.data_as(C.POINTER(C.c_double)), 1,
                     z.ctypes.data_as(C.POINTER(C.c_double)), z.ctypes.data_as(C.POINTER(C.c_double)),
                     np.zeros(6).ctypes.data_as(C.POINTER(C.c_double)),
                     np.zeros(8).ctypes.data_as(C.POINTER(C.c_double)), 0, 0)
    assert st == 1 and b"row index" in L.gp_last_error()
    assert not h.value


def test_init_ta_matches_oracle():
    from gaspacho2_amd import lib
    from oracle import oracle
    assert np.array_equal(lib.init_ta_with_per(16, 4), oracle.init_ta_with_per(16, 4))


def test_encoder_create_rejects_bad_shapes_without_device():
    from gaspacho2_amd import lib
    L = lib.load()
    h = C.c_void_p()
    for args, msg in (((2000, 500, 256, 10, 1024), b"multiples of 64"),
                      ((2000, 512, 256, 10, 1000), b"batch"),
                      ((2000, 512, 256, 40, 1024), b"Q <= 32")):
        st = L.gp_enc_create(C.byref(h), *args, 0, 0)
        assert st == 1 and msg in L.gp_last_error(), L.gp_last_error()
        assert not h.value
