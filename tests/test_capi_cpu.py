"""Host-side checks of the C-ABI library that need no GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT

import variantdbscan_amd as vd
from variantdbscan_amd import _lib


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "vdbscan.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(libVDBSCAN|vdb_[a-z_]+)\s*\(", text)))


def test_library_exists_and_loads():
    assert os.path.exists(_lib.LIB_PATH)
    L = _lib.load()
    assert L.vdb_version() >= 10000


def test_every_declared_symbol_is_exported():
    decl = _declared_symbols()
    assert "libVDBSCAN" in decl and "vdb_cluster" in decl
    L = ctypes.CDLL(_lib.LIB_PATH)
    for name in decl:
        assert hasattr(L, name), name
    assert set(decl) == set(_lib.EXPORTED)


def test_reference_signature_binds_like_the_reference_wrapper():
    # the reference's ctypes argtypes (src/variantdbscan.py:66-69)
    import numpy.ctypeslib as npct
    from ctypes import c_bool, c_int, c_uint
    lib = npct.load_library(os.path.basename(_lib.LIB_PATH), os.path.dirname(_lib.LIB_PATH))
    a1d = npct.ndpointer(dtype=ctypes.c_double, ndim=1, flags="CONTIGUOUS")
    u1d = npct.ndpointer(dtype=ctypes.c_uint, ndim=1, flags="CONTIGUOUS")
    lib.libVDBSCAN.restype = c_int
    lib.libVDBSCAN.argtypes = [a1d, a1d, c_uint, a1d, u1d, c_uint, c_int, u1d, c_bool]
    # empty dataset: returns 0 without touching the device
    e = np.zeros(0)
    assert lib.libVDBSCAN(e, e, 0, np.array([1.0]), np.array([4], dtype=np.uint32), 1, 70,
                          np.zeros(0, dtype=np.uint32), False) == 0


def test_invalid_mbbsize_rejected_without_device():
    L = _lib.load()
    x = np.zeros(3)
    r = np.zeros(3, dtype=np.uint32)
    rc = L.libVDBSCAN(ctypes.c_void_p(x.ctypes.data), ctypes.c_void_p(x.ctypes.data), 3,
                      ctypes.c_void_p(np.array([1.0]).ctypes.data),
                      ctypes.c_void_p(np.array([4], dtype=np.uint32).ctypes.data), 1, 0,
                      ctypes.c_void_p(r.ctypes.data), False)
    assert rc == 1
    assert "MBBsize" in _lib.last_error()


def test_wrapper_rejects_length_mismatch():
    with pytest.raises(ValueError):
        vd.VariantDBSCAN([1.0], [4], [1.0, 2.0], [1.0])
    with pytest.raises(ValueError):
        vd.VariantDBSCAN([1.0, 2.0], [4], [1.0], [1.0])


def test_datasets_are_seeded():
    from variantdbscan_amd import datasets
    a = datasets.plain_10k(seed=3)
    b = datasets.plain_10k(seed=3)
    assert (a[0] == b[0]).all() and (a[1] == b[1]).all()
    x, y = a
    assert len(x) == 10_000
    assert x.min() >= -90 and x.max() <= 90 and y.min() >= -180 and y.max() <= 180
    s = datasets.multi_sizes()
    assert len(s) == 256 and s.min() >= 1000 and s.max() <= 500_000


def test_reference_package_names_resolve():
    # the reference's package and module names (python_package/setup.py,
    # src/vtest.py) give the drop-in call
    import importlib
    import os
    import sys
    import variantdbscan_amd
    compat = os.path.join(os.path.dirname(variantdbscan_amd.__file__), "compat")
    sys.path.insert(0, compat)
    try:
        for mod in ("vdbscan", "vdbscan.variantdbscan", "variantdbscan"):
            m = importlib.import_module(mod)
            assert m.VariantDBSCAN is variantdbscan_amd.VariantDBSCAN, mod
    finally:
        sys.path.remove(compat)


def _cluster_rc(opt):
    L = _lib.load()
    x = np.array([0.0, 1.0, 2.0])
    out = np.zeros(3, dtype=np.uint32)
    eps = np.array([1.0])
    mp = np.array([2], dtype=np.uint32)
    return L.vdb_cluster(_lib._ptr(x), _lib._ptr(x), 3, _lib._ptr(eps), _lib._ptr(mp), 1,
                         _lib._ptr(out), 0, ctypes.byref(opt), None)


def test_options_checked_before_device_work():
    # unknown values are refused before any device call
    O = _lib.VdbOptions
    assert _cluster_rc(O(7, 0, 0, 0, 0, 0)) == 1
    assert _cluster_rc(O(0, 2, 0, 0, 0, 0)) == 1
    assert _cluster_rc(O(0, 0, 0, 0, 0, 2)) == 1
    assert "border" in _lib.last_error()


def test_options_struct_matches_header():
    # the ctypes mirror has the header's fields, in order
    text = open(os.path.join(ROOT, "include", "vdbscan.h")).read()
    body = re.search(r"typedef struct vdb_options \{(.*?)\} vdb_options;", text, re.S).group(1)
    names = re.findall(r"\b([a-z_]+);", body)
    assert names == [f for f, _ in _lib.VdbOptions._fields_]
