"""Drop-in Python surface of the reference (src/variantdbscan.py:32-109).

``VariantDBSCAN(eps_array, mp_array, x_array, y_array, mbbsize=70,
verbose=False)`` takes the same arguments, converts them the same way
(float64 / uint32, C-contiguous), calls ``libVDBSCAN`` of the HIP library,
and returns the same shapes: ``(N,)`` uint32 for one variant, the transposed
``(N, V)`` view otherwise. 0 = noise.

Differences, all extensions: the library is loaded from this package's
``lib/`` directory instead of ``/usr/local/lib/``; a nonzero return code of
``libVDBSCAN`` (invalid input, device failure) raises instead of returning
zeros; mismatched x/y lengths raise (the reference reads past the shorter
array).
"""
import ctypes

import numpy as np

from . import _lib


def VariantDBSCAN(eps_array, mp_array, x_array, y_array, mbbsize=70, verbose=False,
                  reuse=False):
    """reuse=True (extension): a multi-variant call returns the labels of the
    reference's greedy cluster reuse as it runs with one search thread (see
    VDB_REUSE_REFERENCE in include/vdbscan.h); the default gives every
    variant its one-variant labels."""
    libvdbscan = _lib.load()
    # float64 / uint32, C-contiguous (the conversions of src/variantdbscan.py:74-96)
    c_eps_array = np.atleast_1d(np.ascontiguousarray(eps_array, dtype=np.float64))
    c_mp_array = np.atleast_1d(np.ascontiguousarray(mp_array, dtype=np.uint32))
    c_x_array = np.atleast_1d(np.ascontiguousarray(x_array, dtype=np.float64))
    c_y_array = np.atleast_1d(np.ascontiguousarray(y_array, dtype=np.float64))
    if len(c_x_array) != len(c_y_array):
        raise ValueError("x_array and y_array differ in length")
    if len(c_eps_array) != len(c_mp_array):
        raise ValueError("eps_array and mp_array differ in length")

    if reuse:
        if int(mbbsize) <= 0:
            raise _lib.VdbError("MBBsize must be positive")
        results = _lib.cluster(c_x_array, c_y_array, c_eps_array, c_mp_array, reuse=True,
                               verbose=verbose).reshape(-1)
        if len(c_eps_array) > 1:
            results = results.reshape([len(c_eps_array), len(c_x_array)]).T
        return results
    results = np.zeros(len(c_x_array) * len(c_eps_array), dtype=np.uint32)  # :98
    res = libvdbscan.libVDBSCAN(
        ctypes.c_void_p(c_x_array.ctypes.data), ctypes.c_void_p(c_y_array.ctypes.data),
        len(c_x_array), ctypes.c_void_p(c_eps_array.ctypes.data),
        ctypes.c_void_p(c_mp_array.ctypes.data), len(c_eps_array), int(mbbsize),
        ctypes.c_void_p(results.ctypes.data), bool(verbose))
    if res != 0:
        raise _lib.VdbError("libVDBSCAN failed (%d): %s" % (res, _lib.last_error()))

    if len(c_eps_array) > 1:                          # :106-107
        results = results.reshape([len(c_eps_array), len(c_x_array)]).T
    return results
