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


def VariantDBSCAN(eps_array, mp_array, x_array, y_array, mbbsize=70, verbose=False):
    libvdbscan = _lib.load()

    def convert_type(in_array, new_dtype):          # src/variantdbscan.py:74-87
        ret_array = in_array
        if not isinstance(in_array, np.ndarray):
            ret_array = np.array(in_array, dtype=new_dtype)
        elif in_array.dtype != new_dtype:
            ret_array = np.array(ret_array, dtype=new_dtype)
        if not ret_array.flags["C_CONTIGUOUS"]:
            ret_array = np.ascontiguousarray(ret_array)
        return ret_array

    c_eps_array = np.atleast_1d(convert_type(eps_array, np.float64))
    c_mp_array = np.atleast_1d(convert_type(mp_array, np.uint32))
    c_x_array = np.atleast_1d(convert_type(x_array, np.float64))
    c_y_array = np.atleast_1d(convert_type(y_array, np.float64))
    if len(c_x_array) != len(c_y_array):
        raise ValueError("x_array and y_array differ in length")
    if len(c_eps_array) != len(c_mp_array):
        raise ValueError("eps_array and mp_array differ in length")

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
