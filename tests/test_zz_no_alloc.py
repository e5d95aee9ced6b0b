"""vdb_cluster_noalloc (the allocation-free entry point) against vdb_cluster.

Two calls in a row through the C ABI, the second with a smaller input, each
compared bit for bit with vdb_cluster on the same device arrays and options
(the fused per-variant engine). The largest call takes the maximum point count
at a small eps (about one cell per point: the most scratch a call can need).
"""
import ctypes

import numpy as np
import pytest

import variantdbscan_amd as vd
from variantdbscan_amd import _lib, datasets

pytestmark = pytest.mark.gpu

MAX_POINTS = 1 << 18


def _bind():
    L = _lib.load()
    vp = ctypes.c_void_p
    L.vdb_cluster_noalloc.argtypes = [vp, vp, ctypes.c_uint32, vp, vp, ctypes.c_uint32, vp,
                                      ctypes.POINTER(_lib.VdbOptions),
                                      ctypes.POINTER(_lib.VdbStats)]
    L.vdb_cluster_noalloc.restype = ctypes.c_int
    return L


class _Hip:
    """Device buffers through the HIP runtime the library itself links (found
    by its soname among the loaded objects), so the test needs no framework."""

    def __init__(self):
        _lib.load()
        H = ctypes.CDLL("libamdhip64.so.7")
        vp = ctypes.c_void_p
        H.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
        H.hipFree.argtypes = [vp]
        H.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
        for f in (H.hipMalloc, H.hipFree, H.hipMemcpy, H.hipDeviceSynchronize):
            f.restype = ctypes.c_int
        self.H = H
        self.bufs = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.H.hipMalloc(ctypes.byref(p), max(nbytes, 256)) == 0
        self.bufs.append(p)
        return p.value

    def put(self, a):
        a = np.ascontiguousarray(a)
        d = self.alloc(a.nbytes)
        assert self.H.hipMemcpy(d, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return d

    def get(self, d, count, dtype):
        a = np.empty(count, dtype=dtype)
        assert self.H.hipMemcpy(a.ctypes.data, d, a.nbytes, 2) == 0  # device to host
        return a

    def free(self):
        for p in self.bufs:
            self.H.hipFree(p)
        self.bufs = []


def _run(L, x, y, eps, mp, precision=vd.PREC_FP64, border=vd.BORDER_REFERENCE):
    hip = _Hip()
    n = len(x)
    e = np.ascontiguousarray(eps, dtype=np.float64)
    m = np.ascontiguousarray(mp, dtype=np.uint32)
    dx, dy = hip.put(np.asarray(x, np.float64)), hip.put(np.asarray(y, np.float64))
    dgot = hip.put(np.full(len(e) * n, 7, dtype=np.uint32))
    dref = hip.alloc(4 * len(e) * n)
    opt = _lib.VdbOptions(vd.PATH_FUSED, precision, 0, 0, 0, border)
    st = _lib.VdbStats()
    rc = L.vdb_cluster_noalloc(ctypes.c_void_p(dx), ctypes.c_void_p(dy), n,
                               ctypes.c_void_p(e.ctypes.data), ctypes.c_void_p(m.ctypes.data),
                               len(e), ctypes.c_void_p(dgot), ctypes.byref(opt), ctypes.byref(st))
    assert rc == 0, _lib.last_error()
    assert st.path_used == vd.PATH_FUSED
    vd.cluster_device(dx, dy, n, e, m, dref, path=vd.PATH_FUSED, precision=precision,
                      cells_per_eps=2, border=border)
    assert hip.H.hipDeviceSynchronize() == 0
    got, ref = hip.get(dgot, len(e) * n, np.uint32), hip.get(dref, len(e) * n, np.uint32)
    hip.free()
    return got, ref


@pytest.mark.parametrize("prec,border", [(vd.PREC_FP64, vd.BORDER_REFERENCE),
                                         (vd.PREC_FP32_RECHECK, vd.BORDER_STANDARD)],
                         ids=["fp64-ref", "fp32-std"])
def test_noalloc_equals_vdb_cluster_twice(prec, border):
    L = _bind()
    # first call: the maximum size, small eps (a cell per point or so)
    x, y = datasets.aniso(MAX_POINTS, seed=71)
    got, ref = _run(L, x, y, [0.05, 0.3, 0.3], [2, 4, 9], prec, border)
    assert (got == ref).all()
    # second call on the same scratch: a smaller input, larger eps
    x, y = datasets.plain_10k(seed=72, n=30_000)
    got, ref = _run(L, x, y, [0.8, 0.2], [5, 3], prec, border)
    assert (got == ref).all()


def test_noalloc_rejects_oversize_and_bad_options():
    L = _bind()
    opt = _lib.VdbOptions(vd.PATH_FUSED, 0, 0, 0, 0, 0)
    e = np.array([0.5])
    m = np.array([4], dtype=np.uint32)
    # size checks come before any device work: fake non-null device pointers
    rc = L.vdb_cluster_noalloc(ctypes.c_void_p(256), ctypes.c_void_p(256), MAX_POINTS + 1,
                               ctypes.c_void_p(e.ctypes.data), ctypes.c_void_p(m.ctypes.data), 1,
                               ctypes.c_void_p(256), ctypes.byref(opt), None)
    assert rc == 5 and b"262144" in L.vdb_last_error()
    bad = _lib.VdbOptions(vd.PATH_BATCH, 0, 0, 0, 0, 0)
    rc = L.vdb_cluster_noalloc(ctypes.c_void_p(256), ctypes.c_void_p(256), 10,
                               ctypes.c_void_p(e.ctypes.data), ctypes.c_void_p(m.ctypes.data), 1,
                               ctypes.c_void_p(256), ctypes.byref(bad), None)
    assert rc == 1
