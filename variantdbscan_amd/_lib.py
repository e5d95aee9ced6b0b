"""ctypes bindings of libSharedVDBSCAN.so (declared in include/vdbscan.h).

The library is the HIP build in ``variantdbscan_amd/lib``. There is no CPU
fallback: if the library is missing or a call fails, an exception is raised.
"""
import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(PKG_DIR, "lib")
LIB_NAME = "libSharedVDBSCAN.so"
LIB_PATH = os.path.join(LIB_DIR, LIB_NAME)

DEVICE_PTRS = 1
REUSE_REFERENCE = 2

PATH_AUTO, PATH_FUSED, PATH_TABLE, PATH_BATCH = 0, 1, 2, 3
PREC_FP64, PREC_FP32_RECHECK = 0, 1
BORDER_REFERENCE, BORDER_STANDARD = 0, 1

EXPORTED = ("libVDBSCAN", "vdb_last_error", "vdb_version", "vdb_set_stream", "vdb_release",
            "vdb_cluster", "vdb_rank_order", "vdb_neighbor_counts", "vdb_neighbor_table",
            "vdb_neighbor_table_digest", "vdb_cluster_multi", "vdb_read_xy_csv",
            "vdb_read_lonlat_csv", "vdb_read_variants_csv", "vdb_write_labels_csv", "vdb_free",
            "vdb_time_kernel", "vdb_kernel_time", "vdb_tuning", "vdb_cluster_noalloc",
            "vdb_batch_info")


class VdbOptions(ctypes.Structure):
    _fields_ = [("path", ctypes.c_int), ("precision", ctypes.c_int),
                ("table_budget_bytes", ctypes.c_uint64), ("verbose", ctypes.c_int),
                ("cells_per_eps", ctypes.c_int), ("border", ctypes.c_int)]


class VdbStats(ctypes.Structure):
    _fields_ = [("ms_total", ctypes.c_double), ("ms_prepare", ctypes.c_double),
                ("ms_count", ctypes.c_double), ("ms_scan", ctypes.c_double),
                ("ms_fill", ctypes.c_double), ("ms_uf", ctypes.c_double),
                ("ms_border", ctypes.c_double), ("ms_labels", ctypes.c_double),
                ("pairs", ctypes.c_uint64), ("rechecked", ctypes.c_uint64),
                ("candidates", ctypes.c_uint64), ("tiles", ctypes.c_uint32),
                ("path_used", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class VdbError(RuntimeError):
    pass


_lib = None


def load():
    """Load (once) the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("%s not found: build it with __graft_entry__.build() or "
                          "make -C variantdbscan_amd/csrc" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    L.libVDBSCAN.argtypes = [vp, vp, ctypes.c_uint, vp, vp, ctypes.c_uint, ctypes.c_int, vp,
                             ctypes.c_bool]
    L.libVDBSCAN.restype = ctypes.c_int
    L.vdb_last_error.argtypes = []
    L.vdb_last_error.restype = ctypes.c_char_p
    L.vdb_version.restype = ctypes.c_int
    L.vdb_set_stream.argtypes = [vp]
    L.vdb_set_stream.restype = ctypes.c_int
    L.vdb_release.restype = ctypes.c_int
    L.vdb_cluster.argtypes = [vp, vp, ctypes.c_uint32, vp, vp, ctypes.c_uint32, vp,
                              ctypes.c_uint32, ctypes.POINTER(VdbOptions),
                              ctypes.POINTER(VdbStats)]
    L.vdb_cluster.restype = ctypes.c_int
    L.vdb_rank_order.argtypes = [vp, vp, ctypes.c_uint32, vp, ctypes.c_uint32]
    L.vdb_rank_order.restype = ctypes.c_int
    L.vdb_neighbor_counts.argtypes = [vp, vp, ctypes.c_uint32, ctypes.c_double, vp,
                                      ctypes.c_uint32, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_uint64)]
    L.vdb_neighbor_counts.restype = ctypes.c_int
    L.vdb_neighbor_table.argtypes = [vp, vp, ctypes.c_uint32, ctypes.c_double, vp, vp, vp,
                                     ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64),
                                     ctypes.c_uint32, ctypes.c_int, ctypes.POINTER(VdbStats)]
    L.vdb_neighbor_table.restype = ctypes.c_int
    u64p = ctypes.POINTER(ctypes.c_uint64)
    L.vdb_neighbor_table_digest.argtypes = [vp, vp, ctypes.c_uint32, ctypes.c_double,
                                            ctypes.c_uint32, ctypes.c_int, u64p, u64p, u64p]
    L.vdb_neighbor_table_digest.restype = ctypes.c_int
    L.vdb_cluster_multi.argtypes = [vp, vp, vp, ctypes.c_uint32, vp, vp, ctypes.c_uint32, vp,
                                    ctypes.c_uint32, ctypes.POINTER(VdbOptions),
                                    ctypes.POINTER(VdbStats)]
    L.vdb_cluster_multi.restype = ctypes.c_int
    dpp = ctypes.POINTER(ctypes.POINTER(ctypes.c_double))
    upp = ctypes.POINTER(ctypes.POINTER(ctypes.c_uint32))
    u32p = ctypes.POINTER(ctypes.c_uint32)
    for nm in ("vdb_read_xy_csv", "vdb_read_lonlat_csv"):
        getattr(L, nm).argtypes = [ctypes.c_char_p, dpp, dpp, u32p]
        getattr(L, nm).restype = ctypes.c_int
    L.vdb_read_variants_csv.argtypes = [ctypes.c_char_p, dpp, upp, u32p]
    L.vdb_read_variants_csv.restype = ctypes.c_int
    L.vdb_write_labels_csv.argtypes = [ctypes.c_char_p, vp, vp, ctypes.c_uint32, vp,
                                       ctypes.c_uint32]
    L.vdb_write_labels_csv.restype = ctypes.c_int
    L.vdb_free.argtypes = [vp]
    L.vdb_free.restype = None
    L.vdb_time_kernel.argtypes = [ctypes.c_char_p]
    L.vdb_time_kernel.restype = ctypes.c_int
    L.vdb_kernel_time.argtypes = [ctypes.POINTER(ctypes.c_double),
                                  ctypes.POINTER(ctypes.c_uint64)]
    L.vdb_kernel_time.restype = ctypes.c_int
    L.vdb_tuning.argtypes = [ctypes.c_char_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
    L.vdb_tuning.restype = ctypes.c_int
    L.vdb_batch_info.argtypes = [ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint32]
    L.vdb_batch_info.restype = ctypes.c_int
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        msg = load().vdb_last_error()
        raise VdbError("vdbscan error %d: %s" % (rc, msg.decode() if msg else ""))


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def cluster(x, y, eps, minpts, path=PATH_AUTO, precision=PREC_FP64, verbose=False,
            table_budget_bytes=0, return_stats=False, cells_per_eps=0, border=BORDER_REFERENCE,
            reuse=False):
    """Labels (V, n) for host arrays; retArr layout of libVDBSCAN.
    cells_per_eps (1..8) sets the per-variant grid's cells per eps; 0 lets
    the library choose. border=BORDER_REFERENCE gives the reference's uint32
    labels (0 = noise, clusters 1..); BORDER_STANDARD gives int32 labels with
    textbook DBSCAN borders (-1 = noise, clusters 0.. by smallest core index,
    scikit-learn's labelling). reuse=True gives the labels of the
    reference's greedy cluster reuse with one search thread
    (VDB_REUSE_REFERENCE)."""
    L = load()
    x, y = _f64(x), _f64(y)
    eps = _f64(np.atleast_1d(eps))
    mp = _u32(np.atleast_1d(minpts))
    if len(x) != len(y) or len(eps) != len(mp):
        raise ValueError("length mismatch")
    n, v = len(x), len(eps)
    out = np.zeros(n * v, dtype=np.uint32)
    opt = VdbOptions(path, precision, table_budget_bytes, 1 if verbose else 0, cells_per_eps,
                     border)
    st = VdbStats()
    _check(L.vdb_cluster(_ptr(x), _ptr(y), n, _ptr(eps), _ptr(mp), v, _ptr(out),
                         REUSE_REFERENCE if reuse else 0, ctypes.byref(opt), ctypes.byref(st)))
    lab = out.reshape(v, n)
    if border == BORDER_STANDARD:
        lab = lab.view(np.int32)
    return (lab, st.as_dict()) if return_stats else lab


def cluster_device(x_ptr, y_ptr, n, eps, minpts, labels_ptr, path=PATH_AUTO,
                   precision=PREC_FP64, stats=True, cells_per_eps=0, border=BORDER_REFERENCE):
    """Device-resident inputs/outputs (raw device pointers, e.g. tensor.data_ptr())."""
    L = load()
    eps = _f64(np.atleast_1d(eps))
    mp = _u32(np.atleast_1d(minpts))
    opt = VdbOptions(path, precision, 0, 0, cells_per_eps, border)
    st = VdbStats()
    _check(L.vdb_cluster(ctypes.c_void_p(x_ptr), ctypes.c_void_p(y_ptr), n, _ptr(eps),
                         _ptr(mp), len(eps), ctypes.c_void_p(labels_ptr), DEVICE_PTRS,
                         ctypes.byref(opt), ctypes.byref(st) if stats else None))
    return st.as_dict()


def time_kernel(name):
    """Start timing every launch of the named library kernel (None stops)."""
    _check(load().vdb_time_kernel(name.encode() if name else None))


def kernel_time():
    """(ms, launches) of the timed kernel since the last read."""
    ms, n = ctypes.c_double(0), ctypes.c_uint64(0)
    _check(load().vdb_kernel_time(ctypes.byref(ms), ctypes.byref(n)))
    return ms.value, n.value


def tuning():
    """The autotuner's state: {"cache": path, "tuned_in_process": N,
    "entries": {key: choice}}."""
    import json
    n = ctypes.c_uint64(0)
    _check(load().vdb_tuning(None, 0, ctypes.byref(n)))
    buf = ctypes.create_string_buffer(n.value + 1)
    _check(load().vdb_tuning(buf, n.value + 1, ctypes.byref(n)))
    return json.loads(buf.value.decode())


def batch_info():
    """Shape of the last batched run: dict of points, levels, minpts_values,
    list_len, listed_points, cells, cell_pairs, shape, snapshots, and the
    union-find worklist totals fresh_pairs, carried_pairs, test_records."""
    keys = ("points", "levels", "minpts_values", "list_len", "listed_points", "cells",
            "cell_pairs", "shape", "snapshots", "fresh_pairs", "carried_pairs", "test_records")
    a = (ctypes.c_uint64 * len(keys))()
    _check(load().vdb_batch_info(a, len(keys)))
    return {k: int(a[i]) for i, k in enumerate(keys)}


def rank_order(x, y):
    L = load()
    x, y = _f64(x), _f64(y)
    m = np.zeros(len(x), dtype=np.uint32)
    _check(L.vdb_rank_order(_ptr(x), _ptr(y), len(x), _ptr(m), 0))
    return m


def neighbor_counts(x, y, eps, precision=PREC_FP64, return_rechecked=False):
    L = load()
    x, y = _f64(x), _f64(y)
    c = np.zeros(len(x), dtype=np.uint32)
    rc = ctypes.c_uint64(0)
    _check(L.vdb_neighbor_counts(_ptr(x), _ptr(y), len(x), float(eps), _ptr(c), 0,
                                 precision, ctypes.byref(rc)))
    return (c, rc.value) if return_rechecked else c


def neighbor_table(x, y, eps, with_d2=False, precision=PREC_FP64):
    """(offsets uint64[n+1], ids uint32[P], d2 float32[P] or None), caller order."""
    L = load()
    x, y = _f64(x), _f64(y)
    n = len(x)
    off = np.zeros(n + 1, dtype=np.uint64)
    total = ctypes.c_uint64(0)
    st = VdbStats()
    _check(L.vdb_neighbor_table(_ptr(x), _ptr(y), n, float(eps), _ptr(off), None, None, 0,
                                ctypes.byref(total), 0, precision, ctypes.byref(st)))
    cap = max(int(total.value), 1)
    ids = np.zeros(cap, dtype=np.uint32)
    d2 = np.zeros(cap, dtype=np.float32) if with_d2 else None
    _check(L.vdb_neighbor_table(_ptr(x), _ptr(y), n, float(eps), _ptr(off), _ptr(ids),
                                _ptr(d2) if with_d2 else None, cap, ctypes.byref(total), 0,
                                precision, ctypes.byref(st)))
    t = int(total.value)
    return off, ids[:t], (d2[:t] if with_d2 else None)


def last_error():
    msg = load().vdb_last_error()
    return msg.decode() if msg else ""


def table_digest(x, y, eps, precision=PREC_FP64, device=False, n=None):
    """(digest, pairs, rechecked) of the caller-order eps-neighbour table.
    With device=True, x and y are raw device pointers and n the point count."""
    L = load()
    d, p, r = ctypes.c_uint64(0), ctypes.c_uint64(0), ctypes.c_uint64(0)
    if device:
        _check(L.vdb_neighbor_table_digest(ctypes.c_void_p(x), ctypes.c_void_p(y), n,
                                           float(eps), DEVICE_PTRS, precision, ctypes.byref(d),
                                           ctypes.byref(p), ctypes.byref(r)))
    else:
        x, y = _f64(x), _f64(y)
        _check(L.vdb_neighbor_table_digest(_ptr(x), _ptr(y), len(x), float(eps), 0, precision,
                                           ctypes.byref(d), ctypes.byref(p), ctypes.byref(r)))
    return d.value, p.value, r.value


def digest_of_csr(offsets, ids):
    """The same digest computed on the host from a CSR table (numpy, uint64)."""
    offsets = np.asarray(offsets, dtype=np.int64)
    rows = np.repeat(np.arange(len(offsets) - 1, dtype=np.uint64), np.diff(offsets))
    z = (rows << np.uint64(32)) | np.asarray(ids, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = z + np.uint64(0x9e3779b97f4a7c15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xbf58476d1ce4e5b9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94d049bb133111eb)
        z = z ^ (z >> np.uint64(31))
    return int(z.sum(dtype=np.uint64))


def cluster_multi(sets, eps, minpts, path=PATH_AUTO, precision=PREC_FP64, return_stats=False,
                  border=BORDER_REFERENCE):
    """Cluster a list of (x, y) datasets under the same variants in one call.
    Returns a list of (V, n_d) label arrays (uint32; int32 with
    BORDER_STANDARD)."""
    L = load()
    eps = _f64(np.atleast_1d(eps))
    mp = _u32(np.atleast_1d(minpts))
    sizes = np.array([len(s[0]) for s in sets], dtype=np.uint64)
    offs = np.zeros(len(sets) + 1, dtype=np.uint64)
    offs[1:] = np.cumsum(sizes)
    x = _f64(np.concatenate([s[0] for s in sets])) if sets else np.zeros(0)
    y = _f64(np.concatenate([s[1] for s in sets])) if sets else np.zeros(0)
    V = len(eps)
    out = np.zeros(int(offs[-1]) * V, dtype=np.uint32)
    opt = VdbOptions(path, precision, 0, 0, 0, border)
    st = VdbStats()
    _check(L.vdb_cluster_multi(_ptr(x), _ptr(y), _ptr(offs), len(sets), _ptr(eps), _ptr(mp), V,
                               _ptr(out), 0, ctypes.byref(opt), ctypes.byref(st)))
    if border == BORDER_STANDARD:
        out = out.view(np.int32)
    res = [out[int(offs[d]) * V:int(offs[d + 1]) * V].reshape(V, int(sizes[d]))
           for d in range(len(sets))]
    return (res, st.as_dict()) if return_stats else res


def cluster_multi_device(x_ptr, y_ptr, offsets, eps, minpts, labels_ptr, path=PATH_AUTO,
                         precision=PREC_FP64, border=BORDER_REFERENCE):
    """Device-resident variant of cluster_multi (offsets is a host array)."""
    L = load()
    eps = _f64(np.atleast_1d(eps))
    mp = _u32(np.atleast_1d(minpts))
    offs = np.ascontiguousarray(offsets, dtype=np.uint64)
    opt = VdbOptions(path, precision, 0, 0, 0, border)
    st = VdbStats()
    _check(L.vdb_cluster_multi(ctypes.c_void_p(x_ptr), ctypes.c_void_p(y_ptr), _ptr(offs),
                               len(offs) - 1, _ptr(eps), _ptr(mp), len(eps),
                               ctypes.c_void_p(labels_ptr), DEVICE_PTRS, ctypes.byref(opt),
                               ctypes.byref(st)))
    return st.as_dict()


# ------------------------------------------------------------ files (host only)
def _take(L, p, n, dtype):
    """Copy a library-allocated array of n values into numpy and release it."""
    try:
        return np.ctypeslib.as_array(p, shape=(n,)).astype(dtype, copy=True) if n else \
            np.zeros(0, dtype=dtype)
    finally:
        L.vdb_free(ctypes.cast(p, ctypes.c_void_p))


def read_points_csv(path, lonlat=False):
    """(x, y) float64 from an x,y CSV (the reference harness's format) or, with
    lonlat=True, a lon,lat[,TEC] CSV (the GPU prototype's: x = latitude moved
    to [0, 180], y = longitude)."""
    L = load()
    px = ctypes.POINTER(ctypes.c_double)()
    py = ctypes.POINTER(ctypes.c_double)()
    n = ctypes.c_uint32(0)
    fn = L.vdb_read_lonlat_csv if lonlat else L.vdb_read_xy_csv
    _check(fn(os.fsencode(path), ctypes.byref(px), ctypes.byref(py), ctypes.byref(n)))
    return _take(L, px, n.value, np.float64), _take(L, py, n.value, np.float64)


def read_variants_csv(path):
    """(eps float64, minpts uint32) from an eps,minpts CSV, file order."""
    L = load()
    pe = ctypes.POINTER(ctypes.c_double)()
    pm = ctypes.POINTER(ctypes.c_uint32)()
    n = ctypes.c_uint32(0)
    _check(L.vdb_read_variants_csv(os.fsencode(path), ctypes.byref(pe), ctypes.byref(pm),
                                   ctypes.byref(n)))
    return _take(L, pe, n.value, np.float64), _take(L, pm, n.value, np.uint32)


def write_labels_csv(path, x, y, labels):
    """Write x,y,label_0,..,label_{V-1} per point; labels (V, n) uint32."""
    L = load()
    x, y = _f64(x), _f64(y)
    lab = np.ascontiguousarray(np.atleast_2d(labels), dtype=np.uint32)
    if lab.shape[1] != len(x) or len(x) != len(y):
        raise ValueError("length mismatch")
    _check(L.vdb_write_labels_csv(os.fsencode(path), _ptr(x), _ptr(y), len(x), _ptr(lab),
                                  lab.shape[0]))
