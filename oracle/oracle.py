"""ctypes front-end for the CPU oracle and the compiled reference library.

TEST INFRASTRUCTURE ONLY: the checker for the HIP library. Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.

* ``liboracle.so`` (built by ``make`` in this directory from vdb_oracle.c) is
  the C restatement of the reference semantics.
* ``_ref/libSharedVDBSCAN_ref_t<N>.so`` (built by ``make ref`` from the
  reference sources where they lie) is the reference itself; it is driven in a
  child process through ``_ref/ref_runner`` because the reference keeps global
  state between calls (src/DBScan.cpp:51, src/main.cpp:58).
"""
import ctypes
import json
import os
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")

_lib = None


def build():
    """Compile liboracle.so (and the reference build when its sources exist)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        i32p = ctypes.POINTER(ctypes.c_int32)
        i64p = ctypes.POINTER(ctypes.c_int64)
        L.vdbo_bin_order.argtypes = [dp, dp, ctypes.c_uint32, ctypes.c_int, i32p]
        L.vdbo_bin_order.restype = ctypes.c_int
        L.vdbo_threshold.argtypes = [ctypes.c_double]
        L.vdbo_threshold.restype = ctypes.c_double
        L.vdbo_dist2.argtypes = [ctypes.c_double] * 4
        L.vdbo_dist2.restype = ctypes.c_double
        L.vdbo_neighbor_csr.argtypes = [dp, dp, ctypes.c_uint32, ctypes.c_double, i64p, i32p]
        L.vdbo_neighbor_csr.restype = ctypes.c_int64
        L.vdbo_neighbor_counts.argtypes = [dp, dp, ctypes.c_uint32, ctypes.c_double, u32p]
        L.vdbo_neighbor_counts.restype = ctypes.c_int
        for fn in (L.vdbo_dbscan_literal, L.vdbo_dbscan_closed):
            fn.argtypes = [dp, dp, ctypes.c_uint32, ctypes.c_double, ctypes.c_uint32, u32p]
            fn.restype = ctypes.c_int
        L.vdbo_libVDBSCAN.argtypes = [dp, dp, ctypes.c_uint32, dp, u32p, ctypes.c_uint32,
                                      ctypes.c_int, u32p, ctypes.c_int]
        L.vdbo_libVDBSCAN.restype = ctypes.c_int
        L.vdbo_levels_closed.argtypes = [dp, dp, ctypes.c_uint32, ctypes.c_double, u32p,
                                         ctypes.c_uint32, u32p]
        L.vdbo_levels_closed.restype = ctypes.c_int
        L.vdbo_libVDBSCAN_reuse.argtypes = [dp, dp, ctypes.c_uint32, dp, u32p, ctypes.c_uint32,
                                            u32p]
        L.vdbo_libVDBSCAN_reuse.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def threshold(eps):
    return lib().vdbo_threshold(float(eps))


def dist2(px, py, qx, qy):
    return lib().vdbo_dist2(float(px), float(py), float(qx), float(qy))


def bin_order(x, y, num_bins=300):
    x, y = _f64(x), _f64(y)
    m = np.empty(len(x), dtype=np.int32)
    rc = lib().vdbo_bin_order(_p(x, ctypes.c_double), _p(y, ctypes.c_double), len(x),
                              num_bins, _p(m, ctypes.c_int32))
    if rc:
        raise ValueError("bins out of range (rc=%d)" % rc)
    return m


def neighbor_csr(x, y, eps):
    """(offsets int64[n+1], ids int32[P]) in caller order, ids ascending per row."""
    x, y = _f64(x), _f64(y)
    n = len(x)
    off = np.zeros(n + 1, dtype=np.int64)
    total = lib().vdbo_neighbor_csr(_p(x, ctypes.c_double), _p(y, ctypes.c_double), n,
                                    float(eps), _p(off, ctypes.c_int64), None)
    if total < 0:
        raise MemoryError("oracle grid allocation failed")
    ids = np.empty(max(total, 1), dtype=np.int32)
    lib().vdbo_neighbor_csr(_p(x, ctypes.c_double), _p(y, ctypes.c_double), n,
                            float(eps), _p(off, ctypes.c_int64), _p(ids, ctypes.c_int32))
    return off, ids[:total]


def neighbor_counts(x, y, eps):
    x, y = _f64(x), _f64(y)
    c = np.empty(len(x), dtype=np.uint32)
    rc = lib().vdbo_neighbor_counts(_p(x, ctypes.c_double), _p(y, ctypes.c_double), len(x),
                                    float(eps), _p(c, ctypes.c_uint32))
    if rc:
        raise MemoryError("oracle grid allocation failed")
    return c


def dbscan_binned(xb, yb, eps, minpts, closed=False):
    xb, yb = _f64(xb), _f64(yb)
    lab = np.empty(len(xb), dtype=np.uint32)
    fn = lib().vdbo_dbscan_closed if closed else lib().vdbo_dbscan_literal
    rc = fn(_p(xb, ctypes.c_double), _p(yb, ctypes.c_double), len(xb), float(eps),
            int(minpts), _p(lab, ctypes.c_uint32))
    if rc < 0:
        raise RuntimeError("oracle dbscan failed (rc=%d)" % rc)
    return lab


def libvdbscan(x, y, eps, minpts, mbbsize=70, closed=True):
    """Scratch-per-variant labels, retArr layout: shape (V, n) uint32."""
    x, y = _f64(x), _f64(y)
    eps = _f64(np.atleast_1d(eps))
    mp = np.ascontiguousarray(np.atleast_1d(minpts), dtype=np.uint32)
    n, v = len(x), len(eps)
    ret = np.zeros(n * v, dtype=np.uint32)
    rc = lib().vdbo_libVDBSCAN(_p(x, ctypes.c_double), _p(y, ctypes.c_double), n,
                               _p(eps, ctypes.c_double), _p(mp, ctypes.c_uint32), v,
                               int(mbbsize), _p(ret, ctypes.c_uint32), 1 if closed else 0)
    if rc:
        raise ValueError("oracle libVDBSCAN failed (rc=%d)" % rc)
    return ret.reshape(v, n)


def libvdbscan_reuse(x, y, eps, minpts):
    """The reference's multi-variant call with greedy cluster reuse as it runs
    with one search thread (deterministic): retArr layout, shape (V, n)."""
    x, y = _f64(x), _f64(y)
    eps = _f64(np.atleast_1d(eps))
    mp = np.ascontiguousarray(np.atleast_1d(minpts), dtype=np.uint32)
    n, v = len(x), len(eps)
    ret = np.zeros(n * v, dtype=np.uint32)
    rc = lib().vdbo_libVDBSCAN_reuse(_p(x, ctypes.c_double), _p(y, ctypes.c_double), n,
                                     _p(eps, ctypes.c_double), _p(mp, ctypes.c_uint32), v,
                                     _p(ret, ctypes.c_uint32))
    if rc:
        raise ValueError("oracle libVDBSCAN_reuse failed (rc=%d)" % rc)
    return ret.reshape(v, n)


def levels_closed(x, y, eps, minpts):
    """Labels (len(minpts), n) of the variants (eps, m) for every m in minpts,
    retArr layout; the same labels as libvdbscan(x, y, [eps]*k, minpts) but
    without a neighbour table (O(n * k) memory), for full-size workloads."""
    x, y = _f64(x), _f64(y)
    mp = np.ascontiguousarray(np.atleast_1d(minpts), dtype=np.uint32)
    n = len(x)
    out = np.zeros(len(mp) * n, dtype=np.uint32)
    rc = lib().vdbo_levels_closed(_p(x, ctypes.c_double), _p(y, ctypes.c_double), n,
                                  float(eps), _p(mp, ctypes.c_uint32), len(mp),
                                  _p(out, ctypes.c_uint32))
    if rc:
        raise ValueError("oracle levels_closed failed (rc=%d)" % rc)
    return out.reshape(len(mp), n)


def standard_labels(x, y, eps, minpts):
    """Textbook DBSCAN labels (int32, -1 = noise, clusters 0..) under the
    reference's eps predicate, for the library's standard border rule.

    A literal restatement of scikit-learn 1.7's ``dbscan_inner``
    (sklearn/cluster/_dbscan_inner.pyx): visit points in index order; an
    unlabelled core point starts a new cluster, grown depth-first; every
    point reached gets the cluster, core points reached are expanded. The
    neighbourhoods (self included) come from neighbor_csr, so only the
    labelling rule differs from libvdbscan. Core: at least minpts
    neighbours (minpts above INT_MAX: none, as in the reference's int).
    Pure Python over the table: small inputs only.
    """
    off, ids = neighbor_csr(x, y, eps)
    n = len(off) - 1
    cnt = np.diff(off)
    mp = int(minpts)
    core = (cnt >= mp) if mp <= 0x7fffffff else np.zeros(n, dtype=bool)
    lab = np.full(n, -1, dtype=np.int32)
    ids = ids.tolist()
    off = off.tolist()
    core = core.tolist()
    nxt = 0
    for i in range(n):
        if lab[i] != -1 or not core[i]:
            continue
        stack = [i]
        while stack:
            p = stack.pop()
            if lab[p] != -1:
                continue
            lab[p] = nxt
            if core[p]:
                for q in ids[off[p]:off[p + 1]]:
                    if lab[q] == -1:
                        stack.append(q)
        nxt += 1
    return lab


# ------------------------------------------------------- compiled reference

def ref_lib_path(threads=1):
    return os.path.join(REF_DIR, "libSharedVDBSCAN_ref_t%d.so" % threads)


def ref_available(threads=1):
    return os.path.exists(ref_lib_path(threads)) and os.path.exists(
        os.path.join(REF_DIR, "ref_runner"))


def run_reference(x, y, eps, minpts, mbbsize=70, threads=1, timeout=None):
    """Run the compiled reference libVDBSCAN once in a child process.

    Returns (labels uint32 (V, n), seconds of the libVDBSCAN call).
    """
    x, y = _f64(x), _f64(y)
    eps = _f64(np.atleast_1d(eps))
    mp = np.ascontiguousarray(np.atleast_1d(minpts), dtype=np.uint32)
    n, v = len(x), len(eps)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<IIiI", n, v, int(mbbsize), 0))
            f.write(x.tobytes())
            f.write(y.tobytes())
            f.write(eps.tobytes())
            f.write(mp.tobytes())
        r = subprocess.run([os.path.join(REF_DIR, "ref_runner"), ref_lib_path(threads), fin, fout],
                           capture_output=True, text=True, timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError("reference run failed: %s %s" % (r.stdout, r.stderr))
        line = [l for l in r.stdout.splitlines() if l.startswith("{")][-1]
        info = json.loads(line)
        lab = np.fromfile(fout, dtype=np.uint32).reshape(v, n)
    return lab, info["seconds"]
