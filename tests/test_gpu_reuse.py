"""GPU reuse emulation (VDB_REUSE_REFERENCE) against the reference.

tests/golden/reuse_t1.npz holds multi-variant labels of the compiled
reference built with NSEARCHTHREADS=1, where its greedy cluster reuse
(src/schedule.cpp:68-116, src/DBScan.cpp:157-518) is deterministic. The
library's reuse mode must give those labels bit for bit, through every
surface: vdb_cluster with the flag, libVDBSCAN with VDBSCAN_REUSE=1 and the
Python VariantDBSCAN(..., reuse=True). Random sets are checked against the
oracle's literal restatement (oracle.libvdbscan_reuse).
"""
import ctypes
import os
import sys

import numpy as np
import pytest

import oracle
from conftest import GOLDEN

import variantdbscan_amd as vd
from variantdbscan_amd import _lib, datasets

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _fixture():
    g = np.load(os.path.join(GOLDEN, "reuse_t1.npz"))
    names = sorted(set(k.rsplit("_", 1)[0] for k in g.files))
    return g, names


@pytest.mark.parametrize("prec", [vd.PREC_FP64, vd.PREC_FP32_RECHECK], ids=["fp64", "fp32"])
def test_reuse_matches_reference_fixture(prec):
    g, names = _fixture()
    for nm in names:
        x, y, eps, mp = g[nm + "_x"], g[nm + "_y"], g[nm + "_eps"], g[nm + "_minpts"]
        say("reuse fixture", nm, len(x), "points", len(eps), "variants")
        got = vd.cluster(x, y, eps, mp, reuse=True, precision=prec)
        bad = [v for v in range(len(eps)) if not (got[v] == g[nm + "_labels"][v]).all()]
        assert not bad, (nm, bad)


def test_reuse_through_the_drop_in_surfaces(monkeypatch):
    g, _ = _fixture()
    x, y, eps, mp, want = (g["plain10k_x"], g["plain10k_y"], g["plain10k_eps"],
                           g["plain10k_minpts"], g["plain10k_labels"])
    res = vd.VariantDBSCAN(eps, mp, x, y, 70, reuse=True)
    assert res.shape == (len(x), len(eps)) and (res.T == want).all()
    # libVDBSCAN itself, switched by the environment
    L = _lib.load()
    ret = np.zeros(len(x) * len(eps), dtype=np.uint32)
    xs, ys = np.ascontiguousarray(x), np.ascontiguousarray(y)
    es, ms = np.ascontiguousarray(eps), np.ascontiguousarray(mp, dtype=np.uint32)

    def call():
        rc = L.libVDBSCAN(ctypes.c_void_p(xs.ctypes.data), ctypes.c_void_p(ys.ctypes.data),
                          len(x), ctypes.c_void_p(es.ctypes.data), ctypes.c_void_p(ms.ctypes.data),
                          len(eps), 70, ctypes.c_void_p(ret.ctypes.data), False)
        assert rc == 0
        return ret.reshape(len(eps), len(x)).copy()

    monkeypatch.setenv("VDBSCAN_REUSE", "1")
    assert (call() == want).all()
    monkeypatch.delenv("VDBSCAN_REUSE")
    assert (call() == oracle.libvdbscan(x, y, eps, mp)).all()


def test_reuse_random_against_oracle():
    rng = np.random.default_rng(9)
    for trial in range(4):
        n = int(rng.integers(3000, 15000))
        x, y = datasets.aniso(n, seed=int(rng.integers(1, 10**6)),
                              n_blobs=int(rng.integers(4, 40)))
        E = sorted(set(np.round(rng.uniform(0.2, 1.2, 3), 2).tolist()))
        M = sorted(set(rng.integers(2, 40, 3).tolist()))
        eps = np.array([e for e in E for m in M])
        mp = np.array([m for e in E for m in M], dtype=np.uint32)
        p = rng.permutation(len(eps))
        eps, mp = eps[p], mp[p]
        say("reuse random", trial, n, "points", len(eps), "variants")
        want = oracle.libvdbscan_reuse(x, y, eps, mp)
        got = vd.cluster(x, y, eps, mp, reuse=True)
        bad = [v for v in range(len(eps)) if not (got[v] == want[v]).all()]
        assert not bad, (trial, bad)


def test_reuse_options_checked():
    x, y = datasets.plain_10k(seed=3, n=500)
    with pytest.raises(vd._lib.VdbError):
        vd.cluster(x, y, [0.5, 1.0], [4, 4], reuse=True, border=vd.BORDER_STANDARD)
    with pytest.raises(vd._lib.VdbError):
        vd.cluster(x, y, [0.5, np.nan], [4, 4], reuse=True)
    # one variant: reuse never applies, scratch labels
    got = vd.cluster(x, y, [0.7], [4], reuse=True)
    assert (got == oracle.libvdbscan(x, y, [0.7], [4])).all()
