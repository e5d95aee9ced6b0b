"""Uneven multi-dataset batch (vdb_cluster_multi): every dataset's labels
equal a one-dataset call of the reference semantics (the oracle)."""
import sys
import time

import numpy as np
import pytest

import oracle

import variantdbscan_amd as vd
from variantdbscan_amd import datasets

pytestmark = pytest.mark.gpu

EPS = np.array([e for e in datasets.MULTI_EPS for m in datasets.MULTI_MINPTS])
MP = np.array([m for e in datasets.MULTI_EPS for m in datasets.MULTI_MINPTS], dtype=np.uint32)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_multi_small_against_oracle():
    rng = np.random.default_rng(3)
    sets = []
    for i, n in enumerate([1, 37, 1000, 5000, 12000, 3]):
        sets.append(datasets.blobs_isotropic(n, int(rng.integers(5, 30)), 0.5, 0.7, 900 + i)
                    if n > 10 else (rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)))
    got = vd.cluster_multi(sets, EPS, MP)
    for d, (x, y) in enumerate(sets):
        single = vd.cluster(x, y, EPS, MP)
        assert (got[d] == single).all(), d
        if len(x) >= 2:
            assert (got[d] == oracle.libvdbscan(x, y, EPS, MP)).all(), d


def test_multi_edge_datasets():
    # empty datasets between others, one-point sets, a set of identical points,
    # all-negative coordinates, isolated-point and never-core variants
    rng = np.random.default_rng(11)
    x0, y0 = datasets.blobs_isotropic(4000, 12, 0.5, 0.7, 77)
    sets = [
        (np.zeros(0), np.zeros(0)),
        (x0, y0),
        (np.zeros(0), np.zeros(0)),
        (np.array([3.0]), np.array([-4.0])),
        (np.full(50, 2.5), np.full(50, 2.5)),
        (-np.abs(x0[:900]) - 1.0, -np.abs(y0[:900]) - 1.0),
        (rng.uniform(0, 2, 300), rng.uniform(0, 2, 300)),
        (np.zeros(0), np.zeros(0)),
    ]
    eps = np.array([0.3, 0.3, 0.6, 1.0, -1.0, np.nan, 0.45, 0.45], dtype=np.float64)
    mp = np.array([4, 0, 8, 16, 0, 3, 2**31 + 5, 1], dtype=np.uint32)
    got = vd.cluster_multi(sets, eps, mp)
    fused = vd.cluster_multi(sets, eps, mp, path=vd.PATH_FUSED)
    for d, (x, y) in enumerate(sets):
        assert got[d].shape == (len(eps), len(x))
        assert (got[d] == fused[d]).all(), d
        if len(x) == 0:
            continue
        assert (got[d] == vd.cluster(x, y, eps, mp)).all(), d
        if len(x) >= 2 and not np.all(x == x[0]):
            assert (got[d] == oracle.libvdbscan(x, y, eps, mp)).all(), d


def test_multi_256_batch_equals_per_dataset_engine():
    # the one-pass batched run against every dataset clustered on its own by
    # the per-variant engine
    sets = datasets.multi_256()
    a = vd.cluster_multi(sets, EPS, MP)
    b = vd.cluster_multi(sets, EPS, MP, path=vd.PATH_FUSED)
    bad = [d for d in range(len(sets)) if not (a[d] == b[d]).all()]
    assert not bad, bad


def test_multi_256_uneven():
    sets = datasets.multi_256()
    ntot = sum(len(s[0]) for s in sets)
    say("multi_256:", len(sets), "datasets,", ntot, "points")
    t = time.time()
    got, st = vd.cluster_multi(sets, EPS, MP, return_stats=True)
    wall = time.time() - t
    say("datasets x variants / s (device %.1f ms): %.1f" %
        (st["ms_total"], len(sets) * len(EPS) / (st["ms_total"] / 1e3)), "wall %.2f s" % wall)
    rng = np.random.default_rng(0)
    small = [d for d in range(len(sets)) if len(sets[d][0]) <= 15_000]
    for d in rng.choice(small, size=min(6, len(small)), replace=False):
        x, y = sets[d]
        assert (got[d] == oracle.libvdbscan(x, y, EPS, MP)).all(), d
