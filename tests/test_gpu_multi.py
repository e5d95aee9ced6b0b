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
