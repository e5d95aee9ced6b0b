"""Batched engine at the boundaries of its pair sort. The cell pairs of all
minpts values of a pass are sorted by (minpts value, eps level): J*(K+1)
buckets. Up to 16 buckets the one-pass bucket sort uses 4-bit keys, up to
128 7-bit keys, beyond that the radix sort runs. Every case must give the
oracle's labels (reference rule) and scikit-learn's (standard rule)."""
import sys

import numpy as np
import pytest

import oracle

import variantdbscan_amd as vd
from variantdbscan_amd import datasets

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def aniso20k():
    return datasets.aniso(20_000, seed=31)


# (eps levels K, minpts values): buckets J*(K+1) = 16, 32, 128, 132
CASES = [(15, (4,)), (15, (4, 9)), (31, (3, 4, 9, 20)), (32, (3, 4, 9, 20))]
IDS = ["16-buckets", "32-buckets", "128-buckets", "132-radix"]


def _grid(levels, minpts):
    E = np.round(np.linspace(0.2, 0.8, levels), 6)
    eps = np.repeat(E, len(minpts))
    mp = np.tile(np.array(minpts, dtype=np.uint32), levels)
    return eps, mp


@pytest.mark.parametrize("levels,minpts", CASES, ids=IDS)
def test_batch_levels_reference_rule(aniso20k, levels, minpts):
    x, y = aniso20k
    eps, mp = _grid(levels, minpts)
    say("aniso 20k, %d eps levels x %d minpts, batch" % (levels, len(minpts)))
    got, st = vd.cluster(x, y, eps, mp, path=vd.PATH_BATCH, return_stats=True)
    assert st["path_used"] == vd.PATH_BATCH
    want = oracle.libvdbscan(x, y, eps, mp)
    for v in range(len(eps)):
        assert (got[v] == want[v]).all(), (levels, v)


@pytest.mark.parametrize("levels,minpts", CASES, ids=IDS)
def test_batch_levels_standard_rule(aniso20k, levels, minpts):
    from sklearn.cluster import DBSCAN
    x, y = aniso20k
    eps, mp = _grid(levels, minpts)
    say("aniso 20k, %d eps levels x %d minpts, batch, standard rule" % (levels, len(minpts)))
    got = vd.cluster(x, y, eps, mp, path=vd.PATH_BATCH, border=vd.BORDER_STANDARD)
    X = np.column_stack([x, y])
    # every third variant (scikit-learn takes a while per call)
    for v in range(0, len(eps), 3):
        want = DBSCAN(eps=eps[v], min_samples=int(mp[v]), algorithm="kd_tree").fit(X).labels_
        assert (got[v] == want).all(), (levels, v)
