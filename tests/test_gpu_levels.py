"""Batched engine at the boundary of its pair-bucket sort: up to 15 eps levels
the cell pairs are bucketed by level in one counting pass, from 16 levels on
by the radix sort. Both must give the oracle's labels (reference rule) and
scikit-learn's (standard rule)."""
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


@pytest.mark.parametrize("levels", [15, 16])
def test_batch_levels_reference_rule(aniso20k, levels):
    x, y = aniso20k
    E = np.round(np.linspace(0.2, 0.8, levels), 6)
    eps = np.repeat(E, 2)
    mp = np.tile(np.array([4, 9], dtype=np.uint32), levels)
    say("aniso 20k, %d eps levels x 2 minpts, batch" % levels)
    got, st = vd.cluster(x, y, eps, mp, path=vd.PATH_BATCH, return_stats=True)
    assert st["path_used"] == vd.PATH_BATCH
    want = oracle.libvdbscan(x, y, eps, mp)
    for v in range(len(eps)):
        assert (got[v] == want[v]).all(), (levels, v)


@pytest.mark.parametrize("levels", [15, 16])
def test_batch_levels_standard_rule(aniso20k, levels):
    from sklearn.cluster import DBSCAN
    x, y = aniso20k
    E = np.round(np.linspace(0.2, 0.8, levels), 6)
    eps = np.repeat(E, 2)
    mp = np.tile(np.array([4, 9], dtype=np.uint32), levels)
    say("aniso 20k, %d eps levels x 2 minpts, batch, standard rule" % levels)
    got = vd.cluster(x, y, eps, mp, path=vd.PATH_BATCH, border=vd.BORDER_STANDARD)
    X = np.column_stack([x, y])
    for v in range(len(eps)):
        want = DBSCAN(eps=eps[v], min_samples=int(mp[v]), algorithm="kd_tree").fit(X).labels_
        assert (got[v] == want).all(), (levels, v)
