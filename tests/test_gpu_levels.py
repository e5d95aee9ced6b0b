"""Batched engine at the boundaries of its pair sort. The cell pairs of all
minpts values of a pass are ordered by (minpts value, eps level): J*(K+1)
buckets. Up to 128 buckets the eligibility pass counts each tile's pairs per
bucket and one scatter writes the records (levels matched with 4-bit keys
up to 16 levels, 7-bit keys beyond; more than 8 minpts values take two
eligibility planes); beyond 128 buckets the radix sort runs. Every case
must give the oracle's labels (reference rule) and scikit-learn's
(standard rule)."""
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


# (eps levels K, minpts values): buckets J*(K+1) = 16, 32, 128, 128 (16
# minpts values: two planes), 132
CASES = [(15, (4,)), (15, (4, 9)), (31, (3, 4, 9, 20)),
         (7, (3, 4, 5, 6, 7, 8, 9, 10, 12, 14, 16, 18, 20, 24, 28, 32)), (32, (3, 4, 9, 20))]
IDS = ["16-buckets", "32-buckets", "128-buckets", "128-buckets-16-minpts", "132-radix"]


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


# Shapes of the one-kernel label pass: minpts values in groups of four (here
# 3, 5 and 6 values: partial groups), 8 or 16 level tiles (K <= 8, 9..15),
# requests that leave (level, minpts) pairs out, and repeated variants (one
# staged label, several output rows).
LAYOUTS = {
    "k6-j6-full": ([0.2, 0.3, 0.4, 0.5, 0.6, 0.7], [3, 4, 6, 9, 12, 20], None),
    "k9-j3-full": ([0.2, 0.25, 0.3, 0.35, 0.4, 0.5, 0.6, 0.7, 0.8], [4, 9, 15], None),
    "k8-j5-sparse": ([0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8, 0.9], [3, 5, 8, 13, 21], 3),
    "k4-j2-repeats": ([0.25, 0.4, 0.55, 0.7], [4, 10], "repeat"),
}


def _layout(name):
    E, M, mode = LAYOUTS[name]
    eps = np.array([e for e in E for m in M])
    mp = np.array([m for e in E for m in M], dtype=np.uint32)
    if mode == "repeat":
        # every variant twice, the copies in reverse order
        eps = np.concatenate([eps, eps[::-1]])
        mp = np.concatenate([mp, mp[::-1]])
    elif mode:
        keep = np.arange(len(eps)) % mode != 1  # drop every third pair
        eps, mp = eps[keep], mp[keep]
    return eps, mp


@pytest.mark.parametrize("name", sorted(LAYOUTS))
@pytest.mark.parametrize("border", [vd.BORDER_REFERENCE, vd.BORDER_STANDARD], ids=["ref", "std"])
def test_batch_label_layouts(aniso20k, name, border):
    x, y = aniso20k
    eps, mp = _layout(name)
    say("aniso 20k, label layout %s, %d variants, border %d" % (name, len(eps), border))
    got, st = vd.cluster(x, y, eps, mp, path=vd.PATH_BATCH, border=border, return_stats=True)
    assert st["path_used"] == vd.PATH_BATCH
    if border == vd.BORDER_REFERENCE:
        want = oracle.libvdbscan(x, y, eps, mp)
        for v in range(len(eps)):
            assert (got[v] == want[v]).all(), (name, v, float(eps[v]), int(mp[v]))
        return
    # standard rule: the oracle is pure Python, so every third variant (and
    # each distinct one once)
    seen = {}
    for v in range(0, len(eps), 3):
        key = (float(eps[v]), int(mp[v]))
        if key not in seen:
            seen[key] = oracle.standard_labels(x, y, key[0], key[1])
        assert (got[v] == seen[key]).all(), (name, v) + key
