"""Standard border rule (VDB_BORDER_STANDARD): textbook DBSCAN labels from
every engine, checked against scikit-learn's DBSCAN on random data and
against the oracle's restatement of its labelling (over the reference's exact
eps predicate) on the ULP-adversarial fixtures."""
import os
import sys

import numpy as np
import pytest

import oracle
from conftest import GOLDEN

import variantdbscan_amd as vd
from variantdbscan_amd import datasets

pytestmark = pytest.mark.gpu

STD = vd.BORDER_STANDARD
PATHS = [vd.PATH_FUSED, vd.PATH_TABLE, vd.PATH_BATCH]
PRECS = [vd.PREC_FP64, vd.PREC_FP32_RECHECK]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def sklearn_labels(x, y, eps, mp):
    from sklearn.cluster import DBSCAN
    return DBSCAN(eps=eps, min_samples=int(mp), algorithm="kd_tree").fit(
        np.column_stack([x, y])).labels_


@pytest.fixture(scope="module")
def aniso40k():
    x, y = datasets.aniso(40_000, seed=14)
    variants = [(0.2, 4), (0.2, 16), (0.45, 8), (0.7, 32), (1.3, 60)]
    eps = np.array([v[0] for v in variants])
    mp = np.array([v[1] for v in variants], dtype=np.uint32)
    want = np.stack([sklearn_labels(x, y, e, m) for e, m in variants])
    return x, y, eps, mp, want


def test_standard_against_sklearn(aniso40k):
    x, y, eps, mp, want = aniso40k
    for path in PATHS:
        for prec in PRECS:
            say("aniso 40k standard border, path", path, "prec", prec)
            got, st = vd.cluster(x, y, eps, mp, path=path, precision=prec, border=STD,
                                 return_stats=True)
            assert got.dtype == np.int32
            assert st["path_used"] == path
            for v in range(len(eps)):
                assert (got[v] == want[v]).all(), (path, prec, v)


def test_standard_auto_path_and_cells(aniso40k):
    x, y, eps, mp, want = aniso40k
    say("aniso 40k standard border, auto path")
    got, st = vd.cluster(x, y, eps, mp, border=STD, return_stats=True)
    assert st["path_used"] == vd.PATH_BATCH
    assert (got == want).all()
    for k in (1, 4):
        say("aniso 40k standard border, cells_per_eps", k)
        got = vd.cluster(x, y, eps[:3], mp[:3], path=vd.PATH_FUSED, border=STD, cells_per_eps=k)
        assert (got == want[:3]).all(), k


def test_standard_same_clusters_as_reference_rule(aniso40k):
    # core points fall into the same clusters under both rules; the standard
    # rule labels every point the reference rule labels, and more borders
    x, y, eps, mp, _ = aniso40k
    say("aniso 40k both rules")
    ref = vd.cluster(x, y, eps, mp)
    std = vd.cluster(x, y, eps, mp, border=STD)
    more = 0
    for v in range(len(eps)):
        core = oracle.neighbor_counts(x, y, eps[v]) >= mp[v]
        pairs = set(zip(ref[v][core].tolist(), std[v][core].tolist()))
        assert len(pairs) == len(set(ref[v][core].tolist())) == len(set(std[v][core].tolist()))
        assert (std[v][ref[v] > 0] >= 0).all()
        more += int((ref[v] == 0).sum() - (std[v] == -1).sum())
    assert more > 0


def test_standard_edge_fixtures():
    # ULP-adversarial circles, axis pairs, lattices, duplicates: the oracle
    # restates scikit-learn's labelling over the reference's exact predicate
    e = np.load(os.path.join(GOLDEN, "edge_cases.npz"))
    names = sorted(set(n.rsplit("_", 1)[0] for n in e.files))
    for nm in names:
        x, y = e[nm + "_x"], e[nm + "_y"]
        want = [oracle.standard_labels(x, y, ev, mv)
                for ev, mv in zip(e[nm + "_eps"], e[nm + "_minpts"])]
        for path in PATHS:
            for prec in PRECS:
                for k in (1, 2, 4):
                    say("edge set", nm, "standard border, path", path, "prec", prec, "cells", k)
                    lab = vd.cluster(x, y, e[nm + "_eps"], e[nm + "_minpts"], path=path,
                                     precision=prec, cells_per_eps=k, border=STD)
                    for v in range(len(want)):
                        assert (lab[v] == want[v]).all(), (nm, v, path, prec, k)


def test_standard_small_and_degenerate():
    # single point, all duplicates, minpts 0 and 1, minpts above INT_MAX
    cases = [
        (np.array([1.0]), np.array([2.0])),
        (np.full(300, 3.25), np.full(300, -7.5)),
        (np.array([0.0, 0.5, 1.0, 5.0, 5.2, 9.0]), np.zeros(6)),
    ]
    variants = [(0.6, 0), (0.6, 1), (0.6, 2), (0.6, 3), (0.6, 2 ** 31), (0.1, 300)]
    eps = np.array([v[0] for v in variants])
    mp = np.array([v[1] for v in variants], dtype=np.uint32)
    for i, (x, y) in enumerate(cases):
        say("degenerate case", i, "standard border")
        got = vd.cluster(x, y, eps, mp, border=STD)
        for v in range(len(eps)):
            want = oracle.standard_labels(x, y, eps[v], mp[v])
            assert (got[v] == want).all(), (i, v)


def test_standard_multi_datasets():
    rng = np.random.default_rng(5)
    sets = []
    for d in range(5):
        n = int(rng.integers(200, 6000))
        sets.append(datasets.plain_10k(seed=40 + d) if d == 0 else
                    (rng.normal(rng.uniform(-60, 60), 1.0, n), rng.normal(rng.uniform(-60, 60), 1.0, n)))
    eps = np.array([0.25, 0.5, 0.5])
    mp = np.array([4, 8, 30], dtype=np.uint32)
    for path in (vd.PATH_AUTO, vd.PATH_FUSED, vd.PATH_TABLE):
        say("multi datasets standard border, path", path)
        got = vd.cluster_multi(sets, eps, mp, path=path, border=STD)
        for d, (x, y) in enumerate(sets):
            assert got[d].dtype == np.int32
            for v in range(len(eps)):
                assert (got[d][v] == sklearn_labels(x, y, eps[v], mp[v])).all(), (path, d, v)


def test_standard_batched_64_variants():
    # the batched engine on 64 variants (eps 0.2..0.9 x minpts 4..48) of 30k
    # anisotropic points, with eps < 0 and NaN variants (no neighbours) mixed in
    x, y = datasets.aniso(30_000, seed=23)
    E, M = datasets.BATCH64_EPS, datasets.BATCH64_MINPTS
    eps = np.array([e for e in E for m in M] + [-1.0, np.nan])
    mp = np.array([m for e in E for m in M] + [0, 3], dtype=np.uint32)
    say("aniso 30k x 66 variants standard border, batch")
    got, st = vd.cluster(x, y, eps, mp, path=vd.PATH_BATCH, border=STD, return_stats=True)
    assert st["path_used"] == vd.PATH_BATCH
    for v in range(64):
        assert (got[v] == sklearn_labels(x, y, eps[v], mp[v])).all(), v
    assert (got[64] == np.arange(len(x))).all()   # every point its own cluster
    assert (got[65] == -1).all()
    say("aniso 30k x 66 variants standard border, fused")
    assert (vd.cluster(x, y, eps, mp, path=vd.PATH_FUSED, border=STD) == got).all()
