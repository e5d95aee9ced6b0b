"""Pins the CPU oracle (oracle/vdb_oracle.c) before it is trusted as checker.

Each check is chosen so a plausible mistake in the restatement (wrong visit
order, fused arithmetic, dropped self pair, border rule off by one) fails it.
"""
import math
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN


def _golden_csv():
    d = np.loadtxt(os.path.join(GOLDEN, "test_dataset.csv"), delimiter=",")
    g = np.loadtxt(os.path.join(GOLDEN, "test_output.csv"), delimiter=",")
    return d[:, 0], d[:, 1], g[:, 2].astype(np.uint32), g[:, 3].astype(np.uint32)


@pytest.mark.parametrize("closed", [False, True])
def test_golden_csv_bit_exact(closed):
    # src/test_output.csv: variants (1.0, 10) and (0.5, 12) of src/variant_example.py
    x, y, l1, l2 = _golden_csv()
    lab = oracle.libvdbscan(x, y, [1.0, 0.5], [10, 12], closed=closed)
    assert (lab[0] == l1).all()
    assert (lab[1] == l2).all()
    # documented label counts of the golden vector
    assert np.bincount(l1).tolist() == [22, 98, 105]
    assert np.bincount(l2).tolist() == [59, 87, 70, 9]


def test_visit_order_matters():
    # caller order instead of the binned order does not reproduce the golden
    # vector: the border rule depends on the visit order
    x, y, l1, l2 = _golden_csv()
    lab1 = oracle.dbscan_binned(x, y, 1.0, 10)
    lab2 = oracle.dbscan_binned(x, y, 0.5, 12)
    assert not ((lab1 == l1).all() and (lab2 == l2).all())


def test_threshold_closed_form():
    rng = np.random.default_rng(3)
    for eps in list(rng.uniform(1e-6, 10, 300)) + [0.1, 0.2, 0.5, 1.0, 1e-300, 1e300, 3.0]:
        t = oracle.threshold(eps)
        assert math.sqrt(t) <= eps
        nt = np.nextafter(t, np.inf)
        assert not math.isfinite(nt) or math.sqrt(nt) > eps
    assert oracle.threshold(0.0) == 0.0
    assert oracle.threshold(-1.0) == -1.0
    assert oracle.threshold(float("nan")) == -1.0
    assert oracle.threshold(float("inf")) == float("inf")


def test_dist2_is_unfused():
    # a case where fma(dx,dx,dy*dy) differs from the separately rounded sum
    rng = np.random.default_rng(11)
    found = 0
    for _ in range(20000):
        px, py, qx, qy = rng.uniform(-180, 180, 4)
        dx, dy = px - qx, py - qy
        ref = dx * dx + dy * dy          # numpy float64: separately rounded
        assert oracle.dist2(px, py, qx, qy) == ref
        fused = math.fma(dx, dx, dy * dy) if hasattr(math, "fma") else ref
        found += fused != ref
    # python < 3.13 has no math.fma; the equality above is the pin


def test_neighbor_csr_matches_brute_force():
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.normal(0, 0.3, 300), rng.uniform(-3, 3, 200)])
    y = np.concatenate([rng.normal(0, 0.3, 300), rng.uniform(-3, 3, 200)])
    for eps in [0.05, 0.2, 0.7]:
        off, ids = oracle.neighbor_csr(x, y, eps)
        # brute force with the reference formula, no threshold trick
        d = np.sqrt((x[:, None] - x[None, :]) ** 2 + (y[:, None] - y[None, :]) ** 2)
        adj = d <= eps
        for i in range(len(x)):
            want = np.nonzero(adj[i])[0]
            got = ids[off[i]:off[i + 1]]
            assert (got == want).all()
        assert (oracle.neighbor_counts(x, y, eps) == adj.sum(1)).all()


def test_bin_order_hand_cases():
    # three bins along x; stable inside a bin
    x = np.array([5.0, 0.0, 10.0, 0.1, 5.1])
    y = np.array([0.0, 0.0, 0.0, 0.0, 0.0])
    m = oracle.bin_order(x, y)
    # y all zero: min_y=0, max_y=DBL_MIN (src/main.cpp:468), width tiny but > 0
    assert m.tolist() == [1, 3, 0, 4, 2]
    # row-major over (xbin, ybin): same xbin, y decides
    x = np.array([1.0, 1.0, 1.0, 2.0])
    y = np.array([3.0, 1.0, 2.0, 0.0])
    assert oracle.bin_order(x, y).tolist() == [1, 2, 0, 3]


def test_bin_order_all_negative_quirk():
    # max starts at DBL_MIN, so an all-negative axis has width (DBL_MIN-min)*1.01
    x = np.array([-10.0, -5.0, -1.0])
    y = np.array([-3.0, -2.0, -1.0])
    m = oracle.bin_order(x, y)
    w = (np.finfo(float).tiny - (-10.0)) * 1.01
    bins = np.floor(((x + 10.0) / w) * 300)
    assert np.all(np.diff(bins[m]) >= 0)
    assert m.tolist() == [0, 1, 2]


def test_closed_form_equals_literal_random():
    rng = np.random.default_rng(9)
    for trial in range(6):
        n = int(rng.integers(50, 1500))
        x = np.concatenate([rng.normal(0, 0.5, n // 2), rng.uniform(-4, 4, n - n // 2)])
        y = np.concatenate([rng.normal(0, 0.5, n // 2), rng.uniform(-4, 4, n - n // 2)])
        for eps, mp in [(0.1, 3), (0.3, 6), (0.6, 20), (0.05, 1), (0.2, 0)]:
            a = oracle.libvdbscan(x, y, [eps], [mp], closed=False)
            b = oracle.libvdbscan(x, y, [eps], [mp], closed=True)
            assert (a == b).all()


def test_minpts_wraps_like_int():
    x, y, _, _ = _golden_csv()
    lab = oracle.libvdbscan(x, y, [1.0], [2 ** 31])   # int(minpts) < 0: nothing is core
    assert (lab == 0).all()
    lab = oracle.libvdbscan(x, y, [1.0], [0])         # every point is core
    assert (lab > 0).all()


def test_fixture_plain10k_against_oracle(golden_dir):
    g = np.load(os.path.join(golden_dir, "plain_10k.npz"))
    for v in range(len(g["plain10k_eps"])):
        o = oracle.libvdbscan(g["plain10k_x"], g["plain10k_y"], [g["plain10k_eps"][v]],
                              [g["plain10k_minpts"][v]])[0]
        assert (o == g["plain10k_labels"][v]).all()


def test_fixture_multi_variant_core_partition(golden_dir):
    # the reference's multi-variant call may reuse clusters (timing dependent):
    # only the partition of core points is invariant
    g = np.load(os.path.join(golden_dir, "plain_10k.npz"))
    for v in range(len(g["plain10k_eps"])):
        a = g["plain10k_labels"][v]
        b = g["plain10k_multi_labels"][v]
        counts = oracle.neighbor_counts(g["plain10k_x"], g["plain10k_y"], g["plain10k_eps"][v])
        core = counts >= g["plain10k_minpts"][v]
        pairs = set(zip(a[core].tolist(), b[core].tolist()))
        assert len(pairs) == len(set(a[core].tolist())) == len(set(b[core].tolist()))


def test_fixture_edge_cases_against_oracle(golden_dir):
    e = np.load(os.path.join(golden_dir, "edge_cases.npz"))
    names = sorted(set(k.rsplit("_", 1)[0] for k in e.files))
    assert len(names) >= 6
    for nm in names:
        for v in range(len(e[nm + "_eps"])):
            for closed in (False, True):
                o = oracle.libvdbscan(e[nm + "_x"], e[nm + "_y"], [e[nm + "_eps"][v]],
                                      [e[nm + "_minpts"][v]], closed=closed)[0]
                assert (o == e[nm + "_labels"][v]).all(), (nm, v, closed)


def test_levels_closed_equals_libvdbscan(golden_dir):
    # the table-free multi-minpts form against the per-variant restatements,
    # on the golden fixtures (their labels come from the compiled reference)
    # and on random blobs at several eps, minpts including 0, 1 and > INT_MAX
    g = np.load(os.path.join(golden_dir, "plain_10k.npz"))
    x, y = g["plain10k_x"], g["plain10k_y"]
    for eps in sorted(set(g["plain10k_eps"].tolist())):
        sel = np.flatnonzero(g["plain10k_eps"] == eps)
        mp = g["plain10k_minpts"][sel]
        got = oracle.levels_closed(x, y, eps, mp)
        for r, v in enumerate(sel):
            assert (got[r] == g["plain10k_labels"][v]).all(), (eps, v)
    e = np.load(os.path.join(golden_dir, "edge_cases.npz"))
    for nm in sorted(set(k.rsplit("_", 1)[0] for k in e.files)):
        for v in range(len(e[nm + "_eps"])):
            got = oracle.levels_closed(e[nm + "_x"], e[nm + "_y"], e[nm + "_eps"][v],
                                       [e[nm + "_minpts"][v]])[0]
            assert (got == e[nm + "_labels"][v]).all(), (nm, v)
    rng = np.random.default_rng(33)
    for trial in range(3):
        n = int(rng.integers(1000, 5000))
        x = np.concatenate([rng.normal(rng.uniform(-50, 50), 0.5, n // 2),
                            rng.uniform(-60, 60, n - n // 2)])
        y = np.concatenate([rng.normal(rng.uniform(-50, 50), 0.5, n // 2),
                            rng.uniform(-60, 60, n - n // 2)])
        mps = [0, 1, 3, 8, 17, 0x80000000]
        for eps in (0.05, 0.3, 1.1):
            got = oracle.levels_closed(x, y, eps, mps)
            for r, m in enumerate(mps):
                lit = oracle.libvdbscan(x, y, [eps], [m], closed=False)[0]
                assert (got[r] == lit).all(), (trial, eps, m)


@pytest.mark.skipif(not oracle.ref_available(1), reason="oracle/_ref not built")
def test_oracle_against_compiled_reference_random():
    rng = np.random.default_rng(21)
    for trial in range(4):
        n = int(rng.integers(500, 4000))
        x = np.concatenate([rng.normal(rng.uniform(-50, 50), 0.4, n // 2),
                            rng.uniform(-60, 60, n - n // 2)])
        y = np.concatenate([rng.normal(rng.uniform(-50, 50), 0.4, n // 2),
                            rng.uniform(-60, 60, n - n // 2)])
        for eps, mp in [(0.2, 4), (0.5, 8), (1.5, 12)]:
            ref, _ = oracle.run_reference(x, y, [eps], [mp])
            assert (oracle.libvdbscan(x, y, [eps], [mp]) == ref).all()


def test_reuse_oracle_against_reference_fixture(golden_dir):
    # tests/golden/reuse_t1.npz: multi-variant labels of the compiled reference
    # built with NSEARCHTHREADS=1, where its greedy reuse is deterministic
    g = np.load(os.path.join(golden_dir, "reuse_t1.npz"))
    names = sorted(set(k.rsplit("_", 1)[0] for k in g.files))
    assert len(names) >= 4
    differs = 0
    for nm in names:
        x, y, eps, mp = g[nm + "_x"], g[nm + "_y"], g[nm + "_eps"], g[nm + "_minpts"]
        got = oracle.libvdbscan_reuse(x, y, eps, mp)
        assert (got == g[nm + "_labels"]).all(), nm
        differs += int((got != oracle.libvdbscan(x, y, eps, mp)).any(axis=1).sum())
    # the fixtures exercise the reuse path: many variants differ from scratch
    assert differs >= 10


@pytest.mark.skipif(not oracle.ref_available(1), reason="oracle/_ref not built")
def test_reuse_oracle_against_compiled_reference_random():
    rng = np.random.default_rng(5)
    for trial in range(3):
        n = int(rng.integers(2000, 6000))
        c = rng.uniform(-20, 20, (8, 2))[rng.integers(0, 8, n // 2)]
        x = np.concatenate([rng.normal(c[:, 0], 0.3), rng.uniform(-25, 25, n - n // 2)])
        y = np.concatenate([rng.normal(c[:, 1], 0.3), rng.uniform(-25, 25, n - n // 2)])
        eps = rng.choice([0.2, 0.35, 0.5, 0.8], 6)
        mp = rng.choice([2, 4, 7, 15, 30], 6).astype(np.uint32)
        ref, _ = oracle.run_reference(x, y, eps, mp, threads=1)
        assert (oracle.libvdbscan_reuse(x, y, eps, mp) == ref).all(), trial


# ------------------------------------------------- standard border rule oracle

def _sklearn_labels(x, y, eps, mp):
    from sklearn.cluster import DBSCAN
    X = np.column_stack([x, y])
    return DBSCAN(eps=eps, min_samples=mp, algorithm="kd_tree").fit(X).labels_


def test_standard_labels_hand_case():
    # two chains of 4 core points (minpts 4, eps 1) with a border point 0.9
    # from the end of each; caller order: chain Y, the border, chain X, a
    # far point. The border joins the lower-numbered cluster (Y = 0); the
    # far point is noise.
    x = np.array([2.7, 3.0, 3.3, 3.6, 1.8, 0.0, 0.3, 0.6, 0.9, 50.0])
    y = np.zeros_like(x)
    lab = oracle.standard_labels(x, y, 1.0, 4)
    assert lab.tolist() == [0, 0, 0, 0, 0, 1, 1, 1, 1, -1]
    # the same chains with X first in caller order: the border joins X
    perm = [5, 6, 7, 8, 4, 0, 1, 2, 3, 9]
    lab = oracle.standard_labels(x[perm], y[perm], 1.0, 4)
    assert lab.tolist() == [0, 0, 0, 0, 0, 1, 1, 1, 1, -1]
    assert (oracle.standard_labels(x, y, 1.0, 2 ** 31) == -1).all()
    assert (oracle.standard_labels(x, y, 1.0, 0) >= 0).all()


def test_standard_labels_against_sklearn():
    # scikit-learn's DBSCAN is the definition of the standard rule; on random
    # data no pair sits within an ulp of eps, so its eps test (squared
    # distance against eps^2) and the reference's agree
    rng = np.random.default_rng(31)
    for trial in range(4):
        n = int(rng.integers(300, 3000))
        x = np.concatenate([rng.normal(0, 0.4, n // 2), rng.uniform(-3, 3, n - n // 2)])
        y = np.concatenate([rng.normal(0.5, 0.3, n // 2), rng.uniform(-3, 3, n - n // 2)])
        for eps, mp in [(0.1, 3), (0.2, 8), (0.35, 20), (0.05, 1), (0.6, 60)]:
            want = _sklearn_labels(x, y, eps, mp)
            got = oracle.standard_labels(x, y, eps, mp)
            assert (got == want).all(), (trial, eps, mp)


def test_standard_and_reference_rules_share_clusters():
    # same core points and clusters; a point the reference labels is labelled
    # by the standard rule too, and only border points can differ in kind
    rng = np.random.default_rng(32)
    n = 2500
    x = np.concatenate([rng.normal(0, 0.3, n // 2), rng.uniform(-3, 3, n - n // 2)])
    y = np.concatenate([rng.normal(0, 0.3, n // 2), rng.uniform(-3, 3, n - n // 2)])
    for eps, mp in [(0.1, 4), (0.2, 12), (0.3, 30)]:
        ref = oracle.libvdbscan(x, y, [eps], [mp])[0]
        std = oracle.standard_labels(x, y, eps, mp)
        core = oracle.neighbor_counts(x, y, eps) >= mp
        pairs = set(zip(ref[core].tolist(), std[core].tolist()))
        assert len(pairs) == len(set(ref[core].tolist())) == len(set(std[core].tolist()))
        assert (std[ref > 0] >= 0).all()
        # the rules differ here: borders the reference leaves as noise
        assert (std == -1).sum() < (ref == 0).sum()
