"""GPU parity: the HIP library through its C ABI against the oracle and the
reference's golden outputs. Integer outputs must match bit for bit."""
import os
import sys

import numpy as np
import pytest

import oracle
from conftest import GOLDEN

import variantdbscan_amd as vd
from variantdbscan_amd import datasets

pytestmark = pytest.mark.gpu

PATHS = [vd.PATH_FUSED, vd.PATH_TABLE, vd.PATH_BATCH]
PRECS = [vd.PREC_FP64, vd.PREC_FP32_RECHECK]
PATH_IDS = ["fused", "table", "batch"]
PREC_IDS = ["fp64", "fp32"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _golden_csv():
    d = np.loadtxt(os.path.join(GOLDEN, "test_dataset.csv"), delimiter=",")
    g = np.loadtxt(os.path.join(GOLDEN, "test_output.csv"), delimiter=",")
    return d[:, 0], d[:, 1], g[:, 2].astype(np.uint32), g[:, 3].astype(np.uint32)


def test_golden_csv_drop_in():
    x, y, l1, l2 = _golden_csv()
    say("golden csv via VariantDBSCAN")
    res = vd.VariantDBSCAN(np.array([1, 0.5]), np.array([10, 12]), x, y, 70, verbose=False)
    assert res.shape == (len(x), 2) and res.dtype == np.uint32
    assert (res[:, 0] == l1).all()
    assert (res[:, 1] == l2).all()
    one = vd.VariantDBSCAN(np.array([0.5]), np.array([12]), x, y)
    assert one.shape == (len(x),) and (one == l2).all()


@pytest.mark.parametrize("path", PATHS, ids=PATH_IDS)
@pytest.mark.parametrize("prec", PRECS, ids=PREC_IDS)
def test_golden_csv_paths(path, prec):
    x, y, l1, l2 = _golden_csv()
    say("golden csv path", path, "prec", prec)
    lab = vd.cluster(x, y, [1.0, 0.5], [10, 12], path=path, precision=prec)
    assert (lab[0] == l1).all()
    assert (lab[1] == l2).all()


@pytest.mark.parametrize("path", PATHS, ids=PATH_IDS)
def test_plain10k_fixture(path):
    g = np.load(os.path.join(GOLDEN, "plain_10k.npz"))
    x, y = g["plain10k_x"], g["plain10k_y"]
    say("plain10k multi-variant call, path", path)
    lab = vd.cluster(x, y, g["plain10k_eps"], g["plain10k_minpts"], path=path)
    for v in range(len(g["plain10k_eps"])):
        assert (lab[v] == g["plain10k_labels"][v]).all(), v


@pytest.mark.parametrize("path", PATHS, ids=PATH_IDS)
def test_edge_case_fixtures(path):
    e = np.load(os.path.join(GOLDEN, "edge_cases.npz"))
    names = sorted(set(k.rsplit("_", 1)[0] for k in e.files))
    for nm in names:
        say("edge set", nm, "path", path)
        for prec in PRECS:
            lab = vd.cluster(e[nm + "_x"], e[nm + "_y"], e[nm + "_eps"], e[nm + "_minpts"],
                             path=path, precision=prec)
            for v in range(len(e[nm + "_eps"])):
                assert (lab[v] == e[nm + "_labels"][v]).all(), (nm, v, prec)


def test_rank_order_matches_oracle():
    rng = np.random.default_rng(1)
    x, y, _, _ = _golden_csv()
    say("rank order")
    assert (vd.rank_order(x, y) == oracle.bin_order(x, y)).all()
    # one point: the reference divides 0 by 0 while binning (undefined); this
    # library bins a zero-width axis to bin 0
    assert vd.rank_order([2.0], [3.0]).tolist() == [0]
    for n in [2, 17, 5000, 200_000]:
        xx = rng.uniform(-90, 90, n)
        yy = rng.uniform(-180, 180, n)
        if n > 100:
            xx[: n // 3] = np.round(xx[: n // 3], 1)   # many exact ties inside bins
        assert (vd.rank_order(xx, yy) == oracle.bin_order(xx, yy)).all(), n
    xx = -50 - rng.random(1000)                          # all negative (DBL_MIN max)
    yy = -10 - rng.random(1000)
    assert (vd.rank_order(xx, yy) == oracle.bin_order(xx, yy)).all()


@pytest.mark.parametrize("prec", PRECS, ids=PREC_IDS)
def test_neighbor_counts_match_oracle(prec):
    x, y = datasets.aniso(60_000, seed=12)
    for eps in [0.05, 0.3, 0.9]:
        say("counts eps", eps, "prec", prec)
        want = oracle.neighbor_counts(x, y, eps)
        got, rechecked = vd.neighbor_counts(x, y, eps, precision=prec, return_rechecked=True)
        assert (got == want).all()
        if prec == vd.PREC_FP64:
            assert rechecked == 0


def test_neighbor_table_matches_oracle():
    x, y = datasets.aniso(20_000, seed=13)
    for eps in [0.1, 0.4]:
        for prec in PRECS:
            say("table eps", eps, "prec", prec)
            off, ids, d2 = vd.neighbor_table(x, y, eps, with_d2=True, precision=prec)
            ooff, oids = oracle.neighbor_csr(x, y, eps)
            assert (off.astype(np.int64) == ooff).all()
            for i in range(0, len(x), 97):
                row = np.sort(ids[off[i]:off[i + 1]])
                assert (row == oids[ooff[i]:ooff[i + 1]]).all()
            # full multiset check through a row-sorted key
            rows = np.repeat(np.arange(len(x)), np.diff(off.astype(np.int64)))
            key = rows.astype(np.int64) * len(x) + ids
            orows = np.repeat(np.arange(len(x)), np.diff(ooff))
            okey = orows.astype(np.int64) * len(x) + oids
            assert (np.sort(key) == np.sort(okey)).all()
            # float32 squared distances: within float rounding of the binary64 value
            j = ids.astype(np.int64)
            ref = (x[rows] - x[j]) ** 2 + (y[rows] - y[j]) ** 2
            assert np.allclose(d2, ref, rtol=2e-7, atol=1e-30)


@pytest.mark.parametrize("path", PATHS, ids=PATH_IDS)
def test_random_against_oracle(path):
    x, y = datasets.aniso(40_000, seed=14)
    variants = [(0.2, 4), (0.2, 16), (0.45, 8), (0.7, 32), (0.7, 4)]
    eps = [v[0] for v in variants]
    mp = [v[1] for v in variants]
    say("random 40k path", path)
    want = oracle.libvdbscan(x, y, eps, mp)
    got = vd.cluster(x, y, eps, mp, path=path)
    for v in range(len(variants)):
        assert (got[v] == want[v]).all(), variants[v]


@pytest.mark.parametrize("path", PATHS, ids=PATH_IDS)
def test_degenerate_inputs(path):
    say("degenerate inputs path", path)
    one = vd.cluster([1.5], [2.5], [0.5, 0.5], [1, 2], path=path)
    assert one[:, 0].tolist() == [1, 0]
    # all coordinates equal and positive (reference: 0/0 in binning)
    x = np.full(50, 3.0)
    y = np.full(50, 7.0)
    lab = vd.cluster(x, y, [0.0, 1.0], [50, 51], path=path)
    assert (lab[0] == 1).all() and (lab[1] == 0).all()
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, 2000)
    y = rng.uniform(-1, 1, 2000)
    for eps, mp in [(0.0, 1), (-1.0, 0), (float("inf"), 3), (0.1, 0), (0.1, 2 ** 31),
                    (1e300, 5), (1e-9, 1)]:
        want = oracle.libvdbscan(x, y, [eps], [mp])[0]
        got = vd.cluster(x, y, [eps], [mp], path=path)[0]
        assert (got == want).all(), (eps, mp)


def test_nonfinite_rejected():
    x = np.array([0.0, np.nan, 1.0])
    y = np.zeros(3)
    with pytest.raises(vd._lib.VdbError):
        vd.cluster(x, y, [1.0], [2])


def test_repeatable():
    x, y = datasets.aniso(100_000, seed=15)
    eps = [0.3, 0.6]
    mp = [5, 20]
    a = vd.cluster(x, y, eps, mp)
    b = vd.cluster(x, y, eps, mp)
    assert (a == b).all()
