"""Per-variant kernels at every grid resolution (cells per eps 1..4): the
windows, the trimmed rows and the ranges accepted without a distance test
must give the labels of the oracle, bit for bit, in FP64 and in FP32 with the
binary64 recheck."""
import os
import sys

import numpy as np
import pytest

import oracle
from conftest import GOLDEN

import variantdbscan_amd as vd
from variantdbscan_amd import datasets

pytestmark = pytest.mark.gpu

CELLS = [1, 2, 3, 4, 6, 8]
PATHS = [vd.PATH_FUSED, vd.PATH_TABLE]
PRECS = [vd.PREC_FP64, vd.PREC_FP32_RECHECK]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def aniso40k():
    x, y = datasets.aniso(40_000, seed=14)
    variants = [(0.2, 4), (0.2, 16), (0.45, 8), (0.7, 32), (0.7, 4), (1.3, 60)]
    eps = np.array([v[0] for v in variants])
    mp = np.array([v[1] for v in variants], dtype=np.uint32)
    return x, y, eps, mp, oracle.libvdbscan(x, y, eps, mp)


@pytest.mark.parametrize("k", CELLS)
def test_cells_against_oracle(aniso40k, k):
    x, y, eps, mp, want = aniso40k
    for path in PATHS:
        for prec in PRECS:
            say("aniso 40k cells_per_eps", k, "path", path, "prec", prec)
            got = vd.cluster(x, y, eps, mp, path=path, precision=prec, cells_per_eps=k)
            for v in range(len(eps)):
                assert (got[v] == want[v]).all(), (k, path, prec, v)


@pytest.mark.parametrize("k", CELLS)
def test_cells_edge_fixtures(k):
    # ULP-adversarial circles, axis pairs, lattices, duplicates
    e = np.load(os.path.join(GOLDEN, "edge_cases.npz"))
    names = sorted(set(n.rsplit("_", 1)[0] for n in e.files))
    for nm in names:
        for path in PATHS:
            for prec in PRECS:
                say("edge set", nm, "cells_per_eps", k, "path", path, "prec", prec)
                lab = vd.cluster(e[nm + "_x"], e[nm + "_y"], e[nm + "_eps"],
                                 e[nm + "_minpts"], path=path, precision=prec, cells_per_eps=k)
                for v in range(len(e[nm + "_eps"])):
                    assert (lab[v] == e[nm + "_labels"][v]).all(), (nm, v, k, path, prec)


def test_cells_dense_clump():
    # one dense clump: whole cells fall inside eps of each other, so most
    # pairs take the accept-without-test route at 3 and 4 cells per eps
    rng = np.random.default_rng(8)
    x = np.concatenate([rng.uniform(10.0, 10.4, 30_000), rng.uniform(-90, 90, 5_000)])
    y = np.concatenate([rng.uniform(20.0, 20.4, 30_000), rng.uniform(-180, 180, 5_000)])
    eps = np.array([0.05, 0.3, 0.3])
    mp = np.array([40, 5, 3000], dtype=np.uint32)
    want = oracle.libvdbscan(x, y, eps, mp)
    for k in CELLS:
        for path in PATHS:
            say("dense clump cells_per_eps", k, "path", path)
            got = vd.cluster(x, y, eps, mp, path=path, cells_per_eps=k)
            for v in range(len(eps)):
                assert (got[v] == want[v]).all(), (k, path, v)
