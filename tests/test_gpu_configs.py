"""GPU checks at the benchmark workloads' full sizes, where the CPU oracle is
too slow: two independent engines must agree bit for bit, and the FP32
screening with binary64 recheck must give the binary64 table exactly."""
import sys

import numpy as np
import pytest

import oracle

import variantdbscan_amd as vd
from variantdbscan_amd import datasets
from variantdbscan_amd import _lib

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_digest_matches_oracle_table():
    # pins the device digest against the oracle's table hashed on the host
    x, y = datasets.aniso(20_000, seed=31)
    for eps in [0.1, 0.35]:
        off, ids = oracle.neighbor_csr(x, y, eps)
        want = _lib.digest_of_csr(off, ids)
        for prec in [vd.PREC_FP64, vd.PREC_FP32_RECHECK]:
            d, p, r = _lib.table_digest(x, y, eps, precision=prec)
            assert p == len(ids)
            assert d == want, (eps, prec)


def _core_consistency(x, y, eps, mp, labels):
    counts = vd.neighbor_counts(x, y, eps)
    core = counts >= mp
    assert (labels[core] > 0).all()
    # each cluster holds a core point
    cl = np.unique(labels[labels > 0])
    assert np.isin(cl, np.unique(labels[core])).all()
    return core


def test_skew8m_fp32_recheck_table_identical():
    # 8M points, 5% in one 1x1 deg clump; eps 0.3
    x, y = datasets.skew_8m()
    say("skew_8m generated", len(x))
    d64, p64, r64 = _lib.table_digest(x, y, 0.3, precision=vd.PREC_FP64)
    d32, p32, r32 = _lib.table_digest(x, y, 0.3, precision=vd.PREC_FP32_RECHECK)
    say("pairs", p64, "rechecked", r32, "fraction %.3e" % (r32 / max(p64, 1)))
    assert p64 == p32 and d64 == d32
    assert r64 == 0 and r32 > 0
    c64 = vd.neighbor_counts(x, y, 0.3, precision=vd.PREC_FP64)
    c32 = vd.neighbor_counts(x, y, 0.3, precision=vd.PREC_FP32_RECHECK)
    assert (c64 == c32).all()
    assert int(c64.sum(dtype=np.uint64)) == p64


def test_skew8m_engines_agree():
    x, y = datasets.skew_8m()
    a = vd.cluster(x, y, [0.3], [10], path=vd.PATH_FUSED)[0]
    b = vd.cluster(x, y, [0.3], [10], path=vd.PATH_BATCH)[0]
    say("skew_8m clusters", int(a.max()), "noise", int((a == 0).sum()))
    assert (a == b).all()
    _core_consistency(x, y, 0.3, 10, a)


def test_aniso2m_single_variant_engines_agree():
    x, y = datasets.aniso_2m()
    a = vd.cluster(x, y, [0.5], [8], path=vd.PATH_FUSED)[0]
    b = vd.cluster(x, y, [0.5], [8], path=vd.PATH_BATCH)[0]
    c = vd.cluster(x, y, [0.5], [8], path=vd.PATH_TABLE, precision=vd.PREC_FP32_RECHECK)[0]
    assert (a == b).all() and (a == c).all()
    _core_consistency(x, y, 0.5, 8, a)


def test_aniso2m_64_variants_batch_equals_per_variant():
    x, y = datasets.aniso_2m()
    E, M = datasets.BATCH64_EPS, datasets.BATCH64_MINPTS
    eps = np.array([e for e in E for m in M])
    mp = np.array([m for e in E for m in M], dtype=np.uint32)
    a = vd.cluster(x, y, eps, mp, path=vd.PATH_BATCH)
    b = vd.cluster(x, y, eps, mp, path=vd.PATH_FUSED)
    bad = [v for v in range(len(eps)) if not (a[v] == b[v]).all()]
    assert not bad, bad
