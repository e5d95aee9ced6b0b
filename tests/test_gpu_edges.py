"""GPU edge cases of the precision and batching machinery.

* FP32 screening with binary64 recheck outside float32's normal range:
  coordinates scaled by 1e-24 (squared differences underflow float32) and
  by 1e30 (they overflow it). Counts and tables must equal the oracle's
  binary64 ones exactly.
* The batched engine's cell-pair test lists: when they overflow (forced here
  with VDB_BATCH_TEST_CAP), the group falls back to the per-variant engine
  and the labels are still the reference's.
"""
import sys

import numpy as np
import pytest

import oracle

import variantdbscan_amd as vd
from variantdbscan_amd import _lib, datasets

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("scale", [1e-24, 1.0, 1e30], ids=["tiny", "unit", "huge"])
def test_fp32_recheck_counts_at_extreme_scales(scale):
    x, y = datasets.aniso(6000, seed=41)
    x, y = x * scale, y * scale
    eps = 0.4 * scale
    say("fp32 recheck scale", scale)
    want = oracle.neighbor_counts(x, y, eps)
    c64 = vd.neighbor_counts(x, y, eps, precision=vd.PREC_FP64)
    c32, rechecked = vd.neighbor_counts(x, y, eps, precision=vd.PREC_FP32_RECHECK,
                                        return_rechecked=True)
    assert (c64 == want).all()
    assert (c32 == want).all()
    pairs = int(want.sum(dtype=np.int64))
    say("pairs", pairs, "rechecked", rechecked)
    if scale == 1.0:
        # inside float32's range only pairs near the threshold are rechecked
        assert 0 < rechecked < pairs // 10
    else:
        # outside it every tested candidate is decided in binary64 (pairs in
        # cell ranges accepted without a test are never tested)
        assert rechecked > pairs // 4


@pytest.mark.parametrize("scale", [1e-24, 1e30], ids=["tiny", "huge"])
def test_fp32_recheck_table_at_extreme_scales(scale):
    x, y = datasets.aniso(6000, seed=42)
    x, y = x * scale, y * scale
    eps = 0.3 * scale
    off, ids = oracle.neighbor_csr(x, y, eps)
    want = _lib.digest_of_csr(off, ids)
    d, p, _ = _lib.table_digest(x, y, eps, precision=vd.PREC_FP32_RECHECK)
    assert p == len(ids) and d == want


def test_fp32_recheck_labels_at_tiny_scale():
    x, y = datasets.aniso(6000, seed=43)
    x, y = x * 1e-24, y * 1e-24
    eps, mp = [0.3e-24, 0.5e-24], [4, 8]
    want = oracle.libvdbscan(x, y, eps, mp)
    for path in (vd.PATH_FUSED, vd.PATH_TABLE):
        got = vd.cluster(x, y, eps, mp, path=path, precision=vd.PREC_FP32_RECHECK)
        assert (got == want).all(), path


def test_batch_test_list_overflow_falls_back(monkeypatch):
    x, y = datasets.aniso(20_000, seed=44)
    eps = np.array([0.2, 0.3, 0.45, 0.6])
    mp = np.array([4, 8, 4, 16], dtype=np.uint32)
    want = oracle.libvdbscan(x, y, eps, mp)
    say("batch without a forced cap")
    got, st = vd.cluster(x, y, eps, mp, path=vd.PATH_BATCH, return_stats=True)
    assert st["path_used"] == vd.PATH_BATCH
    assert (got == want).all()
    # one test record of room: the first level with two tests overflows
    monkeypatch.setenv("VDB_BATCH_TEST_CAP", "1")
    say("batch with VDB_BATCH_TEST_CAP=1")
    got, st = vd.cluster(x, y, eps, mp, path=vd.PATH_BATCH, return_stats=True)
    assert st["path_used"] == vd.PATH_FUSED
    assert (got == want).all()
    got = vd.cluster(x, y, eps, mp, path=vd.PATH_BATCH, border=vd.BORDER_STANDARD)
    for v in range(len(eps)):
        assert (got[v] == oracle.standard_labels(x, y, eps[v], mp[v])).all(), v
    monkeypatch.delenv("VDB_BATCH_TEST_CAP")
    got, st = vd.cluster(x, y, eps, mp, path=vd.PATH_BATCH, return_stats=True)
    assert st["path_used"] == vd.PATH_BATCH and (got == want).all()
