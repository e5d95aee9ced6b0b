"""GPU labels at the benchmark workloads' real sizes against committed digests.

tests/golden/full_size_digests.json (written by make_full_digests.py) holds,
per variant, the SHA-256 of the uint32 labels in caller order, the cluster
and noise counts and the labels at 4096 fixed positions:

* aniso_2m_500k (first 500,000 points of the 2M-point set, 64 variants) and
  skew_8m_1m (first 1,000,000 points of the 8M-point set): labels of the
  compiled reference, one single-variant call per variant;
* aniso_2m (2M points, 64 variants: the batched workload, which includes the
  eps 0.5 / minpts 8 single-variant workload) and skew_8m (8M points, eps
  0.3 / minpts 10): the oracle's table-free form, which matched the
  reference on both subsamples.

Every engine that takes the workload is compared, variant by variant.
"""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN

import variantdbscan_amd as vd
from variantdbscan_amd import datasets

pytestmark = pytest.mark.gpu

DIGESTS = os.path.join(GOLDEN, "full_size_digests.json")
SAMPLES = os.path.join(GOLDEN, "full_size_samples.npz")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _load():
    if not os.path.exists(DIGESTS):
        pytest.skip("full-size digests not generated")
    d = json.load(open(DIGESTS))
    return d, np.load(SAMPLES)


def _points(name):
    if name.startswith("aniso_2m"):
        x, y = datasets.aniso_2m()
    else:
        x, y = datasets.skew_8m()
    n = {"aniso_2m_500k": 500_000, "skew_8m_1m": 1_000_000}.get(name)
    return (x[:n], y[:n]) if n else (x, y)


def _sample_index(n, d):
    rng = np.random.Generator(np.random.PCG64(d["sample_seed"]))
    return np.sort(rng.choice(n, min(d["sample_size"], n), replace=False))


def _check(name, keys, labels, d, smp):
    w = d["workloads"][name]
    idx = _sample_index(w["n"], d)
    bad = []
    for r, k in enumerate(keys):
        lab = np.ascontiguousarray(labels[r], dtype="<u4")
        want = w["variants"][k]
        got = hashlib.sha256(lab.tobytes()).hexdigest()
        if got != want["sha256"]:
            s = smp[(name + "/" + k).replace("/", "_")]
            bad.append((k, int(lab.max()), want["clusters"],
                        int(np.count_nonzero(lab == 0)), want["noise"],
                        int(np.count_nonzero(lab[idx] != s))))
    assert not bad, bad


def _grid(w):
    keys = sorted(w["variants"])
    eps = np.array([w["variants"][k]["eps"] for k in keys])
    mp = np.array([w["variants"][k]["minpts"] for k in keys], dtype=np.uint32)
    return keys, eps, mp


@pytest.mark.parametrize("name", ["aniso_2m_500k", "aniso_2m"])
def test_batched_64_variants_match_digests(name):
    d, smp = _load()
    w = d["workloads"][name]
    x, y = _points(name)
    keys, eps, mp = _grid(w)
    say(name, "batched engine,", len(keys), "variants, source:", w["source"])
    lab, st = vd.cluster(x, y, eps, mp, path=vd.PATH_AUTO, return_stats=True)
    assert st["path_used"] == vd.PATH_BATCH
    _check(name, keys, lab, d, smp)


def test_aniso2m_per_variant_engines_match_digests():
    # the single-variant workload (eps 0.5, minpts 8) and the extreme eps
    # levels through the per-variant engines
    d, smp = _load()
    w = d["workloads"]["aniso_2m"]
    x, y = _points("aniso_2m")
    keys = ["eps0.50/mp8", "eps0.20/mp4", "eps0.90/mp48"]
    eps = np.array([w["variants"][k]["eps"] for k in keys])
    mp = np.array([w["variants"][k]["minpts"] for k in keys], dtype=np.uint32)
    for path, prec in ((vd.PATH_FUSED, vd.PREC_FP64), (vd.PATH_TABLE, vd.PREC_FP32_RECHECK)):
        say("aniso_2m per-variant path", path, "precision", prec)
        lab = vd.cluster(x, y, eps, mp, path=path, precision=prec)
        _check("aniso_2m", keys, lab, d, smp)


@pytest.mark.parametrize("name", ["skew_8m_1m", "skew_8m"])
def test_skew_match_digests(name):
    d, smp = _load()
    x, y = _points(name)
    keys = ["eps0.30/mp10"]
    for path, prec in ((vd.PATH_AUTO, vd.PREC_FP64), (vd.PATH_FUSED, vd.PREC_FP32_RECHECK),
                       (vd.PATH_TABLE, vd.PREC_FP32_RECHECK)):
        say(name, "path", path, "precision", prec)
        lab = vd.cluster(x, y, [0.3], [10], path=path, precision=prec)
        _check(name, keys, lab, d, smp)
