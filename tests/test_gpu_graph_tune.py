"""Graph replays and the persisted autotuner of the batched engine.

The batched engine runs its launch sequence after the last host read
directly on a workload's first call, captures it into a HIP graph on the
second and replays the graph afterwards. Every replay must give the labels of
a run without graphs (VDB_NO_GRAPH=1), with other eager library work and
synchronisations between the replays. The autotuner's choices go to a JSON
cache that a later process reads instead of tuning again.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle

import variantdbscan_amd as vd
from variantdbscan_amd import _lib, datasets

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _grid():
    E, M = [0.2, 0.35, 0.5, 0.8], [4, 9, 20, 48]
    eps = np.array([e for e in E for m in M])
    mp = np.array([m for e in E for m in M], dtype=np.uint32)
    return eps, mp


@pytest.mark.parametrize("border", [vd.BORDER_REFERENCE, vd.BORDER_STANDARD], ids=["ref", "std"])
def test_graph_replays_match_direct_runs(monkeypatch, border):
    x, y = datasets.aniso(60_000, seed=51)
    x2, y2 = datasets.aniso(20_000, seed=52)
    eps, mp = _grid()
    monkeypatch.setenv("VDB_NO_GRAPH", "1")
    want = vd.cluster(x, y, eps, mp, path=vd.PATH_BATCH, border=border)
    monkeypatch.delenv("VDB_NO_GRAPH")
    other = vd.cluster(x2, y2, [0.3], [5], path=vd.PATH_FUSED, border=border)
    for i in range(6):
        say("graph call", i)
        got, st = vd.cluster(x, y, eps, mp, path=vd.PATH_BATCH, border=border,
                             return_stats=True)
        assert st["path_used"] == vd.PATH_BATCH
        assert (got == want).all(), i
        # eager work and a synchronisation between replays
        assert (vd.cluster(x2, y2, [0.3], [5], path=vd.PATH_FUSED, border=border) == other).all()
    if border == vd.BORDER_REFERENCE:
        assert (want[5] == oracle.libvdbscan(x, y, [eps[5]], [mp[5]])[0]).all()


def test_graph_follows_new_points_of_the_same_shape():
    # same sizes, different coordinates: a replay must read the new points
    eps, mp = _grid()
    x, y = datasets.aniso(30_000, seed=61)
    for k in range(3):
        xs, ys = x + 0.01 * k, y - 0.02 * k
        got = vd.cluster(xs, ys, eps, mp, path=vd.PATH_BATCH)
        want = oracle.libvdbscan(xs, ys, eps[:4], mp[:4])
        assert (got[:4] == want).all(), k


_CHILD = """
import json, sys
sys.path.insert(0, %r)
import variantdbscan_amd as vd
from variantdbscan_amd import _lib, datasets
import hashlib
import numpy as np
x, y = datasets.aniso(200_000, seed=53)
E, M = [0.2, 0.3, 0.45], [4, 12, 32]
eps = np.array([e for e in E for m in M]); mp = np.array([m for e in E for m in M], dtype=np.uint32)
lab = vd.cluster(x, y, eps, mp, path=vd.PATH_BATCH)
t = _lib.tuning()
t["labels"] = hashlib.sha256(lab.tobytes()).hexdigest()
print(json.dumps(t))
""" % ROOT


def _child(cache):
    env = dict(os.environ, VDB_TUNE_CACHE=str(cache))
    r = subprocess.run([sys.executable, "-c", _CHILD], env=env, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_tuning_cache_persists_across_processes(tmp_path):
    cache = tmp_path / "tune.json"
    say("first process: tunes")
    a = _child(cache)
    assert a["cache"] == str(cache) and cache.exists()
    keys = [k for k in a["entries"] if ":batch:" in k]
    assert keys and a["tuned_in_process"] >= 1
    say("second process: reads the cache")
    b = _child(cache)
    assert b["tuned_in_process"] == 0
    assert b["entries"][keys[0]] == a["entries"][keys[0]]
    assert a["labels"] == b["labels"]
