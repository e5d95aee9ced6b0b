"""The C++ entry points (include/vdbscan.hpp), driven like src/main.cpp drives
the reference's DBScan/Schedule classes, give the labels of the C ABI."""
import os
import struct
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

import variantdbscan_amd as vd

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "tests", "cpp", "api_check")


def run_api(x, y, eps, mp, scratch=False):
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<II", len(x), len(eps)))
            f.write(np.asarray(x, np.float64).tobytes())
            f.write(np.asarray(y, np.float64).tobytes())
            f.write(np.asarray(eps, np.float64).tobytes())
            f.write(np.asarray(mp, np.uint32).tobytes())
        r = subprocess.run([EXE, fin, fout] + (["scratch"] if scratch else []),
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        return np.fromfile(fout, dtype=np.uint32).reshape(len(eps), len(x))


def test_cpp_api_golden():
    assert os.path.exists(EXE), "build tests/cpp first (__graft_entry__.build())"
    d = np.loadtxt(os.path.join(GOLDEN, "test_dataset.csv"), delimiter=",")
    g = np.loadtxt(os.path.join(GOLDEN, "test_output.csv"), delimiter=",")
    lab = run_api(d[:, 0], d[:, 1], [1.0, 0.5], [10, 12])
    assert (lab[0] == g[:, 2]).all() and (lab[1] == g[:, 3]).all()


def test_cpp_api_scratch_matches_c_abi():
    # every variant through algDBScanParallel: one batched call for all of
    # them, the one-variant labels of the reference
    z = np.load(os.path.join(GOLDEN, "plain_10k.npz"))
    x, y = z["plain10k_x"], z["plain10k_y"]
    eps, mp = z["plain10k_eps"], z["plain10k_minpts"]
    lab = run_api(x, y, eps, mp, scratch=True)
    assert (lab == z["plain10k_labels"]).all()
    assert (lab == vd.cluster(x, y, eps, mp)).all()


def test_cpp_api_reuse_loop_matches_reference():
    # the reference's own loop (src/main.cpp:212-251) over one thread: reused
    # variants through assignPointsToPredefinedCluster give the reference's
    # NSEARCHTHREADS=1 multi-variant labels (tests/golden/reuse_t1.npz)
    g = np.load(os.path.join(GOLDEN, "reuse_t1.npz"))
    for nm in sorted(set(k.rsplit("_", 1)[0] for k in g.files)):
        lab = run_api(g[nm + "_x"], g[nm + "_y"], g[nm + "_eps"], g[nm + "_minpts"])
        bad = [v for v in range(len(lab)) if not (lab[v] == g[nm + "_labels"][v]).all()]
        assert not bad, (nm, bad)
