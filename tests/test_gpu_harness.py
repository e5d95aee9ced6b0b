"""The reference's own test program (src/c_test_prog.cpp, README Appendix A),
compiled unmodified from the reference sources into oracle/_ref/c_test_prog
and linked against this library (oracle/Makefile), run on the GPU.

It dlopens libSharedVDBSCAN.so, reads test_dataset.csv from its working
directory, clusters (eps 1.0, minpts 10) and (eps 0.5, minpts 12) through
libVDBSCAN and prints "point, x, y, cluster" per variant. Its cluster ids
must be the golden output's (src/test_output.csv), and every variant's block
must list every point once.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

HARNESS = os.path.join(oracle.REF_DIR, "c_test_prog")


@pytest.mark.skipif(not os.path.exists(HARNESS), reason="oracle/_ref/c_test_prog not built")
def test_reference_harness_against_golden_output():
    r = subprocess.run([HARNESS], cwd=GOLDEN, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "Importing dataset: test_dataset.csv" in r.stdout
    assert "Verbose mode on" in r.stdout  # the harness runs with verbose=true
    g = np.loadtxt(os.path.join(GOLDEN, "test_output.csv"), delimiter=",")
    blocks = re.split(r"Variant: (\d+), Epsilon: ([0-9.]+), minpts: (\d+)", r.stdout)
    got = {}
    for i in range(1, len(blocks), 4):
        v = int(blocks[i])
        rows = re.findall(r"^(\d+), (-?[0-9.]+), (-?[0-9.]+), (\d+)$", blocks[i + 3], re.M)
        got[v] = (float(blocks[i + 1]), int(blocks[i + 2]),
                  np.array([[float(a), float(b), float(c), int(d)] for a, b, c, d in rows]))
    assert sorted(got) == [0, 1]
    assert got[0][:2] == (1.0, 10) and got[1][:2] == (0.5, 12)
    for v in (0, 1):
        rows = got[v][2]
        assert len(rows) == len(g)
        assert (rows[:, 0] == np.arange(len(g))).all()
        # %f prints 6 decimals of the coordinates
        assert np.allclose(rows[:, 1], g[:, 0], atol=5e-7)
        assert np.allclose(rows[:, 2], g[:, 1], atol=5e-7)
        assert (rows[:, 3].astype(np.uint32) == g[:, 2 + v].astype(np.uint32)).all(), v
