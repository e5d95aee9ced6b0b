"""The vdbscan command-line program on the reference's own dataset: labels
must equal src/test_output.csv (eps 1.0 / minpts 10 and eps 0.5 / minpts 12,
the variants of src/c_test_prog.cpp:129-135)."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import GOLDEN

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "variantdbscan_amd", "bin", "vdbscan")

pytestmark = pytest.mark.gpu


def run(args):
    print("vdbscan", " ".join(args), flush=True)
    sys.stdout.flush()
    r = subprocess.run([CLI] + args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


def golden():
    g = np.loadtxt(os.path.join(GOLDEN, "test_output.csv"), delimiter=",")
    return g[:, 2].astype(np.uint32), g[:, 3].astype(np.uint32)


def test_cli_labels_equal_golden(tmp_path):
    pts = os.path.join(GOLDEN, "test_dataset.csv")
    out = tmp_path / "labels.csv"
    run(["--points", pts, "--eps", "1.0,0.5", "--minpts", "10,12", "--out", str(out)])
    d = np.loadtxt(str(out), delimiter=",")
    l1, l2 = golden()
    src = np.loadtxt(pts, delimiter=",")
    assert np.array_equal(d[:, :2], src)
    assert (d[:, 2] == l1).all() and (d[:, 3] == l2).all()


def test_cli_variants_file_and_harness_text(tmp_path):
    pts = os.path.join(GOLDEN, "test_dataset.csv")
    var = tmp_path / "variants.csv"
    var.write_text("1.0,10\n0.5,12\n")
    text = run(["--points", pts, "--variants", str(var), "--harness", "--path", "table"])
    lines = [ln for ln in text.split("\n") if ln]
    assert lines[0] == "Variant: 0, Epsilon: 1.000000, minpts: 10"
    first = lines[1].split(", ")
    assert first[0] == "0" and first[1] == "6.689900" and first[2] == "-0.131300"
    l1, l2 = golden()
    got1 = [int(ln.split(", ")[3]) for ln in lines[1:226]]
    assert lines[226] == "Variant: 1, Epsilon: 0.500000, minpts: 12"
    got2 = [int(ln.split(", ")[3]) for ln in lines[227:452]]
    assert got1 == l1.tolist() and got2 == l2.tolist()


def test_cli_standard_border_labels(tmp_path):
    # --standard: -1 = noise, clusters 0.., as scikit-learn's DBSCAN labels
    import oracle
    pts = os.path.join(GOLDEN, "test_dataset.csv")
    out = tmp_path / "labels.csv"
    run(["--points", pts, "--eps", "1.0,0.5", "--minpts", "10,12", "--standard", "--out",
         str(out)])
    d = np.loadtxt(str(out), delimiter=",")
    src = np.loadtxt(pts, delimiter=",")
    for v, (eps, mp) in enumerate([(1.0, 10), (0.5, 12)]):
        want = oracle.standard_labels(src[:, 0], src[:, 1], eps, mp)
        assert (d[:, 2 + v] == want).all(), v
    assert (d[:, 2:] == -1).any()


def test_reference_example_runs_unchanged():
    # src/vtest.py's body, importing `variantdbscan` as the reference's
    # examples do (integer coordinates and minpts, as there)
    import oracle
    compat = os.path.join(ROOT, "variantdbscan_amd", "compat")
    code = ("import numpy as np\n"
            "from variantdbscan import VariantDBSCAN\n"
            "x = np.array([3, 2, 8, 1])\n"
            "y = np.array([3, 2, 8, 1])\n"
            "eps = np.array([np.sqrt(2)+0.01, np.sqrt(2)-0.01])\n"
            "mp = np.array([2,2])\n"
            "res = VariantDBSCAN(eps,mp,x,y)\n"
            "import json\n"
            "print(json.dumps(res.tolist()))\n")
    env = dict(os.environ, PYTHONPATH=compat + os.pathsep + ROOT)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                       env=env, timeout=300)
    assert r.returncode == 0, r.stderr
    got = np.array(json.loads(r.stdout.strip().split("\n")[-1]), dtype=np.uint32)
    x = np.array([3.0, 2.0, 8.0, 1.0])
    want = oracle.libvdbscan(x, x, [np.sqrt(2) + 0.01, np.sqrt(2) - 0.01], [2, 2])
    assert np.array_equal(got, want.T)
