"""Dataset and variant files (host code of the library; no GPU needed).

x,y CSV: the reference harness's import (src/c_test_prog.cpp:86-104);
lon,lat[,TEC] CSV: the GPU prototype's importDataset
(src/gpu/import_dataset.cpp:63-134); eps,minpts CSV: importDBScanInstances
(src/gpu/import_dataset.cpp:232-262). Values are pinned against numpy's
correctly rounded parsing, known answers of the latitude move and the
atof/atoi rules."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

from variantdbscan_amd import _lib


def test_xy_csv_matches_numpy_on_the_reference_dataset():
    path = os.path.join(GOLDEN, "test_dataset.csv")
    x, y = _lib.read_points_csv(path)
    d = np.loadtxt(path, delimiter=",")
    assert len(x) == 225
    assert np.array_equal(x, d[:, 0]) and np.array_equal(y, d[:, 1])


def test_xy_csv_line_rules(tmp_path):
    p = tmp_path / "pts.csv"
    # CRLF, trailing fields, blank and comma-less lines (skipped), atof
    # prefixes, no final newline
    p.write_bytes(b"1.5,2.5\r\n\n3,4,99\njunk\n 5e-1 ,-7x\n0x10,.25")
    x, y = _lib.read_points_csv(str(p))
    assert x.tolist() == [1.5, 3.0, 0.5, 16.0]
    assert y.tolist() == [2.5, 4.0, -7.0, 0.25]


def test_lonlat_known_answers(tmp_path):
    lats = [90.0, -90.0, 0.0, 45.0, -45.0, 30.0, -60.0]
    want = [0.0, 180.0, 90.0, 45.0, 135.0, 60.0, 150.0]
    p = tmp_path / "tec.csv"
    p.write_text("".join("%r,%r,50.0\n" % (10.0 * i, lat) for i, lat in enumerate(lats)))
    x, y = _lib.read_points_csv(str(p), lonlat=True)
    assert x.tolist() == want
    assert y.tolist() == [10.0 * i for i in range(len(lats))]


def test_lonlat_rounding_follows_the_prototype(tmp_path):
    # the prototype scales through a ratio: (90 - lat) / 90 * 90, not 90 - lat;
    # both roundings are visible for these latitudes
    rng = np.random.default_rng(5)
    lats = np.concatenate([rng.uniform(-90, 90, 2000), [1e-300, -1e-300, 89.99999999]])
    p = tmp_path / "tec.csv"
    p.write_text("".join("0,%r,1\n" % float(v) for v in lats))
    x, _ = _lib.read_points_csv(str(p), lonlat=True)
    pos = lats > 0
    ratio = np.where(pos, ((90.0 - lats) / 90.0) * 90.0, ((-lats) / 90.0) * 90.0 + 90.0)
    assert np.array_equal(x, ratio)
    # the ratio form differs from plain 90 - lat for some inputs (so a
    # simplified move would fail the comparison above)
    assert (x[pos] != 90.0 - lats[pos]).any()
    assert ((x >= 0) & (x <= 180)).all()


def test_variants_csv(tmp_path):
    p = tmp_path / "var.csv"
    p.write_text("1.0,10\n0.5, 12\n\n2.5,-1\n0.25,7.9\n")
    eps, mp = _lib.read_variants_csv(str(p))
    assert eps.tolist() == [1.0, 0.5, 2.5, 0.25]
    # atoi into the reference's unsigned field: -1 wraps, 7.9 truncates
    assert mp.tolist() == [10, 12, 2 ** 32 - 1, 7]


def test_labels_csv_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    x = rng.normal(size=50)
    y = rng.normal(size=50)
    lab = rng.integers(0, 9, size=(3, 50)).astype(np.uint32)
    p = tmp_path / "out.csv"
    _lib.write_labels_csv(str(p), x, y, lab)
    d = np.loadtxt(str(p), delimiter=",")
    assert np.array_equal(d[:, 0], x) and np.array_equal(d[:, 1], y)
    assert np.array_equal(d[:, 2:].T.astype(np.uint32), lab)


def test_missing_file_is_an_error(tmp_path):
    with pytest.raises(_lib.VdbError, match="cannot open"):
        _lib.read_points_csv(str(tmp_path / "nope.csv"))
    with pytest.raises(_lib.VdbError, match="cannot open"):
        _lib.read_variants_csv(str(tmp_path / "nope.csv"))


def test_cli_usage_errors_need_no_device(tmp_path):
    # argument errors end before any device call: exit 2 with the usage text,
    # which names every option (the standard border rule's included)
    import subprocess
    cli = os.path.join(os.path.dirname(_lib.LIB_PATH), os.pardir, "bin", "vdbscan")
    for args in (["--bogus"], ["--points"], ["--points", "x.csv"],
                 ["--points", "x.csv", "--eps", "0.5"], ["--path", "sideways"]):
        r = subprocess.run([cli] + args, capture_output=True, text=True, timeout=60)
        assert r.returncode == 2, args
        assert "usage: vdbscan" in r.stderr
    r = subprocess.run([cli, "--bogus"], capture_output=True, text=True, timeout=60)
    for opt in ("--points", "--lonlat", "--variants", "--eps", "--minpts", "--out",
                "--harness", "--path", "--fp32", "--standard", "--verbose"):
        assert opt in r.stderr, opt
