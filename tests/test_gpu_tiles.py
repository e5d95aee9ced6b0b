"""Tiled CSR tables: when the table does not fit the device-memory budget the
per-variant table engine builds and consumes it tile by tile; labels must not
change. Covers a pooled table buffer kept from an earlier, smaller table."""
import sys

import numpy as np
import pytest

import oracle

import variantdbscan_amd as vd
from variantdbscan_amd import datasets

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_tiles_over_a_smaller_held_buffer():
    x, y = datasets.aniso(30_000, seed=21)
    vd._lib.load().vdb_release()
    # the first call leaves a pooled table buffer sized for a small table
    _, st1 = vd.cluster(x, y, [0.1], [5], path=vd.PATH_TABLE, return_stats=True)
    p1 = int(st1["pairs"])
    # a larger table under a budget of 1.5x that buffer: tiles must fit the
    # buffer actually held, not the budget
    eps = 0.3
    lab, st2 = vd.cluster(x, y, [eps], [5], path=vd.PATH_TABLE,
                          table_budget_bytes=int(4 * p1 * 1.5), return_stats=True)
    say("held buffer", p1, "ids; table", st2["pairs"], "ids in", st2["tiles"], "tiles")
    assert st2["pairs"] > 2 * p1 and st2["tiles"] > 1
    want = oracle.libvdbscan(x, y, [eps], [5])[0]
    assert (lab[0] == want).all()


def test_tiles_small_budget_fp64_and_fp32():
    x, y = datasets.aniso(30_000, seed=22)
    eps, mp = [0.25, 0.4], [4, 12]
    want = oracle.libvdbscan(x, y, eps, mp)
    for prec in (vd.PREC_FP64, vd.PREC_FP32_RECHECK):
        lab, st = vd.cluster(x, y, eps, mp, path=vd.PATH_TABLE, precision=prec,
                             table_budget_bytes=4 << 20, return_stats=True)
        say("prec", prec, "tiles", st["tiles"])
        assert st["tiles"] > 2
        assert (lab == want).all()
