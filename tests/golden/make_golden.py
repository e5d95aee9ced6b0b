"""Write the golden label fixtures of tests/golden/ from the compiled reference.

Runs in the build container only (it needs oracle/_ref, built by
``make -C oracle ref`` from the reference sources). Every label array below is
the output of the reference's own ``libVDBSCAN`` (one variant per process, the
reference's exact scratch semantics), so the fixtures pin results without
the reference being present at test time.

Usage: python tests/golden/make_golden.py [--reuse-only]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle  # noqa: E402
from variantdbscan_amd import datasets  # noqa: E402


def ref_single(x, y, eps, mp):
    lab, _ = oracle.run_reference(x, y, [eps], [mp], threads=1)
    return lab[0]


def ref_multi(x, y, eps, mp):
    lab, _ = oracle.run_reference(x, y, eps, mp, threads=8)
    return lab


def edge_sets():
    """Small adversarial datasets: name -> (x, y, [(eps, minpts), ...])."""
    sets = {}
    rng = np.random.default_rng(7)
    # lattice with non-representable spacing: many pairs sit at d == eps
    i, j = np.meshgrid(np.arange(20), np.arange(20), indexing="ij")
    x = (i * 0.1).ravel()
    y = (j * 0.1).ravel()
    sets["lattice_0p1"] = (x, y, [(0.1, 3), (0.1, 5), (0.1 * np.sqrt(2), 5), (0.2, 9)])
    # integer lattice, eps exactly representable
    i, j = np.meshgrid(np.arange(15), np.arange(15), indexing="ij")
    sets["lattice_int"] = (i.ravel().astype(float), j.ravel().astype(float),
                           [(1.0, 5), (1.0, 4), (np.sqrt(2.0), 9), (2.0, 13)])
    # exact duplicates and eps = 0
    base = rng.uniform(-5, 5, (30, 2))
    dup = base[rng.integers(0, 30, 90)]
    xy = np.concatenate([base, dup])
    sets["duplicates"] = (xy[:, 0], xy[:, 1], [(0.0, 2), (0.0, 4), (0.5, 3)])
    # all coordinates negative: max starts at DBL_MIN in binDataset
    xy = rng.normal(0, 1, (400, 2)) - 50.0
    sets["all_negative"] = (xy[:, 0], xy[:, 1], [(0.3, 4), (0.6, 8)])
    # points on circles of radius eps around centres, nudged by -2..+2 ulps
    cx = rng.uniform(-60, 60, 24)
    cy = rng.uniform(-150, 150, 24)
    eps = 0.37
    pts = []
    for k in range(24):
        pts.append((cx[k], cy[k]))
        for t in rng.uniform(0, 2 * np.pi, 12):
            px, py = cx[k] + eps * np.cos(t), cy[k] + eps * np.sin(t)
            for u in range(int(rng.integers(-2, 3))):
                px = np.nextafter(px, np.inf)
            pts.append((px, py))
    xy = np.array(pts)
    sets["ulp_circles"] = (xy[:, 0], xy[:, 1], [(eps, 2), (eps, 5), (eps, 13)])
    # axis-aligned pairs at distance ~eps (dy = 0): the case where the
    # reference's box prefilter and the distance test meet
    xs, ys = [], []
    for k in range(60):
        x0 = rng.uniform(-80, 80)
        y0 = rng.uniform(-170, 170)
        xs += [x0, x0 + 0.25, np.nextafter(x0 + 0.25, np.inf), np.nextafter(x0 + 0.25, -np.inf)]
        ys += [y0, y0, y0, y0]
    sets["axis_pairs"] = (np.array(xs), np.array(ys), [(0.25, 2), (0.25, 3)])
    return sets


def reuse_sets():
    """Multi-variant datasets for the reuse emulation: name -> (x, y, eps, minpts).
    Variant orders are shuffled so the reference's sort decides the schedule."""
    sets = {}
    x, y = datasets.plain_10k(seed=1)
    sets["plain10k"] = (x, y, np.array([1.0, 0.5, 0.5, 1.0, 1.5, 2.0]),
                        np.array([4, 4, 12, 10, 8, 30], dtype=np.uint32))
    rng = np.random.default_rng(11)
    for k, n in enumerate([8_000, 20_000]):
        x, y = datasets.aniso(n, seed=100 + k, n_blobs=12 + 10 * k)
        E = [0.25, 0.4, 0.7, 1.1][: 3 + k]
        M = [3, 6, 12, 24][: 3 + k]
        eps = np.array([e for e in E for m in M])
        mp = np.array([m for e in E for m in M], dtype=np.uint32)
        p = rng.permutation(len(eps))
        sets["aniso%dk" % (n // 1000)] = (x, y, eps[p], mp[p])
    # equal variants, minpts 0 and 1, and minpts above INT_MAX (negative int)
    x, y = datasets.blobs_isotropic(6_000, 8, 0.4, 0.8, seed=77)
    sets["special"] = (x, y, np.array([0.5, 0.5, 0.8, 0.3, 0.8, 0.5]),
                       np.array([6, 6, 0, 1, 0x80000000, 3], dtype=np.uint32))
    return sets


def write_reuse_fixture():
    """tests/golden/reuse_t1.npz: multi-variant labels of the reference built
    with NSEARCHTHREADS=1 (its greedy reuse is then deterministic)."""
    out = {}
    for name, (x, y, eps, mp) in reuse_sets().items():
        lab, _ = oracle.run_reference(x, y, eps, mp, threads=1)
        out[name + "_x"] = np.asarray(x, dtype=np.float64)
        out[name + "_y"] = np.asarray(y, dtype=np.float64)
        out[name + "_eps"] = eps
        out[name + "_minpts"] = mp
        out[name + "_labels"] = lab
    np.savez_compressed(os.path.join(HERE, "reuse_t1.npz"), **out)
    print("wrote reuse_t1.npz")


def main():
    if not oracle.ref_available(1) or not oracle.ref_available(8):
        raise SystemExit("build oracle/_ref first: make -C oracle ref")
    if "--reuse-only" in sys.argv:
        write_reuse_fixture()
        return
    write_reuse_fixture()
    out = {}
    # the shipped 225-point example is kept as CSV (test_dataset.csv / test_output.csv)
    x, y = datasets.plain_10k(seed=1)
    variants = [(1.0, 4), (0.5, 4), (0.5, 12), (1.0, 10), (1.5, 8), (2.0, 30)]
    out["plain10k_x"] = x
    out["plain10k_y"] = y
    out["plain10k_eps"] = np.array([v[0] for v in variants])
    out["plain10k_minpts"] = np.array([v[1] for v in variants], dtype=np.uint32)
    out["plain10k_labels"] = np.stack([ref_single(x, y, e, m) for e, m in variants])
    out["plain10k_multi_labels"] = ref_multi(x, y, out["plain10k_eps"], out["plain10k_minpts"])
    np.savez_compressed(os.path.join(HERE, "plain_10k.npz"), **out)

    edges = {}
    for name, (x, y, vs) in edge_sets().items():
        edges[name + "_x"] = np.asarray(x, dtype=np.float64)
        edges[name + "_y"] = np.asarray(y, dtype=np.float64)
        edges[name + "_eps"] = np.array([v[0] for v in vs], dtype=np.float64)
        edges[name + "_minpts"] = np.array([v[1] for v in vs], dtype=np.uint32)
        edges[name + "_labels"] = np.stack([ref_single(x, y, e, m) for e, m in vs])
    np.savez_compressed(os.path.join(HERE, "edge_cases.npz"), **edges)
    print("wrote plain_10k.npz, edge_cases.npz")


if __name__ == "__main__":
    main()
