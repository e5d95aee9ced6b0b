"""Label digests pinning the benchmark workloads at their real sizes.

Runs in the build container only (it needs oracle/_ref, which ``make -C
oracle ref`` builds from the reference sources, and oracle/liboracle.so).

Two sources, and where they overlap they must agree bit for bit:

* the compiled reference: one single-variant ``libVDBSCAN`` call per
  variant, each in its own process (oracle/_ref/ref_runner), so every label
  is the reference's deterministic scratch label (src/main.cpp:50-299 with
  numVariants = 1). Its expansion keeps every core point's whole neighbour
  list on one stack (src/DBScan.cpp:555-596 via copyVect), so its memory
  grows with the pair count: about 1.2 GB at 200,000 points and eps 0.9,
  several hundred GB for the 2M-point workload. It runs on stated
  subsamples: the first 500,000 points of the 2M-point set (64 variants)
  and the first 1,000,000 points of the 8M-point set;
* the oracle's table-free form (oracle.levels_closed, O(n) memory per
  minpts value), which this script first checks against the reference on
  both subsamples, then runs at full size: the 2M-point set under all 64
  variants and the 8M-point set at eps 0.3 / minpts 10.

The labels are too large to commit (2M x 64 x 4 B = 512 MB), so per variant
the fixture holds the SHA-256 of the uint32 labels in caller order (the
retArr row of src/main.cpp:273-284, little endian), the cluster count
(largest label), the noise count (label 0), the source, and the labels at
4096 fixed positions (full_size_samples.npz), which say where a mismatch
sits.

Usage: python tests/golden/make_full_digests.py [--ref-workers 5] [--cache DIR]
A finished variant is kept in the cache directory (default /tmp/vdb_full_ref),
so an interrupted run resumes.
"""
import argparse
import hashlib
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle  # noqa: E402
from variantdbscan_amd import datasets  # noqa: E402

SAMPLE_SEED = 20261203
SAMPLE_SIZE = 4096


def sample_index(n):
    rng = np.random.Generator(np.random.PCG64(SAMPLE_SEED))
    return np.sort(rng.choice(n, min(SAMPLE_SIZE, n), replace=False)).astype(np.int64)


def digest(labels):
    lab = np.ascontiguousarray(labels, dtype="<u4")
    return {"sha256": hashlib.sha256(lab.tobytes()).hexdigest(),
            "clusters": int(lab.max()) if len(lab) else 0,
            "noise": int(np.count_nonzero(lab == 0))}


def workloads():
    """name -> (x, y, [(eps, [minpts...]), ...], description)"""
    x, y = datasets.aniso_2m()
    sx, sy = datasets.skew_8m()
    grid = [(e, list(datasets.BATCH64_MINPTS)) for e in datasets.BATCH64_EPS]
    return {
        "aniso_2m": (x, y, grid, "2,000,000 points, datasets.aniso_2m() (seed 2), 64 variants"),
        "aniso_2m_500k": (x[:500_000], y[:500_000], grid,
                          "first 500,000 points of datasets.aniso_2m(), 64 variants"),
        "skew_8m": (sx, sy, [(0.3, [10])], "8,000,000 points, datasets.skew_8m() (seed 4)"),
        "skew_8m_1m": (sx[:1_000_000], sy[:1_000_000], [(0.3, [10])],
                       "first 1,000,000 points of datasets.skew_8m()"),
    }


REFERENCE_RUNS = ("aniso_2m_500k", "skew_8m_1m")


def vkey(eps, mp):
    return "eps%.2f/mp%d" % (eps, mp)


def cached(cache, name, tag, eps, mps, fn):
    """labels (len(mps), n) from the cache or fn()"""
    path = os.path.join(cache, "%s_%s_eps%.4f_%s.npy" % (name, tag, eps,
                                                        "-".join(str(m) for m in mps)))
    if os.path.exists(path):
        return np.load(path), None
    t0 = time.time()
    lab = fn()
    np.save(path + ".tmp.npy", lab)
    os.replace(path + ".tmp.npy", path)
    return lab, time.time() - t0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref-workers", type=int, default=5)
    ap.add_argument("--oracle-workers", type=int, default=8)
    ap.add_argument("--cache", default="/tmp/vdb_full_ref")
    a = ap.parse_args()
    if not oracle.ref_available(1):
        raise SystemExit("build oracle/_ref first: make -C oracle ref")
    os.makedirs(a.cache, exist_ok=True)
    W = workloads()
    t0 = time.time()

    # reference: one variant per process, largest eps first
    ref_tasks = [(nm, e, m) for nm in REFERENCE_RUNS for e, ms in W[nm][2] for m in ms]
    ref_tasks.sort(key=lambda t: -t[1])

    def run_ref(t):
        nm, e, m = t
        x, y = W[nm][0], W[nm][1]
        lab, sec = cached(a.cache, nm, "ref", e, [m],
                          lambda: oracle.run_reference(x, y, [e], [m], threads=1)[0])
        print("reference %-14s %-14s %s" % (nm, vkey(e, m),
                                            "cached" if sec is None else "%.1f s" % sec),
              flush=True)
        return (nm, vkey(e, m)), lab[0]

    # oracle: one eps (all its minpts values) per task
    orc_tasks = [(nm, e, ms) for nm in W for e, ms in W[nm][2]]
    orc_tasks.sort(key=lambda t: -(len(W[t[0]][0]) * t[1] * t[1]))

    def run_orc(t):
        nm, e, ms = t
        x, y = W[nm][0], W[nm][1]
        lab, sec = cached(a.cache, nm, "oracle", e, ms,
                          lambda: oracle.levels_closed(x, y, e, ms))
        print("oracle    %-14s eps %.2f     %s" % (nm, e,
                                                  "cached" if sec is None else "%.1f s" % sec),
              flush=True)
        return [((nm, vkey(e, m)), lab[r]) for r, m in enumerate(ms)]

    with ThreadPoolExecutor(a.ref_workers) as ex:
        ref = dict(ex.map(run_ref, ref_tasks))
    with ThreadPoolExecutor(a.oracle_workers) as ex:
        orc = {}
        for part in ex.map(run_orc, orc_tasks):
            orc.update(part)

    # where both ran they must agree
    agree = {}
    for k, lab in ref.items():
        same = bool(np.array_equal(lab, orc[k]))
        agree.setdefault(k[0], []).append(same)
        if not same:
            print("MISMATCH reference vs oracle", k, flush=True)
    out = {"sample_seed": SAMPLE_SEED, "sample_size": SAMPLE_SIZE, "workloads": {}}
    samples = {}
    for nm, (x, y, grid, desc) in W.items():
        n = len(x)
        idx = sample_index(n)
        var = {}
        for e, ms in grid:
            for m in ms:
                k = (nm, vkey(e, m))
                lab = ref[k] if k in ref else orc[k]
                rec = digest(lab)
                rec.update({"eps": e, "minpts": m})
                var[vkey(e, m)] = rec
                samples[nm + "/" + vkey(e, m)] = lab[idx]
        src = ("compiled reference (oracle/_ref, one variant per call); oracle "
               "levels_closed identical on %d/%d variants" % (sum(agree[nm]), len(agree[nm]))
               if nm in agree else
               "oracle levels_closed (pinned to the reference on the subsamples)")
        out["workloads"][nm] = {"n": n, "description": desc, "source": src,
                                "reference_agrees": all(agree[nm]) if nm in agree else None,
                                "variants": var}
    with open(os.path.join(HERE, "full_size_digests.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    np.savez_compressed(os.path.join(HERE, "full_size_samples.npz"),
                        **{k.replace("/", "_"): v for k, v in samples.items()})
    ok = all(all(v) for v in agree.values())
    print("wrote full_size_digests.json, full_size_samples.npz (%.0f s); reference and "
          "oracle agree: %s" % (time.time() - t0, ok))
    if not ok:
        raise SystemExit(1)


if __name__ == "__main__":
    main()
