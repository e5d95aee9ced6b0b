"""Seeded synthetic point clouds for the benchmark workloads.

All generators return float64 arrays ``(x, y)`` with x = latitude in
[-90, 90] and y = longitude in [-180, 180], drawn from
``numpy.random.Generator(PCG64(seed))``. Blob samples that fall outside the
box are rejected and redrawn. The concatenated points are shuffled so the
caller order carries no blob structure.

Workloads (sizes as the benchmark names them):

* ``plain_10k``      10,000 points, 70% from 20 isotropic blobs (sigma 0.5 deg,
                     equal weights), 30% uniform; clustered at eps 1.0, minpts 4.
* ``aniso_2m``       2,000,000 points, 60% from 200 anisotropic blobs
                     (sigma_x, sigma_y lognormal, median 0.3 deg, log-sd 0.5,
                     rotation U[0, pi), Zipf s=1.1 blob weights), 40% uniform.
* ``skew_8m``        8,000,000 points, 50% uniform, 45% from 500 isotropic
                     blobs (sigma 0.2 deg), 5% uniform inside one 1 x 1 deg box.
* ``multi_256``      256 datasets, sizes lognormal (median 40k, log-sd 1.15)
                     clipped to [1k, 500k], each shaped like ``plain_10k`` with
                     its own seed and a blob count drawn from {5..100}.
"""
import numpy as np

LAT = (-90.0, 90.0)
LON = (-180.0, 180.0)

# variant grids of the batched workloads
BATCH64_EPS = [round(0.2 + 0.1 * i, 10) for i in range(8)]
BATCH64_MINPTS = [4, 6, 8, 12, 16, 24, 32, 48]
MULTI_EPS = [0.25, 0.5, 0.75, 1.0]
MULTI_MINPTS = [4, 8, 16, 32]


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def _uniform(rng, n, lat=LAT, lon=LON):
    x = rng.uniform(lat[0], lat[1], n)
    y = rng.uniform(lon[0], lon[1], n)
    return x, y


def _inside(x, y):
    return (x >= LAT[0]) & (x <= LAT[1]) & (y >= LON[0]) & (y <= LON[1])


def _blob(rng, n, cx, cy, sx, sy, theta):
    """n samples of a rotated anisotropic Gaussian, redrawn until in the box."""
    if n <= 0:
        return np.zeros(0), np.zeros(0)
    xs, ys = [], []
    need = n
    c, s = np.cos(theta), np.sin(theta)
    while need > 0:
        u = rng.normal(0.0, sx, need)
        v = rng.normal(0.0, sy, need)
        x = cx + c * u - s * v
        y = cy + s * u + c * v
        ok = _inside(x, y)
        xs.append(x[ok])
        ys.append(y[ok])
        need -= int(ok.sum())
    return np.concatenate(xs)[:n], np.concatenate(ys)[:n]


def _shuffle(rng, x, y):
    p = rng.permutation(len(x))
    return np.ascontiguousarray(x[p]), np.ascontiguousarray(y[p])


def blobs_isotropic(n, n_blobs, sigma, frac_blob, seed):
    """Equal-weight isotropic blobs plus uniform noise (the ``plain_10k`` shape)."""
    rng = _rng(seed)
    nb = int(round(n * frac_blob))
    counts = rng.multinomial(nb, np.full(n_blobs, 1.0 / n_blobs))
    cx, cy = _uniform(rng, n_blobs)
    xs, ys = [], []
    for k in range(n_blobs):
        bx, by = _blob(rng, int(counts[k]), cx[k], cy[k], sigma, sigma, 0.0)
        xs.append(bx)
        ys.append(by)
    ux, uy = _uniform(rng, n - nb)
    xs.append(ux)
    ys.append(uy)
    return _shuffle(rng, np.concatenate(xs), np.concatenate(ys))


def plain_10k(seed=1, n=10_000):
    return blobs_isotropic(n, 20, 0.5, 0.70, seed)


def aniso(n, seed=2, n_blobs=200, frac_blob=0.60, median_sigma=0.3, log_sd=0.5, zipf_s=1.1):
    rng = _rng(seed)
    nb = int(round(n * frac_blob))
    w = np.arange(1, n_blobs + 1, dtype=np.float64) ** (-zipf_s)
    w /= w.sum()
    counts = rng.multinomial(nb, w)
    cx, cy = _uniform(rng, n_blobs)
    sx = median_sigma * np.exp(rng.normal(0.0, log_sd, n_blobs))
    sy = median_sigma * np.exp(rng.normal(0.0, log_sd, n_blobs))
    th = rng.uniform(0.0, np.pi, n_blobs)
    xs, ys = [], []
    for k in range(n_blobs):
        bx, by = _blob(rng, int(counts[k]), cx[k], cy[k], sx[k], sy[k], th[k])
        xs.append(bx)
        ys.append(by)
    ux, uy = _uniform(rng, n - nb)
    xs.append(ux)
    ys.append(uy)
    return _shuffle(rng, np.concatenate(xs), np.concatenate(ys))


def aniso_2m(seed=2):
    return aniso(2_000_000, seed)


def skew(n, seed=4, n_blobs=500, sigma=0.2, frac_uniform=0.50, frac_blob=0.45):
    rng = _rng(seed)
    nu = int(round(n * frac_uniform))
    nb = int(round(n * frac_blob))
    nc = n - nu - nb
    counts = rng.multinomial(nb, np.full(n_blobs, 1.0 / n_blobs))
    cx, cy = _uniform(rng, n_blobs)
    xs, ys = [], []
    for k in range(n_blobs):
        bx, by = _blob(rng, int(counts[k]), cx[k], cy[k], sigma, sigma, 0.0)
        xs.append(bx)
        ys.append(by)
    ux, uy = _uniform(rng, nu)
    xs.append(ux)
    ys.append(uy)
    # one dense 1 x 1 deg clump
    c0x, c0y = rng.uniform(-80.0, 79.0), rng.uniform(-170.0, 169.0)
    kx, ky = _uniform(rng, nc, (c0x, c0x + 1.0), (c0y, c0y + 1.0))
    xs.append(kx)
    ys.append(ky)
    return _shuffle(rng, np.concatenate(xs), np.concatenate(ys))


def skew_8m(seed=4):
    return skew(8_000_000, seed)


def multi_sizes(n_sets=256, seed=5, median=40_000, log_sd=1.15, lo=1_000, hi=500_000):
    rng = _rng(seed)
    s = np.exp(np.log(median) + rng.normal(0.0, log_sd, n_sets))
    return np.clip(np.round(s), lo, hi).astype(np.int64)


def multi_256(seed=5, n_sets=256):
    """List of (x, y) for the uneven multi-dataset batch."""
    sizes = multi_sizes(n_sets, seed)
    rng = _rng(seed + 1)
    blob_counts = rng.integers(5, 101, n_sets)
    out = []
    for i in range(n_sets):
        out.append(blobs_isotropic(int(sizes[i]), int(blob_counts[i]), 0.5, 0.70,
                                   seed * 100_003 + i))
    return out
