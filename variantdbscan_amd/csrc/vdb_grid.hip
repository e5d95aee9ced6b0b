// Reference visit order (binning) and grid construction.
#include "vdb_grid.h"
#include "vdb_prims.h"

#include <cfloat>
#include <cmath>
#include <cstring>
#include <algorithm>
#include <vector>

namespace vdb {

// the largest t >= 0 with sqrt(t) <= eps, by bisection over the bits
// (doubles >= 0 order like their bits)
static double threshold_bisect(double eps)
{
    uint64_t lo = 0, hi = 0x7fefffffffffffffull;
    double dhi;
    memcpy(&dhi, &hi, 8);
    if (std::sqrt(dhi) <= eps) return dhi;
    while (hi - lo > 1) { // sqrt(lo) <= eps < sqrt(hi)
        uint64_t mid = lo + (hi - lo) / 2;
        double dm;
        memcpy(&dm, &mid, 8);
        if (std::sqrt(dm) <= eps) lo = mid; else hi = mid;
    }
    double t;
    memcpy(&t, &lo, 8);
    return t;
}

double threshold_for_eps(double eps)
{
    if (!(eps >= 0.0)) return -1.0;
    if (std::isinf(eps)) return INFINITY;
    // sqrt is correctly rounded, hence monotone: the answer lies a few ulps
    // from eps*eps, so step there (the bisection takes a case that would
    // need more steps); the host calls this for every variant several times
    // per call
    double t = eps * eps;
    if (t > DBL_MAX) t = DBL_MAX;
    int steps = 0;
    while (std::sqrt(t) > eps && steps < 64) {
        t = std::nextafter(t, 0.0);
        steps++;
    }
    while (steps < 64) {
        double u = std::nextafter(t, INFINITY);
        if (std::isinf(u) || std::sqrt(u) > eps) break;
        t = u;
        steps++;
    }
    return steps < 64 ? t : threshold_bisect(eps);
}

namespace {

constexpr int kBlock = 256;

struct BinParams {
    double xoff, xw, yoff, yw;
    int xdeg, ydeg; // dimension has zero width: every point in bin 0
};

// Extent of every dataset in two passes. Pass 1: block b of dataset d
// (blocks bbase[d] .. bbase[d+1]-1) reduces kExtChunk points of it to
// part[b] = {min x, max x, min y, max y} and counts non-finite coordinates.
// Pass 2: one wave per dataset reduces that dataset's partials.
constexpr uint32_t kExtChunk = 8 * kBlock;

__device__ __forceinline__ void ext_merge(double *m, const double *o)
{
    m[0] = o[0] < m[0] ? o[0] : m[0];
    m[1] = o[1] > m[1] ? o[1] : m[1];
    m[2] = o[2] < m[2] ? o[2] : m[2];
    m[3] = o[3] > m[3] ? o[3] : m[3];
}

__device__ __forceinline__ void ext_wave_reduce(double *m)
{
    for (int o = kWave / 2; o > 0; o >>= 1) {
        double t[4];
        for (int q = 0; q < 4; q++) t[q] = __shfl_xor(m[q], o);
        ext_merge(m, t);
    }
}

__global__ void __launch_bounds__(kBlock)
    k_seg_extent_part(const double *x, const double *y, const uint32_t *seg,
                      const uint32_t *bbase, uint32_t nsets, double *part, uint32_t *pbad)
{
    uint32_t d = seg_of(bbase, nsets, blockIdx.x);
    uint32_t a = seg[d] + (blockIdx.x - bbase[d]) * kExtChunk;
    uint32_t e = seg[d + 1];
    e = e - a > kExtChunk ? a + kExtChunk : e;
    double m[4] = {DBL_MAX, -DBL_MAX, DBL_MAX, -DBL_MAX};
    uint32_t bad = 0;
    for (uint32_t i = a + threadIdx.x; i < e; i += kBlock) {
        double u = x[i], v = y[i];
        bad += !(isfinite(u) && isfinite(v));
        double o[4] = {u, u, v, v};
        ext_merge(m, o);
    }
    ext_wave_reduce(m);
    for (int o = kWave / 2; o > 0; o >>= 1) bad += __shfl_xor(bad, o);
    __shared__ double sm[kBlock / kWave][4];
    __shared__ uint32_t sb[kBlock / kWave];
    int w = threadIdx.x / kWave;
    if (lane_id() == 0) {
        for (int q = 0; q < 4; q++) sm[w][q] = m[q];
        sb[w] = bad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < kBlock / kWave; q++) {
            ext_merge(m, sm[q]);
            bad += sb[q];
        }
        for (int q = 0; q < 4; q++) part[4 * (size_t)blockIdx.x + q] = m[q];
        pbad[blockIdx.x] = bad; // non-finite coordinates of the block
    }
}

// a wave per dataset: its extent, and its count of non-finite coordinates
__global__ void k_seg_extent_final(const double *part, const uint32_t *pbad,
                                   const uint32_t *bbase, uint32_t nsets, double *ext,
                                   uint32_t *nonfinite)
{
    uint32_t d = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    if (d >= nsets) return;
    double m[4] = {DBL_MAX, -DBL_MAX, DBL_MAX, -DBL_MAX};
    uint32_t bad = 0;
    for (uint32_t b = bbase[d] + lane_id(); b < bbase[d + 1]; b += kWave) {
        ext_merge(m, part + 4 * (size_t)b);
        bad += pbad[b];
    }
    ext_wave_reduce(m);
    for (int o = kWave / 2; o > 0; o >>= 1) bad += __shfl_xor(bad, o);
    if (lane_id() == 0) {
        for (int q = 0; q < 4; q++) ext[4 * (size_t)d + q] = m[q];
        nonfinite[d] = bad;
    }
}

// Bin parameters of every dataset from its extent (thread per dataset), as
// src/main.cpp:465-527 computes them: min from DBL_MAX, max from DBL_MIN
// (the smallest positive double), width (max - min) * 1.01, offset -min for
// a positive min and |min| otherwise. A zero width (all coordinates equal
// and positive divides 0 by 0 in the reference, which then aborts binning and
// reads past its mapping) maps the whole dimension to bin 0.
__global__ void k_bin_params(const double *ext, uint32_t nsets, BinParams *bps, uint32_t *bad)
{
    uint32_t d = blockIdx.x * kBlock + threadIdx.x;
    if (d == 0) *bad = 0u; // the bin check of k_bin_keys, which runs next
    if (d >= nsets) return;
    const double *e = ext + 4 * (size_t)d;
    double min_x = e[0] < DBL_MAX ? e[0] : DBL_MAX;
    double max_x = e[1] > DBL_MIN ? e[1] : DBL_MIN;
    double min_y = e[2] < DBL_MAX ? e[2] : DBL_MAX;
    double max_y = e[3] > DBL_MIN ? e[3] : DBL_MIN;
    BinParams bp;
    bp.xw = __dmul_rn(__dsub_rn(max_x, min_x), 1.01);
    bp.yw = __dmul_rn(__dsub_rn(max_y, min_y), 1.01);
    bp.xoff = (min_x > 0) ? -min_x : fabs(min_x);
    bp.yoff = (min_y > 0) ? -min_y : fabs(min_y);
    bp.xdeg = !(bp.xw > 0.0);
    bp.ydeg = !(bp.yw > 0.0);
    bps[d] = bp;
}

__global__ void k_bin_keys(const double *x, const double *y, uint32_t n,
                           const BinParams *bps, const uint32_t *seg, uint32_t nsets,
                           uint32_t *keys, uint32_t *vals, uint32_t *bad)
{
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    uint32_t d = seg_of(seg, nsets, i);
    BinParams bp = bps[d];
    // floor((((x + xoffset) / x_width) * numBins)   src/main.cpp:539-540
    double fx = bp.xdeg ? 0.0 : floor(__dmul_rn(__ddiv_rn(__dadd_rn(x[i], bp.xoff), bp.xw), 300.0));
    double fy = bp.ydeg ? 0.0 : floor(__dmul_rn(__ddiv_rn(__dadd_rn(y[i], bp.yoff), bp.yw), 300.0));
    if (!(fx >= 0.0 && fx <= 299.0 && fy >= 0.0 && fy <= 299.0)) {
        atomicOr(bad, 1u);
        fx = 0.0;
        fy = 0.0;
    }
    // dataset-major: each dataset's points stay contiguous, in its own binned order
    keys[i] = d * 90000u + (uint32_t)fx * 300u + (uint32_t)fy;
    vals[i] = i;
}

__global__ void k_gather_xy(const double *x, const double *y, const uint32_t *idx, uint32_t n,
                            double *ox, double *oy)
{
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    uint32_t j = idx[i];
    ox[i] = x[j];
    oy[i] = y[j];
}

__global__ void k_cell_keys(const double *x, const double *y, uint32_t n, const uint32_t *seg,
                            uint32_t nsets, const double *sgx0, const double *sgy0, double w,
                            const uint32_t *sncx, const uint32_t *sncy, const uint32_t *skbase,
                            uint32_t *keys, uint32_t *vals)
{
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    uint32_t d = seg_of(seg, nsets, i);
    uint32_t ncx = sncx[d], ncy = sncy[d];
    double fx = floor(__ddiv_rn(__dsub_rn(x[i], sgx0[d]), w));
    double fy = floor(__ddiv_rn(__dsub_rn(y[i], sgy0[d]), w));
    fx = fx < 0.0 ? 0.0 : (fx > (double)(ncx - 1) ? (double)(ncx - 1) : fx);
    fy = fy < 0.0 ? 0.0 : (fy > (double)(ncy - 1) ? (double)(ncy - 1) : fy);
    keys[i] = skbase[d] + (uint32_t)fx * ncy + (uint32_t)fy;
    vals[i] = i;
}

__global__ void k_head_flags(const uint32_t *keys, uint32_t n, uint32_t *flags)
{
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    flags[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

// excl = exclusive scan of head flags: cell index of pos i is excl[i]+flag[i]-1
// kflag (may be null): the occupied keys are marked, for a dense key table
__global__ void k_cells(const uint32_t *keys, const uint32_t *flags, const uint32_t *excl,
                        uint32_t n, uint32_t *cell_of_pos, uint32_t *cell_key,
                        uint32_t *cell_start, const uint32_t *pos_rank, uint32_t *rank_pos,
                        uint32_t *kflag)
{
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    uint32_t c = excl[i] + flags[i] - 1;
    cell_of_pos[i] = c;
    rank_pos[pos_rank[i]] = i;
    if (flags[i]) {
        cell_key[c] = keys[i];
        cell_start[c] = i;
        if (kflag) kflag[keys[i]] = 1u;
    }
}

__global__ void k_cell_tail(uint32_t *cell_start, uint32_t ncell, uint32_t n, uint32_t *nbig)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        cell_start[ncell] = n;
        *nbig = 0; // the box pass's list of large cells
    }
}

// Bounding box of every cell's points. Thread per cell for the common small
// cell (its points read 8 at a time); cells holding more than kBoxSerial
// points go to a list that k_cell_box_big reduces, a wave per cell.
constexpr uint32_t kBoxSerial = 16;
constexpr int kBoxUnroll = 8;
__global__ void k_cell_box(const double *xs, const double *ys, const uint32_t *cell_start,
                           uint32_t ncell, double *box, uint32_t *big, uint32_t *nbig)
{
    uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    uint32_t a = 0, b = 0;
    if (c < ncell) {
        a = cell_start[c];
        b = cell_start[c + 1];
    }
    bool large = c < ncell && b - a > kBoxSerial;
    if (c < ncell && !large) {
        double x0 = DBL_MAX, x1 = -DBL_MAX, y0 = DBL_MAX, y1 = -DBL_MAX;
        // 8 points per round of loads, indices clamped into the cell (a
        // cell is never empty), so the loads go out together
        for (uint32_t p0 = a; p0 < b; p0 += kBoxUnroll) {
            double u[kBoxUnroll], v[kBoxUnroll];
#pragma unroll
            for (int t = 0; t < kBoxUnroll; t++) {
                uint32_t p = p0 + t < b ? p0 + t : b - 1;
                u[t] = xs[p];
                v[t] = ys[p];
            }
#pragma unroll
            for (int t = 0; t < kBoxUnroll; t++) {
                x0 = u[t] < x0 ? u[t] : x0;
                x1 = u[t] > x1 ? u[t] : x1;
                y0 = v[t] < y0 ? v[t] : y0;
                y1 = v[t] > y1 ? v[t] : y1;
            }
        }
        box[4 * (size_t)c + 0] = x0;
        box[4 * (size_t)c + 1] = x1;
        box[4 * (size_t)c + 2] = y0;
        box[4 * (size_t)c + 3] = y1;
    }
    uint64_t m = ballot(large);
    if (m) {
        int leader = __builtin_ctzll(m);
        uint32_t base = 0;
        if (lane_id() == leader) base = atomicAdd(nbig, (uint32_t)popc64(m));
        base = __shfl(base, leader);
        if (large) big[base + mask_rank(m)] = c;
    }
}

// the listed large cells, a wave per cell over a grid sized for the GPU
__global__ void k_cell_box_big(const double *xs, const double *ys, const uint32_t *cell_start,
                               const uint32_t *big, const uint32_t *nbig, double *box)
{
    const uint32_t nb = *nbig;
    for (uint32_t w = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave; w < nb;
         w += gridDim.x * (kBlock / kWave)) {
        uint32_t cc = big[w];
        uint32_t ca = cell_start[cc], cb = cell_start[cc + 1];
        double x0 = DBL_MAX, x1 = -DBL_MAX, y0 = DBL_MAX, y1 = -DBL_MAX;
        // kBoxUnroll strides per step, all loads issued before the first
        // min/max: a dense cell (thousands of points) is a long walk
        uint32_t p = ca + lane_id();
        for (; p + (kBoxUnroll - 1) * kWave < cb; p += kBoxUnroll * kWave) {
            double u[kBoxUnroll], v[kBoxUnroll];
#pragma unroll
            for (int t = 0; t < kBoxUnroll; t++) {
                u[t] = xs[p + t * kWave];
                v[t] = ys[p + t * kWave];
            }
#pragma unroll
            for (int t = 0; t < kBoxUnroll; t++) {
                x0 = u[t] < x0 ? u[t] : x0;
                x1 = u[t] > x1 ? u[t] : x1;
                y0 = v[t] < y0 ? v[t] : y0;
                y1 = v[t] > y1 ? v[t] : y1;
            }
        }
        for (; p < cb; p += kWave) {
            double u = xs[p], v = ys[p];
            x0 = u < x0 ? u : x0;
            x1 = u > x1 ? u : x1;
            y0 = v < y0 ? v : y0;
            y1 = v > y1 ? v : y1;
        }
        for (int o = kWave / 2; o > 0; o >>= 1) {
            double t;
            t = __shfl_xor(x0, o); x0 = t < x0 ? t : x0;
            t = __shfl_xor(x1, o); x1 = t > x1 ? t : x1;
            t = __shfl_xor(y0, o); y0 = t < y0 ? t : y0;
            t = __shfl_xor(y1, o); y1 = t > y1 ? t : y1;
        }
        if (lane_id() == 0) {
            box[4 * (size_t)cc + 0] = x0;
            box[4 * (size_t)cc + 1] = x1;
            box[4 * (size_t)cc + 2] = y0;
            box[4 * (size_t)cc + 3] = y1;
        }
    }
}

// largest cell population: block-level maximum, one atomic per block
__global__ void __launch_bounds__(kBlock)
    k_max_pop(const uint32_t *cell_start, uint32_t ncell, uint32_t *max_pop)
{
    uint32_t m = 0;
    for (uint32_t c = blockIdx.x * kBlock + threadIdx.x; c < ncell; c += gridDim.x * kBlock) {
        uint32_t p = cell_start[c + 1] - cell_start[c];
        m = p > m ? p : m;
    }
    for (int o = kWave / 2; o > 0; o >>= 1) {
        uint32_t t = __shfl_xor(m, o);
        m = t > m ? t : m;
    }
    __shared__ uint32_t sm[kBlock / kWave];
    if (lane_id() == 0) sm[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBlock / kWave; w++) m = sm[w] > m ? sm[w] : m;
        atomicMax(max_pop, m);
    }
}

// Smallest i in [0, n] with a[i] >= v (a ascending), searched outward from
// hint with doubling steps, then by bisection: a window row's cells lie near
// the centre cell in key order, so this reads O(log distance) keys instead
// of O(log n).
__device__ __forceinline__ uint32_t lower_bound_near(const uint32_t *a, uint32_t n, uint32_t v,
                                                     uint32_t hint)
{
    uint32_t lo, hi;
    if (hint < n && a[hint] < v) {
        // answer in (hint, n]
        lo = hint + 1;
        hi = n;
        uint64_t off = 1;
        while (true) {
            uint64_t p = (uint64_t)hint + off;
            if (p >= n) break;
            if (a[p] >= v) {
                hi = (uint32_t)p;
                break;
            }
            lo = (uint32_t)p + 1;
            off <<= 1;
        }
    } else {
        // answer in [0, min(hint, n)]
        hi = hint < n ? hint : n;
        lo = 0;
        uint32_t h = hi;
        uint64_t off = 1;
        while (off <= h) {
            uint32_t p = h - (uint32_t)off;
            if (a[p] < v) {
                lo = p + 1;
                break;
            }
            hi = p;
            off <<= 1;
        }
    }
    while (lo < hi) {
        uint32_t m = lo + ((hi - lo) >> 1);
        if (a[m] < v) lo = m + 1;
        else hi = m;
    }
    return lo;
}

// window rows: for cell c and row offset d in [-R, R], the cells of grid row
// cx+d whose cy lies in [cy-R, cy+R] form one contiguous run of the sorted keys
// rows stay inside the cell's own dataset grid. A thread takes one row offset
// for kRowRun consecutive cells: each search starts from the previous cell's
// answer, a few positions away, instead of galloping across up to R columns.
constexpr uint32_t kRowRun = 4;
__global__ void k_row_ranges(const uint32_t *cell_key, uint32_t ncell, const uint32_t *sncx,
                             const uint32_t *sncy, const uint32_t *skbase, uint32_t nsets, int R,
                             uint32_t *row_lo, uint32_t *row_hi, uint32_t *cell_ds)
{
    // 64-bit thread index: runs x window rows exceeds 2^32 on large grids
    uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    uint32_t W = 2 * R + 1;
    uint32_t nrun = (ncell + kRowRun - 1) / kRowRun;
    if (t >= (uint64_t)nrun * W) return;
    uint32_t run = (uint32_t)(t / W);
    int dr = (int)(t % W) - R;
    uint32_t c0 = run * kRowRun;
    uint32_t c1 = c0 + kRowRun < ncell ? c0 + kRowRun : ncell;
    uint32_t hint = c0;
    for (uint32_t c = c0; c < c1; c++) {
        uint32_t key = cell_key[c];
        uint32_t ds = seg_of(skbase, nsets, key);
        if (dr == 0) cell_ds[c] = ds;
        uint32_t ncx = sncx[ds], ncy = sncy[ds], kb = skbase[ds];
        uint32_t local = key - kb;
        int64_t cx = local / ncy, cy = local % ncy;
        int64_t rx = cx + dr;
        uint32_t lo = 0, hi = 0;
        if (rx >= 0 && rx < (int64_t)ncx) {
            int64_t y0 = cy - R < 0 ? 0 : cy - R;
            int64_t y1 = cy + R > (int64_t)ncy - 1 ? (int64_t)ncy - 1 : cy + R;
            uint64_t k0 = kb + (uint64_t)(rx * ncy + y0);
            uint64_t k1 = kb + (uint64_t)(rx * ncy + y1);
            lo = lower_bound_near(cell_key, ncell, (uint32_t)k0, hint);
            hi = k1 >= 0xffffffffull ? ncell
                                     : lower_bound_near(cell_key, ncell, (uint32_t)(k1 + 1), lo);
            hint = lo;
        }
        row_lo[(size_t)c * W + (dr + R)] = lo;
        row_hi[(size_t)c * W + (dr + R)] = hi;
    }
}

// The same row ranges from a dense table over the key space: krank[k] = the
// number of non-empty cells whose key is below k (an exclusive scan of the
// occupied keys, which k_cells marks), so a row's run is [krank[k0],
// krank[k1 + 1]). A thread per (cell, row), no search: for key spaces that
// are a small multiple of the cell count. With row_pts, the run's points too.
__global__ void k_row_ranges_dense(const uint32_t *cell_key, uint32_t ncell, const uint32_t *sncx,
                                   const uint32_t *sncy, const uint32_t *skbase, uint32_t nsets,
                                   int R, const uint32_t *krank, const uint32_t *cell_start,
                                   uint32_t *row_lo, uint32_t *row_hi, uint32_t *cell_ds,
                                   uint2 *row_pts)
{
    uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    uint32_t W = 2 * R + 1;
    if (t >= (uint64_t)ncell * W) return;
    uint32_t c = (uint32_t)(t / W);
    int dr = (int)(t - (uint64_t)c * W) - R;
    uint32_t key = cell_key[c];
    uint32_t ds = seg_of(skbase, nsets, key);
    if (dr == 0) cell_ds[c] = ds;
    uint32_t ncx = sncx[ds], ncy = sncy[ds], kb = skbase[ds];
    uint32_t local = key - kb;
    int64_t cx = local / ncy, cy = local % ncy;
    int64_t rx = cx + dr;
    uint32_t lo = 0, hi = 0;
    if (rx >= 0 && rx < (int64_t)ncx) {
        int64_t y0 = cy - R < 0 ? 0 : cy - R;
        int64_t y1 = cy + R > (int64_t)ncy - 1 ? (int64_t)ncy - 1 : cy + R;
        uint64_t k0 = kb + (uint64_t)(rx * ncy + y0);
        uint64_t k1 = kb + (uint64_t)(rx * ncy + y1);
        lo = krank[k0];
        hi = krank[k1 + 1];
    }
    row_lo[t] = lo;
    row_hi[t] = hi;
    if (row_pts) {
        uint2 r = make_uint2(0u, 0u);
        if (lo < hi) {
            r.x = cell_start[lo];
            r.y = cell_start[hi] - r.x;
        }
        row_pts[t] = r;
    }
}

// Point range (start, count) of every window row, from the row's cells
__global__ void k_row_points(const uint32_t *row_lo, const uint32_t *row_hi,
                             const uint32_t *cell_start, size_t nrows, uint2 *row_pts)
{
    size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= nrows) return;
    uint32_t lo = row_lo[t], hi = row_hi[t];
    uint2 r = make_uint2(0u, 0u);
    if (lo < hi) {
        r.x = cell_start[lo];
        r.y = cell_start[hi] - r.x;
    }
    row_pts[t] = r;
}

std::string slot(const char *tag, const char *name) { return std::string(tag) + "." + name; }

} // namespace

void single_segment(Points &P)
{
    P.nsets = 1;
    P.hseg.assign({0u, P.n});
}

int build_rank_order(Context &c, Points &P, const char *tag)
{
    uint32_t n = P.n;
    if (P.hseg.size() != (size_t)P.nsets + 1) single_segment(P);
    uint32_t ns = P.nsets;
    // dataset bounds and the first extent block of each dataset, in one
    // upload: seg[0..ns], then bbase[0..ns]
    std::vector<uint32_t> up(2 * ((size_t)ns + 1), 0);
    std::copy(P.hseg.begin(), P.hseg.end(), up.begin());
    uint32_t *blk = up.data() + ns + 1;
    for (uint32_t d = 0; d < ns; d++)
        blk[d + 1] = blk[d] + (uint32_t)ceil_div(P.hseg[d + 1] - P.hseg[d], kExtChunk);
    uint32_t *segb;
    VDB_TRY(c.get(slot(tag, "seg").c_str(), up.size(), &segb));
    VDB_HIP(hipMemcpyAsync(segb, up.data(), sizeof(uint32_t) * up.size(), hipMemcpyHostToDevice,
                           c.stream()));
    P.seg = segb;
    uint32_t *bbase = segb + ns + 1;
    // what comes back to the host, in one region: the extents (4 doubles a
    // dataset), the non-finite counts (one a dataset), the bin check
    const size_t cbytes = 32 * (size_t)ns + 4 * (size_t)ns + 4;
    unsigned char *check;
    double *part = nullptr;
    uint32_t *pbad = nullptr;
    VDB_TRY(c.get(slot(tag, "check").c_str(), cbytes, &check));
    double *ext = (double *)check;
    uint32_t *nonfin = (uint32_t *)(check + 32 * (size_t)ns);
    uint32_t *bad = nonfin + ns;
    VDB_TRY(c.get(slot(tag, "extpart").c_str(), (size_t)std::max(blk[ns], 1u) * 4, &part));
    VDB_TRY(c.get(slot(tag, "extbad").c_str(), (size_t)std::max(blk[ns], 1u), &pbad));
    if (blk[ns])
        k_seg_extent_part<<<blk[ns], kBlock, 0, c.stream()>>>(P.x, P.y, P.seg, bbase, ns, part,
                                                             pbad);
    VDB_HIP(hipGetLastError());
    k_seg_extent_final<<<ceil_div(ns, kBlock / kWave), kBlock, 0, c.stream()>>>(
        part, pbad, bbase, ns, ext, nonfin);
    VDB_HIP(hipGetLastError());
    // bin parameters on the device: the extent is read back only once, with
    // the non-finite counts and the bin check, after the sort
    BinParams *dbp;
    VDB_TRY(c.get(slot(tag, "binparams").c_str(), ns, &dbp));
    k_bin_params<<<ceil_div(ns, kBlock), kBlock, 0, c.stream()>>>(ext, ns, dbp, bad);
    VDB_HIP(hipGetLastError());

    // the sort's own buffers: its result stays where the last pass left it
    uint32_t *keys, *vals, *tk, *tv;
    VDB_TRY(c.get(slot(tag, "bk").c_str(), n, &keys));
    VDB_TRY(c.get(slot(tag, "bv").c_str(), n, &vals));
    VDB_TRY(c.get(slot(tag, "btk").c_str(), n, &tk));
    VDB_TRY(c.get(slot(tag, "btv").c_str(), n, &tv));
    unsigned g = ceil_div(n, kBlock);
    k_bin_keys<<<g, kBlock, 0, c.stream()>>>(P.x, P.y, n, dbp, P.seg, ns, keys, vals, bad);
    VDB_HIP(hipGetLastError());
    uint32_t *sk, *sv;
    VDB_TRY(radix_sort_pairs(c, keys, vals, tk, tv, n, bits_for((uint64_t)ns * 90000), &sk,
                             &sv));
    P.mapping = sv;
    VDB_TRY(c.get(slot(tag, "xb").c_str(), n, &P.xb));
    VDB_TRY(c.get(slot(tag, "yb").c_str(), n, &P.yb));
    k_gather_xy<<<g, kBlock, 0, c.stream()>>>(P.x, P.y, P.mapping, n, P.xb, P.yb);
    VDB_HIP(hipGetLastError());
    // one read back: extents, non-finite count, bin check (a non-finite
    // coordinate or an overflowing bin only made the keys meaningless: every
    // access above stays in bounds)
    char *h = nullptr;
    VDB_TRY(c.host_scratch(cbytes, (void **)&h));
    VDB_HIP(hipMemcpyAsync(h, check, cbytes, hipMemcpyDeviceToHost, c.stream()));
    VDB_TRY(c.wait(c.stream()));
    uint32_t hnf = 0, hb = 0;
    for (uint32_t d = 0; d < ns; d++) {
        uint32_t v;
        memcpy(&v, h + 32 * (size_t)ns + 4 * (size_t)d, 4);
        hnf += v;
    }
    memcpy(&hb, h + 36 * (size_t)ns, 4);
    if (hnf) {
        set_error("non-finite coordinate in input");
        return kErrNonFinite;
    }
    if (hb) {
        set_error("bins out of range (coordinate arithmetic overflowed)");
        return kErrUnsupported;
    }
    P.hext.assign((const double *)h, (const double *)h + 4 * (size_t)ns);
    P.ext[0] = DBL_MAX;
    P.ext[1] = -DBL_MAX;
    P.ext[2] = DBL_MAX;
    P.ext[3] = -DBL_MAX;
    for (uint32_t d = 0; d < ns; d++) {
        const double *e = &P.hext[4 * (size_t)d];
        if (P.hseg[d + 1] > P.hseg[d]) {
            P.ext[0] = std::min(P.ext[0], e[0]);
            P.ext[1] = std::max(P.ext[1], e[1]);
            P.ext[2] = std::min(P.ext[2], e[2]);
            P.ext[3] = std::max(P.ext[3], e[3]);
        }
    }
    if (P.ext[0] > P.ext[1]) {
        P.ext[0] = P.ext[1] = P.ext[2] = P.ext[3] = 0.0;
    }
    return kOk;
}

// Thread per (cell, window row). Rows are runs of consecutive cells whose
// point boxes are ordered along the row, so trimming from both ends removes
// every cell whose box lies beyond eps of the centre cell's box; the first
// run of cells wholly within eps becomes the full range.
__global__ void k_classify_rows(const double *box, uint32_t ncell, int W, double T,
                                uint32_t *row_lo, uint32_t *row_hi, uint32_t *row_f0,
                                uint32_t *row_f1)
{
    // 64-bit thread index: cells x window rows exceeds 2^32 on large grids
    uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= (uint64_t)ncell * (uint32_t)W) return;
    uint32_t c = (uint32_t)(t / (uint32_t)W);
    const double *A = box + 4 * (size_t)c;
    uint32_t lo = row_lo[t], hi = row_hi[t];
    while (lo < hi && boxes_dist2_min(A, box + 4 * (size_t)lo) > T) lo++;
    while (hi > lo && boxes_dist2_min(A, box + 4 * (size_t)(hi - 1)) > T) hi--;
    uint32_t f0 = lo;
    while (f0 < hi && boxes_dist2_max(A, box + 4 * (size_t)f0) > T) f0++;
    uint32_t f1 = f0;
    while (f1 < hi && boxes_dist2_max(A, box + 4 * (size_t)f1) <= T) f1++;
    if (f0 == f1) f0 = f1 = hi;
    row_lo[t] = lo;
    row_hi[t] = hi;
    row_f0[t] = f0;
    row_f1[t] = f1;
}

int classify_rows(Context &c, Grid &G, double T, const char *tag)
{
    size_t W = 2 * (size_t)G.R + 1;
    VDB_TRY(c.get(slot(tag, "row_f0").c_str(), G.ncell * W, &G.row_f0));
    VDB_TRY(c.get(slot(tag, "row_f1").c_str(), G.ncell * W, &G.row_f1));
    if (G.ncell == 0) return kOk;
    k_classify_rows<<<ceil_div(G.ncell * W, kBlock), kBlock, 0, c.stream()>>>(
        G.cell_box, G.ncell, (int)W, T, G.row_lo, G.row_hi, G.row_f0, G.row_f1);
    VDB_HIP(hipGetLastError());
    return kOk;
}

int build_grid(Context &c, const Points &P, double w_min, int R, const char *tag, Grid &G,
               bool *enlarged)
{
    uint32_t n = P.n;
    uint32_t ns = P.nsets;
    double spanmax = 0.0;
    for (uint32_t d = 0; d < ns; d++) {
        if (P.hseg[d + 1] == P.hseg[d]) continue;
        const double *e = &P.hext[4 * (size_t)d];
        spanmax = std::fmax(spanmax, std::fmax(e[1] - e[0], e[3] - e[2]));
    }
    double w = w_min;
    *enlarged = false;
    if (!(w > 0.0) || !std::isfinite(w)) {
        w = std::fmax(spanmax, 1.0) * 2.0;
        *enlarged = true;
    }
    // every axis stays below 2^16 cells (so one dataset's keys fit 32 bits);
    // with several datasets the side then grows until all keys fit together
    if (w < spanmax / 65000.0) {
        w = spanmax / 65000.0;
        *enlarged = true;
    }
    std::vector<double> gx0(ns), gy0(ns);
    std::vector<uint32_t> ncx(ns), ncy(ns), kbase(ns + 1);
    for (int attempt = 0;; attempt++) {
        uint64_t total = 0;
        bool ok = true;
        for (uint32_t d = 0; d < ns; d++) {
            const double *e = &P.hext[4 * (size_t)d];
            bool empty = P.hseg[d + 1] == P.hseg[d];
            double sx = empty ? 0.0 : e[1] - e[0], sy = empty ? 0.0 : e[3] - e[2];
            gx0[d] = empty ? 0.0 : e[0];
            gy0[d] = empty ? 0.0 : e[2];
            double fx = empty ? 1.0 : std::floor(sx / w) + 2;
            double fy = empty ? 1.0 : std::floor(sy / w) + 2;
            if (fx > 65536.0 || fy > 65536.0) ok = false;
            ncx[d] = ok ? (uint32_t)fx : 1;
            ncy[d] = ok ? (uint32_t)fy : 1;
            kbase[d] = (uint32_t)std::min<uint64_t>(total, 0xffffffffull);
            total += (uint64_t)ncx[d] * ncy[d];
        }
        kbase[ns] = (uint32_t)std::min<uint64_t>(total, 0xffffffffull);
        if (ok && total < 0xffffffffull) break;
        w *= ok ? std::sqrt((double)total / 4.0e9) * 1.05 : 1.5;
        *enlarged = true;
        if (attempt > 64) {
            set_error("grid key space");
            return kErrUnsupported;
        }
    }
    G.w = w;
    G.gx0 = gx0[0];
    G.gy0 = gy0[0];
    G.ncx = ncx[0];
    G.ncy = ncy[0];
    G.R = R;
    G.npts = n;
    G.nsets = ns;
    uint64_t nkeys = (uint64_t)kbase[ns];
    // the per-dataset grid parameters in one upload: origins (2 x 8 ns
    // bytes), then cells per axis (2 x 4 ns) and key bases (4 (ns + 1))
    const size_t pbytes = 16 * (size_t)ns + 8 * (size_t)ns + 4 * ((size_t)ns + 1);
    std::vector<unsigned char> hp(pbytes);
    memcpy(hp.data(), gx0.data(), 8 * (size_t)ns);
    memcpy(hp.data() + 8 * (size_t)ns, gy0.data(), 8 * (size_t)ns);
    memcpy(hp.data() + 16 * (size_t)ns, ncx.data(), 4 * (size_t)ns);
    memcpy(hp.data() + 20 * (size_t)ns, ncy.data(), 4 * (size_t)ns);
    memcpy(hp.data() + 24 * (size_t)ns, kbase.data(), 4 * ((size_t)ns + 1));
    unsigned char *dp;
    VDB_TRY(c.get(slot(tag, "params").c_str(), pbytes, &dp));
    VDB_HIP(hipMemcpyAsync(dp, hp.data(), pbytes, hipMemcpyHostToDevice, c.stream()));
    G.sgx0 = (double *)dp;
    G.sgy0 = (double *)(dp + 8 * (size_t)ns);
    G.sncx = (uint32_t *)(dp + 16 * (size_t)ns);
    G.sncy = (uint32_t *)(dp + 20 * (size_t)ns);
    G.skbase = (uint32_t *)(dp + 24 * (size_t)ns);

    // the sort's own buffers: its result stays where the last pass left it
    uint32_t *keys, *vals, *tk, *tv, *flags, *excl;
    VDB_TRY(c.get(slot(tag, "gk").c_str(), n, &keys));
    VDB_TRY(c.get(slot(tag, "gv").c_str(), n, &vals));
    VDB_TRY(c.get(slot(tag, "gtk").c_str(), n, &tk));
    VDB_TRY(c.get(slot(tag, "gtv").c_str(), n, &tv));
    unsigned g = ceil_div(n, kBlock);
    k_cell_keys<<<g, kBlock, 0, c.stream()>>>(P.xb, P.yb, n, P.seg, ns, G.sgx0, G.sgy0, w,
                                              G.sncx, G.sncy, G.skbase, keys, vals);
    VDB_HIP(hipGetLastError());
    uint32_t *sk, *sv;
    VDB_TRY(radix_sort_pairs(c, keys, vals, tk, tv, n, bits_for(nkeys), &sk, &sv));
    G.pos_rank = sv;
    uint32_t *skeys = sk;
    VDB_TRY(c.get(slot(tag, "xs").c_str(), n, &G.xs));
    VDB_TRY(c.get(slot(tag, "ys").c_str(), n, &G.ys));
    k_gather_xy<<<g, kBlock, 0, c.stream()>>>(P.xb, P.yb, G.pos_rank, n, G.xs, G.ys);
    VDB_HIP(hipGetLastError());

    VDB_TRY(c.get(slot(tag, "flags").c_str(), n, &flags));
    VDB_TRY(c.get(slot(tag, "excl").c_str(), n, &excl));
    k_head_flags<<<g, kBlock, 0, c.stream()>>>(skeys, n, flags);
    VDB_HIP(hipGetLastError());
    uint32_t ncell = 0;
    VDB_TRY(compact_count(c, flags, excl, n, &ncell));
    G.ncell = ncell;
    VDB_TRY(c.get(slot(tag, "cell_of_pos").c_str(), n, &G.cell_of_pos));
    VDB_TRY(c.get(slot(tag, "rank_pos").c_str(), n, &G.rank_pos));
    VDB_TRY(c.get(slot(tag, "cell_key").c_str(), ncell, &G.cell_key));
    VDB_TRY(c.get(slot(tag, "cell_start").c_str(), (size_t)ncell + 1, &G.cell_start));
    // window rows from a dense key table when the key space is at most 16
    // keys per cell (not in the allocation-free mode, whose scratch is
    // small); k_cells marks the occupied keys
    const bool dense = !c.in_arena() && ncell &&
                       nkeys <= std::max<uint64_t>(16ull * ncell, (uint64_t)1 << 22);
    uint32_t *kflag = nullptr;
    if (dense) {
        VDB_TRY(c.get(slot(tag, "kflag").c_str(), nkeys + 1, &kflag));
        VDB_HIP(hipMemsetAsync(kflag, 0, sizeof(uint32_t) * (nkeys + 1), c.stream()));
    }
    k_cells<<<g, kBlock, 0, c.stream()>>>(skeys, flags, excl, n, G.cell_of_pos, G.cell_key,
                                          G.cell_start, G.pos_rank, G.rank_pos, kflag);
    VDB_HIP(hipGetLastError());
    uint32_t *bigc, *nbig;
    VDB_TRY(c.get(slot(tag, "bigcells").c_str(), std::max<uint32_t>(ncell, 1), &bigc));
    VDB_TRY(c.get(slot(tag, "nbig").c_str(), 1, &nbig));
    k_cell_tail<<<1, 64, 0, c.stream()>>>(G.cell_start, ncell, n, nbig);
    VDB_HIP(hipGetLastError());

    VDB_TRY(c.get(slot(tag, "cell_box").c_str(), (size_t)ncell * 4, &G.cell_box));
    if (ncell) {
        k_cell_box<<<ceil_div(ncell, kBlock), kBlock, 0, c.stream()>>>(
            G.xs, G.ys, G.cell_start, ncell, G.cell_box, bigc, nbig);
        VDB_HIP(hipGetLastError());
        k_cell_box_big<<<std::min(ceil_div(ncell, kBlock / kWave), 1024u), kBlock, 0,
                         c.stream()>>>(G.xs, G.ys, G.cell_start, bigc, nbig, G.cell_box);
        VDB_HIP(hipGetLastError());
    }

    size_t W = 2 * (size_t)R + 1;
    VDB_TRY(c.get(slot(tag, "row_lo").c_str(), ncell * W, &G.row_lo));
    VDB_TRY(c.get(slot(tag, "row_hi").c_str(), ncell * W, &G.row_hi));
    VDB_TRY(c.get(slot(tag, "cell_ds").c_str(), ncell, &G.cell_ds));
    G.row_pts = nullptr;
    if (G.want_row_pts)
        VDB_TRY(c.get(slot(tag, "row_pts").c_str(), std::max<size_t>(ncell * W, 1), &G.row_pts));
    if (dense) {
        uint32_t *krank;
        VDB_TRY(c.get(slot(tag, "krank").c_str(), nkeys + 1, &krank));
        VDB_TRY(scan_exclusive_u32(c, kflag, krank, nkeys + 1, nullptr));
        k_row_ranges_dense<<<ceil_div((uint64_t)ncell * W, kBlock), kBlock, 0, c.stream()>>>(
            G.cell_key, ncell, G.sncx, G.sncy, G.skbase, ns, R, krank, G.cell_start, G.row_lo,
            G.row_hi, G.cell_ds, G.row_pts);
        VDB_HIP(hipGetLastError());
        return kOk;
    }
    k_row_ranges<<<ceil_div(ceil_div(ncell, kRowRun) * W, kBlock), kBlock, 0, c.stream()>>>(
        G.cell_key, ncell, G.sncx, G.sncy, G.skbase, ns, R, G.row_lo, G.row_hi, G.cell_ds);
    VDB_HIP(hipGetLastError());
    if (G.row_pts && ncell) {
        k_row_points<<<ceil_div(ncell * W, kBlock), kBlock, 0, c.stream()>>>(
            G.row_lo, G.row_hi, G.cell_start, ncell * W, G.row_pts);
        VDB_HIP(hipGetLastError());
    }
    return kOk;
}

int grid_max_cell(Context &c, const Grid &G, uint32_t *out)
{
    uint32_t *maxpop;
    VDB_TRY(c.get("grid.maxpop", 1, &maxpop));
    VDB_HIP(hipMemsetAsync(maxpop, 0, sizeof(uint32_t), c.stream()));
    if (G.ncell)
        k_max_pop<<<std::min<unsigned>(ceil_div(G.ncell, kBlock), 512), kBlock, 0, c.stream()>>>(
            G.cell_start, G.ncell, maxpop);
    VDB_HIP(hipGetLastError());
    void *h = nullptr;
    VDB_TRY(c.host_scratch(sizeof(uint32_t), &h));
    VDB_HIP(hipMemcpyAsync(h, maxpop, sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream()));
    VDB_HIP(hipStreamSynchronize(c.stream()));
    memcpy(out, h, sizeof(uint32_t));
    return kOk;
}

} // namespace vdb
