// Reference visit order (binning) and grid construction.
#include "vdb_grid.h"
#include "vdb_prims.h"

#include <cfloat>
#include <cmath>
#include <cstring>

namespace vdb {

double threshold_for_eps(double eps)
{
    if (!(eps >= 0.0)) return -1.0;
    if (std::isinf(eps)) return INFINITY;
    uint64_t lo = 0, hi = 0x7fefffffffffffffull;
    double dhi;
    memcpy(&dhi, &hi, 8);
    if (std::sqrt(dhi) <= eps) return dhi;
    while (hi - lo > 1) { // sqrt(lo) <= eps < sqrt(hi); doubles >= 0 order like their bits
        uint64_t mid = lo + (hi - lo) / 2;
        double dm;
        memcpy(&dm, &mid, 8);
        if (std::sqrt(dm) <= eps) lo = mid; else hi = mid;
    }
    double t;
    memcpy(&t, &lo, 8);
    return t;
}

namespace {

constexpr int kBlock = 256;

struct BinParams {
    double xoff, xw, yoff, yw;
    int xdeg, ydeg; // dimension has zero width: every point in bin 0
};

__global__ void k_bin_keys(const double *x, const double *y, uint32_t n, BinParams bp,
                           uint32_t *keys, uint32_t *vals, uint32_t *bad)
{
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    // floor((((x + xoffset) / x_width) * numBins)   src/main.cpp:539-540
    double fx = bp.xdeg ? 0.0 : floor(__dmul_rn(__ddiv_rn(__dadd_rn(x[i], bp.xoff), bp.xw), 300.0));
    double fy = bp.ydeg ? 0.0 : floor(__dmul_rn(__ddiv_rn(__dadd_rn(y[i], bp.yoff), bp.yw), 300.0));
    if (!(fx >= 0.0 && fx <= 299.0 && fy >= 0.0 && fy <= 299.0)) {
        atomicOr(bad, 1u);
        fx = 0.0;
        fy = 0.0;
    }
    keys[i] = (uint32_t)fx * 300u + (uint32_t)fy;
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

__global__ void k_cell_keys(const double *x, const double *y, uint32_t n, double gx0,
                            double gy0, double w, uint32_t ncx, uint32_t ncy, uint32_t *keys,
                            uint32_t *vals)
{
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    double fx = floor(__ddiv_rn(__dsub_rn(x[i], gx0), w));
    double fy = floor(__ddiv_rn(__dsub_rn(y[i], gy0), w));
    fx = fx < 0.0 ? 0.0 : (fx > (double)(ncx - 1) ? (double)(ncx - 1) : fx);
    fy = fy < 0.0 ? 0.0 : (fy > (double)(ncy - 1) ? (double)(ncy - 1) : fy);
    keys[i] = (uint32_t)fx * ncy + (uint32_t)fy;
    vals[i] = i;
}

__global__ void k_head_flags(const uint32_t *keys, uint32_t n, uint32_t *flags)
{
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    flags[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

// excl = exclusive scan of head flags: cell index of pos i is excl[i]+flag[i]-1
__global__ void k_cells(const uint32_t *keys, const uint32_t *flags, const uint32_t *excl,
                        uint32_t n, uint32_t *cell_of_pos, uint32_t *cell_key,
                        uint32_t *cell_start, const uint32_t *pos_rank, uint32_t *rank_pos)
{
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    uint32_t c = excl[i] + flags[i] - 1;
    cell_of_pos[i] = c;
    rank_pos[pos_rank[i]] = i;
    if (flags[i]) {
        cell_key[c] = keys[i];
        cell_start[c] = i;
    }
}

__global__ void k_cell_tail(uint32_t *cell_start, uint32_t ncell, uint32_t n)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) cell_start[ncell] = n;
}

// bbox of each cell from its points; one wave per cell
__global__ void k_cell_box(const double *xs, const double *ys, const uint32_t *cell_start,
                           uint32_t ncell, double *box)
{
    uint32_t c = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    if (c >= ncell) return;
    uint32_t a = cell_start[c], b = cell_start[c + 1];
    double x0 = DBL_MAX, x1 = -DBL_MAX, y0 = DBL_MAX, y1 = -DBL_MAX;
    for (uint32_t p = a + lane_id(); p < b; p += kWave) {
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
        box[4 * (size_t)c + 0] = x0;
        box[4 * (size_t)c + 1] = x1;
        box[4 * (size_t)c + 2] = y0;
        box[4 * (size_t)c + 3] = y1;
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

__device__ __forceinline__ uint32_t lower_bound_u32(const uint32_t *a, uint32_t n, uint32_t v)
{
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t m = (lo + hi) >> 1;
        if (a[m] < v) lo = m + 1; else hi = m;
    }
    return lo;
}

// window rows: for cell c and row offset d in [-R, R], the cells of grid row
// cx+d whose cy lies in [cy-R, cy+R] form one contiguous run of the sorted keys
__global__ void k_row_ranges(const uint32_t *cell_key, uint32_t ncell, uint32_t ncx,
                             uint32_t ncy, int R, uint32_t *row_lo, uint32_t *row_hi)
{
    uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    uint32_t W = 2 * R + 1;
    if (t >= ncell * W) return;
    uint32_t c = t / W;
    int d = (int)(t % W) - R;
    uint32_t key = cell_key[c];
    int64_t cx = key / ncy, cy = key % ncy;
    int64_t rx = cx + d;
    uint32_t lo = 0, hi = 0;
    if (rx >= 0 && rx < (int64_t)ncx) {
        int64_t y0 = cy - R < 0 ? 0 : cy - R;
        int64_t y1 = cy + R > (int64_t)ncy - 1 ? (int64_t)ncy - 1 : cy + R;
        uint32_t k0 = (uint32_t)(rx * ncy + y0);
        uint32_t k1 = (uint32_t)(rx * ncy + y1);
        lo = lower_bound_u32(cell_key, ncell, k0);
        hi = lower_bound_u32(cell_key, ncell, k1 + 1);
        if (k1 == 0xffffffffu) hi = ncell;
    }
    row_lo[t] = lo;
    row_hi[t] = hi;
}

std::string slot(const char *tag, const char *name) { return std::string(tag) + "." + name; }

} // namespace

int build_rank_order(Context &c, Points &P, const char *tag)
{
    uint32_t n = P.n;
    double *ext = nullptr;
    VDB_TRY(c.get(slot(tag, "ext").c_str(), 4, &ext));
    VDB_TRY(extent_f64(c, P.x, P.y, n, ext));
    void *h = nullptr;
    VDB_TRY(c.host_scratch(4 * sizeof(double), &h));
    VDB_HIP(hipMemcpyAsync(h, ext, 4 * sizeof(double), hipMemcpyDeviceToHost, c.stream()));
    VDB_HIP(hipStreamSynchronize(c.stream()));
    memcpy(P.ext, h, 4 * sizeof(double));
    for (int i = 0; i < 4; i++)
        if (!std::isfinite(P.ext[i])) {
            set_error("non-finite coordinate in input");
            return kErrNonFinite;
        }
    // src/main.cpp:465-492: min from DBL_MAX, max from DBL_MIN (smallest
    // positive double), width = (max - min) * 1.01
    double min_x = P.ext[0] < DBL_MAX ? P.ext[0] : DBL_MAX;
    double max_x = P.ext[1] > DBL_MIN ? P.ext[1] : DBL_MIN;
    double min_y = P.ext[2] < DBL_MAX ? P.ext[2] : DBL_MAX;
    double max_y = P.ext[3] > DBL_MIN ? P.ext[3] : DBL_MIN;
    BinParams bp;
    bp.xw = (max_x - min_x) * 1.01;
    bp.yw = (max_y - min_y) * 1.01;
    bp.xoff = (min_x > 0) ? -min_x : std::fabs(min_x); // :513-527
    bp.yoff = (min_y > 0) ? -min_y : std::fabs(min_y);
    // zero width (all coordinates equal and positive) divides 0 by 0 in the
    // reference, which then aborts binning and reads past its mapping; here the
    // whole dimension maps to bin 0
    bp.xdeg = !(bp.xw > 0.0);
    bp.ydeg = !(bp.yw > 0.0);

    uint32_t *keys, *vals, *tk, *tv, *bad;
    VDB_TRY(c.get(slot(tag, "bk").c_str(), n, &keys));
    VDB_TRY(c.get(slot(tag, "bv").c_str(), n, &vals));
    VDB_TRY(c.get("sort.tk", n, &tk));
    VDB_TRY(c.get("sort.tv", n, &tv));
    VDB_TRY(c.get(slot(tag, "bad").c_str(), 1, &bad));
    VDB_HIP(hipMemsetAsync(bad, 0, sizeof(uint32_t), c.stream()));
    unsigned g = ceil_div(n, kBlock);
    k_bin_keys<<<g, kBlock, 0, c.stream()>>>(P.x, P.y, n, bp, keys, vals, bad);
    VDB_HIP(hipGetLastError());
    uint32_t *sk, *sv;
    VDB_TRY(radix_sort_pairs(c, keys, vals, tk, tv, n, 17, &sk, &sv));
    VDB_TRY(c.get(slot(tag, "mapping").c_str(), n, &P.mapping));
    VDB_HIP(hipMemcpyAsync(P.mapping, sv, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice,
                           c.stream()));
    VDB_TRY(c.get(slot(tag, "xb").c_str(), n, &P.xb));
    VDB_TRY(c.get(slot(tag, "yb").c_str(), n, &P.yb));
    k_gather_xy<<<g, kBlock, 0, c.stream()>>>(P.x, P.y, P.mapping, n, P.xb, P.yb);
    VDB_HIP(hipGetLastError());
    uint32_t hb = 0;
    VDB_HIP(hipMemcpyAsync(h, bad, sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream()));
    VDB_HIP(hipStreamSynchronize(c.stream()));
    memcpy(&hb, h, sizeof(uint32_t));
    if (hb) {
        set_error("bins out of range (coordinate arithmetic overflowed)");
        return kErrUnsupported;
    }
    return kOk;
}

int build_grid(Context &c, const Points &P, double w_min, int R, const char *tag, Grid &G,
               bool *enlarged)
{
    uint32_t n = P.n;
    double spanx = P.ext[1] - P.ext[0], spany = P.ext[3] - P.ext[2];
    double w = w_min;
    *enlarged = false;
    if (!(w > 0.0) || !std::isfinite(w)) {
        w = std::fmax(std::fmax(spanx, spany), 1.0) * 2.0;
        *enlarged = true;
    }
    // keep both axes below 2^16 cells so keys fit 32 bits
    double wmin_keys = std::fmax(spanx, spany) / 65000.0;
    if (w < wmin_keys) {
        w = wmin_keys;
        *enlarged = true;
    }
    G.w = w;
    G.gx0 = P.ext[0];
    G.gy0 = P.ext[2];
    G.ncx = (uint32_t)std::floor(spanx / w) + 2;
    G.ncy = (uint32_t)std::floor(spany / w) + 2;
    G.R = R;
    G.npts = n;
    uint64_t nkeys = (uint64_t)G.ncx * G.ncy;

    uint32_t *keys, *vals, *tk, *tv, *flags, *excl;
    VDB_TRY(c.get(slot(tag, "gk").c_str(), n, &keys));
    VDB_TRY(c.get(slot(tag, "gv").c_str(), n, &vals));
    VDB_TRY(c.get("sort.tk", n, &tk));
    VDB_TRY(c.get("sort.tv", n, &tv));
    unsigned g = ceil_div(n, kBlock);
    k_cell_keys<<<g, kBlock, 0, c.stream()>>>(P.xb, P.yb, n, G.gx0, G.gy0, w, G.ncx, G.ncy,
                                              keys, vals);
    VDB_HIP(hipGetLastError());
    uint32_t *sk, *sv;
    VDB_TRY(radix_sort_pairs(c, keys, vals, tk, tv, n, bits_for(nkeys), &sk, &sv));
    VDB_TRY(c.get(slot(tag, "pos_rank").c_str(), n, &G.pos_rank));
    VDB_HIP(hipMemcpyAsync(G.pos_rank, sv, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice,
                           c.stream()));
    uint32_t *skeys;
    VDB_TRY(c.get(slot(tag, "skeys").c_str(), n, &skeys));
    VDB_HIP(hipMemcpyAsync(skeys, sk, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice,
                           c.stream()));
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
    k_cells<<<g, kBlock, 0, c.stream()>>>(skeys, flags, excl, n, G.cell_of_pos, G.cell_key,
                                          G.cell_start, G.pos_rank, G.rank_pos);
    VDB_HIP(hipGetLastError());
    k_cell_tail<<<1, 64, 0, c.stream()>>>(G.cell_start, ncell, n);
    VDB_HIP(hipGetLastError());

    uint32_t *maxpop;
    VDB_TRY(c.get(slot(tag, "maxpop").c_str(), 1, &maxpop));
    VDB_HIP(hipMemsetAsync(maxpop, 0, sizeof(uint32_t), c.stream()));
    VDB_TRY(c.get(slot(tag, "cell_box").c_str(), (size_t)ncell * 4, &G.cell_box));
    k_cell_box<<<ceil_div(ncell, kBlock / kWave), kBlock, 0, c.stream()>>>(
        G.xs, G.ys, G.cell_start, ncell, G.cell_box);
    VDB_HIP(hipGetLastError());
    k_max_pop<<<std::min<unsigned>(ceil_div(ncell, kBlock), 512), kBlock, 0, c.stream()>>>(
        G.cell_start, ncell, maxpop);
    VDB_HIP(hipGetLastError());

    size_t W = 2 * (size_t)R + 1;
    VDB_TRY(c.get(slot(tag, "row_lo").c_str(), ncell * W, &G.row_lo));
    VDB_TRY(c.get(slot(tag, "row_hi").c_str(), ncell * W, &G.row_hi));
    k_row_ranges<<<ceil_div(ncell * W, kBlock), kBlock, 0, c.stream()>>>(
        G.cell_key, ncell, G.ncx, G.ncy, R, G.row_lo, G.row_hi);
    VDB_HIP(hipGetLastError());

    void *h = nullptr;
    VDB_TRY(c.host_scratch(sizeof(uint32_t), &h));
    VDB_HIP(hipMemcpyAsync(h, maxpop, sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream()));
    VDB_HIP(hipStreamSynchronize(c.stream()));
    memcpy(&G.max_cell, h, sizeof(uint32_t));
    return kOk;
}

} // namespace vdb
