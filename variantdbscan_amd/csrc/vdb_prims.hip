// Device-wide primitives: scans, stable LSD radix sort, extent reduction.
#include "vdb_prims.h"

#include <cfloat>

namespace vdb {

namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems; // 2048

template <typename T> __device__ __forceinline__ T wave_incl_scan(T x)
{
    int lane = lane_id();
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        T y = __shfl_up(x, o, kWave);
        if (lane >= o) x += y;
    }
    return x;
}

// exclusive scan across the block; sm must hold NW+1 elements
template <typename T, int NW>
__device__ __forceinline__ T block_excl_scan(T v, T *sm, T *block_total)
{
    int lane = lane_id(), w = threadIdx.x / kWave;
    T inc = wave_incl_scan(v);
    if (lane == kWave - 1) sm[w] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        T run = 0;
        for (int i = 0; i < NW; i++) {
            T t = sm[i];
            sm[i] = run;
            run += t;
        }
        sm[NW] = run;
    }
    __syncthreads();
    T r = sm[w] + inc - v;
    *block_total = sm[NW];
    __syncthreads();
    return r;
}

// Scan inputs: a plain array, or the population counts of an array of
// 32-bit words (the scan then numbers the set bits).
struct PopcSrc {
    const uint32_t *w;
    __device__ uint32_t operator[](size_t i) const { return (uint32_t)__popc(w[i]); }
};

template <typename Src, typename Out>
__global__ void __launch_bounds__(kScanBlock) k_tile_sums(Src in, size_t n, Out *sums)
{
    __shared__ Out sm[kScanBlock / kWave + 1];
    size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    Out s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; i++)
        if (base + i < n) s += (Out)in[base + i];
    Out tot;
    block_excl_scan<Out, kScanBlock / kWave>(s, sm, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// tile_off == nullptr: single tile, writes total
template <typename Src, typename Out>
__global__ void __launch_bounds__(kScanBlock)
    k_tile_scan(Src in, Out *out, size_t n, const Out *tile_off, Out *total)
{
    __shared__ Out sm[kScanBlock / kWave + 1];
    size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    Out v[kScanItems];
    Out s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        v[i] = (base + i < n) ? (Out)in[base + i] : (Out)0;
        s += v[i];
    }
    Out tot;
    Out run = block_excl_scan<Out, kScanBlock / kWave>(s, sm, &tot);
    if (tile_off) run += tile_off[blockIdx.x];
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        if (base + i < n) out[base + i] = run;
        run += v[i];
    }
    if (!tile_off && total && threadIdx.x == 0) *total = tot;
}

// Up to kScanFewTiles tiles, the second pass needs no scan of the tile sums:
// each block adds up the sums of the tiles before it (at most that many
// loads), and the last block writes the total.
constexpr unsigned kScanFewTiles = 1024;
template <typename Src, typename Out>
__global__ void __launch_bounds__(kScanBlock)
    k_tile_scan_few(Src in, Out *out, size_t n, const Out *sums, Out *total)
{
    __shared__ Out sm[kScanBlock / kWave + 1];
    __shared__ Out red[kScanBlock / kWave];
    Out part = 0;
    for (unsigned t = threadIdx.x; t < blockIdx.x; t += kScanBlock) part += sums[t];
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) part += __shfl_xor(part, o, kWave);
    if (lane_id() == 0) red[threadIdx.x / kWave] = part;
    __syncthreads();
    Out off = 0;
#pragma unroll
    for (int w = 0; w < kScanBlock / kWave; w++) off += red[w];
    size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    Out v[kScanItems];
    Out s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        v[i] = (base + i < n) ? (Out)in[base + i] : (Out)0;
        s += v[i];
    }
    Out tot;
    Out run = block_excl_scan<Out, kScanBlock / kWave>(s, sm, &tot) + off;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        if (base + i < n) out[base + i] = run;
        run += v[i];
    }
    if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *total = off + tot;
}

template <typename Src, typename Out>
int scan_impl(Context &c, Src in, Out *out, size_t n, Out *total_dev, int level)
{
    if (n == 0) {
        if (total_dev) VDB_HIP(hipMemsetAsync(total_dev, 0, sizeof(Out), c.stream()));
        return kOk;
    }
    if (n <= (size_t)kScanTile) {
        k_tile_scan<Src, Out><<<1, kScanBlock, 0, c.stream()>>>(in, out, n, nullptr, total_dev);
        VDB_HIP(hipGetLastError());
        return kOk;
    }
    size_t tiles = (n + kScanTile - 1) / kScanTile;
    char na[32], nb[32];
    snprintf(na, sizeof na, "scan.L%d.a.%d", level, (int)sizeof(Out));
    snprintf(nb, sizeof nb, "scan.L%d.b.%d", level, (int)sizeof(Out));
    Out *sums = nullptr, *offs = nullptr;
    VDB_TRY(c.get(na, tiles, &sums));
    k_tile_sums<Src, Out><<<(unsigned)tiles, kScanBlock, 0, c.stream()>>>(in, n, sums);
    VDB_HIP(hipGetLastError());
    if (tiles <= kScanFewTiles) {
        k_tile_scan_few<Src, Out><<<(unsigned)tiles, kScanBlock, 0, c.stream()>>>(in, out, n, sums,
                                                                                  total_dev);
        VDB_HIP(hipGetLastError());
        return kOk;
    }
    VDB_TRY(c.get(nb, tiles, &offs));
    VDB_TRY((scan_impl<const Out *, Out>(c, sums, offs, tiles, total_dev, level + 1)));
    k_tile_scan<Src, Out><<<(unsigned)tiles, kScanBlock, 0, c.stream()>>>(in, out, n, offs, nullptr);
    VDB_HIP(hipGetLastError());
    return kOk;
}

// ------------------------------------------------------------ radix sort
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;
constexpr int kSortBlock = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortBlock * kSortItems; // 4096
constexpr int kSortWaves = kSortBlock / kWave;

__global__ void __launch_bounds__(kSortBlock)
    k_radix_hist(const uint32_t *keys, size_t n, int shift, uint32_t *hist, unsigned tiles)
{
    __shared__ uint32_t h[kSortWaves][kRadix];
    for (int i = threadIdx.x; i < kSortWaves * kRadix; i += kSortBlock) (&h[0][0])[i] = 0;
    __syncthreads();
    int w = threadIdx.x / kWave;
    size_t base = (size_t)blockIdx.x * kSortTile;
    for (int i = 0; i < kSortItems; i++) {
        size_t idx = base + (size_t)i * kSortBlock + threadIdx.x;
        if (idx < n) atomicAdd(&h[w][(keys[idx] >> shift) & (kRadix - 1)], 1u);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < kRadix; d += kSortBlock) {
        uint32_t s = 0;
        for (int q = 0; q < kSortWaves; q++) s += h[q][d];
        hist[(size_t)d * tiles + blockIdx.x] = s; // digit-major for the scan
    }
}

__global__ void __launch_bounds__(kSortBlock)
    k_radix_scatter(const uint32_t *keys, const uint32_t *vals, uint32_t *okeys,
                    uint32_t *ovals, size_t n, int shift, const uint32_t *offs,
                    unsigned tiles)
{
    __shared__ uint32_t base_d[kRadix];
    __shared__ uint32_t wcnt[kSortWaves][kRadix];
    for (int d = threadIdx.x; d < kRadix; d += kSortBlock)
        base_d[d] = offs[(size_t)d * tiles + blockIdx.x];
    for (int i = threadIdx.x; i < kSortWaves * kRadix; i += kSortBlock) (&wcnt[0][0])[i] = 0;
    __syncthreads();
    int w = threadIdx.x / kWave;
    size_t tbase = (size_t)blockIdx.x * kSortTile;
    for (int it = 0; it < kSortItems; it++) {
        size_t idx = tbase + (size_t)it * kSortBlock + threadIdx.x;
        bool valid = idx < n;
        uint32_t k = valid ? keys[idx] : 0u;
        uint32_t v = valid ? vals[idx] : 0u;
        // digit 256 marks an invalid slot so it never matches a real digit
        uint32_t d = valid ? ((k >> shift) & (kRadix - 1)) : (uint32_t)kRadix;
        uint64_t peers = ~0ull;
#pragma unroll
        for (int b = 0; b <= kRadixBits; b++) {
            bool bit = (d >> b) & 1u;
            uint64_t bal = ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        int rank_w = mask_rank(peers);
        if (valid && rank_w == 0) wcnt[w][d] = (uint32_t)popc64(peers);
        __syncthreads();
        if (valid) {
            uint32_t pre = 0;
            for (int q = 0; q < w; q++) pre += wcnt[q][d];
            uint32_t dst = base_d[d] + pre + (uint32_t)rank_w;
            okeys[dst] = k;
            ovals[dst] = v;
        }
        __syncthreads();
        for (int dd = threadIdx.x; dd < kRadix; dd += kSortBlock) {
            uint32_t s = 0;
            for (int q = 0; q < kSortWaves; q++) {
                s += wcnt[q][dd];
                wcnt[q][dd] = 0;
            }
            base_d[dd] += s;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- extent
constexpr int kExtBlock = 256;

__global__ void __launch_bounds__(kExtBlock)
    k_extent_partial(const double *x, const double *y, size_t n, double *part)
{
    double mnx = DBL_MAX, mxx = -DBL_MAX, mny = DBL_MAX, mxy = -DBL_MAX;
    for (size_t i = (size_t)blockIdx.x * kExtBlock + threadIdx.x; i < n;
         i += (size_t)gridDim.x * kExtBlock) {
        double a = x[i], b = y[i];
        mnx = a < mnx ? a : mnx;
        mxx = a > mxx ? a : mxx;
        mny = b < mny ? b : mny;
        mxy = b > mxy ? b : mxy;
    }
    for (int o = kWave / 2; o > 0; o >>= 1) {
        double t;
        t = __shfl_xor(mnx, o); mnx = t < mnx ? t : mnx;
        t = __shfl_xor(mxx, o); mxx = t > mxx ? t : mxx;
        t = __shfl_xor(mny, o); mny = t < mny ? t : mny;
        t = __shfl_xor(mxy, o); mxy = t > mxy ? t : mxy;
    }
    __shared__ double sm[kExtBlock / kWave][4];
    int w = threadIdx.x / kWave;
    if (lane_id() == 0) {
        sm[w][0] = mnx; sm[w][1] = mxx; sm[w][2] = mny; sm[w][3] = mxy;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < kExtBlock / kWave; q++) {
            mnx = sm[q][0] < mnx ? sm[q][0] : mnx;
            mxx = sm[q][1] > mxx ? sm[q][1] : mxx;
            mny = sm[q][2] < mny ? sm[q][2] : mny;
            mxy = sm[q][3] > mxy ? sm[q][3] : mxy;
        }
        double *o = part + 4 * (size_t)blockIdx.x;
        o[0] = mnx; o[1] = mxx; o[2] = mny; o[3] = mxy;
    }
}

__global__ void k_extent_final(const double *part, int nparts, double *ext)
{
    if (threadIdx.x != 0) return;
    double mnx = DBL_MAX, mxx = -DBL_MAX, mny = DBL_MAX, mxy = -DBL_MAX;
    for (int i = 0; i < nparts; i++) {
        const double *p = part + 4 * (size_t)i;
        mnx = p[0] < mnx ? p[0] : mnx;
        mxx = p[1] > mxx ? p[1] : mxx;
        mny = p[2] < mny ? p[2] : mny;
        mxy = p[3] > mxy ? p[3] : mxy;
    }
    ext[0] = mnx; ext[1] = mxx; ext[2] = mny; ext[3] = mxy;
}

} // namespace

int bits_for(uint64_t m)
{
    int b = 0;
    while (b < 64 && ((uint64_t)1 << b) < m) b++;
    return b;
}

int scan_exclusive_u32(Context &c, const uint32_t *in, uint32_t *out, size_t n,
                       uint32_t *total_dev)
{
    return scan_impl<const uint32_t *, uint32_t>(c, in, out, n, total_dev, 0);
}

int scan_popc_u32(Context &c, const uint32_t *words, uint32_t *out, size_t n)
{
    return scan_impl<PopcSrc, uint32_t>(c, PopcSrc{words}, out, n, nullptr, 0);
}

int scan_exclusive_u32_u64(Context &c, const uint32_t *in, uint64_t *out, size_t n,
                           uint64_t *total_dev)
{
    return scan_impl<const uint32_t *, uint64_t>(c, in, out, n, total_dev, 0);
}

int compact_count(Context &c, const uint32_t *flags, uint32_t *out, size_t n,
                  uint32_t *total_host)
{
    uint32_t *tot = nullptr;
    VDB_TRY(c.get("compact.total", 1, &tot));
    VDB_TRY(scan_exclusive_u32(c, flags, out, n, tot));
    void *h = nullptr;
    VDB_TRY(c.host_scratch(sizeof(uint32_t), &h));
    VDB_HIP(hipMemcpyAsync(h, tot, sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream()));
    VDB_TRY(c.wait(c.stream()));
    *total_host = *(uint32_t *)h;
    return kOk;
}

int radix_sort_pairs(Context &c, uint32_t *keys, uint32_t *vals, uint32_t *tmp_keys,
                     uint32_t *tmp_vals, size_t n, int key_bits, uint32_t **keys_out,
                     uint32_t **vals_out)
{
    uint32_t *ka = keys, *va = vals, *kb = tmp_keys, *vb = tmp_vals;
    if (n > 1 && key_bits > 0) {
        unsigned tiles = ceil_div(n, kSortTile);
        uint32_t *hist = nullptr, *offs = nullptr;
        VDB_TRY(c.get("radix.hist", (size_t)tiles * kRadix, &hist));
        VDB_TRY(c.get("radix.offs", (size_t)tiles * kRadix, &offs));
        for (int shift = 0; shift < key_bits; shift += kRadixBits) {
            k_radix_hist<<<tiles, kSortBlock, 0, c.stream()>>>(ka, n, shift, hist, tiles);
            VDB_HIP(hipGetLastError());
            VDB_TRY(scan_exclusive_u32(c, hist, offs, (size_t)tiles * kRadix, nullptr));
            k_radix_scatter<<<tiles, kSortBlock, 0, c.stream()>>>(ka, va, kb, vb, n, shift,
                                                                  offs, tiles);
            VDB_HIP(hipGetLastError());
            uint32_t *t = ka; ka = kb; kb = t;
            t = va; va = vb; vb = t;
        }
    }
    *keys_out = ka;
    *vals_out = va;
    return kOk;
}


int extent_f64(Context &c, const double *x, const double *y, size_t n, double *ext_dev)
{
    int blocks = (int)std::min<size_t>(1024, (n + kExtBlock - 1) / kExtBlock);
    if (blocks < 1) blocks = 1;
    double *part = nullptr;
    VDB_TRY(c.get("extent.part", (size_t)blocks * 4, &part));
    k_extent_partial<<<blocks, kExtBlock, 0, c.stream()>>>(x, y, n, part);
    VDB_HIP(hipGetLastError());
    k_extent_final<<<1, 64, 0, c.stream()>>>(part, blocks, ext_dev);
    VDB_HIP(hipGetLastError());
    return kOk;
}

} // namespace vdb
