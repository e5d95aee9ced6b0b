// Per-variant neighbour counting, neighbour table and clustering kernels.
//
// Work decomposition: a wavefront takes one item = up to kOrigins consecutive
// origin points of one grid cell. The 64 lanes sweep the candidate points of
// the cell's 3x3 window (three contiguous runs of the cell-sorted arrays, one
// per grid row), 64 candidates per step, and every lane tests its candidate
// against all origins of the item; the origin coordinates are wave-uniform.
// A ballot per origin gives the accepted lanes, so counts are popcounts and
// table rows are written as contiguous runs (coalesced stores).
#include "vdb_cluster.h"
#include "vdb_prims.h"

#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

namespace vdb {

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr uint32_t kNone = 0xffffffffu;
constexpr int kDefaultCellsPerEps = 2;

struct GridView {
    const double *xs, *ys;
    const uint32_t *cell_start, *row_lo, *row_hi, *row_f0, *row_f1;
    int W;
    uint32_t n;
};

GridView view(const Grid &G, uint32_t n)
{
    return GridView{G.xs,     G.ys,     G.cell_start,   G.row_lo, G.row_hi,
                    G.row_f0, G.row_f1, 2 * G.R + 1,    n};
}

// Row r of cell c's window as cell-sorted positions: candidates to test in
// [a, f0) and [f1, e), candidates accepted without a test in [f0, f1).
struct RowSeg {
    uint32_t a, f0, f1, e;
};

__device__ __forceinline__ RowSeg row_seg(const GridView &g, uint32_t c, int r)
{
    size_t t = (size_t)c * g.W + r;
    uint32_t lo = g.row_lo[t], hi = g.row_hi[t];
    RowSeg s;
    s.a = g.cell_start[lo];
    s.e = lo < hi ? g.cell_start[hi] : s.a;
    s.f0 = s.f1 = s.e;
    if (g.row_f0 && lo < hi) {
        s.f0 = g.cell_start[g.row_f0[t]];
        s.f1 = g.cell_start[g.row_f1[t]];
    }
    return s;
}

// 64-candidate blocks of positions [a, e): lane l holds a + 64*i + l; lanes
// past e see NaN coordinates (never accepted).
template <class F>
__device__ __forceinline__ void run_blocks(const GridView &g, uint32_t a, uint32_t e, F &&f)
{
    for (uint32_t q0 = a; q0 < e; q0 += kWave) {
        uint32_t q = q0 + lane_id();
        bool v = q < e;
        uint32_t qc = v ? q : e - 1;
        double qx = g.xs[qc], qy = g.ys[qc];
        qx = v ? qx : __longlong_as_double(0x7ff8000000000000ll);
        qy = v ? qy : __longlong_as_double(0x7ff8000000000000ll);
        f(q, v, qx, qy);
    }
}

// Visit the window of cell c in 64-candidate blocks. With `half`, only rows
// below the centre row and the centre row up to (excluding) qmax are visited:
// every unordered pair inside the window is then seen from one side.
template <class F>
__device__ __forceinline__ void window_blocks(const GridView &g, uint32_t c, bool half,
                                              uint32_t qmax, F &&f)
{
    int R = g.W / 2;
    for (int r = 0; r < g.W; r++) {
        int d = r - R;
        if (half && d > 0) break;
        uint32_t lo = g.row_lo[(size_t)c * g.W + r], hi = g.row_hi[(size_t)c * g.W + r];
        if (lo >= hi) continue;
        uint32_t a = g.cell_start[lo], e = g.cell_start[hi];
        if (half && d == 0 && qmax < e) e = qmax;
        for (uint32_t q0 = a; q0 < e; q0 += kWave) {
            uint32_t q = q0 + lane_id();
            bool v = q < e;
            // clamp instead of branching; lanes past the run see NaN (never accepted)
            uint32_t qc = v ? q : e - 1;
            double qx = g.xs[qc], qy = g.ys[qc];
            qx = v ? qx : __longlong_as_double(0x7ff8000000000000ll);
            qy = v ? qy : __longlong_as_double(0x7ff8000000000000ll);
            f(q, v, qx, qy);
        }
    }
}

// Chunks of up to kChunk 64-candidate blocks of positions [a, e) held in
// registers (lane l holds candidates q0 + i*64 + l), for consumers that finish
// one origin over a whole chunk before the next.
constexpr int kChunk = 2;
template <class F>
__device__ __forceinline__ void run_chunks(const GridView &g, uint32_t a, uint32_t e, F &&f)
{
    const double qnan = __longlong_as_double(0x7ff8000000000000ll);
    for (uint32_t q0 = a; q0 < e; q0 += kWave * kChunk) {
        int nblk = (int)((e - q0 + kWave - 1) / kWave);
        nblk = nblk < kChunk ? nblk : kChunk;
        double qx[kChunk], qy[kChunk];
#pragma unroll
        for (int i = 0; i < kChunk; i++) {
            uint32_t q = q0 + i * kWave + lane_id();
            bool v = i < nblk && q < e;
            uint32_t qc = v ? q : e - 1;
            double x = g.xs[qc], y = g.ys[qc];
            qx[i] = v ? x : qnan;
            qy[i] = v ? y : qnan;
        }
        f(q0, nblk, qx, qy);
    }
}

// FP32 screening with an exact FP64 decision inside the band. Coordinates
// are taken relative to the item's first origin r: origins lie within w of r
// and candidates within (R+1)w per axis, so converting a difference to
// float32 and evaluating dx*dx+dy*dy in float32 stays within band of the
// binary64 value s (band_for()).
struct Band {
    float lo, hi; // accept if s32 <= lo, reject if s32 > hi, else recheck
};

__device__ __forceinline__ float dist2_f32(float ox, float oy, float qx, float qy)
{
    float dx = __fsub_rn(ox, qx);
    float dy = __fsub_rn(oy, qy);
    return __fmaf_rn(dx, dx, __fmul_rn(dy, dy));
}

// Wave-uniform values pinned to scalar registers, so the 8 origins of an
// item cost no vector registers.
__device__ __forceinline__ double sgpr_d(double v)
{
    uint64_t u = __double_as_longlong(v);
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)u);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(u >> 32));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
__device__ __forceinline__ uint64_t sgpr_u64(uint64_t v)
{
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint32_t sgpr_u32(uint32_t v)
{
    return __builtin_amdgcn_readfirstlane(v);
}

// ----------------------------------------------------------------- items
__global__ void k_item_counts(const uint32_t *cell_start, uint32_t ncell, uint32_t *cnt)
{
    uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= ncell) return;
    uint32_t n = cell_start[c + 1] - cell_start[c];
    cnt[c] = (n + kOrigins - 1) / kOrigins;
}

__global__ void k_item_fill(const uint32_t *cell_start, const uint32_t *excl, uint32_t ncell,
                            uint32_t *item_cell, uint32_t *item_o0)
{
    uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= ncell) return;
    uint32_t a = cell_start[c], e = cell_start[c + 1], k = excl[c];
    for (uint32_t o = a; o < e; o += kOrigins, k++) {
        item_cell[k] = c;
        item_o0[k] = o;
    }
}

// ------------------------------------------------------------- counting
template <int PREC>
__global__ void __launch_bounds__(kBlock)
    k_count(GridView g, const uint32_t *__restrict__ item_cell,
            const uint32_t *__restrict__ item_o0, uint32_t nitems, double T, Band band,
            uint32_t *__restrict__ counts, unsigned long long *recheck)
{
    uint32_t it = blockIdx.x * kWavesPerBlock + wave_id_in_block();
    if (it >= nitems) return;
    uint32_t c = sgpr_u32(item_cell[it]), o0 = sgpr_u32(item_o0[it]);
    uint32_t nb = sgpr_u32(g.cell_start[c + 1]) - o0;
    nb = nb < kOrigins ? nb : kOrigins;
    const double qnan = __longlong_as_double(0x7ff8000000000000ll);
    double ox[kOrigins], oy[kOrigins];
    float fx[kOrigins], fy[kOrigins];
    uint32_t cnt[kOrigins];
    double rx = sgpr_d(g.xs[o0]), ry = sgpr_d(g.ys[o0]);
#pragma unroll
    for (int b = 0; b < kOrigins; b++) {
        uint32_t ob = o0 + (b < (int)nb ? b : 0);
        // padded origins are NaN: they accept nothing
        ox[b] = b < (int)nb ? sgpr_d(g.xs[ob]) : qnan;
        oy[b] = sgpr_d(g.ys[ob]);
        fx[b] = (float)__dsub_rn(ox[b], rx);
        fy[b] = (float)__dsub_rn(oy[b], ry);
        cnt[b] = 0;
    }
    uint32_t nre = 0, full = 0;
    auto test = [&](uint32_t q, bool v, double qx, double qy) {
        if (PREC == 0) {
#pragma unroll
            for (int b = 0; b < kOrigins; b++)
                cnt[b] += popc64(ballot(dist2_exact(ox[b], oy[b], qx, qy) <= T));
        } else {
            float gx = (float)__dsub_rn(qx, rx), gy = (float)__dsub_rn(qy, ry);
#pragma unroll
            for (int b = 0; b < kOrigins; b++) {
                float s = dist2_f32(fx[b], fy[b], gx, gy);
                bool acc = s <= band.lo;
                uint64_t amb = ballot(!acc && s <= band.hi);
                if (amb) {
                    if ((amb >> lane_id()) & 1ull) acc = dist2_exact(ox[b], oy[b], qx, qy) <= T;
                    nre += popc64(amb);
                }
                cnt[b] += popc64(ballot(acc));
            }
        }
    };
    for (int r = 0; r < g.W; r++) {
        RowSeg sg = row_seg(g, c, r);
        run_blocks(g, sg.a, sg.f0, test);
        full += sg.f1 - sg.f0;
        run_blocks(g, sg.f1, sg.e, test);
    }
    uint32_t mine = 0;
#pragma unroll
    for (int b = 0; b < kOrigins; b++)
        if (lane_id() == b) mine = cnt[b] + full;
    if (lane_id() < (int)nb) counts[o0 + lane_id()] = mine;
    if (PREC == 1 && nre && lane_id() == 0) atomicAdd(recheck, (unsigned long long)nre);
}

// ----------------------------------------------------------------- table
// ids[row_off[p] + k] = idmap[q] (or q when idmap is null) for the k-th
// accepted candidate q of origin p; rows located through rowmap.
// Order-independent digest of a neighbour table: the sum over all (row, id)
// pairs of a 64-bit mix of row*2^32+id (splitmix64 finalizer), mod 2^64.
__device__ __forceinline__ uint64_t pair_mix(uint64_t row, uint64_t id)
{
    uint64_t z = (row << 32) | id;
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

template <int PREC, bool D2, bool DIGEST = false>
__global__ void __launch_bounds__(kBlock)
    k_fill(GridView g, const uint32_t *__restrict__ item_cell,
           const uint32_t *__restrict__ item_o0, uint32_t it0, uint32_t it1, double T,
           Band band, const uint64_t *__restrict__ offsets, uint64_t off_base,
           const uint32_t *__restrict__ rowmap, const uint32_t *__restrict__ idmap,
           uint32_t *__restrict__ ids, float *__restrict__ d2, uint64_t cap,
           uint32_t *overflow)
{
    // block-uniform: do the block's 4 items exist and share one cell?
    uint32_t ib = it0 + blockIdx.x * kWavesPerBlock;
    bool shared = ib + kWavesPerBlock <= it1 &&
                  item_cell[ib] == item_cell[ib + kWavesPerBlock - 1];
    uint32_t it = ib + wave_id_in_block();
    if (!shared && it >= it1) return;
    uint32_t c = sgpr_u32(item_cell[it]), o0 = sgpr_u32(item_o0[it]);
    uint32_t nb = sgpr_u32(g.cell_start[c + 1]) - o0;
    nb = nb < kOrigins ? nb : kOrigins;
    const double qnan = __longlong_as_double(0x7ff8000000000000ll);
    double ox[kOrigins], oy[kOrigins];
    float fx[kOrigins], fy[kOrigins];
    uint64_t run[kOrigins], lim[kOrigins];
    uint32_t rowid[kOrigins];
    double rx = sgpr_d(g.xs[o0]), ry = sgpr_d(g.ys[o0]);
    // FP64 tests read the origins for every candidate: they live in vector
    // registers (broadcast from one lane-parallel load), since the scalar
    // file cannot hold them beside the rows' write state without spilling
    double xl = 0.0, yl = 0.0;
    if (PREC == 0) {
        uint32_t ol = o0 + ((uint32_t)lane_id() < nb ? (uint32_t)lane_id() : 0u);
        xl = g.xs[ol];
        yl = g.ys[ol];
    }
#pragma unroll
    for (int b = 0; b < kOrigins; b++) {
        uint32_t ob = o0 + (b < (int)nb ? b : 0);
        if (PREC == 0) {
            ox[b] = b < (int)nb ? __shfl(xl, b) : qnan;
            oy[b] = __shfl(yl, b);
        } else {
            ox[b] = b < (int)nb ? sgpr_d(g.xs[ob]) : qnan;
            oy[b] = sgpr_d(g.ys[ob]);
        }
        fx[b] = (float)__dsub_rn(ox[b], rx);
        fy[b] = (float)__dsub_rn(oy[b], ry);
        uint32_t row = rowmap ? rowmap[ob] : ob;
        rowid[b] = sgpr_u32(row);
        if (!DIGEST) {
            run[b] = sgpr_u64(offsets[row] - off_base);
            lim[b] = sgpr_u64(offsets[row + 1] - off_base);
        }
    }
    bool over = false;
    uint64_t dig = 0;
    // Row staging (ids only: no D2, no digest). Accepted ids are appended to
    // a 128-slot LDS ring per origin and leave in 64-entry segments aligned in
    // the table, so nearly every global store writes 256 aligned bytes with
    // all lanes active; appending in 88-byte pieces straight to the table
    // costs a third of the write bandwidth. Row b holds cur[b] appended ids,
    // of which fl[b] are written; bnd[b] ends the aligned segment holding id
    // fl[b], and cur[b] < bnd[b] between appends. Id j sits in slot j mod 128.
    constexpr bool STAGE = !D2 && !DIGEST;
    constexpr int kRing = 2 * kWave;
    __shared__ uint32_t ring_all[STAGE ? kWavesPerBlock * kOrigins * kRing : 1];
    uint32_t *ring = ring_all + (STAGE ? wave_id_in_block() * kOrigins * kRing : 0);
    uint32_t cur[kOrigins], fl[kOrigins], bnd[kOrigins], rlen[kOrigins];
#pragma unroll
    for (int b = 0; b < kOrigins; b++) {
        cur[b] = 0;
        fl[b] = 0;
        bnd[b] = STAGE ? kWave - (uint32_t)(run[b] & (kWave - 1)) : 0;
        // writes stop at the row end and at the end of the ids buffer (a row
        // cut short there raises the overflow flag below)
        uint64_t room = cap > run[b] ? cap - run[b] : 0;
        uint64_t len = lim[b] - run[b];
        rlen[b] = STAGE ? (uint32_t)(len < room ? len : room) : 0;
    }
    // write staged ids [fl, upto) of row b: one store into the aligned segment
    auto flush = [&](int b, uint32_t upto) {
        // the ring entries were written by other lanes of this wave; LDS
        // operations of a wave complete in order
        __builtin_amdgcn_wave_barrier();
        // entry fl sits in lane lo of its aligned segment; lanes [lo, hi)
        // take entries fl .. min(upto, rlen) - 1 (hi <= 64: upto <= bnd)
        uint64_t at = run[b] + fl[b];
        uint32_t lo = (uint32_t)at & (kWave - 1);
        uint32_t end = upto < rlen[b] ? upto : rlen[b];
        uint32_t hi = end > fl[b] ? lo + (end - fl[b]) : lo;
        uint32_t l = (uint32_t)lane_id();
        // the table is far larger than the caches and is not read back by
        // this kernel: streaming (non-temporal) stores
        if (l >= lo && l < hi)
            __builtin_nontemporal_store(ring[b * kRing + ((fl[b] - lo + l) & (kRing - 1))],
                                        &ids[(at - lo) + l]);
        fl[b] = upto;
        bnd[b] = upto + kWave;
    };
    // append the lanes of m (at most 64 ids) to row b
    auto append = [&](int b, bool acc, uint64_t m, uint32_t id) {
        if (acc) ring[b * kRing + ((cur[b] + (uint32_t)mask_rank(m)) & (kRing - 1))] = id;
        cur[b] += (uint32_t)popc64(m);
        if (cur[b] >= bnd[b]) flush(b, bnd[b]);
    };
    // origin-major over a register chunk: each origin's accepted ids of the
    // chunk go out as one contiguous run before the next origin's
    auto chunk = [&](uint32_t q0, int nblk, const double *qx, const double *qy,
                     const uint32_t *qid) {
#pragma unroll
        for (int b = 0; b < kOrigins; b++) {
#pragma unroll
            for (int i = 0; i < kChunk; i++) {
                if (i >= nblk) break;
                double s = 0.0;
                bool acc;
                if (PREC == 0) {
                    s = dist2_exact(ox[b], oy[b], qx[i], qy[i]);
                    acc = s <= T;
                } else {
                    float gx = (float)__dsub_rn(qx[i], rx), gy = (float)__dsub_rn(qy[i], ry);
                    float s32 = dist2_f32(fx[b], fy[b], gx, gy);
                    acc = s32 <= band.lo;
                    bool amb = !acc && s32 <= band.hi;
                    if (ballot(amb)) {
                        if (amb) acc = dist2_exact(ox[b], oy[b], qx[i], qy[i]) <= T;
                    }
                    if (D2) s = dist2_exact(ox[b], oy[b], qx[i], qy[i]);
                }
                if (DIGEST) {
                    if (acc) dig += pair_mix(rowid[b], qid[i]);
                    continue;
                }
                uint64_t m = ballot(acc);
                if (m) {
                    if (STAGE) {
                        append(b, acc, m, qid[i]);
                        continue;
                    }
                    uint64_t dst = run[b] + (uint64_t)mask_rank(m);
                    // the count pass sized every row; never write past it
                    bool w = acc && dst < lim[b] && dst < cap;
                    over |= acc && !w;
                    if (w) {
                        ids[dst] = qid[i];
                        if (D2) d2[dst] = (float)s;
                    }
                    run[b] += (uint64_t)popc64(m);
                }
            }
        }
    };
    // candidates accepted without a test: every origin's row gets the run of
    // ids as it stands (no distances, so only without D2)
    auto full = [&](uint32_t a, uint32_t e) {
        if (STAGE && !idmap) {
            // the ids are the positions a..e-1 themselves: top up the open
            // segment of each row through the ring, store the whole aligned
            // segments that follow straight from registers (no LDS round
            // trip), and leave the rest in the ring
#pragma unroll
            for (int b = 0; b < kOrigins; b++) {
                if (b >= (int)nb || a >= e) break;
                uint32_t q = a, left = e - a;
                uint32_t h = bnd[b] - cur[b]; // > 0 between appends
                h = left < h ? left : h;
                append(b, (uint32_t)lane_id() < h, h == (uint32_t)kWave ? ~0ull : (1ull << h) - 1,
                       q + (uint32_t)lane_id());
                q += h;
                left -= h;
                if (left >= (uint32_t)kWave) {
                    // the top-up flushed the open segment: cur = fl and
                    // run + fl is a multiple of 64
                    uint32_t nseg = left / kWave;
                    uint32_t *dst = ids + (run[b] + fl[b]);
                    for (uint32_t s = 0; s < nseg; s++) {
                        uint32_t t = fl[b] + s * kWave + (uint32_t)lane_id();
                        if (t < rlen[b])
                            __builtin_nontemporal_store(q + s * kWave + (uint32_t)lane_id(),
                                                        &dst[s * kWave + lane_id()]);
                    }
                    fl[b] += nseg * kWave;
                    cur[b] = fl[b];
                    bnd[b] = fl[b] + kWave;
                    q += nseg * kWave;
                    left -= nseg * kWave;
                }
                if (left)
                    append(b, (uint32_t)lane_id() < left, (1ull << left) - 1, q + (uint32_t)lane_id());
            }
            return;
        }
        for (uint32_t q0 = a; q0 < e; q0 += kWave * kChunk) {
            int nblk = (int)((e - q0 + kWave - 1) / kWave);
            nblk = nblk < kChunk ? nblk : kChunk;
            uint32_t qid[kChunk];
#pragma unroll
            for (int i = 0; i < kChunk; i++) {
                uint32_t q = q0 + i * kWave + lane_id();
                qid[i] = (idmap && i < nblk) ? idmap[q < e ? q : a] : q;
            }
#pragma unroll
            for (int b = 0; b < kOrigins; b++) {
                if (b >= (int)nb) break;
#pragma unroll
                for (int i = 0; i < kChunk; i++) {
                    if (i >= nblk) break;
                    uint32_t left = e - (q0 + i * kWave);
                    uint32_t cnt = left < (uint32_t)kWave ? left : (uint32_t)kWave;
                    bool v = (uint32_t)lane_id() < cnt;
                    if (DIGEST) {
                        if (v) dig += pair_mix(rowid[b], qid[i]);
                        continue;
                    }
                    if (STAGE) {
                        append(b, v, cnt == (uint32_t)kWave ? ~0ull : (1ull << cnt) - 1, qid[i]);
                        continue;
                    }
                    uint64_t dst = run[b] + (uint64_t)lane_id();
                    bool w = v && dst < lim[b] && dst < cap;
                    over |= v && !w;
                    if (w) ids[dst] = qid[i];
                    run[b] += cnt;
                }
            }
        }
    };
    if (shared) {
        // the block's 4 items share one cell: stage each chunk of the window
        // in LDS once for all 32 origins
        __shared__ double sx[kChunk * kWave], sy[kChunk * kWave];
        __shared__ uint32_t sid[kChunk * kWave];
        const double qnan2 = __longlong_as_double(0x7ff8000000000000ll);
        auto staged = [&](uint32_t a, uint32_t e) {
            for (uint32_t q0 = a; q0 < e; q0 += kWave * kChunk) {
                int nblk = (int)((e - q0 + kWave - 1) / kWave);
                nblk = nblk < kChunk ? nblk : kChunk;
                for (int j = threadIdx.x; j < kChunk * kWave; j += kBlock) {
                    uint32_t q = q0 + j;
                    bool v = q < e;
                    uint32_t qc = v ? q : e - 1;
                    double x = g.xs[qc], y = g.ys[qc];
                    sx[j] = v ? x : qnan2;
                    sy[j] = v ? y : qnan2;
                    sid[j] = idmap ? idmap[qc] : q;
                }
                __syncthreads();
                double qx[kChunk], qy[kChunk];
                uint32_t qid[kChunk];
#pragma unroll
                for (int i = 0; i < kChunk; i++) {
                    qx[i] = sx[i * kWave + lane_id()];
                    qy[i] = sy[i * kWave + lane_id()];
                    qid[i] = sid[i * kWave + lane_id()];
                }
                __syncthreads();
                chunk(q0, nblk, qx, qy, qid);
            }
        };
        for (int r = 0; r < g.W; r++) {
            RowSeg sg = row_seg(g, c, r);
            if (D2) sg.f0 = sg.f1 = sg.e;
            staged(sg.a, sg.f0);
            full(sg.f0, sg.f1);
            staged(sg.f1, sg.e);
        }
    } else {
        auto tested = [&](uint32_t q0, int nblk, const double *qx, const double *qy) {
            // ids of the chunk's candidates, loaded before any store: a load
            // between stores would make every store wait for all earlier ones
            uint32_t qid[kChunk];
#pragma unroll
            for (int i = 0; i < kChunk; i++) {
                uint32_t q = q0 + i * kWave + lane_id();
                qid[i] = (idmap && i < nblk) ? idmap[q < g.n ? q : 0] : q;
            }
            chunk(q0, nblk, qx, qy, qid);
        };
        for (int r = 0; r < g.W; r++) {
            RowSeg sg = row_seg(g, c, r);
            if (D2) sg.f0 = sg.f1 = sg.e;
            run_chunks(g, sg.a, sg.f0, tested);
            full(sg.f0, sg.f1);
            run_chunks(g, sg.f1, sg.e, tested);
        }
    }
    if (STAGE) {
#pragma unroll
        for (int b = 0; b < kOrigins; b++) {
            if (b >= (int)nb) break;
            if (cur[b] > fl[b]) flush(b, cur[b]);
            // the count pass sized every row
            over |= cur[b] != rlen[b];
        }
    }
    if (DIGEST) {
        for (int o = kWave / 2; o > 0; o >>= 1) dig += __shfl_xor(dig, o);
        if (lane_id() == 0) atomicAdd((unsigned long long *)ids, (unsigned long long)dig);
        return;
    }
    if (ballot(over) && lane_id() == 0) atomicOr(overflow, 1u);
}

// ------------------------------------------------------------ union-find
// Parents live in cell-position space; every parent pointer goes to a node
// of smaller binned rank, so a root is the smallest rank of its tree. Loads
// and stores use agent scope so that no wave spins on a stale L1 copy.
__device__ __forceinline__ uint32_t ld_par(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_par(uint32_t *p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ uint32_t uf_find(uint32_t *parent, uint32_t x)
{
    while (true) {
        uint32_t p = ld_par(parent + x);
        if (p == x) return x;
        uint32_t gp = ld_par(parent + p);
        if (gp == p) return p;
        st_par(parent + x, gp); // path halving: gp is an ancestor of x
        x = gp;
    }
}

__device__ void uf_unite(uint32_t *parent, const uint32_t *rank, uint32_t a, uint32_t b)
{
    while (true) {
        a = uf_find(parent, a);
        b = uf_find(parent, b);
        if (a == b) return;
        if (rank[a] > rank[b]) {
            uint32_t t = a;
            a = b;
            b = t;
        }
        // hook the larger-rank root under the smaller one
        if (atomicCAS(parent + b, b, a) == b) return;
    }
}

__global__ void k_core_init(const uint32_t *counts, uint32_t n, uint32_t minpts,
                            uint8_t *core, uint32_t *parent, uint32_t *best)
{
    uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    // minpts above INT_MAX wraps negative in the reference's int and no point
    // is core (src/structs.h:37, src/DBScan.cpp:552)
    bool isc = minpts <= 0x7fffffffu && counts[p] >= minpts;
    core[p] = isc ? 1 : 0;
    parent[p] = p;
    best[p] = kNone;
}

__global__ void k_fill_u32_n(uint32_t *a, uint32_t v, uint32_t n)
{
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void __launch_bounds__(kBlock)
    k_uf_fused(GridView g, const uint32_t *__restrict__ item_cell,
               const uint32_t *__restrict__ item_o0, uint32_t nitems, double T,
               const uint8_t *__restrict__ core, const uint32_t *__restrict__ rank,
               uint32_t *parent)
{
    uint32_t it = blockIdx.x * kWavesPerBlock + wave_id_in_block();
    if (it >= nitems) return;
    uint32_t c = item_cell[it], o0 = item_o0[it];
    uint32_t nb = g.cell_start[c + 1] - o0;
    nb = nb < kOrigins ? nb : kOrigins;
    double ox[kOrigins], oy[kOrigins];
    uint32_t rp[kOrigins];
    uint32_t any = 0;
#pragma unroll
    for (int b = 0; b < kOrigins; b++) {
        uint32_t ob = o0 + (b < (int)nb ? b : nb - 1);
        bool isc = b < (int)nb && core[ob];
        ox[b] = sgpr_d(isc ? g.xs[ob] : __longlong_as_double(0x7ff8000000000000ll));
        oy[b] = sgpr_d(g.ys[ob]);
        rp[b] = isc ? uf_find(parent, ob) : kNone;
        any |= isc;
    }
    if (!any) return;
    uint32_t qmax = o0 + nb;
    // the half window (rows below the centre, the centre row before the
    // item's last origin), row segments accepted without a test included
    auto blockfn = [&](uint32_t q, bool v, double qx, double qy, bool full) {
        bool qc = v && core[q];
        if (!ballot(qc)) return;
        uint32_t pq = qc ? ld_par(parent + q) : 0u;
#pragma unroll
        for (int b = 0; b < kOrigins; b++) {
            if (rp[b] == kNone) continue;
            bool e = qc && (full || dist2_exact(ox[b], oy[b], qx, qy) <= T) && pq != rp[b];
            uint64_t m = ballot(e);
            if (m) {
                if (e) uf_unite(parent, rank, q, o0 + b);
                rp[b] = uf_find(parent, o0 + b);
            }
        }
    };
    int R = g.W / 2;
    for (int r = 0; r <= R; r++) {
        RowSeg sg = row_seg(g, c, r);
        if (r == R) {
            sg.e = sg.e < qmax ? sg.e : qmax;
            sg.f1 = sg.f1 < sg.e ? sg.f1 : sg.e;
            sg.f0 = sg.f0 < sg.f1 ? sg.f0 : sg.f1;
        }
        run_blocks(g, sg.a, sg.f0, [&](uint32_t q, bool v, double qx, double qy) {
            blockfn(q, v, qx, qy, false);
        });
        for (uint32_t q0 = sg.f0; q0 < sg.f1; q0 += kWave) {
            uint32_t q = q0 + lane_id();
            blockfn(q, q < sg.f1, 0.0, 0.0, true);
        }
        run_blocks(g, sg.f1, sg.e, [&](uint32_t q, bool v, double qx, double qy) {
            blockfn(q, v, qx, qy, false);
        });
    }
}

// Root of x without path-halving stores. Used once every union is done:
// concurrent threads then only ever store true roots, so no thread can put
// back an intermediate ancestor over a root another thread just wrote (which
// a halving find could, from a stale read of its path).
__device__ uint32_t uf_root(const uint32_t *parent, uint32_t x)
{
    uint32_t p = ld_par(parent + x);
    while (p != x) {
        x = p;
        p = ld_par(parent + x);
    }
    return x;
}

__global__ void k_compress(uint32_t *parent, const uint8_t *core, uint32_t n)
{
    uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n || !core[p]) return;
    st_par(parent + p, uf_root(parent, p));
}

// border: best[p] = smallest rank[root] among core neighbours of non-core p
// (rank is the binned rank, or any per-root key the caller orders roots by)
__global__ void __launch_bounds__(kBlock)
    k_border_fused(GridView g, const uint32_t *__restrict__ item_cell,
                   const uint32_t *__restrict__ item_o0, uint32_t nitems, double T,
                   const uint8_t *__restrict__ core, const uint32_t *__restrict__ rank,
                   const uint32_t *__restrict__ parent, uint32_t *__restrict__ best)
{
    uint32_t it = blockIdx.x * kWavesPerBlock + wave_id_in_block();
    if (it >= nitems) return;
    uint32_t c = item_cell[it], o0 = item_o0[it];
    uint32_t nb = g.cell_start[c + 1] - o0;
    nb = nb < kOrigins ? nb : kOrigins;
    double ox[kOrigins], oy[kOrigins];
    uint32_t bst[kOrigins];
    uint32_t any = 0;
#pragma unroll
    for (int b = 0; b < kOrigins; b++) {
        uint32_t ob = o0 + (b < (int)nb ? b : nb - 1);
        bool nc = b < (int)nb && !core[ob];
        ox[b] = sgpr_d(nc ? g.xs[ob] : __longlong_as_double(0x7ff8000000000000ll));
        oy[b] = sgpr_d(g.ys[ob]);
        bst[b] = kNone;
        any |= nc;
    }
    if (!any) return;
    window_blocks(g, c, false, 0, [&](uint32_t q, bool v, double qx, double qy) {
        bool qc = v && core[q];
        if (!ballot(qc)) return;
        uint32_t rr = qc ? rank[parent[q]] : kNone;
#pragma unroll
        for (int b = 0; b < kOrigins; b++) {
            uint32_t cand = (qc && dist2_exact(ox[b], oy[b], qx, qy) <= T) ? rr : kNone;
            if (!ballot(cand != kNone)) continue;
            for (int o = kWave / 2; o > 0; o >>= 1) {
                uint32_t t = __shfl_xor(cand, o);
                cand = t < cand ? t : cand;
            }
            bst[b] = cand < bst[b] ? cand : bst[b];
        }
    });
    uint32_t mine = kNone;
#pragma unroll
    for (int b = 0; b < kOrigins; b++)
        if (lane_id() == b) mine = bst[b];
    if (lane_id() < (int)nb && !core[o0 + lane_id()]) best[o0 + lane_id()] = mine;
}

// (C, root) pairs for the reuse emulation: for every non-core position p with
// cl[p] = C != 0, each distinct root (position) among its core neighbours.
// The wave drops a root already emitted for the same origin (the last four
// are remembered); any repeat left is removed by the caller. At most cap
// pairs are stored; *count counts them all.
__global__ void __launch_bounds__(kBlock)
    k_border_pairs(GridView g, const uint32_t *__restrict__ item_cell,
                   const uint32_t *__restrict__ item_o0, uint32_t nitems, double T,
                   const uint8_t *__restrict__ core, const uint32_t *__restrict__ parent,
                   const uint32_t *__restrict__ cl, uint32_t *pairs, uint32_t cap, uint32_t *count)
{
    uint32_t it = blockIdx.x * kWavesPerBlock + wave_id_in_block();
    if (it >= nitems) return;
    uint32_t c = item_cell[it], o0 = item_o0[it];
    uint32_t nb = g.cell_start[c + 1] - o0;
    nb = nb < kOrigins ? nb : kOrigins;
    double ox[kOrigins], oy[kOrigins];
    uint32_t clb[kOrigins], seen[kOrigins][4];
    uint32_t any = 0;
#pragma unroll
    for (int b = 0; b < kOrigins; b++) {
        uint32_t ob = o0 + (b < (int)nb ? b : nb - 1);
        uint32_t cb = b < (int)nb && !core[ob] ? cl[ob] : 0u;
        clb[b] = cb;
        ox[b] = sgpr_d(cb ? g.xs[ob] : __longlong_as_double(0x7ff8000000000000ll));
        oy[b] = sgpr_d(g.ys[ob]);
#pragma unroll
        for (int s = 0; s < 4; s++) seen[b][s] = kNone;
        any |= cb;
    }
    if (!any) return;
    window_blocks(g, c, false, 0, [&](uint32_t q, bool v, double qx, double qy) {
        bool qc = v && core[q];
        if (!ballot(qc)) return;
        uint32_t rr = qc ? parent[q] : kNone;
#pragma unroll
        for (int b = 0; b < kOrigins; b++) {
            uint32_t cand = (qc && dist2_exact(ox[b], oy[b], qx, qy) <= T) ? rr : kNone;
#pragma unroll
            for (int s = 0; s < 4; s++) cand = cand == seen[b][s] ? kNone : cand;
            uint64_t m;
            while ((m = ballot(cand != kNone)) != 0) {
                uint32_t r = (uint32_t)__shfl((int)cand, __builtin_ctzll(m));
                if (cand == r) cand = kNone;
                if (lane_id() == 0) {
                    uint32_t at = atomicAdd(count, 1u);
                    if (at < cap) {
                        pairs[2 * (size_t)at] = clb[b];
                        pairs[2 * (size_t)at + 1] = r;
                    }
                }
                seen[b][3] = seen[b][2];
                seen[b][2] = seen[b][1];
                seen[b][1] = seen[b][0];
                seen[b][0] = r;
            }
        }
    });
}

// --------------------------------------------- border pass over a CSR table
__global__ void __launch_bounds__(kBlock)
    k_border_table(const uint64_t *__restrict__ offsets, uint64_t off_base,
                   const uint32_t *__restrict__ ids, uint32_t p0, uint32_t p1,
                   const uint8_t *__restrict__ core, const uint32_t *__restrict__ rank,
                   const uint32_t *__restrict__ parent, uint32_t *__restrict__ best)
{
    uint32_t p = p0 + blockIdx.x * kWavesPerBlock + wave_id_in_block();
    if (p >= p1 || core[p]) return;
    uint64_t a = offsets[p] - off_base, e = offsets[p + 1] - off_base;
    uint32_t bst = kNone;
    for (uint64_t j0 = a; j0 < e; j0 += kWave) {
        uint64_t j = j0 + lane_id();
        uint32_t q = j < e ? ids[j] : kNone;
        uint32_t cand = (q != kNone && core[q]) ? rank[parent[q]] : kNone;
        bst = cand < bst ? cand : bst;
    }
    for (int o = kWave / 2; o > 0; o >>= 1) {
        uint32_t t = __shfl_xor(bst, o);
        bst = t < bst ? t : bst;
    }
    if (lane_id() == 0) best[p] = bst;
}

// ------------------------------------------------------------ numbering
__global__ void k_root_flags(const uint8_t *core, const uint32_t *parent, const uint32_t *rank,
                             uint32_t n, uint32_t *flag_by_rank)
{
    uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    flag_by_rank[rank[p]] = (core[p] && parent[p] == p) ? 1u : 0u;
}

__global__ void k_labels(const uint8_t *core, const uint32_t *parent, const uint32_t *rank,
                         const uint32_t *best, const uint32_t *cid_by_rank,
                         const uint32_t *mapping, uint32_t n, uint32_t *out)
{
    uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    uint32_t r = rank[p];
    uint32_t lbl;
    if (core[p]) {
        lbl = cid_by_rank[rank[parent[p]]] + 1;
    } else {
        uint32_t b = best[p];
        lbl = (b != kNone && b < r) ? cid_by_rank[b] + 1 : 0u;
    }
    out[mapping[r]] = lbl;
}

// Standard border rule. key[root] = smallest caller index among the core
// points of root's component (key holds NONE on entry; parent is flat). Lanes
// of a wave that share the first lane's root reduce first, so a large
// cluster takes one atomic per wave; the remaining lanes update theirs one
// atomic each.
__global__ void k_root_keys(const uint8_t *core, const uint32_t *parent, const uint32_t *rank,
                            const uint32_t *mapping, uint32_t n, uint32_t *key)
{
    uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    bool valid = p < n && core[p];
    uint32_t root = valid ? parent[p] : kNone;
    uint32_t v = valid ? mapping[rank[p]] : kNone;
    uint64_t todo = ballot(valid);
    if (todo) {
        int leader = __builtin_ctzll(todo);
        uint32_t r = __shfl(root, leader);
        bool same = valid && root == r;
        todo &= ~ballot(same);
        uint32_t m = same ? v : kNone;
        for (int o = kWave / 2; o > 0; o >>= 1) {
            uint32_t t = __shfl_xor(m, o);
            m = t < m ? t : m;
        }
        if (lane_id() == leader) atomicMin(key + r, m);
    }
    if ((todo >> lane_id()) & 1ull) atomicMin(key + root, v);
}

// flag_by_caller[i] = 1 when the point of caller index i is the smallest
// core point (by caller index) of its component; every entry is written
__global__ void k_key_flags(const uint8_t *core, const uint32_t *parent, const uint32_t *rank,
                            const uint32_t *mapping, const uint32_t *key, uint32_t n,
                            uint32_t *flag_by_caller)
{
    uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    uint32_t ci = mapping[rank[p]];
    flag_by_caller[ci] = (core[p] && key[parent[p]] == ci) ? 1u : 0u;
}

// Labels of the standard border rule: clusters 0.. in ascending key, noise
// 0xffffffff (-1 as int32); best[p] is the smallest key among the core
// neighbours of non-core p.
__global__ void k_key_labels(const uint8_t *core, const uint32_t *parent, const uint32_t *rank,
                             const uint32_t *key, const uint32_t *best,
                             const uint32_t *cid_by_caller, const uint32_t *mapping, uint32_t n,
                             uint32_t *out)
{
    uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    uint32_t lbl;
    if (core[p]) {
        lbl = cid_by_caller[key[parent[p]]];
    } else {
        uint32_t b = best[p];
        lbl = b != kNone ? cid_by_caller[b] : kNone;
    }
    out[mapping[rank[p]]] = lbl;
}

__global__ void k_perm_counts(const uint32_t *cnt, const uint32_t *pos_rank,
                              const uint32_t *mapping, uint32_t n, uint32_t *cc, uint32_t *cop)
{
    uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    uint32_t ci = mapping[pos_rank[p]];
    cop[p] = ci;
    cc[ci] = cnt[p];
}

Band band_for(const Grid &G, double T)
{
    // Origins lie within w and candidates within (R+1)w of the reference
    // origin per axis, so |dx|, |dy| <= (R+2)w =: D. The float32 conversions
    // and subtraction leave dx within 2^-23 D, and squares and sum add
    // |s32 - s| <= D^2 * 2^-20.4 in all; the band used is D^2 * 2^-14 / 9
    // (w^2 * 2^-14 for R = 1), widened to the float grid around T.
    double D = (G.R + 2) * G.w;
    double band = D * D * std::ldexp(1.0, -14) / 9.0;
    Band b;
    if (T < 0) {
        b.lo = -1.0f;
        b.hi = -1.0f;
        return b;
    }
    // The bound above is relative error only. It fails where float32 leaves
    // its normal range: differences or squares below FLT_MIN carry an
    // absolute error up to 2^-126 each (a pair 2 eps apart can round to 0
    // at eps = 1e-24), and squares of differences above about 2^60
    // overflow. Outside 2^-100 <= band and 8 D^2 <= 2^120 every candidate is
    // decided in binary64 (accept nothing unchecked, recheck everything).
    if (!(band >= std::ldexp(1.0, -100)) || !(8.0 * D * D <= std::ldexp(1.0, 120))) {
        b.lo = -1.0f;
        b.hi = INFINITY;
        return b;
    }
    double lo = T - band, hi = T + band;
    b.lo = lo > 0 ? std::nextafter((float)lo, 0.0f) : -1.0f;
    if (b.lo > lo) b.lo = std::nextafter(b.lo, -INFINITY);
    b.hi = std::nextafter((float)hi, INFINITY);
    return b;
}

// events are created on the first start(): a pass that is not timed creates
// none
struct EvTimer {
    hipEvent_t a = nullptr, b = nullptr;
    hipStream_t s;
    explicit EvTimer(hipStream_t st) : s(st) {}
    ~EvTimer()
    {
        if (a) hipEventDestroy(a);
        if (b) hipEventDestroy(b);
    }
    void start()
    {
        if (!a) {
            hipEventCreate(&a);
            hipEventCreate(&b);
        }
        hipEventRecord(a, s);
    }
    float stop()
    {
        hipEventRecord(b, s);
        hipEventSynchronize(b);
        float ms = 0;
        hipEventElapsedTime(&ms, a, b);
        return ms;
    }
};

} // namespace

int build_items(Context &c, const Grid &G, const char *tag, Items &I)
{
    uint32_t *cnt, *excl;
    std::string t(tag);
    VDB_TRY(c.get((t + ".icnt").c_str(), G.ncell, &cnt));
    VDB_TRY(c.get((t + ".iexcl").c_str(), G.ncell, &excl));
    k_item_counts<<<ceil_div(G.ncell, kBlock), kBlock, 0, c.stream()>>>(G.cell_start, G.ncell,
                                                                        cnt);
    VDB_HIP(hipGetLastError());
    uint32_t total = 0;
    VDB_TRY(compact_count(c, cnt, excl, G.ncell, &total));
    I.n = total;
    VDB_TRY(c.get((t + ".icell").c_str(), total, &I.cell));
    VDB_TRY(c.get((t + ".io0").c_str(), total, &I.o0));
    k_item_fill<<<ceil_div(G.ncell, kBlock), kBlock, 0, c.stream()>>>(G.cell_start, excl,
                                                                      G.ncell, I.cell, I.o0);
    VDB_HIP(hipGetLastError());
    return kOk;
}

int build_eps_grid(Context &c, const Points &P, double eps, int k, const char *tag, Grid &G)
{
    k = k < 1 ? 1 : (k > kMaxCellsPerEps ? kMaxCellsPerEps : k);
    bool enlarged = false;
    double w = eps * (1.0 + std::ldexp(1.0, -24)) / k;
    VDB_TRY(build_grid(c, P, w, k, tag, G, &enlarged));
    return classify_rows(c, G, threshold_for_eps(eps), tag);
}

int cells_per_eps(const vdb_options *opt)
{
    if (opt && opt->cells_per_eps >= 1 && opt->cells_per_eps <= kMaxCellsPerEps)
        return opt->cells_per_eps;
    return 0;
}

int tune_cells_per_eps(Context &c, const Points &P, double eps, int precision, int requested,
                       int *k_out)
{
    if (requested >= 1 && requested <= kMaxCellsPerEps) {
        *k_out = requested;
        return kOk;
    }
    // below this many points the whole per-variant pass takes well under a
    // millisecond and timings are noise
    constexpr uint32_t kTuneMinPoints = 1u << 17;
    if (P.n < kTuneMinPoints || threshold_for_eps(eps) < 0) {
        *k_out = kDefaultCellsPerEps;
        return kOk;
    }
    auto bits = [](double v) {
        uint64_t u;
        memcpy(&u, &v, sizeof u);
        return (unsigned long long)u;
    };
    // the key names the tuned range (1..kMaxCellsPerEps): a choice cached
    // over a narrower range is tuned again
    char key[160];
    snprintf(key, sizeof key, "cells%d:%u:%d:%llx:%llx:%llx:%llx:%llx", kMaxCellsPerEps, P.n,
             precision, bits(eps),
             bits(P.ext[0]), bits(P.ext[1]), bits(P.ext[2]), bits(P.ext[3]));
    int cached = 0;
    if (c.tuned_get(key, &cached) && cached >= 1 && cached <= kMaxCellsPerEps) {
        *k_out = cached;
        return kOk;
    }
    uint32_t *counts;
    unsigned long long *rc;
    VDB_TRY(c.get("tune.counts", P.n, &counts));
    VDB_TRY(c.get("tune.recheck", 1, &rc));
    // Finer cells test fewer candidates (more of the window is accepted whole)
    // but add rows per window; past the point where the count stops gaining,
    // the fill and the core pass run a little slower on the finer grid. The
    // choice is the coarsest k whose count is within kTuneSlack of the best.
    constexpr float kTuneSlack = 1.05f;
    float ms_k[kMaxCellsPerEps + 1] = {};
    float best_ms = 0;
    for (int k = 1; k <= kMaxCellsPerEps; k++) {
        Grid G;
        Items I;
        EvTimer tm(c.stream());
        tm.start();
        VDB_TRY(build_eps_grid(c, P, eps, k, "g1", G));
        VDB_TRY(build_items(c, G, "g1", I));
        VDB_TRY(count_neighbors(c, G, I, eps, precision, counts, (uint64_t *)rc));
        ms_k[k] = tm.stop();
        if (k == 1 || ms_k[k] < best_ms) best_ms = ms_k[k];
    }
    int best = kMaxCellsPerEps;
    for (int k = 1; k <= kMaxCellsPerEps; k++)
        if (ms_k[k] <= best_ms * kTuneSlack) {
            best = k;
            break;
        }
    c.tuned_put(key, best);
    *k_out = best;
    return kOk;
}

int count_neighbors(Context &c, const Grid &G, const Items &I, double eps, int precision,
                    uint32_t *counts, uint64_t *recheck_dev)
{
    double T = threshold_for_eps(eps);
    if (I.n == 0) return kOk;
    unsigned blocks = ceil_div(I.n, kWavesPerBlock);
    if (precision == 0) {
        VDB_TRY(c.kt_begin("k_count"));
        k_count<0><<<blocks, kBlock, 0, c.stream()>>>(view(G, G.npts), I.cell, I.o0, I.n, T, Band{},
                                                      counts, nullptr);
        VDB_TRY(c.kt_end("k_count"));
    } else {
        VDB_TRY(c.kt_begin("k_count"));
        k_count<1><<<blocks, kBlock, 0, c.stream()>>>(view(G, G.npts), I.cell, I.o0, I.n, T,
                                                      band_for(G, T), counts,
                                                      (unsigned long long *)recheck_dev);
        VDB_TRY(c.kt_end("k_count"));
    }
    VDB_HIP(hipGetLastError());
    return kOk;
}

static int finish_labels(Context &c, const Points &P, const Grid &G, const uint8_t *core,
                         const uint32_t *parent, const uint32_t *best, const uint32_t *key,
                         uint32_t *out)
{
    uint32_t n = P.n;
    unsigned g = ceil_div(n, kBlock);
    uint32_t *flag, *cid;
    VDB_TRY(c.get("lab.flag", n, &flag));
    VDB_TRY(c.get("lab.cid", n, &cid));
    if (key) {
        // standard border rule: clusters numbered by smallest core caller index
        k_key_flags<<<g, kBlock, 0, c.stream()>>>(core, parent, G.pos_rank, P.mapping, key, n,
                                                  flag);
        VDB_HIP(hipGetLastError());
        VDB_TRY(scan_exclusive_u32(c, flag, cid, n, nullptr));
        k_key_labels<<<g, kBlock, 0, c.stream()>>>(core, parent, G.pos_rank, key, best, cid,
                                                   P.mapping, n, out);
        VDB_HIP(hipGetLastError());
        return kOk;
    }
    k_root_flags<<<g, kBlock, 0, c.stream()>>>(core, parent, G.pos_rank, n, flag);
    VDB_HIP(hipGetLastError());
    VDB_TRY(scan_exclusive_u32(c, flag, cid, n, nullptr));
    k_labels<<<g, kBlock, 0, c.stream()>>>(core, parent, G.pos_rank, best, cid, P.mapping, n,
                                           out);
    VDB_HIP(hipGetLastError());
    return kOk;
}

int cluster_variant(Context &c, const Points &P, const Grid &G, const Items &I, double eps,
                    uint32_t minpts, const uint32_t *counts, int mode, int border,
                    size_t table_budget, uint32_t *out, PhaseTimes *pt)
{
    uint32_t n = P.n;
    double T = threshold_for_eps(eps);
    unsigned g = ceil_div(n, kBlock);
    uint8_t *core;
    uint32_t *parent, *best, *key = nullptr;
    VDB_TRY(c.get("cv.core", n, &core));
    VDB_TRY(c.get("cv.parent", n, &parent));
    VDB_TRY(c.get("cv.best", n, &best));
    if (border == VDB_BORDER_STANDARD) VDB_TRY(c.get("cv.key", n, &key));
    // Border kernels take the minimum of bkey[root] over a point's core
    // neighbours: the root's rank (reference rule), or its component's
    // smallest core caller index (standard rule), set once the forest is flat.
    auto root_keys = [&]() -> int {
        if (!key) return kOk;
        VDB_HIP(hipMemsetAsync(key, 0xff, sizeof(uint32_t) * n, c.stream()));
        k_root_keys<<<g, kBlock, 0, c.stream()>>>(core, parent, G.pos_rank, P.mapping, n, key);
        VDB_HIP(hipGetLastError());
        return kOk;
    };
    const uint32_t *bkey = key ? key : G.pos_rank;
    EvTimer tm(c.stream());
    k_core_init<<<g, kBlock, 0, c.stream()>>>(counts, n, minpts, core, parent, best);
    VDB_HIP(hipGetLastError());
    unsigned blocks = ceil_div(I.n, kWavesPerBlock);
    if (mode == 0) {
        if (pt) tm.start();
        if (I.n) {
            VDB_TRY(c.kt_begin("k_uf_fused"));
            k_uf_fused<<<blocks, kBlock, 0, c.stream()>>>(view(G, G.npts), I.cell, I.o0, I.n, T, core,
                                                          G.pos_rank, parent);
            VDB_TRY(c.kt_end("k_uf_fused"));
        }
        VDB_HIP(hipGetLastError());
        k_compress<<<g, kBlock, 0, c.stream()>>>(parent, core, n);
        VDB_HIP(hipGetLastError());
        VDB_TRY(root_keys());
        if (pt) pt->uf += tm.stop(), tm.start();
        if (I.n) {
            VDB_TRY(c.kt_begin("k_border_fused"));
            k_border_fused<<<blocks, kBlock, 0, c.stream()>>>(view(G, G.npts), I.cell, I.o0, I.n, T,
                                                              core, bkey, parent, best);
            VDB_TRY(c.kt_end("k_border_fused"));
        }
        VDB_HIP(hipGetLastError());
        if (pt) pt->border += tm.stop();
    } else {
        // table mode: offsets over positions, then tiles of whole cells
        uint64_t *offsets;
        VDB_TRY(c.get("cv.offsets", (size_t)n + 1, &offsets));
        if (pt) tm.start();
        VDB_TRY(scan_exclusive_u32_u64(c, counts, offsets, n, offsets + n));
        if (pt) pt->scan += tm.stop();
        std::vector<uint64_t> hoff(n + 1);
        std::vector<uint32_t> hstart(G.ncell + 1), hio0(I.n);
        VDB_HIP(hipMemcpyAsync(hoff.data(), offsets, sizeof(uint64_t) * (n + 1),
                               hipMemcpyDeviceToHost, c.stream()));
        VDB_HIP(hipMemcpyAsync(hstart.data(), G.cell_start, sizeof(uint32_t) * (G.ncell + 1),
                               hipMemcpyDeviceToHost, c.stream()));
        VDB_HIP(hipMemcpyAsync(hio0.data(), I.o0, sizeof(uint32_t) * I.n,
                               hipMemcpyDeviceToHost, c.stream()));
        VDB_HIP(hipStreamSynchronize(c.stream()));
        uint64_t P_total = hoff[n];
        if (pt) pt->pairs = P_total;
        uint64_t cap = table_budget / sizeof(uint32_t);
        // largest single cell must fit
        uint32_t *ids = nullptr, *ovf = nullptr;
        VDB_TRY(c.get("cv.ovf", 1, &ovf));
        VDB_HIP(hipMemsetAsync(ovf, 0, sizeof(uint32_t), c.stream()));
        uint64_t want = P_total < cap ? P_total : cap;
        // the slot keeps the largest table seen so far; reuse it whole, and
        // when the table is tiled anyway keep a slot of at least half the
        // budget rather than reallocating hundreds of GB for a few percent
        uint64_t have = c.slot_bytes("cv.ids") / sizeof(uint32_t);
        if (have > want && have <= cap) want = have;
        if (have < want && P_total > have && have >= cap / 2) want = have;
        VDB_TRY(c.get("cv.ids", want ? want : 1, &ids));
        // tiles never exceed the buffer actually held
        cap = want;
        // Core components: the fused half-window union-find (distances
        // recomputed, each unordered pair once) rather than a pass over the
        // table, which would read every row's ids again (95 GB at 2M points,
        // eps 0.5) at gather speed.
        if (pt) tm.start();
        if (I.n) {
            VDB_TRY(c.kt_begin("k_uf_fused"));
            k_uf_fused<<<blocks, kBlock, 0, c.stream()>>>(view(G, G.npts), I.cell, I.o0, I.n, T,
                                                          core, G.pos_rank, parent);
            VDB_TRY(c.kt_end("k_uf_fused"));
        }
        VDB_HIP(hipGetLastError());
        k_compress<<<g, kBlock, 0, c.stream()>>>(parent, core, n);
        VDB_HIP(hipGetLastError());
        VDB_TRY(root_keys());
        if (pt) pt->uf += tm.stop();
        // The table, tile by tile (tiles are runs of items: items are ordered
        // by cell, origins ascending), each tile filled once and its border
        // rows (fewer than minpts entries each) read back for the border pass.
        uint32_t ntiles = 0;
        uint32_t i0 = 0;
        while (i0 < I.n) {
            uint32_t i1 = i0;
            uint64_t base = hoff[hio0[i0]];
            while (i1 < I.n) {
                uint32_t oend = (i1 + 1 < I.n) ? hio0[i1 + 1] : n;
                if (hoff[oend] - base > cap && i1 > i0) break;
                i1++;
            }
            uint32_t pend = i1 < I.n ? hio0[i1] : n;
            uint32_t p0 = hio0[i0];
            if (hoff[pend] - base > cap) {
                set_error("neighbour table tile does not fit the table budget");
                return kErrNoMemory;
            }
            if (pt) tm.start();
            VDB_TRY(c.kt_begin("k_fill"));
            k_fill<0, false><<<ceil_div(i1 - i0, kWavesPerBlock), kBlock, 0, c.stream()>>>(
                view(G, G.npts), I.cell, I.o0, i0, i1, T, Band{}, offsets, base, nullptr, nullptr,
                ids, nullptr, cap, ovf);
            VDB_TRY(c.kt_end("k_fill"));
            VDB_HIP(hipGetLastError());
            if (pt) pt->fill += tm.stop(), tm.start();
            k_border_table<<<ceil_div(pend - p0, kWavesPerBlock), kBlock, 0, c.stream()>>>(
                offsets, base, ids, p0, pend, core, bkey, parent, best);
            VDB_HIP(hipGetLastError());
            if (pt) pt->border += tm.stop();
            ntiles++;
            i0 = i1;
        }
        if (pt) pt->tiles += ntiles;
        uint32_t hov = 0;
        VDB_HIP(hipMemcpyAsync(&hov, ovf, sizeof hov, hipMemcpyDeviceToHost, c.stream()));
        VDB_HIP(hipStreamSynchronize(c.stream()));
        if (hov) {
            set_error("neighbour table fill disagreed with the count pass");
            return kErrDevice;
        }
    }
    if (pt) tm.start();
    VDB_TRY(finish_labels(c, P, G, core, parent, best, key, out));
    if (pt) pt->labels += tm.stop();
    return kOk;
}

namespace {
__global__ void k_pos_to_caller(const uint32_t *v, const uint32_t *pos_rank,
                                const uint32_t *mapping, uint32_t n, uint32_t *out)
{
    uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    out[mapping[pos_rank[p]]] = v[p];
}
} // namespace

int core_forest(Context &c, const Grid &G, const Items &I, double eps, uint32_t minpts,
                const uint32_t *counts, uint8_t *core, uint32_t *parent, uint32_t *best)
{
    uint32_t n = G.npts;
    double T = threshold_for_eps(eps);
    unsigned g = ceil_div(n, kBlock);
    k_core_init<<<g, kBlock, 0, c.stream()>>>(counts, n, minpts, core, parent, best);
    VDB_HIP(hipGetLastError());
    if (I.n) {
        VDB_TRY(c.kt_begin("k_uf_fused"));
        k_uf_fused<<<ceil_div(I.n, kWavesPerBlock), kBlock, 0, c.stream()>>>(
            view(G, n), I.cell, I.o0, I.n, T, core, G.pos_rank, parent);
        VDB_TRY(c.kt_end("k_uf_fused"));
    }
    VDB_HIP(hipGetLastError());
    k_compress<<<g, kBlock, 0, c.stream()>>>(parent, core, n);
    VDB_HIP(hipGetLastError());
    return kOk;
}

int border_min_key(Context &c, const Grid &G, const Items &I, double eps, const uint8_t *core,
                   const uint32_t *parent, const uint32_t *key, uint32_t *best)
{
    uint32_t n = G.npts;
    k_fill_u32_n<<<ceil_div(n, kBlock), kBlock, 0, c.stream()>>>(best, kNone, n);
    VDB_HIP(hipGetLastError());
    if (I.n) {
        VDB_TRY(c.kt_begin("k_border_fused"));
        k_border_fused<<<ceil_div(I.n, kWavesPerBlock), kBlock, 0, c.stream()>>>(
            view(G, n), I.cell, I.o0, I.n, threshold_for_eps(eps), core, key, parent, best);
        VDB_TRY(c.kt_end("k_border_fused"));
    }
    VDB_HIP(hipGetLastError());
    return kOk;
}

int border_pairs(Context &c, const Grid &G, const Items &I, double eps, const uint8_t *core,
                 const uint32_t *parent, const uint32_t *cl_of_pos, uint32_t *pairs,
                 uint32_t cap, uint32_t *count_dev)
{
    k_fill_u32_n<<<1, kBlock, 0, c.stream()>>>(count_dev, 0u, 1);
    VDB_HIP(hipGetLastError());
    if (I.n)
        k_border_pairs<<<ceil_div(I.n, kWavesPerBlock), kBlock, 0, c.stream()>>>(
            view(G, G.npts), I.cell, I.o0, I.n, threshold_for_eps(eps), core, parent, cl_of_pos,
            pairs, cap, count_dev);
    VDB_HIP(hipGetLastError());
    return kOk;
}

int pos_to_caller(Context &c, const Points &P, const Grid &G, const uint32_t *by_pos,
                  uint32_t *by_caller)
{
    k_pos_to_caller<<<ceil_div(P.n, kBlock), kBlock, 0, c.stream()>>>(by_pos, G.pos_rank,
                                                                     P.mapping, P.n, by_caller);
    VDB_HIP(hipGetLastError());
    return kOk;
}

int table_digest(Context &c, const Points &P, const Grid &G, const Items &I, double eps,
                 int precision, uint64_t *digest, uint64_t *pairs, uint64_t *rechecked)
{
    uint32_t n = P.n;
    double T = threshold_for_eps(eps);
    uint32_t *counts, *caller_of_pos;
    VDB_TRY(c.get("nt.counts", n, &counts));
    VDB_TRY(c.get("nt.caller_of_pos", n, &caller_of_pos));
    unsigned long long *acc;
    VDB_TRY(c.get("nt.digest", 2, &acc));
    VDB_HIP(hipMemsetAsync(acc, 0, 2 * sizeof(unsigned long long), c.stream()));
    VDB_TRY(count_neighbors(c, G, I, eps, precision, counts, (uint64_t *)(acc + 1)));
    uint32_t *dummy;
    VDB_TRY(c.get("nt.counts_caller", n, &dummy));
    k_perm_counts<<<ceil_div(n, kBlock), kBlock, 0, c.stream()>>>(counts, G.pos_rank, P.mapping,
                                                                 n, dummy, caller_of_pos);
    VDB_HIP(hipGetLastError());
    uint64_t *tot;
    VDB_TRY(c.get("nt.tot", 1, &tot));
    uint64_t *offs;
    VDB_TRY(c.get("nt.off", (size_t)n + 1, &offs));
    VDB_TRY(scan_exclusive_u32_u64(c, dummy, offs, n, tot));
    if (I.n) {
        unsigned blocks = ceil_div(I.n, kWavesPerBlock);
        Band bd = precision ? band_for(G, T) : Band{};
        if (precision == 0) {
            VDB_TRY(c.kt_begin("k_fill"));
            k_fill<0, false, true><<<blocks, kBlock, 0, c.stream()>>>(
                view(G, G.npts), I.cell, I.o0, 0, I.n, T, bd, nullptr, 0, caller_of_pos, caller_of_pos,
                (uint32_t *)acc, nullptr, 0, nullptr);
            VDB_TRY(c.kt_end("k_fill"));
        } else {
            VDB_TRY(c.kt_begin("k_fill"));
            k_fill<1, false, true><<<blocks, kBlock, 0, c.stream()>>>(
                view(G, G.npts), I.cell, I.o0, 0, I.n, T, bd, nullptr, 0, caller_of_pos, caller_of_pos,
                (uint32_t *)acc, nullptr, 0, nullptr);
            VDB_TRY(c.kt_end("k_fill"));
        }
        VDB_HIP(hipGetLastError());
    }
    unsigned long long h[2];
    VDB_HIP(hipMemcpyAsync(h, acc, sizeof h, hipMemcpyDeviceToHost, c.stream()));
    uint64_t t = 0;
    VDB_HIP(hipMemcpyAsync(&t, tot, sizeof t, hipMemcpyDeviceToHost, c.stream()));
    VDB_HIP(hipStreamSynchronize(c.stream()));
    *digest = h[0];
    *rechecked = h[1];
    *pairs = t;
    return kOk;
}

int neighbor_table(Context &c, const Points &P, const Grid &G, const Items &I, double eps,
                   int precision, uint64_t *offsets, uint32_t *ids, float *d2,
                   uint64_t capacity, uint64_t *total, PhaseTimes *pt)
{
    uint32_t n = P.n;
    double T = threshold_for_eps(eps);
    uint32_t *counts, *counts_caller, *caller_of_pos;
    VDB_TRY(c.get("nt.counts", n, &counts));
    VDB_TRY(c.get("nt.counts_caller", n, &counts_caller));
    VDB_TRY(c.get("nt.caller_of_pos", n, &caller_of_pos));
    unsigned long long *rc;
    VDB_TRY(c.get("nt.recheck", 1, &rc));
    VDB_HIP(hipMemsetAsync(rc, 0, sizeof(unsigned long long), c.stream()));
    EvTimer tm(c.stream());
    if (pt) tm.start();
    VDB_TRY(count_neighbors(c, G, I, eps, precision, counts, (uint64_t *)rc));
    if (pt) pt->count += tm.stop();
    // caller index of each position, and counts in caller order
    k_perm_counts<<<ceil_div(n, kBlock), kBlock, 0, c.stream()>>>(counts, G.pos_rank, P.mapping, n,
                                                            counts_caller, caller_of_pos);
    VDB_HIP(hipGetLastError());
    if (pt) tm.start();
    VDB_TRY(scan_exclusive_u32_u64(c, counts_caller, offsets, n, offsets + n));
    if (pt) pt->scan += tm.stop();
    uint64_t tot = 0;
    VDB_HIP(hipMemcpyAsync(&tot, offsets + n, sizeof(uint64_t), hipMemcpyDeviceToHost,
                           c.stream()));
    VDB_HIP(hipStreamSynchronize(c.stream()));
    *total = tot;
    if (pt) {
        pt->pairs = tot;
        unsigned long long r = 0;
        VDB_HIP(hipMemcpy(&r, rc, sizeof r, hipMemcpyDeviceToHost));
        pt->recheck = r;
    }
    if (tot > capacity) return kOk; // caller sizes its buffers from *total
    if (I.n == 0) return kOk;
    if (pt) tm.start();
    uint32_t *ovf = nullptr;
    VDB_TRY(c.get("nt.ovf", 1, &ovf));
    VDB_HIP(hipMemsetAsync(ovf, 0, sizeof(uint32_t), c.stream()));
    unsigned blocks = ceil_div(I.n, kWavesPerBlock);
    Band bd = precision ? band_for(G, T) : Band{};
    if (precision == 0 && !d2) {
        VDB_TRY(c.kt_begin("k_fill"));
        k_fill<0, false><<<blocks, kBlock, 0, c.stream()>>>(
            view(G, G.npts), I.cell, I.o0, 0, I.n, T, bd, offsets, 0, caller_of_pos, caller_of_pos, ids,
            d2, capacity, ovf);
        VDB_TRY(c.kt_end("k_fill"));
    }
    else if (precision == 0) {
        VDB_TRY(c.kt_begin("k_fill"));
        k_fill<0, true><<<blocks, kBlock, 0, c.stream()>>>(
            view(G, G.npts), I.cell, I.o0, 0, I.n, T, bd, offsets, 0, caller_of_pos, caller_of_pos, ids,
            d2, capacity, ovf);
        VDB_TRY(c.kt_end("k_fill"));
    }
    else if (!d2) {
        VDB_TRY(c.kt_begin("k_fill"));
        k_fill<1, false><<<blocks, kBlock, 0, c.stream()>>>(
            view(G, G.npts), I.cell, I.o0, 0, I.n, T, bd, offsets, 0, caller_of_pos, caller_of_pos, ids,
            d2, capacity, ovf);
        VDB_TRY(c.kt_end("k_fill"));
    } else {
        VDB_TRY(c.kt_begin("k_fill"));
        k_fill<1, true><<<blocks, kBlock, 0, c.stream()>>>(
            view(G, G.npts), I.cell, I.o0, 0, I.n, T, bd, offsets, 0, caller_of_pos, caller_of_pos, ids,
            d2, capacity, ovf);
        VDB_TRY(c.kt_end("k_fill"));
    }
    VDB_HIP(hipGetLastError());
    if (pt) pt->fill += tm.stop();
    uint32_t hov = 0;
    VDB_HIP(hipMemcpyAsync(&hov, ovf, sizeof hov, hipMemcpyDeviceToHost, c.stream()));
    VDB_HIP(hipStreamSynchronize(c.stream()));
    if (hov) {
        set_error("neighbour table fill disagreed with the count pass");
        return kErrDevice;
    }
    return kOk;
}

} // namespace vdb
