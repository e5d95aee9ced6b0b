// Per-variant clustering over a grid of cell side >= eps (window radius 1).
//
// Labels follow the reference's sequential expansion (src/DBScan.cpp:523-609)
// in closed form:
//   * core points (>= minpts eps-neighbours, self included) are grouped into
//     connected components; each component is rooted at its smallest binned
//     rank, and clusters are numbered 1.. in ascending root rank;
//   * a non-core point takes the cluster of the smallest root among its core
//     neighbours, but only when that root's rank is below its own; otherwise
//     the outer loop reached it first and it stays noise (0).
#pragma once

#include "vdb_grid.h"
#include "vdbscan.h"

namespace vdb {

struct PhaseTimes {
    float count = 0, scan = 0, fill = 0, uf = 0, border = 0, labels = 0;
    uint64_t pairs = 0;      // eps-neighbour pairs (self pairs included)
    uint32_t tiles = 0;      // table tiles (table mode)
    uint64_t recheck = 0;    // pairs resolved by the FP64 recheck (FP32 mode)
};

// Work items: (cell, first origin) groups of kOrigins origins of one cell.
struct Items {
    uint32_t n = 0;
    uint32_t *cell = nullptr;
    uint32_t *o0 = nullptr;
};

constexpr int kOrigins = 8;

int build_items(Context &c, const Grid &G, const char *tag, Items &I);

// Grid of the per-variant kernels at eps: cells of side eps*(1+2^-24)/k and a
// window of k cells each way (any pair accepted at eps is at most
// eps*(1+2^-50) apart per axis, and computed cell coordinates are within
// 2^-31 of exact below 2^20 cells per axis, so the pair spans at most k
// cells), with its window rows classified for T(eps).
int build_eps_grid(Context &c, const Points &P, double eps, int k, const char *tag, Grid &G);

// Cells per eps for the per-variant kernels: the option's value when set
// (1..kMaxCellsPerEps), else 0 (let tune_cells_per_eps choose).
constexpr int kMaxCellsPerEps = 8;
int cells_per_eps(const vdb_options *opt);

// Autotuner: the first time a workload (point count, eps, extent, precision)
// is seen, times grid build + count pass at every resolution 1..4 and keeps
// the fastest; later calls reuse the choice. Small inputs skip the timing.
// Uses the "g1" grid slots. requested > 0 is returned unchanged.
int tune_cells_per_eps(Context &c, const Points &P, double eps, int precision, int requested,
                       int *k_out);

// counts[p] = |N_eps(p)| for every position p of G (self included).
// precision 0 = FP64 exact, 1 = FP32 with an FP64 recheck inside the error band.
int count_neighbors(Context &c, const Grid &G, const Items &I, double eps, int precision,
                    uint32_t *counts, uint64_t *recheck_dev);

// Full per-variant clustering; labels written in caller order to out (device).
// mode 0 = fused (no table), 1 = CSR table (count/scan/fill, union-find over
// the table), tiled when the table exceeds table_budget bytes. border =
// VDB_BORDER_REFERENCE gives the labels above; VDB_BORDER_STANDARD numbers
// the same clusters 0.. by their smallest core caller index, gives every
// non-core point with a core neighbour the lowest-numbered adjacent cluster
// and marks noise 0xffffffff.
int cluster_variant(Context &c, const Points &P, const Grid &G, const Items &I, double eps,
                    uint32_t minpts, const uint32_t *counts, int mode, int border,
                    size_t table_budget, uint32_t *out, PhaseTimes *pt);

// Pieces of the fused engine for callers that label differently (the reuse
// emulation). Over positions of G:
//   core_forest: core[p] and the flat core forest (parent[p] = position of
//     the smallest-rank point of p's component, for core p); best[] is
//     cleared.
//   border_min_key: best[p] = smallest key[root] over the core neighbours of
//     non-core p (NONE where there is none, and for core p).
//   border_pairs: for every non-core p with cl_of_pos[p] = C != 0, the
//     pairs (C, root) for the distinct roots among p's core neighbours, as
//     uint32 pairs; at most cap stored, *count_dev = all (repeats possible).
int core_forest(Context &c, const Grid &G, const Items &I, double eps, uint32_t minpts,
                const uint32_t *counts, uint8_t *core, uint32_t *parent, uint32_t *best);
int border_min_key(Context &c, const Grid &G, const Items &I, double eps, const uint8_t *core,
                   const uint32_t *parent, const uint32_t *key, uint32_t *best);
int border_pairs(Context &c, const Grid &G, const Items &I, double eps, const uint8_t *core,
                 const uint32_t *parent, const uint32_t *cl_of_pos, uint32_t *pairs,
                 uint32_t cap, uint32_t *count_dev);

// Scatter a per-position array to caller order.
int pos_to_caller(Context &c, const Points &P, const Grid &G, const uint32_t *by_pos,
                  uint32_t *by_caller);

// Order-independent digest of the caller-order neighbour table (sum over its
// (row, id) pairs of a 64-bit mix, mod 2^64) without storing the table.
int table_digest(Context &c, const Points &P, const Grid &G, const Items &I, double eps,
                 int precision, uint64_t *digest, uint64_t *pairs, uint64_t *rechecked);

// Neighbour table for external callers: rows and ids in caller order (ids
// unsorted inside a row), optional float32 squared distances.
int neighbor_table(Context &c, const Points &P, const Grid &G, const Items &I, double eps,
                   int precision, uint64_t *offsets /*n+1, device*/, uint32_t *ids,
                   float *d2, uint64_t capacity, uint64_t *total, PhaseTimes *pt);

} // namespace vdb
