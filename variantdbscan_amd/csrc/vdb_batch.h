// Batched variants: all (eps, minpts) pairs of one call from a single fine
// grid whose cells are verified cliques at the smallest eps, a distance-sorted
// neighbour list per point truncated at the largest minpts, and union-find over
// cell pairs, carried across eps levels for each minpts.
#pragma once

#include "vdbscan.h"
#include "vdb_grid.h"

namespace vdb {

struct BatchResult {
    uint32_t levels = 0;        // distinct eps values
    uint32_t minpts_values = 0; // distinct minpts values
    uint32_t list_len = 0;      // truncated list length (largest minpts)
    uint64_t listed_points = 0; // points that needed a neighbour list
    uint64_t cell_pairs = 0;
    int shape = 0;              // grid shape the autotuner chose (index)
};

// Variant v of dataset d is written to out + seg[d]*nrows + rows[v]*n_d (for
// one dataset: out + rows[v]*n), labelled by the border rule given
// (VDB_BORDER_REFERENCE or VDB_BORDER_STANDARD). Every eps must be finite and > 0
// or have no neighbour at all (eps < 0, NaN), every minpts <= 64 or above
// INT_MAX, and max eps / min eps <= kBatchEpsRatio. Returns kErrUnsupported
// when the data defeat the engine (cells that are not cliques, too many
// cells); the caller then clusters those variants one by one.
constexpr double kBatchEpsRatio = 16.0;
constexpr uint32_t kBatchMaxMinpts = 64;
bool batchable(double eps, uint32_t minpts);
int cluster_batch(Context &c, const Points &P, const double *eps, const uint32_t *minpts,
                  const uint32_t *rows, uint32_t nvar, uint32_t *out, uint32_t nrows,
                  int border, vdb_stats *st, bool verbose, BatchResult *res);

} // namespace vdb
