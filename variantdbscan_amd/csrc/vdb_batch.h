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
};

// Returns kErrUnsupported when the variant set is outside what the batched
// engine handles (the caller then clusters variant by variant).
int cluster_batch(Context &c, const Points &P, const double *eps, const uint32_t *minpts,
                  uint32_t nvar, int precision, uint32_t *out, vdb_stats *st, bool verbose,
                  BatchResult *res);

} // namespace vdb
