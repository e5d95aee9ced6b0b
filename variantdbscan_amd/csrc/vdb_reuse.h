// Multi-variant calls that reproduce the reference's greedy cluster reuse as
// it runs with one search thread (see vdb_reuse.hip).
#pragma once

#include "vdb_grid.h"

namespace vdb {

// Clusters the device-resident points P under nvar variants into out
// (device, retArr layout: row v holds variant v's labels in caller order),
// with the labels the reference's multi-variant libVDBSCAN gives when built
// with NSEARCHTHREADS=1: variants in compareByEpsilon order, each seeded from
// the last earlier one with minpts >= its own. Reference border rule only;
// eps must be finite.
int cluster_reuse(Context &c, Points &P, const double *eps, const uint32_t *minpts,
                  uint32_t nvar, uint32_t *out, int precision, int cells, bool verbose);

// One reused variant for the C++ classes (DBScan::assignPointsToPredefinedCluster):
// device arrays x, y (n points, caller order), host: the reused variant's labels by
// rank (the binned order of src/main.cpp:459-605) and its cluster count + 1;
// writes this variant's labels by rank and *ncl = clusters + 1. old == NULL
// clusters from scratch. Takes the device's lock.
int reuse_labels(const double *x, const double *y, uint32_t n, double eps, uint32_t minpts,
                 const uint32_t *old, uint32_t old_ncl, uint32_t *lab, uint32_t *ncl);

} // namespace vdb
