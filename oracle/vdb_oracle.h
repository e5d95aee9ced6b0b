/*
 * CPU restatement of VariantDBSCAN's per-variant clustering semantics.
 *
 * TEST INFRASTRUCTURE ONLY. This oracle is the checker for the HIP library in
 * variantdbscan_amd/: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it. The product never links or calls it.
 *
 * Every function cites the reference code it restates (paths relative to the
 * MITHaystack/VariantDBSCAN repository root).
 *
 * Pinning (see the header of vdb_oracle.c for the full list):
 *   - vdbo_bin_order      : against oracle/_ref (mapping recovered from labels) and
 *                           the 225-point golden CSV; hand-computed tiny cases.
 *   - vdbo_threshold      : closed form check sqrt(T)<=eps<sqrt(nextafter(T)).
 *   - vdbo_neighbor_csr   : brute force O(N^2) on small N.
 *   - vdbo_dbscan_literal : src/test_output.csv (both variants, bit-exact) and
 *                           oracle/_ref on seeded datasets.
 *   - vdbo_dbscan_closed  : equals vdbo_dbscan_literal on every test dataset.
 */
#ifndef VDB_ORACLE_H
#define VDB_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* binDataset restatement (src/main.cpp:459-605). mapping[k] = caller index of
 * the point at binned position k. Returns 0, or -1 when a bin is out of range
 * (the reference prints an error and returns early, leaving mapping short). */
int vdbo_bin_order(const double *x, const double *y, uint32_t n, int num_bins,
                   int32_t *mapping);

/* T(eps) = max{ s >= 0 : sqrt_rn(s) <= eps }. The reference accepts a pair
 * when sqrt(dx*dx+dy*dy) <= eps (src/DBScan.cpp:687,699-702); since sqrt_rn is
 * monotone this is s <= T(eps). Returns -1.0 when no s qualifies (eps<0, NaN). */
double vdbo_threshold(double eps);

/* s = fl(fl(dx*dx)+fl(dy*dy)), dx=fl(px-qx), dy=fl(py-qy), no contraction
 * (src/DBScan.cpp:699-702). */
double vdbo_dist2(double px, double py, double qx, double qy);

/* eps-neighbour table in caller index order, self included, ids ascending.
 * offsets has n+1 entries; ids must hold offsets[n] entries (call with
 * ids==NULL first to get the size: returns total pair count, or -1 on error). */
int64_t vdbo_neighbor_csr(const double *x, const double *y, uint32_t n,
                          double eps, int64_t *offsets, int32_t *ids);

/* Per-point neighbour counts (self included). */
int vdbo_neighbor_counts(const double *x, const double *y, uint32_t n,
                         double eps, uint32_t *counts);

/* Literal restatement of DBScan::algDBScanParallel (src/DBScan.cpp:523-609)
 * over points already in binned order. labels[k] for binned position k,
 * 0 = noise. Returns the reference's clusterCnt (clusters + 1). */
int vdbo_dbscan_literal(const double *xb, const double *yb, uint32_t n,
                        double eps, uint32_t minpts, uint32_t *labels);

/* Closed form of the same labels: core components numbered by the rank of
 * their minimum binned index; a non-core point takes the cluster of its
 * smallest adjacent root r only when r < its own binned index. */
int vdbo_dbscan_closed(const double *xb, const double *yb, uint32_t n,
                       double eps, uint32_t minpts, uint32_t *labels);

/* The same labels as vdbo_libVDBSCAN(use_closed=1) for the nmp variants
 * (eps, minpts[j]) of one eps, without storing a neighbour table: one count
 * pass and one union pass over the pairs serve every minpts value, memory is
 * O(n * nmp). labels[j*n + i] = label of caller point i under minpts[j]. For
 * workloads whose pair count rules out a table (or the reference's own
 * expansion stack, which holds every core point's neighbour list). */
int vdbo_levels_closed(const double *x, const double *y, uint32_t n, double eps,
                       const uint32_t *minpts, uint32_t nmp, uint32_t *labels);

/* Full libVDBSCAN restatement (src/main.cpp:50-299) with every variant
 * clustered from scratch (the reference's answer for a single-variant call).
 * retArr[v*n + i] = label of caller point i under caller variant v.
 * use_closed selects vdbo_dbscan_closed instead of the literal expansion. */
int vdbo_libVDBSCAN(const double *x, const double *y, uint32_t n,
                    const double *eps, const uint32_t *minpts, uint32_t nvar,
                    int mbb_size, uint32_t *retArr, int use_closed);

/* Multi-variant libVDBSCAN with the reference's greedy cluster reuse, as the
 * reference runs it with one search thread (NSEARCHTHREADS=1, where it is
 * deterministic): variants in compareByEpsilon order, each one seeded from
 * the last finished variant with minpts >= its own (Schedule::schedGreedy,
 * src/schedule.cpp:68-116), clustered by assignPointsToPredefinedCluster
 * (src/DBScan.cpp:272-416, 157-205, 226-267, 429-518), or from scratch when
 * there is none. Literal restatement (vdb_oracle_reuse.cpp). retArr layout
 * as vdbo_libVDBSCAN. Returns 0, or < 0 on error. */
int vdbo_libVDBSCAN_reuse(const double *x, const double *y, uint32_t n, const double *eps,
                          const uint32_t *minpts, uint32_t nvar, uint32_t *retArr);

/* Grid over points for other oracle files: every j with
 * vdbo_dist2(i, j) <= T, written to out (n entries) when non-NULL. */
void *vdbo_grid_new(const double *x, const double *y, uint32_t n, double eps);
int64_t vdbo_grid_neighbors(const void *g, const double *x, const double *y, uint32_t i,
                            double T, int32_t *out);
void vdbo_grid_free(void *g);

#ifdef __cplusplus
}
#endif
#endif
