/*
 * VariantDBSCAN on AMD Instinct MI355X (gfx950): C ABI of libSharedVDBSCAN.so.
 *
 * The first entry point is the reference library's own symbol, unchanged, so
 * existing callers (ctypes wrappers, dlopen harnesses) load this library in
 * its place. The vdb_* entry points expose the same engine with device
 * pointers, statistics and the neighbour table.
 *
 * All entry points are synchronous, return 0 on success and a nonzero
 * VDB_E* code on failure (vdb_last_error() describes it), and serialise on
 * an internal per-device mutex, so concurrent callers are safe.
 */
#ifndef VDBSCAN_H
#define VDBSCAN_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VDB_OK 0
#define VDB_EINVAL 1       /* bad argument */
#define VDB_ENONFINITE 2   /* NaN or infinite coordinate */
#define VDB_EDEVICE 3      /* HIP runtime failure */
#define VDB_ENOMEM 4       /* allocation failure */
#define VDB_EUNSUPPORTED 5 /* request outside what this build handles */

/*
 * Replaces libVDBSCAN of the reference (src/main.cpp:50; prototype also in
 * src/c_test_prog.h:26; bound by src/variantdbscan.py:66-69).
 * retArr[v*datasetSize + i] = cluster of point i under variant v, 0 = noise,
 * clusters numbered 1.. per variant. Every variant is clustered as the
 * reference clusters a variant from scratch (bit-identical to a one-variant
 * reference call). MBBsize only shaped the reference's R-tree; it must be > 0
 * and is otherwise ignored. verbose prints progress to stdout.
 * Extension: returns nonzero on invalid input (the reference always returns 0).
 */
int libVDBSCAN(double *inputx, double *inputy, unsigned int datasetSize,
               double *inputEpsilon, unsigned int *inputMinpts,
               unsigned int numVariants, int MBBsize, unsigned int *retArr,
               bool verbose);

/* ---------------------------------------------------------------------- */
/* Extended API (no counterpart symbol in the reference).                  */

/* flags */
#define VDB_DEVICE_PTRS 1u /* point / output arrays are device pointers */
/* Multi-variant calls: the labels of the reference's greedy cluster reuse
 * (src/main.cpp:170-251, src/schedule.cpp:68-116, src/DBScan.cpp:157-518) as
 * the reference gives them with one search thread (NSEARCHTHREADS=1, where
 * its schedule is deterministic): variants in compareByEpsilon order, each
 * one seeded from the last earlier variant with minpts >= its own. Without
 * the flag every variant gets its one-variant (scratch) labels. Reference
 * border rule and finite eps only; vdb_cluster only. */
#define VDB_REUSE_REFERENCE 2u

/* path selection for vdb_cluster */
#define VDB_PATH_AUTO 0    /* batched engine when applicable, else fused */
#define VDB_PATH_FUSED 1   /* per variant: count, union-find, border; no table */
#define VDB_PATH_TABLE 2   /* per variant: count/scan/fill CSR table, union-find over it */
#define VDB_PATH_BATCH 3   /* all variants from one distance-sorted truncated table */

#define VDB_PREC_FP64 0      /* every distance test in binary64 */
#define VDB_PREC_FP32_RECHECK 1 /* float32 screening, binary64 inside the error band */

/* Labelling of vdb_cluster / vdb_cluster_multi.
 * VDB_BORDER_REFERENCE: the reference's labels (0 = noise, clusters 1.. in
 *   its visit order, a border point joins a cluster only when that cluster
 *   was found before the outer loop reached the point).
 * VDB_BORDER_STANDARD: textbook DBSCAN labels as int32 values in the uint32
 *   array: -1 (0xffffffff) = noise, clusters 0..k-1 numbered in ascending
 *   order of their smallest core point's caller index, and every non-core
 *   point within eps of a core point joins the lowest-numbered such cluster.
 *   These are the labels of scikit-learn's DBSCAN for the same eps and
 *   min_samples = minpts (it visits points in index order). Core points and
 *   clusters are the same in both modes. No counterpart in the reference;
 *   every path computes it. */
#define VDB_BORDER_REFERENCE 0
#define VDB_BORDER_STANDARD 1

typedef struct vdb_options {
    int path;                   /* VDB_PATH_* */
    int precision;              /* VDB_PREC_* */
    uint64_t table_budget_bytes;/* 0 = most of the free device memory */
    int verbose;
    int cells_per_eps;          /* per-variant grid: cells per eps (1..8), 0 = library's choice */
    int border;                 /* VDB_BORDER_* */
} vdb_options;

typedef struct vdb_stats {
    double ms_total;       /* device time of the whole call (events) */
    double ms_prepare;     /* binning + grids */
    double ms_count;       /* neighbour counting / truncated table */
    double ms_scan;
    double ms_fill;
    double ms_uf;          /* union-find */
    double ms_border;
    double ms_labels;
    uint64_t pairs;        /* eps-neighbour pairs (table path, last variant) */
    uint64_t rechecked;    /* pairs decided by the binary64 recheck */
    uint64_t candidates;   /* distance tests (when measured) */
    uint32_t tiles;        /* table tiles */
    int path_used;
} vdb_stats;

const char *vdb_last_error(void);
int vdb_version(void);

/* Launch all work on this hipStream_t (NULL: the library's own stream). */
int vdb_set_stream(void *hip_stream);
/* Free the pooled device memory of the current device. */
int vdb_release(void);

/* Times the launches of one kernel of the library (its name, e.g. "k_knn",
 * "k_sweep", "k_fill"; NULL or "" stops) with a pair of HIP events on the
 * stream it runs on; vdb_kernel_time returns the summed duration in ms and
 * the launch count since the last read (or since vdb_time_kernel). For
 * roofline figures of the kernel over a measured region. */
int vdb_time_kernel(const char *name);
int vdb_kernel_time(double *ms, uint64_t *launches);

/* Shape of the last batched run on this device, for byte models of its
 * kernels: info[0] points, [1] eps levels, [2] minpts values, [3] list length
 * (largest minpts), [4] points with a truncated list, [5] fine-grid cells,
 * [6] cell pairs, [7] grid shape chosen by the autotuner, [8] requested
 * (eps level, minpts) snapshots, and the union-find worklists summed over
 * levels and minpts values: [9] pairs entering at their level, [10] pairs
 * carried from an earlier level, [11] test records (with unused chunk
 * slots); [9]-[11] are 0 when the minpts values ran in several passes.
 * count entries are written (0 beyond these). No counterpart in the
 * reference. */
int vdb_batch_info(uint64_t *info, uint32_t count);

/* Autotuner state as JSON: {"cache": "<file>", "tuned_in_process": N,
 * "entries": {"<device>:<workload key>": choice, ...}}, the choices read from
 * the tuning cache file and those made in this process (the per-variant
 * grid's cells per eps, the batched engine's grid shape). Writes at most cap
 * bytes (NUL-terminated) and sets *len to the full length. No counterpart in
 * the reference. */
int vdb_tuning(char *buf, uint64_t cap, uint64_t *len);

/* Cluster n points under nvar variants. labels has nvar*n entries laid out
 * as retArr of libVDBSCAN. flags & VDB_DEVICE_PTRS: x, y, labels are device
 * pointers (eps and minpts are always host arrays). opt/stats may be NULL. */
int vdb_cluster(const double *x, const double *y, uint32_t n, const double *eps,
                const uint32_t *minpts, uint32_t nvar, uint32_t *labels,
                uint32_t flags, const vdb_options *opt, vdb_stats *stats);

/* Allocation-free variant of vdb_cluster with VDB_DEVICE_PTRS. No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid allocation). x, y and labels are device pointers, eps and minpts
 * host arrays, labels in retArr layout. Every variant runs on the per-variant
 * fused engine (opt->path VDB_PATH_AUTO or VDB_PATH_FUSED; either border rule
 * and precision; cells_per_eps 0 means 2, no autotuner). All other device
 * memory is a static array of 128 MiB per GPU inside the library, held until
 * the process ends: at most 262,144 points per call (more returns
 * VDB_EUNSUPPORTED, as does a call whose grids outgrow the array). The call
 * allocates no device or pinned memory, creates no events or graphs, and
 * waits for its work before returning. Calls on one GPU share the array
 * (they are serialised; overlapping calls would race on it). stats gets
 * path_used only. Make any other library call once on the device first (it
 * creates the library's stream). */
int vdb_cluster_noalloc(const double *x, const double *y, uint32_t n, const double *eps,
                        const uint32_t *minpts, uint32_t nvar, uint32_t *labels,
                        const vdb_options *opt, vdb_stats *stats);

/* Many independent datasets in one call: dataset d is points
 * [offsets[d], offsets[d+1]) of x, y (offsets has nsets+1 entries, offsets[0]
 * = 0), every dataset clustered under the same nvar variants. Labels of
 * dataset d start at labels + offsets[d]*nvar and are laid out as that
 * dataset's own retArr (nvar rows of its point count). No counterpart in the
 * reference, which takes one dataset per libVDBSCAN call. */
int vdb_cluster_multi(const double *x, const double *y, const uint64_t *offsets,
                      uint32_t nsets, const double *eps, const uint32_t *minpts,
                      uint32_t nvar, uint32_t *labels, uint32_t flags,
                      const vdb_options *opt, vdb_stats *stats);

/* Reference visit order: mapping[k] = caller index of the point at binned
 * position k (src/main.cpp:459-605, 300x300 bins). */
int vdb_rank_order(const double *x, const double *y, uint32_t n, uint32_t *mapping,
                   uint32_t flags);

/* counts[i] = number of points within eps of point i, itself included
 * (the neighbour-set size of src/DBScan.cpp:547). */
int vdb_neighbor_counts(const double *x, const double *y, uint32_t n, double eps,
                        uint32_t *counts, uint32_t flags, int precision,
                        uint64_t *rechecked);

/* eps-neighbour table in CSR form, caller order: offsets[n+1], ids[total],
 * optional d2[total] (float32 squared distances). Ids inside a row are in
 * grid order, not sorted. Replaces the GPU prototype's table construction
 * (src/gpu/GPU.cu:3479-4557 + :4598-4742). If *total exceeds capacity,
 * nothing beyond offsets is written and the call returns 0 with *total set:
 * call again with buffers of that size. */
int vdb_neighbor_table(const double *x, const double *y, uint32_t n, double eps,
                       uint64_t *offsets, uint32_t *ids, float *d2, uint64_t capacity,
                       uint64_t *total, uint32_t flags, int precision, vdb_stats *stats);

/* Order-independent digest of the eps-neighbour table of vdb_neighbor_table
 * (caller order): sum over every (row i, id j) pair, self pairs included, of
 * splitmix64((i << 32) | j), mod 2^64, computed on the device without
 * storing the table. *pairs = table size, *rechecked = pairs decided by the
 * binary64 recheck (VDB_PREC_FP32_RECHECK). No counterpart in the reference:
 * it checks tables too large to hold twice (e.g. FP32+recheck vs FP64). */
int vdb_neighbor_table_digest(const double *x, const double *y, uint32_t n, double eps,
                              uint32_t flags, int precision, uint64_t *digest,
                              uint64_t *pairs, uint64_t *rechecked);

/* ---------------------------------------------------------------------- */
/* Dataset and variant files of the reference's callers (host only: no GPU
 * is touched). Output arrays are malloc'd by the library; release them with
 * vdb_free. Fields parse like atof/atoi (C strtod/strtol), so values equal
 * the reference's; lines without a comma are skipped. */

/* x,y CSV, one point per line (replaces the import loop of the reference's
 * harness, src/c_test_prog.cpp:86-104). */
int vdb_read_xy_csv(const char *path, double **x, double **y, uint32_t *n);

/* lon,lat[,TEC] CSV of the GPU prototype (replaces importDataset,
 * src/gpu/import_dataset.cpp:63-134): x = latitude moved to [0, 180]
 * (90 -> 0, -90 -> 180) with the prototype's arithmetic, y = longitude. */
int vdb_read_lonlat_csv(const char *path, double **x, double **y, uint32_t *n);

/* eps,minpts CSV, one variant per line (replaces importDBScanInstances,
 * src/gpu/import_dataset.cpp:232-262, without its sort: file order is kept). */
int vdb_read_variants_csv(const char *path, double **eps, uint32_t **minpts,
                          uint32_t *nvar);

/* Labels as CSV, one line per point: x,y,label_0,...,label_{nvar-1} (the
 * layout of the reference's src/test_output.csv); labels in retArr layout.
 * path "-" writes to stdout. No counterpart symbol in the reference. */
int vdb_write_labels_csv(const char *path, const double *x, const double *y, uint32_t n,
                         const uint32_t *labels, uint32_t nvar);

/* Releases an array returned by the readers above. */
void vdb_free(void *p);

#ifdef __cplusplus
}
#endif
#endif
