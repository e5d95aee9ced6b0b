/*
 * CPU restatement (oracle) of VariantDBSCAN's per-variant clustering.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, never part of it.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * the shared object built from this file (oracle/liboracle.so).
 *
 * Build: gcc -O2 -ffp-contract=off (oracle/Makefile). FP contraction must stay
 * off: the reference evaluates sqrt((x1-x2)*(x1-x2)+(y1-y2)*(y1-y2)) with
 * separately rounded products (src/DBScan.cpp:699-702).
 *
 * How each function is pinned (tests/test_oracle.py):
 *   vdbo_bin_order       tiny hand-worked inputs (incl. the DBL_MIN max quirk of
 *                        src/main.cpp:467-468); the permutation implied by the
 *                        reference's labels on the golden CSV; oracle/_ref runs.
 *   vdbo_threshold       closed-form property sqrt(T)<=eps<sqrt(next(T)) and
 *                        the special cases eps=0, eps<0, NaN, inf.
 *   vdbo_dist2 / csr     brute force over all pairs with the reference formula
 *                        sqrt(dx*dx+dy*dy)<=eps evaluated directly (no T).
 *   vdbo_dbscan_literal  src/test_output.csv, both variants, bit-exact;
 *                        oracle/_ref single-variant runs on seeded datasets.
 *   vdbo_dbscan_closed   equal to vdbo_dbscan_literal on every dataset tested.
 *   vdbo_libVDBSCAN      oracle/_ref single-variant calls, bit-exact labels.
 *   vdbo_levels_closed   equal to vdbo_libVDBSCAN (literal and closed) on the
 *                        test datasets; equal to oracle/_ref on the 500,000-point
 *                        64-variant and 1,000,000-point skewed workloads
 *                        (tests/golden/make_full_digests.py checks it there).
 *
 * Known divergence from the reference, by construction: the reference only
 * distance-tests points whose R-tree box group (MBBsize consecutive binned
 * points, src/main.cpp:404-436) overlaps the query box [x-eps,x+eps]x[y-eps,y+eps]
 * (src/DBScan.cpp:722-731, src/RTree.h:1613-1626). A pair can pass the distance
 * test while its whole box group misses the query box only when |dx| or |dy|
 * lies within a few ulps of eps; this oracle (and the library) accept such a
 * pair. No test dataset contains one.
 */
#include "vdb_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- helpers */

/* The reference stores minpts in an int (src/structs.h:37, src/DBScan.cpp:69)
 * and compares it with vector::size() (src/DBScan.cpp:552,587): a value above
 * INT_MAX wraps negative, converts to a huge size_t, and no point is core. */
static int is_core(uint64_t count, uint32_t minpts)
{
    if (minpts > 0x7fffffffu) return 0;
    return count >= (uint64_t)minpts;
}

double vdbo_dist2(double px, double py, double qx, double qy)
{
    double dx = px - qx;
    double dy = py - qy;
    double a = dx * dx;
    double b = dy * dy;
    return a + b;
}

double vdbo_threshold(double eps)
{
    if (!(eps >= 0.0)) return -1.0;           /* negative or NaN: nothing passes */
    if (isinf(eps)) return INFINITY;
    /* largest finite s whose bit pattern satisfies sqrt(s) <= eps; positive
     * doubles order like their bit patterns, so bisect on the pattern. */
    uint64_t lo = 0, hi = 0x7fefffffffffffffull; /* 0 .. DBL_MAX */
    double dhi;
    memcpy(&dhi, &hi, 8);
    if (sqrt(dhi) <= eps) return dhi;
    /* invariant: sqrt(lo) <= eps, sqrt(hi) > eps */
    while (hi - lo > 1) {
        uint64_t mid = lo + (hi - lo) / 2;
        double dm;
        memcpy(&dm, &mid, 8);
        if (sqrt(dm) <= eps) lo = mid; else hi = mid;
    }
    double t;
    memcpy(&t, &lo, 8);
    return t;
}

/* ------------------------------------------------------------ binDataset */

int vdbo_bin_order(const double *x, const double *y, uint32_t n, int num_bins,
                   int32_t *mapping)
{
    /* src/main.cpp:465-488: min starts at DBL_MAX, max at DBL_MIN (the smallest
     * positive double, not -DBL_MAX). */
    double min_x = DBL_MAX, min_y = DBL_MAX, max_x = DBL_MIN, max_y = DBL_MIN;
    for (uint32_t i = 0; i < n; i++) {
        if (x[i] < min_x) min_x = x[i];
        if (x[i] > max_x) max_x = x[i];
        if (y[i] < min_y) min_y = y[i];
        if (y[i] > max_y) max_y = y[i];
    }
    double x_width = (max_x - min_x) * 1.01;   /* src/main.cpp:491-492 */
    double y_width = (max_y - min_y) * 1.01;
    double xoffset = (min_x > 0) ? -min_x : fabs(min_x);   /* :513-527 */
    double yoffset = (min_y > 0) ? -min_y : fabs(min_y);

    int64_t nb = (int64_t)num_bins * num_bins;
    int32_t *key = (int32_t *)malloc(sizeof(int32_t) * (n ? n : 1));
    int64_t *cnt = (int64_t *)calloc((size_t)nb + 1, sizeof(int64_t));
    if (!key || !cnt) { free(key); free(cnt); return -2; }
    for (uint32_t i = 0; i < n; i++) {
        double fx = floor(((x[i] + xoffset) / x_width) * num_bins);   /* :539 */
        double fy = floor(((y[i] + yoffset) / y_width) * num_bins);   /* :540 */
        /* :559-563 range check; a NaN or out-of-int value is out of range */
        if (!(fx >= 0.0 && fx <= (double)(num_bins - 1)) ||
            !(fy >= 0.0 && fy <= (double)(num_bins - 1))) {
            free(key); free(cnt);
            return -1;
        }
        key[i] = (int32_t)fx * num_bins + (int32_t)fy;
        cnt[key[i] + 1]++;
    }
    /* :581-595: visit bins row-major over (xbin, ybin), points in insertion
     * order inside a bin -> a stable counting sort on key. */
    for (int64_t b = 0; b < nb; b++) cnt[b + 1] += cnt[b];
    for (uint32_t i = 0; i < n; i++) mapping[cnt[key[i]]++] = (int32_t)i;
    free(key);
    free(cnt);
    return 0;
}

/* ------------------------------------------------------ neighbour search */

/* Uniform grid, cell side cs >= eps, 5x5 stencil. Any accepted pair has
 * |dx|,|dy| <= eps*(1+2^-50) < 2*cs, so their computed cells differ by at most
 * 2 per axis even after rounding of the cell computation. */
typedef struct {
    double minx, miny, cs;
    int64_t ncx, ncy;
    int64_t *start;    /* ncx*ncy+1 */
    int32_t *order;    /* point ids sorted by cell, stable */
    int64_t *cell;     /* cell of each point */
} grid_t;

static void grid_free(grid_t *g)
{
    free(g->start);
    free(g->order);
    free(g->cell);
}

/* cells of side max(eps / k, range / cap); a query then scans R =
 * ceil(eps / side) + 1 cells each way (grid_radius) */
static int grid_build_k(const double *x, const double *y, uint32_t n, double eps, int k,
                        double cap, grid_t *g);

static int grid_build(const double *x, const double *y, uint32_t n, double eps,
                      grid_t *g)
{
    return grid_build_k(x, y, n, eps, 1, 2048.0, g);
}

static int grid_build_k(const double *x, const double *y, uint32_t n, double eps, int k,
                        double cap, grid_t *g)
{
    memset(g, 0, sizeof(*g));
    double minx = INFINITY, maxx = -INFINITY, miny = INFINITY, maxy = -INFINITY;
    for (uint32_t i = 0; i < n; i++) {
        if (x[i] < minx) minx = x[i];
        if (x[i] > maxx) maxx = x[i];
        if (y[i] < miny) miny = y[i];
        if (y[i] > maxy) maxy = y[i];
    }
    if (n == 0) { minx = maxx = miny = maxy = 0.0; }
    double range = fmax(maxx - minx, maxy - miny);
    double cs = fmax(eps / k, range / cap);
    g->minx = minx;
    g->miny = miny;
    if (!(cs > 0.0) || !isfinite(cs) || !isfinite(range)) {
        g->cs = 1.0;
        g->ncx = g->ncy = 1;
    } else {
        g->cs = cs;
        g->ncx = (int64_t)floor((maxx - minx) / cs) + 2;
        g->ncy = (int64_t)floor((maxy - miny) / cs) + 2;
    }
    int64_t nc = g->ncx * g->ncy;
    g->start = (int64_t *)calloc((size_t)nc + 1, sizeof(int64_t));
    g->order = (int32_t *)malloc(sizeof(int32_t) * (n ? n : 1));
    g->cell = (int64_t *)malloc(sizeof(int64_t) * (n ? n : 1));
    if (!g->start || !g->order || !g->cell) { grid_free(g); return -2; }
    for (uint32_t i = 0; i < n; i++) {
        int64_t cx = 0, cy = 0;
        if (g->ncx > 1) {
            cx = (int64_t)floor((x[i] - minx) / g->cs);
            cy = (int64_t)floor((y[i] - miny) / g->cs);
            if (cx < 0) cx = 0;
            if (cx >= g->ncx) cx = g->ncx - 1;
            if (cy < 0) cy = 0;
            if (cy >= g->ncy) cy = g->ncy - 1;
        }
        g->cell[i] = cx * g->ncy + cy;
        g->start[g->cell[i] + 1]++;
    }
    for (int64_t c = 0; c < nc; c++) g->start[c + 1] += g->start[c];
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)nc);
    if (!fill) { grid_free(g); return -2; }
    memcpy(fill, g->start, sizeof(int64_t) * (size_t)nc);
    for (uint32_t i = 0; i < n; i++) g->order[fill[g->cell[i]]++] = (int32_t)i;
    free(fill);
    return 0;
}

/* Visit every j with dist2(i,j) <= T; callback-free: writes ids to out (if
 * non-NULL) and returns the count. */
static int64_t grid_query_r(const grid_t *g, const double *x, const double *y,
                            uint32_t i, double T, int64_t R, int32_t *out)
{
    if (T < 0.0) return 0;
    int64_t cx = g->cell[i] / g->ncy, cy = g->cell[i] % g->ncy;
    int64_t k = 0;
    for (int64_t ax = cx - R; ax <= cx + R; ax++) {
        if (ax < 0 || ax >= g->ncx) continue;
        for (int64_t ay = cy - R; ay <= cy + R; ay++) {
            if (ay < 0 || ay >= g->ncy) continue;
            int64_t c = ax * g->ncy + ay;
            for (int64_t p = g->start[c]; p < g->start[c + 1]; p++) {
                int32_t j = g->order[p];
                if (vdbo_dist2(x[i], y[i], x[j], y[j]) <= T) {
                    if (out) out[k] = j;
                    k++;
                }
            }
        }
    }
    return k;
}

static int64_t grid_query(const grid_t *g, const double *x, const double *y,
                          uint32_t i, double T, int32_t *out)
{
    return grid_query_r(g, x, y, i, T, 2, out);
}

/* window radius covering eps on a grid of side cs: an accepted pair has
 * |dx| <= eps (1 + 2^-50) and floor() moves a coordinate by less than one
 * cell, so ceil(eps / cs) + 1 cells each way hold every neighbour */
static int64_t grid_radius(const grid_t *g, double eps)
{
    if (g->ncx <= 1 && g->ncy <= 1) return 1;
    return (int64_t)ceil(eps / g->cs) + 1;
}

static int cmp_i32(const void *a, const void *b)
{
    int32_t u = *(const int32_t *)a, v = *(const int32_t *)b;
    return (u > v) - (u < v);
}

int64_t vdbo_neighbor_csr(const double *x, const double *y, uint32_t n,
                          double eps, int64_t *offsets, int32_t *ids)
{
    double T = vdbo_threshold(eps);
    grid_t g;
    if (grid_build(x, y, n, eps, &g)) return -1;
    int64_t total = 0;
    if (offsets) offsets[0] = 0;
    for (uint32_t i = 0; i < n; i++) {
        int64_t c = grid_query(&g, x, y, i, T, ids ? ids + total : NULL);
        if (ids) qsort(ids + total, (size_t)c, sizeof(int32_t), cmp_i32);
        total += c;
        if (offsets) offsets[i + 1] = total;
    }
    grid_free(&g);
    return total;
}

int vdbo_neighbor_counts(const double *x, const double *y, uint32_t n,
                         double eps, uint32_t *counts)
{
    double T = vdbo_threshold(eps);
    grid_t g;
    if (grid_build(x, y, n, eps, &g)) return -1;
    for (uint32_t i = 0; i < n; i++)
        counts[i] = (uint32_t)grid_query(&g, x, y, i, T, NULL);
    grid_free(&g);
    return 0;
}

/* ------------------------------------------------- DBSCAN, literal form */

int vdbo_dbscan_literal(const double *xb, const double *yb, uint32_t n,
                        double eps, uint32_t minpts, uint32_t *labels)
{
    double T = vdbo_threshold(eps);
    grid_t g;
    if (grid_build(xb, yb, n, eps, &g)) return -1;
    unsigned char *visited = (unsigned char *)calloc(n ? n : 1, 1);
    int32_t *nb = (int32_t *)malloc(sizeof(int32_t) * (n ? n : 1));
    size_t cap = 1024, top = 0;
    int32_t *stack = (int32_t *)malloc(sizeof(int32_t) * cap);
    if (!visited || !nb || !stack) {
        free(visited); free(nb); free(stack); grid_free(&g);
        return -2;
    }
    for (uint32_t i = 0; i < n; i++) labels[i] = 0;
    uint32_t cluster_cnt = 0;
    for (uint32_t i = 0; i < n; i++) {                 /* src/DBScan.cpp:530 */
        if (visited[i]) continue;                      /* :533-535 */
        visited[i] = 1;                                /* :542 */
        int64_t c = grid_query(&g, xb, yb, i, T, nb);  /* :547 */
        if (!is_core((uint64_t)c, minpts)) {           /* :552-554 */
            labels[i] = 0;
            continue;
        }
        cluster_cnt++;                                 /* :558-560 */
        labels[i] = cluster_cnt;
        top = 0;                                       /* :565 neighborSet=N(i) */
        for (int64_t k = 0; k < c; k++) {
            if (top == cap) { cap *= 2; stack = (int32_t *)realloc(stack, sizeof(int32_t) * cap); }
            stack[top++] = nb[k];
        }
        while (top) {                                  /* :568 */
            int32_t p = stack[top - 1];
            if (visited[p]) { top--; continue; }       /* :573-577 */
            visited[p] = 1;                            /* :580 */
            int64_t cp = grid_query(&g, xb, yb, (uint32_t)p, T, nb);  /* :584 */
            if (is_core((uint64_t)cp, minpts)) {       /* :587-590 append */
                for (int64_t k = 0; k < cp; k++) {
                    if (top == cap) { cap *= 2; stack = (int32_t *)realloc(stack, sizeof(int32_t) * cap); }
                    stack[top++] = nb[k];
                }
            }
            if (labels[p] == 0) labels[p] = cluster_cnt;   /* :592-594 */
        }
    }
    free(visited);
    free(nb);
    free(stack);
    grid_free(&g);
    return (int)cluster_cnt + 1;                       /* :604 */
}

/* ------------------------------------------------- DBSCAN, closed form */

static int32_t uf_find(int32_t *parent, int32_t a)
{
    while (parent[a] != a) {
        parent[a] = parent[parent[a]];
        a = parent[a];
    }
    return a;
}

int vdbo_dbscan_closed(const double *xb, const double *yb, uint32_t n,
                       double eps, uint32_t minpts, uint32_t *labels)
{
    double T = vdbo_threshold(eps);
    grid_t g;
    if (grid_build(xb, yb, n, eps, &g)) return -1;
    unsigned char *core = (unsigned char *)calloc(n ? n : 1, 1);
    int32_t *parent = (int32_t *)malloc(sizeof(int32_t) * (n ? n : 1));
    int32_t *nb = (int32_t *)malloc(sizeof(int32_t) * (n ? n : 1));
    uint32_t *cid = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));
    if (!core || !parent || !nb || !cid) {
        free(core); free(parent); free(nb); free(cid); grid_free(&g);
        return -2;
    }
    for (uint32_t i = 0; i < n; i++) {
        core[i] = (unsigned char)is_core((uint64_t)grid_query(&g, xb, yb, i, T, NULL), minpts);
        parent[i] = (int32_t)i;
    }
    /* core components; the root is kept at the minimum binned index */
    for (uint32_t i = 0; i < n; i++) {
        if (!core[i]) continue;
        int64_t c = grid_query(&g, xb, yb, i, T, nb);
        for (int64_t k = 0; k < c; k++) {
            int32_t j = nb[k];
            if (!core[j]) continue;
            int32_t a = uf_find(parent, (int32_t)i), b = uf_find(parent, j);
            if (a == b) continue;
            if (a < b) parent[b] = a; else parent[a] = b;
        }
    }
    /* number clusters by ascending root index: the reference opens clusters in
     * binned visit order, and a component opens at its first core point */
    uint32_t ncl = 0;
    for (uint32_t i = 0; i < n; i++)
        if (core[i] && uf_find(parent, (int32_t)i) == (int32_t)i) cid[i] = ++ncl;
    for (uint32_t i = 0; i < n; i++) {
        if (core[i]) {
            labels[i] = cid[uf_find(parent, (int32_t)i)];
            continue;
        }
        /* border: the first cluster to reach i is the one with the smallest
         * root; it claims i only if it opens before the outer loop reaches i */
        int64_t c = grid_query(&g, xb, yb, i, T, nb);
        int32_t best = -1;
        for (int64_t k = 0; k < c; k++) {
            int32_t j = nb[k];
            if (!core[j]) continue;
            int32_t r = uf_find(parent, j);
            if (best < 0 || r < best) best = r;
        }
        labels[i] = (best >= 0 && (uint32_t)best < i) ? cid[best] : 0;
    }
    free(core);
    free(parent);
    free(nb);
    free(cid);
    grid_free(&g);
    return (int)ncl + 1;
}

/* ------------------------------------- several minpts at one eps, closed */

int vdbo_levels_closed(const double *x, const double *y, uint32_t n, double eps,
                       const uint32_t *minpts, uint32_t nmp, uint32_t *labels)
{
    if (n == 0 || nmp == 0) return 0;
    double T = vdbo_threshold(eps);
    int32_t *mapping = (int32_t *)malloc(sizeof(int32_t) * n);
    double *xb = (double *)malloc(sizeof(double) * n);
    double *yb = (double *)malloc(sizeof(double) * n);
    uint32_t *cnt = (uint32_t *)malloc(sizeof(uint32_t) * n);
    int32_t *nb = (int32_t *)malloc(sizeof(int32_t) * n);
    int32_t *parent = (int32_t *)malloc(sizeof(int32_t) * (size_t)n * nmp);
    uint32_t *cid = (uint32_t *)malloc(sizeof(uint32_t) * n);
    grid_t g;
    memset(&g, 0, sizeof g);
    int rc = 0;
    if (!mapping || !xb || !yb || !cnt || !nb || !parent || !cid) { rc = -2; goto done; }
    if (vdbo_bin_order(x, y, n, 300, mapping)) { rc = -1; goto done; }   /* src/main.cpp:107 */
    for (uint32_t k = 0; k < n; k++) {
        xb[k] = x[mapping[k]];
        yb[k] = y[mapping[k]];
    }
    if (grid_build_k(xb, yb, n, eps, 4, 8192.0, &g)) { rc = -2; goto done; }
    int64_t R = grid_radius(&g, eps);
    /* neighbour counts, self included (src/DBScan.cpp:547) */
    for (uint32_t i = 0; i < n; i++) cnt[i] = (uint32_t)grid_query_r(&g, xb, yb, i, T, R, NULL);
    for (size_t t = 0; t < (size_t)n * nmp; t++) parent[t] = (int32_t)(t % n);
    /* core components of every minpts value in one pass over the pairs; a
     * root is its component's minimum binned index */
    for (uint32_t i = 0; i < n; i++) {
        int anycore = 0;
        for (uint32_t j = 0; j < nmp; j++) anycore |= is_core(cnt[i], minpts[j]);
        if (!anycore) continue;
        int64_t c = grid_query_r(&g, xb, yb, i, T, R, nb);
        for (uint32_t j = 0; j < nmp; j++) {
            if (!is_core(cnt[i], minpts[j])) continue;
            int32_t *par = parent + (size_t)j * n;
            for (int64_t k = 0; k < c; k++) {
                int32_t q = nb[k];
                /* the predicate is symmetric: each unordered pair once */
                if (q <= (int32_t)i || !is_core(cnt[q], minpts[j])) continue;
                int32_t a = uf_find(par, (int32_t)i), b = uf_find(par, q);
                if (a == b) continue;
                if (a < b) par[b] = a; else par[a] = b;
            }
        }
    }
    for (uint32_t j = 0; j < nmp; j++) {
        int32_t *par = parent + (size_t)j * n;
        uint32_t *out = labels + (size_t)j * n;
        uint32_t ncl = 0;
        for (uint32_t i = 0; i < n; i++)
            if (is_core(cnt[i], minpts[j]) && uf_find(par, (int32_t)i) == (int32_t)i)
                cid[i] = ++ncl;
        for (uint32_t i = 0; i < n; i++) {
            uint32_t lab;
            if (is_core(cnt[i], minpts[j])) {
                lab = cid[uf_find(par, (int32_t)i)];
            } else {
                /* border: smallest adjacent root, only if it opens before i */
                int64_t c = grid_query_r(&g, xb, yb, i, T, R, nb);
                int32_t best = -1;
                for (int64_t k = 0; k < c; k++) {
                    int32_t q = nb[k];
                    if (!is_core(cnt[q], minpts[j])) continue;
                    int32_t r = uf_find(par, q);
                    if (best < 0 || r < best) best = r;
                }
                lab = (best >= 0 && (uint32_t)best < i) ? cid[best] : 0;
            }
            out[mapping[i]] = lab;                               /* src/main.cpp:255-262 */
        }
    }
done:
    free(mapping);
    free(xb);
    free(yb);
    free(cnt);
    free(nb);
    free(parent);
    free(cid);
    grid_free(&g);
    return rc;
}

/* ------------------------------------- grid handle for other oracle files */

void *vdbo_grid_new(const double *x, const double *y, uint32_t n, double eps)
{
    grid_t *g = (grid_t *)malloc(sizeof(grid_t));
    if (!g) return NULL;
    if (grid_build(x, y, n, eps, g)) {
        free(g);
        return NULL;
    }
    return g;
}

int64_t vdbo_grid_neighbors(const void *g, const double *x, const double *y, uint32_t i,
                            double T, int32_t *out)
{
    return grid_query((const grid_t *)g, x, y, i, T, out);
}

void vdbo_grid_free(void *g)
{
    if (!g) return;
    grid_free((grid_t *)g);
    free(g);
}

/* --------------------------------------------------------- libVDBSCAN */

typedef struct {
    double eps;
    int minpts;
    uint32_t vid;
} experiment_t;

/* compareByEpsilon (src/main.cpp:612-620) as a strict weak order */
static int cmp_experiment(const void *a, const void *b)
{
    const experiment_t *u = (const experiment_t *)a, *v = (const experiment_t *)b;
    if (u->eps == v->eps) {
        if (u->minpts != v->minpts) return u->minpts > v->minpts ? -1 : 1;
        return (u->vid > v->vid) - (u->vid < v->vid);
    }
    return u->eps < v->eps ? -1 : 1;
}

int vdbo_libVDBSCAN(const double *x, const double *y, uint32_t n,
                    const double *eps, const uint32_t *minpts, uint32_t nvar,
                    int mbb_size, uint32_t *retArr, int use_closed)
{
    if (mbb_size <= 0) return -3;                 /* src/main.cpp:118 divides by it */
    if (n == 0 || nvar == 0) return 0;
    experiment_t *ex = (experiment_t *)malloc(sizeof(experiment_t) * nvar);
    int32_t *mapping = (int32_t *)malloc(sizeof(int32_t) * n);
    double *xb = (double *)malloc(sizeof(double) * n);
    double *yb = (double *)malloc(sizeof(double) * n);
    uint32_t *lb = (uint32_t *)malloc(sizeof(uint32_t) * n);
    int rc = 0;
    if (!ex || !mapping || !xb || !yb || !lb) { rc = -2; goto done; }
    for (uint32_t v = 0; v < nvar; v++) {          /* :65-75 */
        ex[v].eps = eps[v];
        ex[v].minpts = (int)minpts[v];
        ex[v].vid = v;
    }
    qsort(ex, nvar, sizeof(experiment_t), cmp_experiment);   /* :79 */
    if (vdbo_bin_order(x, y, n, 300, mapping)) { rc = -1; goto done; }  /* :107 */
    for (uint32_t k = 0; k < n; k++) {
        xb[k] = x[mapping[k]];
        yb[k] = y[mapping[k]];
    }
    for (uint32_t v = 0; v < nvar; v++) {
        int r = use_closed
                    ? vdbo_dbscan_closed(xb, yb, n, ex[v].eps, minpts[ex[v].vid], lb)
                    : vdbo_dbscan_literal(xb, yb, n, ex[v].eps, minpts[ex[v].vid], lb);
        if (r < 0) { rc = r; goto done; }
        uint32_t *out = retArr + (size_t)ex[v].vid * n;    /* :273-284 */
        for (uint32_t k = 0; k < n; k++) out[mapping[k]] = lb[k];   /* :255-262 */
    }
done:
    free(ex);
    free(mapping);
    free(xb);
    free(yb);
    free(lb);
    return rc;
}
