// Point layout in HBM.
//
//   caller order  : x[i], y[i]                      (as handed in)
//   rank order    : xb[r], yb[r], mapping[r]         r = position in the reference's
//                                                    binned visit order
//   cell order    : xs[p], ys[p], pos_rank[p]        p = position after a stable sort by
//                                                    grid cell; inside a cell, ranks ascend
//
// A Grid is the cell-order view for one cell side. Only non-empty cells are
// stored (sorted keys + CSR starts), so the footprint is O(n) whatever eps is.
#pragma once

#include "vdb_context.h"

#include <vector>

namespace vdb {

// T(eps) = max{ s >= 0 : sqrt_rn(s) <= eps }, -1 when nothing qualifies.
double threshold_for_eps(double eps);

// Points of one or several independent datasets. Dataset d occupies
// [seg[d], seg[d+1]) both in caller order and in rank order (the binned order
// is computed per dataset); ranks are positions in the concatenated order.
struct Points {
    uint32_t n = 0;
    const double *x = nullptr, *y = nullptr; // caller order (device)
    double *xb = nullptr, *yb = nullptr;     // rank order
    uint32_t *mapping = nullptr;             // rank -> caller index
    double ext[4] = {0, 0, 0, 0};            // extent of all points
    uint32_t nsets = 1;
    std::vector<uint32_t> hseg;              // nsets+1 (host)
    uint32_t *seg = nullptr;                 // nsets+1 (device)
    std::vector<double> hext;                // 4 per dataset (host)
};

// Reference binning (src/main.cpp:459-605) on the device, per dataset.
// P.hseg must be set (nsets+1 host offsets); fills the rank order of P.
// Returns kErrUnsupported when a bin is out of range.
int build_rank_order(Context &c, Points &P, const char *tag);
// One dataset covering all points.
void single_segment(Points &P);

struct Grid {
    double w = 0, gx0 = 0, gy0 = 0;
    uint32_t ncx = 0, ncy = 0; // key = cx*ncy + cy
    uint32_t ncell = 0;        // non-empty cells
    int R = 1;                 // window radius in cells
    double *xs = nullptr, *ys = nullptr;
    uint32_t *pos_rank = nullptr;   // [n]
    uint32_t *rank_pos = nullptr;   // [n] inverse of pos_rank
    uint32_t *cell_of_pos = nullptr;// [n]
    uint32_t *cell_key = nullptr;   // [ncell]
    uint32_t *cell_start = nullptr; // [ncell+1]
    double *cell_box = nullptr;     // [ncell*4] x0,x1,y0,y1 of the cell's points
    uint32_t *row_lo = nullptr;     // [ncell*(2R+1)] first cell of window row
    uint32_t *row_hi = nullptr;     // [ncell*(2R+1)] one past the last
    // with want_row_pts set before build_grid: (first point, point count)
    // of every window row, beside row_lo/row_hi
    bool want_row_pts = false;
    uint2 *row_pts = nullptr;       // [ncell*(2R+1)]
    // after classify_rows (per-variant grids): inside every window row,
    // cells [row_f0, row_f1) are within eps of the whole centre cell (every
    // pair accepted without a test); row_lo/row_hi are then trimmed to cells
    // that can hold a neighbour at all. Null until classified.
    uint32_t *row_f0 = nullptr;
    uint32_t *row_f1 = nullptr;
    uint32_t npts = 0;              // points in the grid
    // per dataset: grid origin, dimensions and first key (keys of dataset d
    // are kbase[d] + cx*ncy[d] + cy)
    uint32_t nsets = 1;
    double *sgx0 = nullptr, *sgy0 = nullptr;
    uint32_t *sncx = nullptr, *sncy = nullptr;
    uint32_t *skbase = nullptr;     // nsets+1
    uint32_t *cell_ds = nullptr;    // [ncell] dataset of each cell
};

// dataset of a position/rank (seg has nsets+1 entries, seg[0] = 0)
__device__ __forceinline__ uint32_t seg_of(const uint32_t *seg, uint32_t nsets, uint32_t p)
{
    uint32_t lo = 0, hi = nsets; // largest d with seg[d] <= p
    while (hi - lo > 1) {
        uint32_t m = (lo + hi) >> 1;
        if (seg[m] <= p) lo = m; else hi = m;
    }
    return lo;
}

// Build a grid of cell side >= w_min over the rank-ordered points. If the key
// space would exceed 32 bits the side is enlarged; *enlarged reports it.
// tag prefixes the pool slot names so several grids can coexist.
int build_grid(Context &c, const Points &P, double w_min, int R, const char *tag,
               Grid &G, bool *enlarged);

// The largest cell population of G (waits for the stream; verbose output).
int grid_max_cell(Context &c, const Grid &G, uint32_t *out);

// Classify the window rows of G for the threshold T (see Grid::row_f0).
int classify_rows(Context &c, Grid &G, double T, const char *tag);

} // namespace vdb
