// Oracle: the reference's multi-variant libVDBSCAN with greedy cluster reuse,
// as it runs with one search thread.
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP library's reuse-emulation
// mode, never part of the library. Built into oracle/liboracle.so.
//
// With NSEARCHTHREADS=1 the reference's OpenMP loop (src/main.cpp:212-251)
// runs its iterations in order on one thread, so Schedule::schedGreedy
// (src/schedule.cpp:68-116) hands out the variants in sorted order and every
// earlier variant has finished: variant e reuses the last variant i < e with
// minpts_i >= minpts_e, or is clustered from scratch when there is none. This
// file restates that control flow and DBScan::assignPointsToPredefinedCluster
// literally, with the oracle's exact eps-predicate in place of the R-trees:
//   * neighbour sets: every point within eps (src/DBScan.cpp:627-702; the
//     R-tree box prefilter divergence of vdb_oracle.c applies here too);
//   * the high-resolution tree query over a cluster's eps-grown box
//     (src/DBScan.cpp:346-350): every point inside the box, bounds included
//     (src/RTree.h:1613-1626 on point rectangles, src/main.cpp:443-450).
// Orders that only change the order of a stack (R-tree result order) do not
// change any label: within one expansion every point reached gets the same
// cluster id. std::sort is called with the reference's comparators on the
// reference's input orders, so ties resolve as they do in the reference
// (same C++ library).
#include "vdb_oracle.h"

#include <algorithm>
#include <cstdlib>
#include <vector>

namespace {

// src/structs.h:35-39
struct Experiment {
    double epsilon;
    int minpts;
    unsigned int variantID;
};

// src/main.cpp:612-620
bool compare_by_epsilon(const Experiment &a, const Experiment &b)
{
    if (a.epsilon == b.epsilon) return a.minpts > b.minpts;
    return a.epsilon < b.epsilon;
}

// src/structs.h:61-64
struct DensityStruct {
    int clusterID;
    double density;
};

// DBScan::comparedensityStructfn (src/DBScan.cpp:210-213)
bool compare_density(const DensityStruct &a, const DensityStruct &b)
{
    return a.density > b.density;
}

const int kNumMinReusePts = 100; // NUMMINREUSEPTS, src/DBScan.cpp:40

// one DBScan instance over the binned points
struct Instance {
    double eps;
    int minpts;
    std::vector<int> clusterIDs; // binned order, 0 = noise
    int clusterCnt = 0;          // getDBScanNumClusters(): clusters + 1 when finished
};

struct Points {
    const double *x, *y;
    uint32_t n;
};

struct Neigh {
    const Points &P;
    void *grid;
    double T;
    std::vector<int32_t> buf;
    Neigh(const Points &p, double eps) : P(p), grid(nullptr), T(vdbo_threshold(eps)), buf(p.n)
    {
        grid = vdbo_grid_new(p.x, p.y, p.n, eps);
    }
    ~Neigh() { vdbo_grid_free(grid); }
    // getNeighboursParallelMPB (src/DBScan.cpp:627-660)
    void query(uint32_t i, std::vector<int> &out)
    {
        int64_t c = vdbo_grid_neighbors(grid, P.x, P.y, i, T, buf.data());
        out.assign(buf.begin(), buf.begin() + c);
    }
};

// setIDsInDist.size() >= minPts with an int minPts (src/DBScan.cpp:552,587):
// a negative value converts to a huge size_t, and no point is core
bool is_core(size_t count, int minpts) { return count >= (size_t)minpts; }

// generateMBBAroundCluster + appendMBBByEpsilon (src/DBScan.cpp:110-147)
void cluster_box(const Points &P, const std::vector<int> &pts, double eps, double *mn,
                 double *mx)
{
    int id = pts[0];
    mn[0] = mx[0] = P.x[id];
    mn[1] = mx[1] = P.y[id];
    for (int q : pts) {
        if (P.x[q] < mn[0]) mn[0] = P.x[q];
        if (P.x[q] > mx[0]) mx[0] = P.x[q];
        if (P.y[q] < mn[1]) mn[1] = P.y[q];
        if (P.y[q] > mx[1]) mx[1] = P.y[q];
    }
    mn[0] = mn[0] - eps;
    mx[0] = mx[0] + eps;
    mn[1] = mn[1] - eps;
    mx[1] = mx[1] + eps;
}

// DBScan::algDBScanParallel (src/DBScan.cpp:523-609)
void from_scratch(const Points &P, Instance &I)
{
    std::vector<uint32_t> lab(P.n);
    int r = vdbo_dbscan_literal(P.x, P.y, P.n, I.eps, (uint32_t)I.minpts, lab.data());
    I.clusterIDs.assign(lab.begin(), lab.end());
    I.clusterCnt = r;
}

// DBScan::DBScanParallelMPBSTUMP (src/DBScan.cpp:157-205)
void stump(Neigh &N, Instance &I, const Instance &old, std::vector<char> &visited,
           std::vector<char> &destroyed, const std::vector<int> &seeds)
{
    std::vector<int> set = seeds, nb;
    while (!set.empty()) {
        int p = set.back();
        if (visited[p]) {
            set.pop_back();
            continue;
        }
        visited[p] = 1;
        N.query((uint32_t)p, nb);
        int oldc = old.clusterIDs[p];
        if (oldc != 0) destroyed[oldc] = 1;
        if (is_core(nb.size(), I.minpts)) set.insert(set.end(), nb.begin(), nb.end());
        if (I.clusterIDs[p] == 0) I.clusterIDs[p] = I.clusterCnt;
    }
}

// DBScan::algDBScanParallelReuseClusterResults (src/DBScan.cpp:429-518)
void remaining(Neigh &N, Instance &I, std::vector<char> &visited)
{
    std::vector<int> set, nb;
    for (uint32_t i = 0; i < N.P.n; i++) {
        if (visited[i]) continue;
        set.clear();
        visited[i] = 1;
        N.query(i, nb);
        if (!is_core(nb.size(), I.minpts)) {
            I.clusterIDs[i] = 0;
            continue;
        }
        I.clusterCnt++;
        I.clusterIDs[i] = I.clusterCnt;
        set = nb;
        while (!set.empty()) {
            int p = set.back();
            if (visited[p]) {
                set.pop_back();
                continue;
            }
            visited[p] = 1;
            N.query((uint32_t)p, nb);
            if (is_core(nb.size(), I.minpts)) set.insert(set.end(), nb.begin(), nb.end());
            if (I.clusterIDs[p] == 0) I.clusterIDs[p] = I.clusterCnt;
        }
    }
    I.clusterCnt++; // cluster 0 is the noise cluster (src/DBScan.cpp:514-517)
}

// DBScan::assignPointsToPredefinedCluster (src/DBScan.cpp:272-416)
void reuse(const Points &P, Instance &I, const Instance &old)
{
    Neigh N(P, I.eps);
    int nother = old.clusterCnt;
    I.clusterCnt = 0;
    I.clusterIDs.assign(P.n, 0);
    std::vector<char> visited(P.n, 0), destroyed(nother > 0 ? nother : 1, 0);
    std::vector<std::vector<int>> clusterArr(nother > 0 ? nother : 1);
    for (uint32_t i = 0; i < P.n; i++) clusterArr[old.clusterIDs[i]].push_back((int)i);
    for (int c = 1; c < nother; c++) std::sort(clusterArr[c].begin(), clusterArr[c].end());
    // setClusterScheduleDensity (src/DBScan.cpp:226-267)
    std::vector<DensityStruct> dens;
    for (int c = 1; c < nother; c++) {
        double mn[2], mx[2];
        cluster_box(P, clusterArr[c], I.eps, mn, mx);
        if ((int)clusterArr[c].size() < kNumMinReusePts) continue;
        DensityStruct d;
        d.clusterID = c;
        d.density = (1.0 * clusterArr[c].size()) / ((mx[0] - mn[0]) * (mx[1] - mn[1]));
        dens.push_back(d);
    }
    std::sort(dens.begin(), dens.end(), compare_density);
    std::vector<int> nb;
    // the densest cluster (schedule[0]) is never reused (src/DBScan.cpp:311)
    for (size_t s = 1; s < dens.size(); s++) {
        int cID = dens[s].clusterID;
        if (destroyed[cID]) continue;
        double mn[2], mx[2];
        cluster_box(P, clusterArr[cID], I.eps, mn, mx);
        I.clusterCnt++;
        for (int q : clusterArr[cID]) {
            I.clusterIDs[q] = I.clusterCnt;
            visited[q] = 1;
        }
        const std::vector<int> &cl = clusterArr[cID];
        std::vector<int> outside; // candidatesNotInCluster
        for (uint32_t q = 0; q < P.n; q++) {
            bool inbox = !(mn[0] > P.x[q] || P.x[q] > mx[0] || mn[1] > P.y[q] || P.y[q] > mx[1]);
            if (inbox && !std::binary_search(cl.begin(), cl.end(), (int)q)) outside.push_back((int)q);
        }
        std::vector<int> seeds; // candidatesToGrowFromInCluster
        for (int q : outside) {
            N.query((uint32_t)q, nb);
            for (int k : nb) {
                if (std::binary_search(cl.begin(), cl.end(), k)) {
                    seeds.push_back(k);
                    visited[k] = 0;
                }
            }
        }
        stump(N, I, old, visited, destroyed, seeds);
    }
    remaining(N, I, visited);
}

} // namespace

extern "C" int vdbo_libVDBSCAN_reuse(const double *x, const double *y, uint32_t n,
                                     const double *eps, const uint32_t *minpts, uint32_t nvar,
                                     uint32_t *retArr)
{
    if (n == 0 || nvar == 0) return 0;
    // src/main.cpp:65-79: experiments in caller order, then sorted
    std::vector<Experiment> ex(nvar);
    for (uint32_t v = 0; v < nvar; v++) {
        ex[v].epsilon = eps[v];
        ex[v].minpts = (int)minpts[v];
        ex[v].variantID = v;
    }
    std::sort(ex.begin(), ex.end(), compare_by_epsilon);
    std::vector<int32_t> mapping(n);
    if (vdbo_bin_order(x, y, n, 300, mapping.data())) return -1; // src/main.cpp:107
    std::vector<double> xb(n), yb(n);
    for (uint32_t k = 0; k < n; k++) {
        xb[k] = x[mapping[k]];
        yb[k] = y[mapping[k]];
    }
    Points P{xb.data(), yb.data(), n};
    std::vector<Instance> inst(nvar);
    for (uint32_t e = 0; e < nvar; e++) {
        inst[e].eps = ex[e].epsilon;
        inst[e].minpts = ex[e].minpts;
        // schedGreedy with one thread: the last finished variant with
        // minpts >= this one's (src/schedule.cpp:104-110)
        int src = -1;
        for (int i = (int)e - 1; i >= 0; i--)
            if (ex[i].minpts >= ex[e].minpts) {
                src = i;
                break;
            }
        if (src < 0) from_scratch(P, inst[e]);
        else reuse(P, inst[e], inst[src]);
    }
    // src/main.cpp:254-284
    for (uint32_t e = 0; e < nvar; e++) {
        uint32_t *out = retArr + (size_t)ex[e].variantID * n;
        for (uint32_t k = 0; k < n; k++) out[mapping[k]] = (uint32_t)inst[e].clusterIDs[k];
    }
    return 0;
}
