// Multi-variant calls with the reference's greedy cluster reuse (opt-in).
//
// The reference's multi-variant libVDBSCAN clusters its variants in
// compareByEpsilon order and seeds each one from a finished variant with
// minpts >= its own (Schedule::schedGreedy, src/schedule.cpp:68-116) through
// DBScan::assignPointsToPredefinedCluster (src/DBScan.cpp:272-416, helpers
// :157-267, remainder :429-518). Which variant is finished depends on thread
// timing, except with one search thread, where variant e reuses the last
// variant i < e with minpts_i >= minpts_e. This file reproduces that
// deterministic schedule.
//
// Closed form of one reused variant (eps, minpts) from an old labelling:
// old core points are core again (eps grew, minpts did not), so every old
// cluster C sits in one new core component N(C). T(C) = {N(C)} plus the
// components that have a core point within eps of a non-core point of C.
// Old clusters of at least 100 points are visited in descending density of
// their eps-grown boxes, the densest skipped (src/DBScan.cpp:226-267, 311).
// Visiting C, in that order: if a component of T(C) was claimed before, C
// is skipped (the reference marks it destroyed when an earlier expansion
// reaches one of its points); otherwise C takes the next label L, and so do
// all points of C, every core point of N(C) (N(C) is claimed) and every
// not yet labelled non-core point within eps of N(C). So a non-core point
// outside the visited clusters takes the label of the earliest claimed
// component it touches. What is left is clustered as from scratch with the
// claimed components and their labelled points removed: components numbered
// L+1.. by their smallest rank, a non-core point joining the smallest root
// below its own rank (src/DBScan.cpp:429-518 keeps the scratch pass's border
// rule). The GPU computes core flags, the core forest, the per-cluster
// sizes, boxes, components and T(C) lists; the host orders the clusters and
// walks the claims (a few thousand clusters), as the reference's schedule
// does; the GPU writes the labels.
#include "vdb_reuse.h"

#include "vdb_cluster.h"
#include "vdb_prims.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <vector>

namespace vdb {

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kNone = 0xffffffffu;
constexpr uint32_t kMinReusePts = 100; // NUMMINREUSEPTS, src/DBScan.cpp:40

// order-preserving uint64 image of a double (no NaN here: inputs are finite)
__device__ __forceinline__ unsigned long long okey(double d)
{
    unsigned long long u = (unsigned long long)__double_as_longlong(d);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

double okey_decode(unsigned long long k)
{
    unsigned long long u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    double d;
    memcpy(&d, &u, sizeof d);
    return d;
}

// Per old cluster C (label at rank order `old`): point count, coordinate
// box, and the root of its new core points (all of them share it). cl[p] =
// C for positions whose old cluster has at least kMinReusePts points (the
// clusters the schedule can visit), else 0, for the T(C) pass.
__global__ void k_old_stats(const uint32_t *old, const uint32_t *pos_rank, const double *xs,
                            const double *ys, const uint8_t *core, const uint32_t *parent,
                            uint32_t n, uint32_t *size, unsigned long long *box,
                            uint32_t *nroot)
{
    uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    uint32_t C = old[pos_rank[p]];
    if (C == 0) return;
    atomicAdd(size + C, 1u);
    unsigned long long kx = okey(xs[p]), ky = okey(ys[p]);
    atomicMin(box + 4 * (size_t)C + 0, kx);
    atomicMax(box + 4 * (size_t)C + 1, kx);
    atomicMin(box + 4 * (size_t)C + 2, ky);
    atomicMax(box + 4 * (size_t)C + 3, ky);
    if (core[p]) nroot[C] = parent[p];
}

__global__ void k_init_stats(uint32_t *size, unsigned long long *box, uint32_t *nroot,
                             uint32_t ncl)
{
    uint32_t C = blockIdx.x * kBlock + threadIdx.x;
    if (C >= ncl) return;
    size[C] = 0;
    box[4 * (size_t)C + 0] = ~0ull;
    box[4 * (size_t)C + 1] = 0ull;
    box[4 * (size_t)C + 2] = ~0ull;
    box[4 * (size_t)C + 3] = 0ull;
    nroot[C] = kNone;
}

__global__ void k_big_cluster_of_pos(const uint32_t *old, const uint32_t *pos_rank,
                                     const uint32_t *size, uint32_t n, uint32_t *cl)
{
    uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    uint32_t C = old[pos_rank[p]];
    cl[p] = (C != 0 && size[C] >= kMinReusePts) ? C : 0u;
}

__global__ void k_fill(uint32_t *a, uint32_t v, uint32_t n)
{
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) a[i] = v;
}

// claim[root position] = label of the visiting cluster that claimed it
__global__ void k_claims(const uint32_t *pairs, uint32_t m, uint32_t *claim)
{
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < m) claim[pairs[2 * i]] = pairs[2 * i + 1];
}

// Unclaimed roots: their rank as the border key (NONE elsewhere), and a flag
// at their rank for the numbering of the remainder's clusters.
__global__ void k_unclaimed(const uint8_t *core, const uint32_t *parent, const uint32_t *claim,
                            const uint32_t *pos_rank, uint32_t n, uint32_t *rankkey,
                            uint32_t *flag_by_rank)
{
    uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    bool root = core[p] && parent[p] == p && claim[p] == kNone;
    rankkey[p] = root ? pos_rank[p] : kNone;
    flag_by_rank[pos_rank[p]] = root ? 1u : 0u;
}

// Labels of the variant, by rank (lab) and in caller order (out).
__global__ void k_reuse_labels(const uint32_t *old, const uint32_t *procL, const uint8_t *core,
                               const uint32_t *parent, const uint32_t *claim,
                               const uint32_t *best_claim, const uint32_t *best_rank,
                               const uint32_t *fid_by_rank, const uint32_t *pos_rank,
                               const uint32_t *mapping, uint32_t nvisited, uint32_t n,
                               uint32_t *lab, uint32_t *out)
{
    uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    uint32_t r = pos_rank[p];
    uint32_t C = old ? old[r] : 0u;
    uint32_t L = 0;
    if (C != 0 && procL[C] != 0) {
        L = procL[C]; // a visited old cluster keeps all its points
    } else if (core[p]) {
        uint32_t root = parent[p];
        L = claim[root] != kNone ? claim[root] : nvisited + 1 + fid_by_rank[pos_rank[root]];
    } else if (best_claim[p] != kNone) {
        L = best_claim[p]; // earliest claimed component within eps
    } else {
        uint32_t b = best_rank[p]; // smallest unclaimed root within eps
        L = (b != kNone && b < r) ? nvisited + 1 + fid_by_rank[b] : 0u;
    }
    lab[r] = L;
    if (out) out[mapping[r]] = L;
}

// src/structs.h:35-39 (only the fields the ordering reads)
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

// src/structs.h:61-64 and DBScan::comparedensityStructfn (src/DBScan.cpp:210-213)
struct DensityStruct {
    int clusterID;
    double density;
};
bool compare_density(const DensityStruct &a, const DensityStruct &b)
{
    return a.density > b.density;
}

// The claim walk on the host. Returns the number of visited clusters; fills
// procL[C] (label of visited cluster C, 0 otherwise) and the claimed roots.
uint32_t walk_claims(double eps, uint32_t ncl, const std::vector<uint32_t> &size,
                     const std::vector<unsigned long long> &box,
                     const std::vector<uint32_t> &nroot,
                     const std::vector<std::vector<uint32_t>> &touch, std::vector<uint32_t> &procL,
                     std::vector<uint32_t> &claims)
{
    // setClusterScheduleDensity (src/DBScan.cpp:226-267): clusters 1.. in
    // order, their boxes grown by eps, those of at least 100 points kept
    std::vector<DensityStruct> dens;
    for (uint32_t C = 1; C < ncl; C++) {
        if (size[C] < kMinReusePts) continue;
        double mn0 = okey_decode(box[4 * C + 0]) - eps, mx0 = okey_decode(box[4 * C + 1]) + eps;
        double mn1 = okey_decode(box[4 * C + 2]) - eps, mx1 = okey_decode(box[4 * C + 3]) + eps;
        DensityStruct d;
        d.clusterID = (int)C;
        d.density = (1.0 * size[C]) / ((mx0 - mn0) * (mx1 - mn1));
        dens.push_back(d);
    }
    std::sort(dens.begin(), dens.end(), compare_density);
    std::unordered_map<uint32_t, uint32_t> claimed;
    procL.assign(ncl, 0);
    uint32_t L = 0;
    for (size_t s = 1; s < dens.size(); s++) { // the densest is never reused
        uint32_t C = (uint32_t)dens[s].clusterID;
        bool destroyed = false;
        for (uint32_t r : touch[C])
            if (claimed.count(r)) {
                destroyed = true;
                break;
            }
        if (destroyed) continue;
        L++;
        procL[C] = L;
        if (nroot[C] != kNone) {
            claimed[nroot[C]] = L;
            claims.push_back(nroot[C]);
            claims.push_back(L);
        }
    }
    return L;
}

// One variant (eps, minpts) of the schedule over the grid G at eps with its
// neighbour counts: labels by rank into lab (device) and, when out is not
// null, in caller order into out (device). old (device, by rank, may be
// null) holds the labels of the variant it reuses, old_ncl that variant's
// getDBScanNumClusters(); null or a variant without clusters clusters from
// scratch. *ncl = clusters + 1; *nvisited = old clusters reused.
int reuse_step(Context &c, const Points &P, const Grid &G, const Items &I,
               const uint32_t *counts, double ev, uint32_t mp, const uint32_t *old,
               uint32_t old_ncl, uint32_t *lab, uint32_t *out, uint32_t *ncl,
               uint32_t *nvisited_out)
{
    uint32_t n = P.n;
    unsigned g = ceil_div(n, kBlock);
    uint32_t *parent, *best, *best2, *claim, *rankkey, *flag, *fid, *cl, *cnt, *nroots_dev;
    uint8_t *core;
    VDB_TRY(c.get("ru.core", n, &core));
    VDB_TRY(c.get("ru.parent", n, &parent));
    VDB_TRY(c.get("ru.best", n, &best));
    VDB_TRY(c.get("ru.best2", n, &best2));
    VDB_TRY(c.get("ru.claim", n, &claim));
    VDB_TRY(c.get("ru.rankkey", n, &rankkey));
    VDB_TRY(c.get("ru.flag", n, &flag));
    VDB_TRY(c.get("ru.fid", n, &fid));
    VDB_TRY(c.get("ru.cl", n, &cl));
    VDB_TRY(c.get("ru.cnt", 1, &cnt));
    VDB_TRY(c.get("ru.nroots", 1, &nroots_dev));
    VDB_TRY(core_forest(c, G, I, ev, mp, counts, core, parent, best));
    k_fill<<<g, kBlock, 0, c.stream()>>>(claim, kNone, n);
    VDB_HIP(hipGetLastError());
    uint32_t *procL_dev = nullptr;
    uint32_t nvisited = 0;
    if (old && old_ncl > 1) {
        uint32_t nc = old_ncl;
        uint32_t *size, *nroot;
        unsigned long long *box;
        VDB_TRY(c.get("ru.size", nc, &size));
        VDB_TRY(c.get("ru.nroot", nc, &nroot));
        VDB_TRY(c.get("ru.box", 4 * (size_t)nc, &box));
        VDB_TRY(c.get("ru.procL", nc, &procL_dev));
        k_init_stats<<<ceil_div(nc, kBlock), kBlock, 0, c.stream()>>>(size, box, nroot, nc);
        VDB_HIP(hipGetLastError());
        k_old_stats<<<g, kBlock, 0, c.stream()>>>(old, G.pos_rank, G.xs, G.ys, core, parent, n,
                                                  size, box, nroot);
        VDB_HIP(hipGetLastError());
        k_big_cluster_of_pos<<<g, kBlock, 0, c.stream()>>>(old, G.pos_rank, size, n, cl);
        VDB_HIP(hipGetLastError());
        // T(C) pairs: counted first, then stored
        uint32_t npairs = 0;
        VDB_TRY(border_pairs(c, G, I, ev, core, parent, cl, nullptr, 0, cnt));
        VDB_HIP(hipMemcpyAsync(&npairs, cnt, 4, hipMemcpyDeviceToHost, c.stream()));
        VDB_HIP(hipStreamSynchronize(c.stream()));
        uint32_t *pairs;
        VDB_TRY(c.get("ru.pairs", 2 * (size_t)(npairs ? npairs : 1), &pairs));
        VDB_TRY(border_pairs(c, G, I, ev, core, parent, cl, pairs, npairs, cnt));
        std::vector<uint32_t> hsize(nc), hnroot(nc), hp(2 * (size_t)npairs);
        std::vector<unsigned long long> hbox(4 * (size_t)nc);
        VDB_HIP(hipMemcpyAsync(hsize.data(), size, 4 * (size_t)nc, hipMemcpyDeviceToHost,
                               c.stream()));
        VDB_HIP(hipMemcpyAsync(hnroot.data(), nroot, 4 * (size_t)nc, hipMemcpyDeviceToHost,
                               c.stream()));
        VDB_HIP(hipMemcpyAsync(hbox.data(), box, 32 * (size_t)nc, hipMemcpyDeviceToHost,
                               c.stream()));
        if (npairs)
            VDB_HIP(hipMemcpyAsync(hp.data(), pairs, 8 * (size_t)npairs, hipMemcpyDeviceToHost,
                                   c.stream()));
        VDB_HIP(hipStreamSynchronize(c.stream()));
        std::vector<std::vector<uint32_t>> touch(nc);
        for (uint32_t C = 1; C < nc; C++)
            if (hnroot[C] != kNone) touch[C].push_back(hnroot[C]);
        for (uint32_t i = 0; i < npairs; i++) touch[hp[2 * i]].push_back(hp[2 * i + 1]);
        std::vector<uint32_t> procL, claims;
        nvisited = walk_claims(ev, nc, hsize, hbox, hnroot, touch, procL, claims);
        VDB_HIP(hipMemcpyAsync(procL_dev, procL.data(), 4 * (size_t)nc, hipMemcpyHostToDevice,
                               c.stream()));
        if (!claims.empty()) {
            uint32_t *dcl;
            VDB_TRY(c.get("ru.claims", claims.size(), &dcl));
            VDB_HIP(hipMemcpyAsync(dcl, claims.data(), 4 * claims.size(), hipMemcpyHostToDevice,
                                   c.stream()));
            uint32_t m = (uint32_t)(claims.size() / 2);
            k_claims<<<ceil_div(m, kBlock), kBlock, 0, c.stream()>>>(dcl, m, claim);
            VDB_HIP(hipGetLastError());
        }
        // the host vectors above are read by the copies: wait for them
        VDB_HIP(hipStreamSynchronize(c.stream()));
    } else {
        old = nullptr; // from scratch: nothing to reuse
    }
    k_unclaimed<<<g, kBlock, 0, c.stream()>>>(core, parent, claim, G.pos_rank, n, rankkey, flag);
    VDB_HIP(hipGetLastError());
    VDB_TRY(scan_exclusive_u32(c, flag, fid, n, nroots_dev));
    VDB_TRY(border_min_key(c, G, I, ev, core, parent, claim, best));
    VDB_TRY(border_min_key(c, G, I, ev, core, parent, rankkey, best2));
    k_reuse_labels<<<g, kBlock, 0, c.stream()>>>(old, procL_dev, core, parent, claim, best, best2,
                                                fid, G.pos_rank, P.mapping, nvisited, n, lab,
                                                out);
    VDB_HIP(hipGetLastError());
    uint32_t nroots = 0;
    VDB_HIP(hipMemcpyAsync(&nroots, nroots_dev, 4, hipMemcpyDeviceToHost, c.stream()));
    VDB_HIP(hipStreamSynchronize(c.stream()));
    *ncl = nvisited + nroots + 1;
    *nvisited_out = nvisited;
    return kOk;
}

// grid and neighbour counts at eps (one per distinct eps of a schedule)
int reuse_grid(Context &c, const Points &P, double eps, int precision, int cells, Grid &G,
               Items &I, uint32_t *counts)
{
    // the FP32 screening counts its binary64 rechecks here (not reported)
    unsigned long long *recheck;
    VDB_TRY(c.get("ru.recheck", 1, &recheck));
    VDB_HIP(hipMemsetAsync(recheck, 0, sizeof(unsigned long long), c.stream()));
    int k = 0;
    VDB_TRY(tune_cells_per_eps(c, P, eps, precision, cells, &k));
    VDB_TRY(build_eps_grid(c, P, eps, k, "g1", G));
    VDB_TRY(build_items(c, G, "g1", I));
    VDB_TRY(count_neighbors(c, G, I, eps, precision, counts, (uint64_t *)recheck));
    return kOk;
}

} // namespace

int cluster_reuse(Context &c, Points &P, const double *eps, const uint32_t *minpts,
                  uint32_t nvar, uint32_t *out, int precision, int cells, bool verbose)
{
    uint32_t n = P.n;
    for (uint32_t v = 0; v < nvar; v++)
        if (!std::isfinite(eps[v])) {
            set_error("the reuse emulation needs finite eps values");
            return kErrInvalid;
        }
    VDB_TRY(build_rank_order(c, P, "pts"));
    // src/main.cpp:65-79: experiments in caller order, then sorted
    std::vector<Experiment> ex(nvar);
    for (uint32_t v = 0; v < nvar; v++) {
        ex[v].epsilon = eps[v];
        ex[v].minpts = (int)minpts[v];
        ex[v].variantID = v;
    }
    std::sort(ex.begin(), ex.end(), compare_by_epsilon);
    uint32_t *labs, *counts;
    VDB_TRY(c.get("ru.labs", (size_t)nvar * n, &labs)); // rank-order labels per sorted variant
    VDB_TRY(c.get("ru.counts", n, &counts));
    std::vector<uint32_t> ncl(nvar, 1); // getDBScanNumClusters() per sorted variant
    Grid G;
    Items I;
    double grid_eps = NAN;
    for (uint32_t e = 0; e < nvar; e++) {
        double ev = ex[e].epsilon;
        if (!(ev == grid_eps)) { // one grid and count pass per distinct eps
            VDB_TRY(reuse_grid(c, P, ev, precision, cells, G, I, counts));
            grid_eps = ev;
        }
        // schedGreedy with one search thread: the last finished variant
        // with minpts >= this one's (src/schedule.cpp:104-110)
        int src = -1;
        for (int i = (int)e - 1; i >= 0; i--)
            if (ex[i].minpts >= ex[e].minpts) {
                src = i;
                break;
            }
        uint32_t nvisited = 0;
        VDB_TRY(reuse_step(c, P, G, I, counts, ev, (uint32_t)ex[e].minpts,
                           src >= 0 ? labs + (size_t)src * n : nullptr, src >= 0 ? ncl[src] : 0,
                           labs + (size_t)e * n, out + (size_t)ex[e].variantID * n, &ncl[e],
                           &nvisited));
        if (verbose)
            printf("\n[vdbscan] reuse: variant %u (eps %g, minpts %d) %s%d, %u clusters reused, "
                   "%u clusters",
                   ex[e].variantID, ev, ex[e].minpts, src >= 0 ? "from sorted variant " : "scratch",
                   src >= 0 ? src : 0, nvisited, ncl[e] - 1);
    }
    return kOk;
}

int reuse_labels(const double *x, const double *y, uint32_t n, double eps, uint32_t minpts,
                 const uint32_t *old, uint32_t old_ncl, uint32_t *lab, uint32_t *ncl)
{
    if (n == 0) {
        *ncl = 1;
        return kOk;
    }
    if (!std::isfinite(eps)) {
        set_error("the reuse emulation needs a finite eps");
        return kErrInvalid;
    }
    Context *cp = nullptr;
    VDB_TRY(get_context(&cp));
    Context &c = *cp;
    std::lock_guard<std::mutex> lock(c.mu);
    uint32_t *dold = nullptr, *dlab, *counts;
    VDB_TRY(c.get("rl.lab", n, &dlab));
    VDB_TRY(c.get("ru.counts", n, &counts));
    if (old) {
        VDB_TRY(c.get("rl.old", n, &dold));
        VDB_HIP(hipMemcpyAsync(dold, old, 4 * (size_t)n, hipMemcpyHostToDevice, c.stream()));
    }
    Points P;
    P.n = n;
    P.x = x;
    P.y = y;
    VDB_TRY(build_rank_order(c, P, "pts"));
    Grid G;
    Items I;
    VDB_TRY(reuse_grid(c, P, eps, VDB_PREC_FP64, 0, G, I, counts));
    uint32_t nvisited = 0;
    VDB_TRY(reuse_step(c, P, G, I, counts, eps, minpts, dold, old_ncl, dlab, nullptr, ncl,
                       &nvisited));
    VDB_HIP(hipMemcpyAsync(lab, dlab, 4 * (size_t)n, hipMemcpyDeviceToHost, c.stream()));
    VDB_HIP(hipStreamSynchronize(c.stream()));
    return kOk;
}

} // namespace vdb
