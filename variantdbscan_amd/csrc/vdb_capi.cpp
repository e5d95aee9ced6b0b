// C ABI of libSharedVDBSCAN.so (declared in include/vdbscan.h).
#include "vdbscan.h"

#include "vdb_batch.h"
#include "vdb_cluster.h"
#include "vdb_prims.h"
#include "vdb_reuse.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <vector>

using namespace vdb;

namespace {

// Phase timer for vdb_stats: events only when stats are asked for.
struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    hipStream_t s;
    Timer(hipStream_t st, bool on) : s(st)
    {
        if (!on) return;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
        hipEventRecord(a, s);
    }
    ~Timer()
    {
        if (a) hipEventDestroy(a);
        if (b) hipEventDestroy(b);
    }
    double ms()
    {
        if (!a || !b) return 0.0;
        hipEventRecord(b, s);
        hipEventSynchronize(b);
        float m = 0;
        hipEventElapsedTime(&m, a, b);
        return m;
    }
};

int upload_points(Context &c, const double *x, const double *y, uint32_t n, uint32_t flags,
                  Points &P)
{
    P.n = n;
    if (flags & VDB_DEVICE_PTRS) {
        P.x = x;
        P.y = y;
        return kOk;
    }
    double *dx, *dy;
    VDB_TRY(c.get("in.x", n, &dx));
    VDB_TRY(c.get("in.y", n, &dy));
    VDB_HIP(hipMemcpyAsync(dx, x, sizeof(double) * n, hipMemcpyHostToDevice, c.stream()));
    VDB_HIP(hipMemcpyAsync(dy, y, sizeof(double) * n, hipMemcpyHostToDevice, c.stream()));
    P.x = dx;
    P.y = dy;
    return kOk;
}

// Table budget: free device memory plus the pooled table slot (reused by the
// next table), minus headroom for the other buffers of the call.
size_t default_budget(Context &c, const vdb_options *opt)
{
    if (opt && opt->table_budget_bytes) return (size_t)opt->table_budget_bytes;
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return (size_t)1 << 32;
    fr += c.slot_bytes("cv.ids");
    size_t keep = (size_t)8 << 30;
    return fr > 2 * keep ? fr - keep : fr / 2;
}

// per-variant paths: one grid per distinct eps, counts shared by its variants
int run_per_variant(Context &c, Points &P, const double *eps, const uint32_t *minpts,
                    const std::vector<uint32_t> &which, uint32_t *labels_dev, int mode,
                    int border, int precision, int cells, size_t budget, vdb_stats *st,
                    bool verbose)
{
    // group by eps (NaN eps compare unequal to themselves: key them by bits)
    std::map<uint64_t, std::vector<uint32_t>> by_eps;
    for (uint32_t v : which) {
        uint64_t bits;
        memcpy(&bits, &eps[v], 8);
        by_eps[bits].push_back(v);
    }
    uint32_t *counts;
    VDB_TRY(c.get("pv.counts", P.n, &counts));
    unsigned long long *rc;
    VDB_TRY(c.get("pv.recheck", 1, &rc));
    VDB_HIP(hipMemsetAsync(rc, 0, sizeof(unsigned long long), c.stream()));
    PhaseTimes pt;
    for (auto &kv : by_eps) {
        double e = eps[kv.second[0]];
        Grid G;
        Items I;
        Timer t0(c.stream(), st != nullptr);
        int k = 0;
        VDB_TRY(tune_cells_per_eps(c, P, e, precision, cells, &k));
        VDB_TRY(build_eps_grid(c, P, e, k, "g1", G));
        VDB_TRY(build_items(c, G, "g1", I));
        if (st) st->ms_prepare += t0.ms();
        Timer t1(c.stream(), st != nullptr);
        VDB_TRY(count_neighbors(c, G, I, e, precision, counts, (uint64_t *)rc));
        if (st) st->ms_count += t1.ms();
        if (verbose) {
            uint32_t largest = 0;
            VDB_TRY(grid_max_cell(c, G, &largest));
            printf("\n[vdbscan] eps=%g grid %ux%u (%d cells per eps), %u non-empty cells, "
                   "largest %u points",
                   e, G.ncx, G.ncy, G.R, G.ncell, largest);
        }
        for (uint32_t v : kv.second) {
            VDB_TRY(cluster_variant(c, P, G, I, e, minpts[v], counts, mode, border, budget,
                                    labels_dev + (size_t)v * P.n, st ? &pt : nullptr));
        }
    }
    if (st) {
        st->ms_scan += pt.scan;
        st->ms_fill += pt.fill;
        st->ms_uf += pt.uf;
        st->ms_border += pt.border;
        st->ms_labels += pt.labels;
        st->pairs = pt.pairs;
        st->tiles += pt.tiles;
        unsigned long long r = 0;
        VDB_HIP(hipMemcpy(&r, rc, sizeof r, hipMemcpyDeviceToHost));
        st->rechecked = r;
    }
    return kOk;
}

// Clusters the device-resident points P (x, y set; several datasets when
// P.nsets > 1, with P.hseg their starts) under the variants `which` of the
// nvar given into out (device): dataset d's block out + hseg[d]*nvar is its
// own retArr, nvar rows of n_d labels, and variant v fills row v. Variants
// the batched engine takes run for all datasets at once; the rest, and groups
// the data make it refuse, run dataset by dataset. Returns the path used in
// *used_out.
int cluster_points(Context &c, Points &P, const char *tag, const double *eps,
                   const uint32_t *minpts, uint32_t nvar, const std::vector<uint32_t> &which,
                   uint32_t *out, const vdb_options *opt, vdb_stats *st, int *used_out)
{
    int path = opt ? opt->path : VDB_PATH_AUTO;
    int precision = opt ? opt->precision : VDB_PREC_FP64;
    int border = opt ? opt->border : VDB_BORDER_REFERENCE;
    bool verbose = opt && opt->verbose;
    Timer tprep(c.stream(), st != nullptr);
    VDB_TRY(build_rank_order(c, P, tag));
    if (st) st->ms_prepare += tprep.ms();
    int used = path;
    // variants left for the per-variant engine (all of them unless batched)
    std::vector<uint32_t> rest;
    if (path == VDB_PATH_AUTO || path == VDB_PATH_BATCH) {
        // batchable variants, grouped so that max eps / min eps stays within
        // the fine grid's window; each group is one batched run
        std::vector<uint32_t> bat;
        for (uint32_t v : which) (batchable(eps[v], minpts[v]) ? bat : rest).push_back(v);
        std::sort(bat.begin(), bat.end(), [&](uint32_t a, uint32_t b) {
            bool ia = threshold_for_eps(eps[a]) < 0, ib = threshold_for_eps(eps[b]) < 0;
            if (ia != ib) return ia;
            return eps[a] < eps[b] || (eps[a] == eps[b] && a < b);
        });
        size_t g0 = 0;
        bool any_batched = false;
        while (g0 < bat.size()) {
            size_t g1 = g0;
            double lo = -1.0;
            while (g1 < bat.size()) {
                double e = eps[bat[g1]];
                if (threshold_for_eps(e) >= 0) {
                    if (lo < 0) lo = e;
                    else if (e / lo > kBatchEpsRatio) break;
                }
                g1++;
            }
            std::vector<double> ge;
            std::vector<uint32_t> gm, gr;
            for (size_t i = g0; i < g1; i++) {
                ge.push_back(eps[bat[i]]);
                gm.push_back(minpts[bat[i]]);
                gr.push_back(bat[i]);
            }
            BatchResult br;
            int s = cluster_batch(c, P, ge.data(), gm.data(), gr.data(), (uint32_t)ge.size(), out,
                                  nvar, border, st, verbose, &br);
            if (s == kOk) {
                any_batched = true;
            } else if (s == kErrUnsupported) {
                rest.insert(rest.end(), gr.begin(), gr.end());
            } else {
                return s;
            }
            g0 = g1;
        }
        used = any_batched ? VDB_PATH_BATCH : VDB_PATH_FUSED;
    } else {
        rest = which;
    }
    if (rest.empty()) {
        *used_out = used;
        return kOk;
    }
    std::sort(rest.begin(), rest.end());
    if (P.nsets > 1) {
        // one dataset at a time: its own batched run first, then per variant
        for (uint32_t d = 0; d < P.nsets; d++) {
            uint32_t nd = P.hseg[d + 1] - P.hseg[d];
            if (nd == 0) continue;
            Points Q;
            Q.n = nd;
            Q.x = P.x + P.hseg[d];
            Q.y = P.y + P.hseg[d];
            int u = 0;
            VDB_TRY(cluster_points(c, Q, "pts1", eps, minpts, nvar, rest,
                                   out + (size_t)P.hseg[d] * nvar, opt, st, &u));
            if (u == VDB_PATH_BATCH) used = u;
        }
    } else {
        VDB_TRY(run_per_variant(c, P, eps, minpts, rest, out, path == VDB_PATH_TABLE ? 1 : 0,
                                border, precision, cells_per_eps(opt), default_budget(c, opt),
                                st, verbose));
    }
    *used_out = used;
    return kOk;
}

// Options are checked before any device work.
int check_options(const vdb_options *opt)
{
    if (!opt) return kOk;
    if (opt->path < VDB_PATH_AUTO || opt->path > VDB_PATH_BATCH ||
        opt->precision < VDB_PREC_FP64 || opt->precision > VDB_PREC_FP32_RECHECK ||
        opt->border < VDB_BORDER_REFERENCE || opt->border > VDB_BORDER_STANDARD) {
        set_error("unknown path, precision or border option");
        return kErrInvalid;
    }
    return kOk;
}

std::vector<uint32_t> all_variants(uint32_t nvar)
{
    std::vector<uint32_t> w(nvar);
    for (uint32_t v = 0; v < nvar; v++) w[v] = v;
    return w;
}

int cluster_impl(const double *x, const double *y, uint32_t n, const double *eps,
                 const uint32_t *minpts, uint32_t nvar, uint32_t *labels, uint32_t flags,
                 const vdb_options *opt, vdb_stats *st)
{
    if (st) memset(st, 0, sizeof(*st));
    if (n == 0 || nvar == 0) return kOk;
    if (!x || !y || !eps || !minpts || !labels) {
        set_error("NULL array argument");
        return kErrInvalid;
    }
    if ((uint64_t)n * nvar > 0xffffffffull) {
        set_error("datasetSize*numVariants exceeds 2^32 labels");
        return kErrInvalid;
    }
    VDB_TRY(check_options(opt));
    if ((flags & VDB_REUSE_REFERENCE) && opt && opt->border != VDB_BORDER_REFERENCE) {
        set_error("the reuse emulation gives the reference's labels only (border rule 0)");
        return kErrInvalid;
    }
    Context *cp = nullptr;
    VDB_TRY(get_context(&cp));
    Context &c = *cp;
    std::lock_guard<std::mutex> lock(c.mu);
    Timer total(c.stream(), st != nullptr);
    Points P;
    VDB_TRY(upload_points(c, x, y, n, flags, P));
    uint32_t *out = labels;
    if (!(flags & VDB_DEVICE_PTRS)) VDB_TRY(c.get("out.labels", (size_t)n * nvar, &out));
    int used = 0;
    if (flags & VDB_REUSE_REFERENCE) {
        // the reference's greedy reuse: per variant, fused-engine kernels
        VDB_TRY(cluster_reuse(c, P, eps, minpts, nvar, out, opt ? opt->precision : VDB_PREC_FP64,
                              cells_per_eps(opt), opt && opt->verbose));
        used = VDB_PATH_FUSED;
    } else {
        VDB_TRY(cluster_points(c, P, "pts", eps, minpts, nvar, all_variants(nvar), out, opt, st,
                               &used));
    }
    if (!(flags & VDB_DEVICE_PTRS)) {
        VDB_HIP(hipMemcpyAsync(labels, out, sizeof(uint32_t) * (size_t)n * nvar,
                               hipMemcpyDeviceToHost, c.stream()));
    }
    VDB_HIP(hipStreamSynchronize(c.stream()));
    if (st) {
        st->ms_total = total.ms();
        st->path_used = used;
    }
    return kOk;
}

} // namespace

extern "C" {

const char *vdb_last_error(void) { return last_error(); }

int vdb_version(void) { return 10000; }

int vdb_set_stream(void *s)
{
    Context *c = nullptr;
    VDB_TRY(get_context(&c));
    std::lock_guard<std::mutex> lock(c->mu);
    c->set_stream((hipStream_t)s);
    return kOk;
}

int vdb_release(void)
{
    Context *c = nullptr;
    VDB_TRY(get_context(&c));
    std::lock_guard<std::mutex> lock(c->mu);
    c->release_all();
    return kOk;
}

int vdb_time_kernel(const char *name)
{
    Context *c = nullptr;
    VDB_TRY(get_context(&c));
    std::lock_guard<std::mutex> lock(c->mu);
    c->kt_select(name);
    return kOk;
}

int vdb_kernel_time(double *ms, uint64_t *launches)
{
    if (!ms || !launches) {
        set_error("NULL output argument");
        return kErrInvalid;
    }
    Context *c = nullptr;
    VDB_TRY(get_context(&c));
    std::lock_guard<std::mutex> lock(c->mu);
    return c->kt_read(ms, launches);
}

int vdb_batch_info(uint64_t *info, uint32_t count)
{
    if (!info && count) {
        set_error("NULL output argument");
        return kErrInvalid;
    }
    Context *c = nullptr;
    VDB_TRY(get_context(&c));
    std::lock_guard<std::mutex> lock(c->mu);
    for (uint32_t i = 0; i < count; i++) info[i] = i < (uint32_t)kBatchInfo ? c->last_batch[i] : 0;
    return kOk;
}

int vdb_tuning(char *buf, uint64_t cap, uint64_t *len)
{
    Context *c = nullptr;
    VDB_TRY(get_context(&c));
    std::lock_guard<std::mutex> lock(c->mu);
    std::string js = "{\"cache\": \"" + c->tune_cache_path() +
                     "\", \"tuned_in_process\": " + std::to_string(c->tuned_in_process) +
                     ", \"entries\": " + c->tuned_json() + "}";
    if (len) *len = js.size();
    if (buf && cap) {
        size_t m = std::min<size_t>(cap - 1, js.size());
        memcpy(buf, js.data(), m);
        buf[m] = 0;
    }
    return kOk;
}

int vdb_cluster(const double *x, const double *y, uint32_t n, const double *eps,
                const uint32_t *minpts, uint32_t nvar, uint32_t *labels, uint32_t flags,
                const vdb_options *opt, vdb_stats *stats)
{
    return cluster_impl(x, y, n, eps, minpts, nvar, labels, flags, opt, stats);
}

int vdb_cluster_multi(const double *x, const double *y, const uint64_t *offsets,
                      uint32_t nsets, const double *eps, const uint32_t *minpts, uint32_t nvar,
                      uint32_t *labels, uint32_t flags, const vdb_options *opt,
                      vdb_stats *st)
{
    if (st) memset(st, 0, sizeof(*st));
    if (nsets == 0 || nvar == 0) return kOk;
    if (!x || !y || !offsets || !eps || !minpts || !labels) {
        set_error("NULL array argument");
        return kErrInvalid;
    }
    uint64_t ntot = offsets[nsets];
    for (uint32_t d = 0; d < nsets; d++)
        if (offsets[d + 1] < offsets[d]) {
            set_error("dataset offsets must not decrease");
            return kErrInvalid;
        }
    if (ntot * nvar > 0xffffffffull || offsets[0] != 0) {
        set_error("offsets[0] must be 0 and total points*variants below 2^32");
        return kErrInvalid;
    }
    VDB_TRY(check_options(opt));
    if (ntot == 0) return kOk;
    Context *cp = nullptr;
    VDB_TRY(get_context(&cp));
    Context &c = *cp;
    std::lock_guard<std::mutex> lock(c.mu);
    Timer total(c.stream(), st != nullptr);
    Points all;
    VDB_TRY(upload_points(c, x, y, (uint32_t)ntot, flags, all));
    uint32_t *out = labels;
    if (!(flags & VDB_DEVICE_PTRS)) VDB_TRY(c.get("out.labels", (size_t)ntot * nvar, &out));
    int used = 0;
    // groups of datasets whose binned keys (90000 bins each) fit 32 bits
    const uint32_t kGroup = 32768;
    std::vector<uint32_t> which = all_variants(nvar);
    for (uint32_t d0 = 0; d0 < nsets; d0 += kGroup) {
        uint32_t ng = std::min(kGroup, nsets - d0);
        Points G;
        G.n = (uint32_t)(offsets[d0 + ng] - offsets[d0]);
        if (G.n == 0) continue;
        G.x = all.x + offsets[d0];
        G.y = all.y + offsets[d0];
        G.nsets = ng;
        G.hseg.resize((size_t)ng + 1);
        for (uint32_t d = 0; d <= ng; d++)
            G.hseg[d] = (uint32_t)(offsets[d0 + d] - offsets[d0]);
        int u = 0;
        VDB_TRY(cluster_points(c, G, "pts", eps, minpts, nvar, which,
                               out + (size_t)offsets[d0] * nvar, opt, st, &u));
        if (u == VDB_PATH_BATCH || used == 0) used = u;
    }
    if (!(flags & VDB_DEVICE_PTRS))
        VDB_HIP(hipMemcpyAsync(labels, out, sizeof(uint32_t) * (size_t)ntot * nvar,
                               hipMemcpyDeviceToHost, c.stream()));
    VDB_HIP(hipStreamSynchronize(c.stream()));
    if (st) {
        st->ms_total = total.ms();
        st->path_used = used;
    }
    return kOk;
}

int vdb_rank_order(const double *x, const double *y, uint32_t n, uint32_t *mapping,
                   uint32_t flags)
{
    if (n == 0) return kOk;
    if (!x || !y || !mapping) {
        set_error("NULL array argument");
        return kErrInvalid;
    }
    Context *cp = nullptr;
    VDB_TRY(get_context(&cp));
    Context &c = *cp;
    std::lock_guard<std::mutex> lock(c.mu);
    Points P;
    VDB_TRY(upload_points(c, x, y, n, flags, P));
    VDB_TRY(build_rank_order(c, P, "pts"));
    VDB_HIP(hipMemcpyAsync(mapping, P.mapping, sizeof(uint32_t) * n,
                           (flags & VDB_DEVICE_PTRS) ? hipMemcpyDeviceToDevice
                                                     : hipMemcpyDeviceToHost,
                           c.stream()));
    VDB_HIP(hipStreamSynchronize(c.stream()));
    return kOk;
}

int vdb_neighbor_counts(const double *x, const double *y, uint32_t n, double eps,
                        uint32_t *counts, uint32_t flags, int precision, uint64_t *rechecked)
{
    if (n == 0) return kOk;
    if (!x || !y || !counts) {
        set_error("NULL array argument");
        return kErrInvalid;
    }
    Context *cp = nullptr;
    VDB_TRY(get_context(&cp));
    Context &c = *cp;
    std::lock_guard<std::mutex> lock(c.mu);
    Points P;
    VDB_TRY(upload_points(c, x, y, n, flags, P));
    VDB_TRY(build_rank_order(c, P, "pts"));
    Grid G;
    Items I;
    int k = 0;
    VDB_TRY(tune_cells_per_eps(c, P, eps, precision, 0, &k));
    VDB_TRY(build_eps_grid(c, P, eps, k, "g1", G));
    VDB_TRY(build_items(c, G, "g1", I));
    uint32_t *cpos, *ccal;
    unsigned long long *rc;
    VDB_TRY(c.get("nc.pos", n, &cpos));
    VDB_TRY(c.get("nc.caller", n, &ccal));
    VDB_TRY(c.get("nc.recheck", 1, &rc));
    VDB_HIP(hipMemsetAsync(rc, 0, sizeof(unsigned long long), c.stream()));
    VDB_TRY(count_neighbors(c, G, I, eps, precision, cpos, (uint64_t *)rc));
    uint32_t *dst = (flags & VDB_DEVICE_PTRS) ? counts : ccal;
    VDB_TRY(pos_to_caller(c, P, G, cpos, dst));
    if (!(flags & VDB_DEVICE_PTRS))
        VDB_HIP(hipMemcpyAsync(counts, ccal, sizeof(uint32_t) * n, hipMemcpyDeviceToHost,
                               c.stream()));
    VDB_HIP(hipStreamSynchronize(c.stream()));
    if (rechecked) {
        unsigned long long r = 0;
        VDB_HIP(hipMemcpy(&r, rc, sizeof r, hipMemcpyDeviceToHost));
        *rechecked = r;
    }
    return kOk;
}

int vdb_neighbor_table(const double *x, const double *y, uint32_t n, double eps,
                       uint64_t *offsets, uint32_t *ids, float *d2, uint64_t capacity,
                       uint64_t *total, uint32_t flags, int precision, vdb_stats *st)
{
    if (st) memset(st, 0, sizeof(*st));
    if (!total) {
        set_error("NULL total");
        return kErrInvalid;
    }
    *total = 0;
    if (n == 0) {
        if (offsets) offsets[0] = 0;
        return kOk;
    }
    if (!x || !y || !offsets) {
        set_error("NULL array argument");
        return kErrInvalid;
    }
    Context *cp = nullptr;
    VDB_TRY(get_context(&cp));
    Context &c = *cp;
    std::lock_guard<std::mutex> lock(c.mu);
    Timer ttot(c.stream(), st != nullptr);
    Points P;
    VDB_TRY(upload_points(c, x, y, n, flags, P));
    Timer tprep(c.stream(), st != nullptr);
    VDB_TRY(build_rank_order(c, P, "pts"));
    Grid G;
    Items I;
    int k = 0;
    VDB_TRY(tune_cells_per_eps(c, P, eps, precision, 0, &k));
    VDB_TRY(build_eps_grid(c, P, eps, k, "g1", G));
    VDB_TRY(build_items(c, G, "g1", I));
    if (st) st->ms_prepare = tprep.ms();
    bool dev = flags & VDB_DEVICE_PTRS;
    uint64_t *doff = offsets;
    uint32_t *dids = ids;
    float *dd2 = d2;
    if (!dev) {
        VDB_TRY(c.get("nt.off", (size_t)n + 1, &doff));
        // host output: stage through the pool, sized by the capacity given
        if (capacity && ids) VDB_TRY(c.get("nt.ids", capacity, &dids));
        if (capacity && d2) VDB_TRY(c.get("nt.d2", capacity, &dd2));
    }
    PhaseTimes pt;
    uint64_t tot = 0;
    VDB_TRY(neighbor_table(c, P, G, I, eps, precision, doff, (ids && capacity) ? dids : nullptr,
                           (d2 && capacity) ? dd2 : nullptr,
                           (ids && capacity) ? capacity : 0, &tot, &pt));
    *total = tot;
    if (!dev) {
        VDB_HIP(hipMemcpyAsync(offsets, doff, sizeof(uint64_t) * ((size_t)n + 1),
                               hipMemcpyDeviceToHost, c.stream()));
        if (tot <= capacity && ids && tot)
            VDB_HIP(hipMemcpyAsync(ids, dids, sizeof(uint32_t) * tot, hipMemcpyDeviceToHost,
                                   c.stream()));
        if (tot <= capacity && d2 && tot)
            VDB_HIP(hipMemcpyAsync(d2, dd2, sizeof(float) * tot, hipMemcpyDeviceToHost,
                                   c.stream()));
    }
    VDB_HIP(hipStreamSynchronize(c.stream()));
    if (st) {
        st->ms_count = pt.count;
        st->ms_scan = pt.scan;
        st->ms_fill = pt.fill;
        st->pairs = tot;
        st->rechecked = pt.recheck;
        st->ms_total = ttot.ms();
        st->path_used = VDB_PATH_TABLE;
    }
    return kOk;
}

int vdb_neighbor_table_digest(const double *x, const double *y, uint32_t n, double eps,
                              uint32_t flags, int precision, uint64_t *digest, uint64_t *pairs,
                              uint64_t *rechecked)
{
    if (!digest || !pairs || !rechecked) {
        set_error("NULL output argument");
        return kErrInvalid;
    }
    *digest = *pairs = *rechecked = 0;
    if (n == 0) return kOk;
    if (!x || !y) {
        set_error("NULL array argument");
        return kErrInvalid;
    }
    Context *cp = nullptr;
    VDB_TRY(get_context(&cp));
    Context &c = *cp;
    std::lock_guard<std::mutex> lock(c.mu);
    Points P;
    VDB_TRY(upload_points(c, x, y, n, flags, P));
    VDB_TRY(build_rank_order(c, P, "pts"));
    Grid G;
    Items I;
    int k = 0;
    VDB_TRY(tune_cells_per_eps(c, P, eps, precision, 0, &k));
    VDB_TRY(build_eps_grid(c, P, eps, k, "g1", G));
    VDB_TRY(build_items(c, G, "g1", I));
    return table_digest(c, P, G, I, eps, precision, digest, pairs, rechecked);
}

int libVDBSCAN(double *inputx, double *inputy, unsigned int datasetSize, double *inputEpsilon,
               unsigned int *inputMinpts, unsigned int numVariants, int MBBsize,
               unsigned int *retArr, bool verbose)
{
    if (verbose) printf("Verbose mode on"); // src/main.cpp:52-53
    if (MBBsize <= 0) {
        // the reference divides by MBBsize (src/main.cpp:118)
        set_error("MBBsize must be positive");
        return kErrInvalid;
    }
    vdb_options opt;
    memset(&opt, 0, sizeof opt);
    opt.path = VDB_PATH_AUTO;
    opt.precision = VDB_PREC_FP64;
    opt.verbose = verbose ? 1 : 0;
    vdb_stats st;
    // VDBSCAN_REUSE=1 in the environment: the reference's greedy reuse with
    // one search thread (VDB_REUSE_REFERENCE) instead of scratch labels
    const char *ru = getenv("VDBSCAN_REUSE");
    uint32_t flags = (ru && ru[0] == '1') ? VDB_REUSE_REFERENCE : 0u;
    int s = vdb_cluster(inputx, inputy, datasetSize, inputEpsilon, inputMinpts, numVariants,
                        retArr, flags, &opt, &st);
    if (verbose) {
        if (s == 0)
            printf("\n[vdbscan] %u points x %u variants, path %d, %.3f ms device time\n",
                   datasetSize, numVariants, st.path_used, st.ms_total);
        else
            printf("\n[vdbscan] error %d: %s\n", s, last_error());
        fflush(stdout);
    }
    return s;
}

} // extern "C"
