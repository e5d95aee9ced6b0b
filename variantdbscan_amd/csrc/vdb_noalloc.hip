// vdb_cluster_noalloc: per-variant clustering that allocates nothing, for
// callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid allocation).
//
// Every device buffer of the call beyond the caller's inputs and outputs is a
// slice of `scratch`, a file-scope array of kNoAllocScratchBytes per GPU,
// held until the process ends. In host code the array's name is only a
// host-side stand-in: its device address is looked up once per device with
// hipGetSymbolAddress (kept in that device's Context) and only that address
// is passed to kernels and copies. The engine is the per-variant fused one
// (grid, count, union-find, border, labels); it runs no autotuner, records no
// events and builds no graphs. Calls on one GPU share the array and keep the
// last call's contents (every kernel initialises what it reads); library
// calls on a device are serialised by its lock and this one waits for its
// work before returning, so two calls never overlap on the array (calls that
// overlapped on one GPU would race on it).
#include "vdbscan.h"

#include "vdb_cluster.h"

#include <cstring>
#include <map>
#include <vector>

namespace {

// 128 MiB per GPU: enough for kNoAllocMaxPoints points at any eps (the
// largest per-point need is the rank order, the grid with its classified
// window rows, counts and the union-find arrays, about 300 bytes a point)
constexpr size_t kNoAllocScratchBytes = (size_t)128 << 20;
constexpr uint32_t kNoAllocMaxPoints = 1u << 18;

} // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

using namespace vdb;

namespace {

// the device address of this device's scratch array (set up on first use)
int scratch_of(Context &c, void **out)
{
    if (!c.noalloc_scratch) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return kErrDevice;
        }
        c.noalloc_scratch = p;
    }
    *out = c.noalloc_scratch;
    return kOk;
}

struct ArenaScope {
    Context &c;
    ArenaScope(Context &cx, void *base) : c(cx) { c.arena_begin(base, kNoAllocScratchBytes); }
    ~ArenaScope() { c.arena_end(); }
};

int noalloc_impl(Context &c, const double *x, const double *y, uint32_t n, const double *eps,
                 const uint32_t *minpts, uint32_t nvar, uint32_t *labels, const vdb_options *opt)
{
    int precision = opt ? opt->precision : VDB_PREC_FP64;
    int border = opt ? opt->border : VDB_BORDER_REFERENCE;
    int k = cells_per_eps(opt);
    if (k == 0) k = 2; // no autotuner here: the library's default resolution
    void *base = nullptr;
    VDB_TRY(scratch_of(c, &base));
    ArenaScope arena(c, base);
    Points P;
    P.n = n;
    P.x = x;
    P.y = y;
    VDB_TRY(build_rank_order(c, P, "pts"));
    uint32_t *counts;
    unsigned long long *rc;
    VDB_TRY(c.get("na.counts", n, &counts));
    VDB_TRY(c.get("na.recheck", 1, &rc));
    VDB_HIP(hipMemsetAsync(rc, 0, sizeof(unsigned long long), c.stream()));
    // one grid and count pass per distinct eps (keyed by its bits)
    std::map<uint64_t, std::vector<uint32_t>> by_eps;
    for (uint32_t v = 0; v < nvar; v++) {
        uint64_t b;
        memcpy(&b, &eps[v], 8);
        by_eps[b].push_back(v);
    }
    for (auto &kv : by_eps) {
        double e = eps[kv.second[0]];
        Grid G;
        Items I;
        VDB_TRY(build_eps_grid(c, P, e, k, "g1", G));
        VDB_TRY(build_items(c, G, "g1", I));
        VDB_TRY(count_neighbors(c, G, I, e, precision, counts, (uint64_t *)rc));
        for (uint32_t v : kv.second)
            VDB_TRY(cluster_variant(c, P, G, I, e, minpts[v], counts, 0, border, 0,
                                    labels + (size_t)v * n, nullptr));
    }
    VDB_HIP(hipStreamSynchronize(c.stream()));
    return kOk;
}

} // namespace

extern "C" int vdb_cluster_noalloc(const double *x, const double *y, uint32_t n,
                                   const double *eps, const uint32_t *minpts, uint32_t nvar,
                                   uint32_t *labels, const vdb_options *opt, vdb_stats *stats)
{
    if (stats) memset(stats, 0, sizeof(*stats));
    if (n == 0 || nvar == 0) return kOk;
    if (!x || !y || !eps || !minpts || !labels) {
        set_error("NULL array argument");
        return kErrInvalid;
    }
    if (n > kNoAllocMaxPoints) {
        set_error("vdb_cluster_noalloc takes at most 262144 points (its static scratch is "
                  "134217728 bytes per GPU)");
        return kErrUnsupported;
    }
    if ((uint64_t)n * nvar > 0xffffffffull) {
        set_error("datasetSize*numVariants exceeds 2^32 labels");
        return kErrInvalid;
    }
    if (opt && ((opt->path != VDB_PATH_AUTO && opt->path != VDB_PATH_FUSED) ||
                opt->precision < VDB_PREC_FP64 || opt->precision > VDB_PREC_FP32_RECHECK ||
                opt->border < VDB_BORDER_REFERENCE || opt->border > VDB_BORDER_STANDARD ||
                opt->cells_per_eps < 0 || opt->cells_per_eps > kMaxCellsPerEps)) {
        set_error("vdb_cluster_noalloc runs the fused per-variant engine: path AUTO or FUSED, "
                  "known precision and border options, cells_per_eps 0..8");
        return kErrInvalid;
    }
    Context *cp = nullptr;
    VDB_TRY(get_context(&cp));
    Context &c = *cp;
    std::lock_guard<std::mutex> lock(c.mu);
    VDB_TRY(noalloc_impl(c, x, y, n, eps, minpts, nvar, labels, opt));
    if (stats) stats->path_used = VDB_PATH_FUSED;
    return kOk;
}
