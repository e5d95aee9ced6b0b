#include "vdb_context.h"

#include <chrono>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <sys/stat.h>
#include <unistd.h>
#include <vector>

namespace vdb {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }
const char *last_error() { return g_last_error.c_str(); }

Context::Context(int device) : device_(device)
{
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess)
        dev_tag_ = std::string(prop.gcnArchName) + ":" + std::to_string(prop.multiProcessorCount) +
                   ":";
    else
        dev_tag_ = "unknown:";
    // default priority on purpose: every stream of this library stays at the
    // default priority
    if (hipStreamCreateWithFlags(&own_stream_, hipStreamNonBlocking) != hipSuccess)
        own_stream_ = nullptr;
    stream_ = own_stream_;
}

Context::~Context()
{
    release_all(); // also drops the graphs
    for (hipEvent_t e : kt_ev_) hipEventDestroy(e);
    if (host_) hipHostFree(host_);
    for (hipEvent_t e : aux_ev_)
        if (e) hipEventDestroy(e);
    if (wait_ev_) hipEventDestroy(wait_ev_);
    if (aux_) hipStreamDestroy(aux_);
    if (own_stream_) hipStreamDestroy(own_stream_);
}

int Context::aux_stream(hipStream_t *out)
{
    if (!aux_) {
        // default priority, like the library's own stream
        VDB_HIP(hipStreamCreateWithFlags(&aux_, hipStreamNonBlocking));
        for (hipEvent_t &e : aux_ev_) VDB_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    *out = aux_;
    return kOk;
}

void Context::arena_begin(void *base, size_t bytes)
{
    arena_ = (char *)base;
    arena_cap_ = bytes;
    arena_used_ = 0;
    arena_map_.clear();
}

size_t Context::arena_end()
{
    size_t used = arena_used_;
    arena_ = nullptr;
    arena_cap_ = arena_used_ = 0;
    arena_map_.clear();
    return used;
}

int Context::get(const char *name, size_t bytes, void **out)
{
    if (arena_) {
        Buffer &a = arena_map_[name];
        if (a.bytes < bytes || !a.ptr) {
            size_t want = ((bytes < 256 ? 256 : bytes) + 255) & ~(size_t)255;
            if (want > arena_cap_ - arena_used_) {
                set_error(std::string("the allocation-free entry point's scratch (") +
                          std::to_string(arena_cap_) + " bytes) is too small for this call ('" +
                          name + "' needs " + std::to_string(want) + " more bytes)");
                return kErrUnsupported;
            }
            a.ptr = arena_ + arena_used_;
            a.bytes = want;
            arena_used_ += want;
        }
        *out = a.ptr;
        return kOk;
    }
    Buffer &b = pool_[name];
    if (b.bytes < bytes) {
        if (b.ptr) {
            VDB_HIP(hipStreamSynchronize(stream_));
            VDB_HIP(hipFree(b.ptr));
            b.ptr = nullptr;
            b.bytes = 0;
        }
        size_t want = bytes < 256 ? 256 : bytes;
        epoch_++;
        hipError_t e = hipMalloc(&b.ptr, want);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            b.ptr = nullptr;
            set_error(std::string("device allocation of ") + std::to_string(want) +
                      " bytes for '" + name + "' failed: " + hipGetErrorString(e));
            return kErrNoMemory;
        }
        b.bytes = want;
    }
    *out = b.ptr;
    return kOk;
}

void Context::release(const char *name)
{
    auto it = pool_.find(name);
    if (it == pool_.end()) return;
    epoch_++;
    if (it->second.ptr) {
        hipStreamSynchronize(stream_);
        hipFree(it->second.ptr);
    }
    pool_.erase(it);
}

void Context::release_all()
{
    if (stream_) hipStreamSynchronize(stream_);
    drop_graphs();
    epoch_++;
    for (auto &kv : pool_)
        if (kv.second.ptr) hipFree(kv.second.ptr);
    pool_.clear();
}

size_t Context::pooled_bytes() const
{
    size_t s = 0;
    for (auto &kv : pool_) s += kv.second.bytes;
    return s;
}

size_t Context::slot_bytes(const char *name) const
{
    auto it = pool_.find(name);
    return it == pool_.end() ? 0 : it->second.bytes;
}

int Context::host_scratch(size_t bytes, void **out)
{
    if (arena_) {
        if (bytes > sizeof host_static_) {
            set_error("host staging beyond the allocation-free entry point's static buffer");
            return kErrUnsupported;
        }
        *out = host_static_;
        return kOk;
    }
    if (host_bytes_ < bytes) {
        if (host_) {
            VDB_HIP(hipStreamSynchronize(stream_));
            hipHostFree(host_);
            host_ = nullptr;
        }
        size_t want = bytes < 4096 ? 4096 : bytes;
        VDB_HIP(hipHostMalloc(&host_, want, hipHostMallocDefault));
        host_bytes_ = want;
    }
    *out = host_;
    return kOk;
}

void Context::kt_select(const char *name)
{
    kt_name_ = name ? name : "";
    kt_used_ = 0;
}

int Context::kt_begin(const char *name)
{
    if (!kt_on(name)) return kOk;
    if (kt_used_ + 2 > kt_ev_.size()) {
        for (int i = 0; i < 2; i++) {
            hipEvent_t e;
            VDB_HIP(hipEventCreate(&e));
            kt_ev_.push_back(e);
        }
    }
    VDB_HIP(hipEventRecord(kt_ev_[kt_used_], stream_));
    return kOk;
}

int Context::kt_end(const char *name)
{
    if (!kt_on(name)) return kOk;
    VDB_HIP(hipEventRecord(kt_ev_[kt_used_ + 1], stream_));
    kt_used_ += 2;
    return kOk;
}

int Context::kt_read(double *ms, uint64_t *launches)
{
    double tot = 0;
    if (kt_used_) VDB_HIP(hipEventSynchronize(kt_ev_[kt_used_ - 1]));
    for (size_t i = 0; i + 1 < kt_used_; i += 2) {
        float m = 0;
        VDB_HIP(hipEventElapsedTime(&m, kt_ev_[i], kt_ev_[i + 1]));
        tot += m;
    }
    *ms = tot;
    *launches = kt_used_ / 2;
    kt_used_ = 0;
    return kOk;
}

std::string Context::tune_cache_path() const
{
    const char *e = getenv("VDB_TUNE_CACHE");
    if (e && e[0]) return strcmp(e, "off") == 0 ? std::string() : std::string(e);
    const char *x = getenv("XDG_CACHE_HOME");
    if (x && x[0]) return std::string(x) + "/variantdbscan_amd/tune.json";
    const char *h = getenv("HOME");
    if (h && h[0]) return std::string(h) + "/.cache/variantdbscan_amd/tune.json";
    return std::string();
}

// entries of a cache file: lines  "key": value  (the format tuned_put writes)
static void read_tune_file(const std::string &path, std::unordered_map<std::string, int> &m)
{
    FILE *f = fopen(path.c_str(), "r");
    if (!f) return;
    char line[1024];
    while (fgets(line, sizeof line, f)) {
        char *a = strchr(line, '"');
        char *b = a ? strchr(a + 1, '"') : nullptr;
        char *colon = b ? strchr(b + 1, ':') : nullptr;
        if (!colon) continue;
        char *end = nullptr;
        long v = strtol(colon + 1, &end, 10);
        if (end == colon + 1) continue;
        m[std::string(a + 1, b)] = (int)v;
    }
    fclose(f);
}

void Context::tuned_load()
{
    if (tuned_loaded_) return;
    tuned_loaded_ = true;
    std::string path = tune_cache_path();
    if (!path.empty()) read_tune_file(path, tuned_);
}

bool Context::tuned_get(const std::string &key, int *value)
{
    tuned_load();
    auto it = tuned_.find(dev_tag_ + key);
    if (it == tuned_.end()) return false;
    *value = it->second;
    return true;
}

std::string Context::tuned_json()
{
    tuned_load();
    std::string out = "{";
    bool first = true;
    for (auto &kv : tuned_) {
        out += first ? "\n" : ",\n";
        first = false;
        out += " \"" + kv.first + "\": " + std::to_string(kv.second);
    }
    out += "\n}\n";
    return out;
}

void Context::tuned_put(const std::string &key, int value)
{
    tuned_load();
    tuned_in_process++;
    tuned_[dev_tag_ + key] = value;
    std::string path = tune_cache_path();
    if (path.empty()) return;
    // merge with what other processes wrote since, then replace the file
    std::unordered_map<std::string, int> disk;
    read_tune_file(path, disk);
    for (auto &kv : tuned_) disk[kv.first] = kv.second;
    tuned_ = disk;
    for (size_t i = path.find('/', 1); i != std::string::npos; i = path.find('/', i + 1))
        mkdir(path.substr(0, i).c_str(), 0755);
    std::string tmp = path + ".tmp" + std::to_string((long)getpid());
    FILE *f = fopen(tmp.c_str(), "w");
    if (!f) return; // an unwritable cache only costs tuning again later
    std::string js = tuned_json();
    fwrite(js.data(), 1, js.size(), f);
    if (fclose(f) == 0) rename(tmp.c_str(), path.c_str());
    else remove(tmp.c_str());
}

void Context::drop_graphs()
{
    for (auto &g : graphs_) hipGraphExecDestroy(g.second);
    graphs_.clear();
    graph_pending_ = 0;
}

// key of a launch sequence over the pool as it stands
static uint64_t graph_key(uint64_t shape, uint64_t epoch)
{
    uint64_t z = shape ^ (epoch * 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 31)) * 0xbf58476d1ce4e5b9ull;
    return z ^ (z >> 29);
}

void Context::graph_mark(uint64_t shape) { graph_pending_ = graph_key(shape, epoch_); }

int Context::graph_run(uint64_t shape, const std::function<int()> &seq)
{
    // the NULL stream (a framework's default stream) cannot be captured
    if (!stream_) return seq();
    auto key_now = [&]() { return graph_key(shape, epoch_); };
    uint64_t key = key_now();
    for (auto &g : graphs_)
        if (g.first == key) {
            VDB_HIP(hipGraphLaunch(g.second, stream_));
            return kOk;
        }
    if (key != graph_pending_) {
        VDB_TRY(seq());
        graph_pending_ = key_now(); // the pool as this run left it
        return kOk;
    }
    // the same sequence over the same buffers a second time: capture it
    uint64_t e0 = epoch_;
    if (hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        graph_pending_ = 0;
        return seq();
    }
    int s = seq();
    hipGraph_t g = nullptr;
    hipError_t ce = hipStreamEndCapture(stream_, &g);
    hipGraphExec_t ex = nullptr;
    bool ok = s == kOk && ce == hipSuccess && g && epoch_ == e0;
    if (ok && hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) != hipSuccess) ok = false;
    if (g) hipGraphDestroy(g);
    if (!ok) {
        // run it directly instead (a failed capture leaves nothing queued)
        (void)hipGetLastError();
        graph_pending_ = 0;
        return seq();
    }
    if (graphs_.size() >= 4) {
        hipGraphExecDestroy(graphs_.front().second);
        graphs_.erase(graphs_.begin());
    }
    graphs_.push_back({key, ex});
    VDB_HIP(hipGraphLaunch(ex, stream_));
    return kOk;
}

int Context::sync()
{
    VDB_HIP(hipStreamSynchronize(stream_));
    return kOk;
}

int Context::wait(hipStream_t s)
{
    if (arena_) { // the allocation-free mode creates nothing
        VDB_HIP(hipStreamSynchronize(s));
        return kOk;
    }
    if (!wait_ev_) VDB_HIP(hipEventCreateWithFlags(&wait_ev_, hipEventDisableTiming));
    VDB_HIP(hipEventRecord(wait_ev_, s));
    return wait_event(wait_ev_);
}

int Context::wait_event(hipEvent_t e)
{
    auto t0 = std::chrono::steady_clock::now();
    for (;;) {
        hipError_t r = hipEventQuery(e);
        if (r == hipSuccess) return kOk;
        if (r != hipErrorNotReady) VDB_HIP(r);
        if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(20)) break;
    }
    VDB_HIP(hipEventSynchronize(e));
    return kOk;
}

int get_context(Context **out)
{
    static std::mutex m;
    static std::vector<std::unique_ptr<Context>> ctxs;
    int dev = 0;
    VDB_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> g(m);
    if ((int)ctxs.size() <= dev) ctxs.resize(dev + 1);
    if (!ctxs[dev]) {
        ctxs[dev].reset(new Context(dev));
        if (!ctxs[dev]->stream()) {
            ctxs[dev].reset();
            set_error("hipStreamCreate failed");
            return kErrDevice;
        }
    }
    *out = ctxs[dev].get();
    return kOk;
}

} // namespace vdb
