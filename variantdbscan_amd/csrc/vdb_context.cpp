#include "vdb_context.h"

#include <memory>
#include <vector>

namespace vdb {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }
const char *last_error() { return g_last_error.c_str(); }

Context::Context(int device) : device_(device)
{
    // default priority on purpose: every stream of this library stays at the
    // default priority
    if (hipStreamCreateWithFlags(&own_stream_, hipStreamNonBlocking) != hipSuccess)
        own_stream_ = nullptr;
    stream_ = own_stream_;
}

Context::~Context()
{
    release_all();
    if (host_) hipHostFree(host_);
    if (own_stream_) hipStreamDestroy(own_stream_);
}

int Context::get(const char *name, size_t bytes, void **out)
{
    Buffer &b = pool_[name];
    if (b.bytes < bytes) {
        if (b.ptr) {
            VDB_HIP(hipStreamSynchronize(stream_));
            VDB_HIP(hipFree(b.ptr));
            b.ptr = nullptr;
            b.bytes = 0;
        }
        size_t want = bytes < 256 ? 256 : bytes;
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
    if (it->second.ptr) {
        hipStreamSynchronize(stream_);
        hipFree(it->second.ptr);
    }
    pool_.erase(it);
}

void Context::release_all()
{
    if (stream_) hipStreamSynchronize(stream_);
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

int Context::sync()
{
    VDB_HIP(hipStreamSynchronize(stream_));
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
