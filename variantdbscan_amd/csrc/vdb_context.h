// Per-device state: the stream every kernel is launched on and a grow-only
// pool of named device buffers that persists across calls, so a caller that
// clusters the same dataset repeatedly allocates device memory only once.
#pragma once

#include "vdb_common.h"

#include <mutex>
#include <string>
#include <unordered_map>

namespace vdb {

struct Buffer {
    void *ptr = nullptr;
    size_t bytes = 0;
};

class Context {
public:
    explicit Context(int device);
    ~Context();

    int device() const { return device_; }
    hipStream_t stream() const { return stream_; }
    // Launch on a caller-owned stream (e.g. the framework's current stream);
    // nullptr restores the library's own stream.
    void set_stream(hipStream_t s) { stream_ = s ? s : own_stream_; }

    // Grow-only named buffer: returns a pointer to at least `bytes` bytes.
    // Contents are not preserved when the slot grows.
    int get(const char *name, size_t bytes, void **out);
    template <typename T> int get(const char *name, size_t count, T **out)
    {
        void *p = nullptr;
        int s = get(name, count * sizeof(T), &p);
        *out = (T *)p;
        return s;
    }
    void release(const char *name);
    void release_all();
    size_t pooled_bytes() const;
    size_t slot_bytes(const char *name) const;

    // Pinned host staging for small device->host reads.
    int host_scratch(size_t bytes, void **out);

    int sync();

    std::mutex mu; // serialises library calls on this device

    // autotuner choices, keyed by workload (see tune_cells_per_eps)
    std::unordered_map<std::string, int> tuned;

private:
    int device_;
    hipStream_t own_stream_ = nullptr;
    hipStream_t stream_ = nullptr;
    std::unordered_map<std::string, Buffer> pool_;
    void *host_ = nullptr;
    size_t host_bytes_ = 0;
};

// Context of the current HIP device (created on first use).
int get_context(Context **out);

} // namespace vdb
