// Per-device state: the stream every kernel is launched on and a grow-only
// pool of named device buffers that persists across calls, so a caller that
// clusters the same dataset repeatedly allocates device memory only once.
#pragma once

#include "vdb_common.h"

#include <functional>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

namespace vdb {

// entries of vdb_batch_info
constexpr int kBatchInfo = 12;

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

    // A second stream of the library (default priority, created on first
    // use) for work that may run beside the current stream's, with events
    // (no timing) to order the two: aux_event(i), i < kAuxEvents.
    static constexpr int kAuxEvents = 3;
    int aux_stream(hipStream_t *out);
    hipEvent_t aux_event(int i) const { return aux_ev_[i]; }

    // Allocation-free mode (vdb_cluster_noalloc): between arena_begin and
    // arena_end, get() carves named slices of the given device region instead
    // of allocating (same name, same slice while it is large enough; a larger
    // request takes a new slice), host_scratch() returns a static host
    // buffer, and per-kernel timing records nothing. A request beyond the
    // region fails with kErrUnsupported. arena_end returns the bytes used.
    void arena_begin(void *base, size_t bytes);
    size_t arena_end();
    bool in_arena() const { return arena_ != nullptr; }
    void *noalloc_scratch = nullptr; // this device's static scratch array

    int sync();
    // Host waits of a call's read-backs: spin on an event instead of a
    // blocking wait, which may sleep and wake tens of microseconds after the
    // work is done (the GPU idles until the host launches again); after
    // about 20 ms of spinning the wait blocks.
    int wait(hipStream_t s);
    int wait_event(hipEvent_t e);

    std::mutex mu; // serialises library calls on this device

    // Autotuner choices, keyed by workload. They persist in a JSON file
    // (one "key": value per line) so that later processes skip the tuning:
    // $VDB_TUNE_CACHE, else $XDG_CACHE_HOME/variantdbscan_amd/tune.json,
    // else $HOME/.cache/variantdbscan_amd/tune.json; VDB_TUNE_CACHE=off
    // keeps them in memory only. Keys carry the device's architecture and
    // CU count.
    bool tuned_get(const std::string &key, int *value);
    void tuned_put(const std::string &key, int value);
    std::string tuned_json();
    uint32_t tuned_in_process = 0; // tunings run by this process
    uint64_t last_batch[kBatchInfo] = {0}; // vdb_batch_info of the last batched run
    std::string tune_cache_path() const;

    // Per-kernel timing: while a kernel name is selected, every launch of
    // that kernel is bracketed by two events on the launch stream, and
    // kt_read() sums their durations. Events are created once and reused.
    void kt_select(const char *name);
    bool kt_active() const { return !kt_name_.empty(); }
    bool kt_on(const char *name) const
    {
        return !arena_ && !kt_name_.empty() && kt_name_ == name;
    }
    int kt_begin(const char *name);
    int kt_end(const char *name);
    int kt_read(double *ms, uint64_t *launches);

    // Launch sequences replayed from cached HIP graphs. `seq` launches on
    // stream() with no host read and no allocation; `shape` identifies
    // everything its launches depend on besides the pooled buffers. The
    // first call of a (shape, pool state) runs seq directly (the pool then
    // holds every buffer seq uses), the second captures it into a graph,
    // later ones replay the graph. Any allocation moves the pool epoch, so a
    // graph never outlives the buffers it was captured with.
    int graph_run(uint64_t shape, const std::function<int()> &seq);
    // the sequence of `shape` has just run directly (as graph_run's first call)
    void graph_mark(uint64_t shape);
    uint64_t pool_epoch() const { return epoch_; }
    void drop_graphs();

private:
    int device_;
    hipStream_t own_stream_ = nullptr;
    hipStream_t stream_ = nullptr;
    hipStream_t aux_ = nullptr;
    hipEvent_t aux_ev_[kAuxEvents] = {};
    hipEvent_t wait_ev_ = nullptr;
    std::unordered_map<std::string, Buffer> pool_;
    void *host_ = nullptr;
    size_t host_bytes_ = 0;
    std::unordered_map<std::string, int> tuned_;
    bool tuned_loaded_ = false;
    std::string dev_tag_; // "<arch>:<CUs>:" prefix of tuning keys
    void tuned_load();
    char *arena_ = nullptr;
    size_t arena_cap_ = 0, arena_used_ = 0;
    std::unordered_map<std::string, Buffer> arena_map_;
    alignas(64) unsigned char host_static_[256];
    uint64_t epoch_ = 1;            // bumped by every pool allocation and release
    uint64_t graph_pending_ = 0;    // key of the last sequence run directly
    std::vector<std::pair<uint64_t, hipGraphExec_t>> graphs_; // oldest first
    std::string kt_name_;
    std::vector<hipEvent_t> kt_ev_; // pairs (begin, end) of recorded launches
    size_t kt_used_ = 0;            // events recorded since the last read
};

// Context of the current HIP device (created on first use).
int get_context(Context **out);

} // namespace vdb
