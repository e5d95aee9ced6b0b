#include "vdb_batch.h"

namespace vdb {

int cluster_batch(Context &, const Points &, const double *, const uint32_t *, uint32_t, int,
                  uint32_t *, vdb_stats *, bool, BatchResult *)
{
    set_error("batched engine not built yet");
    return kErrUnsupported;
}

} // namespace vdb
