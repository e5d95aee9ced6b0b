// Device-wide primitives written for wave64: exclusive scans, a stable LSD
// radix sort of (u32 key, u32 value) pairs and float64 extent reductions.
#pragma once

#include "vdb_context.h"

namespace vdb {

// out[i] = sum(in[0..i)) ; *total_dev (device pointer, may be null) = sum(in).
int scan_exclusive_u32(Context &c, const uint32_t *in, uint32_t *out, size_t n,
                       uint32_t *total_dev);
// out[i] = set bits in words[0..i).
int scan_popc_u32(Context &c, const uint32_t *words, uint32_t *out, size_t n);
int scan_exclusive_u32_u64(Context &c, const uint32_t *in, uint64_t *out, size_t n,
                           uint64_t *total_dev);

// Stable sort of n pairs by the low `key_bits` bits of keys. On return
// *keys_out / *vals_out point at whichever of (keys, tmp_keys) holds the
// result. Keys above 2^key_bits are not allowed.
int radix_sort_pairs(Context &c, uint32_t *keys, uint32_t *vals, uint32_t *tmp_keys,
                     uint32_t *tmp_vals, size_t n, int key_bits, uint32_t **keys_out,
                     uint32_t **vals_out);

// ext_dev[0..3] = (min x, max x, min y, max y) over n points (n >= 1).
int extent_f64(Context &c, const double *x, const double *y, size_t n, double *ext_dev);

// Exclusive scan of flags -> out (u32); returns the total on the host.
int compact_count(Context &c, const uint32_t *flags, uint32_t *out, size_t n,
                  uint32_t *total_host);

int bits_for(uint64_t max_value_exclusive);

} // namespace vdb
