// SPDX-License-Identifier: GPL-3.0-or-later
// Device-wide scan and stable LSD radix sort built from 64-lane wavefront
// ballot / prefix primitives (replacing cub::DeviceScan::InclusiveSum and
// cub::DeviceRadixSort::SortPairs at splat_forward.cu:127-137, 197-210).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace dgmr {

// --- exclusive scan of uint32 ---------------------------------------------
// scratch bytes needed to scan n items
size_t scan_scratch_bytes(int64_t n);
// out[i] = sum(in[0..i-1]); *total (device) = sum(in) when total != nullptr.
// inclusive == true gives out[i] = sum(in[0..i]).  in may alias out.
hipError_t scan_u32(const uint32_t* in, uint32_t* out, uint32_t* total, int64_t n, bool inclusive, void* scratch,
    hipStream_t stream);

// --- stable LSD radix sort of (uint32 or uint16 key, uint32 value) pairs --
// Sorts bits [begin_bit, end_bit) in 8-bit digits.  Pass p reads buffer
// (p % 2 == 0 ? A : B) and writes the other one; the sorted result ends in A
// when the number of passes is even, in B when it is odd.  n may be given on
// the device (n_dev != nullptr, capacity n_max; min(*n_dev, n_max) items are
// sorted) so no host sync is needed.
template <class KeyT>
struct SortBuffersT {
    KeyT* keys_a;
    uint32_t* vals_a;
    KeyT* keys_b;
    uint32_t* vals_b;
};
using SortBuffers = SortBuffersT<uint32_t>;
// 16-bit keys (end_bit <= 16): a quarter less traffic per pass for keys that
// fit, such as tile ids
using SortBuffers16 = SortBuffersT<uint16_t>;
int radix_passes(int begin_bit, int end_bit);
size_t radix_scratch_bytes(int64_t n_max);
// The last pass may write, instead of each sorted key, the view of its value
// (value / n_per_view; culled_view for the key 0xffffffff when culled_view is
// not ~0): the keys of a following stable pass on the view.  n_per_view 0
// keeps the keys.
struct KeyOut {
    uint32_t n_per_view = 0;
    uint32_t culled_view = 0xffffffffu;
};
hipError_t radix_sort_pairs(const SortBuffers& buf, int64_t n_max, const uint32_t* n_dev, int begin_bit, int end_bit,
    void* scratch, hipStream_t stream, KeyOut last_pass_keys = KeyOut {});
hipError_t radix_sort_pairs(const SortBuffers16& buf, int64_t n_max, const uint32_t* n_dev, int begin_bit, int end_bit,
    void* scratch, hipStream_t stream);

// --- stable compaction of (key, value) pairs whose key is not skip_key ----
// Keeps input order.  *count (device) receives the number kept; block_counts
// and block_offsets hold compact_blocks(n) entries; scratch holds
// scan_scratch_bytes(compact_blocks(n)) bytes.
int64_t compact_blocks(int64_t n);
// vals_in may be null: the values are then the input indices
hipError_t compact_pairs(const uint32_t* keys_in, const uint32_t* vals_in, int64_t n, uint32_t skip_key, uint32_t* keys_out,
    uint32_t* vals_out, uint32_t* count, uint32_t* block_counts, uint32_t* block_offsets, void* scratch, hipStream_t stream);

} // namespace dgmr
