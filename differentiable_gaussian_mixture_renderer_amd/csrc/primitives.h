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
// One pass over the input with decoupled look-back between tiles (2 048
// items, or 8 192 from 262 144 items up).  One launch may scan two arrays of
// the same length.  scratch bytes needed to scan n items of `arrays` arrays:
size_t scan_scratch_bytes(int64_t n, int arrays = 1);
// The scan's ticket counter and per-tile status words inside its scratch,
// which must be zero when the scan starts: the scan clears them with a small
// launch of its own unless the caller passes cleared = true, having had the
// kernel that writes the scan's input call scan_clear from a grid of at
// least n_status threads.
struct ScanState {
    uint32_t* ticket;
    uint64_t* status;
    int64_t n_status; // status words (0: the scan is one tile and keeps no state)
};
ScanState scan_state(void* scratch, int64_t n, int arrays = 1);
// called by every thread of the producer's grid, before the scan's launch
__device__ __forceinline__ void scan_clear(const ScanState& r)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i == 0)
        *r.ticket = 0;
    if (i < r.n_status)
        r.status[i] = 0;
}
// out[i] = sum(in[0..i-1]); *total (device) = sum(in) when total != nullptr.
// inclusive == true gives out[i] = sum(in[0..i]).  in may alias out.
struct ScanArray {
    const uint32_t* in;
    uint32_t* out;
    uint32_t* total;
    bool inclusive;
};
hipError_t scan_u32_arrays(const ScanArray* arrays, int count, int64_t n, void* scratch, hipStream_t stream, bool cleared = false);
hipError_t scan_u32(const uint32_t* in, uint32_t* out, uint32_t* total, int64_t n, bool inclusive, void* scratch,
    hipStream_t stream, bool cleared = false);

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
