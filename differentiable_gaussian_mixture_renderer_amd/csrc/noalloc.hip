// SPDX-License-Identifier: GPL-3.0-or-later
// dgmr_sort_pairs_noalloc: the radix sort of dgmr_sort_pairs with its scratch
// (the per-block digit histograms and the scan's tile sums) in a file-scope
// device array instead of a caller workspace, for callers that cannot
// allocate device memory.  The array exists once per GPU for the life of the
// process; host code reaches it only through the address hipGetSymbolAddress
// returns for the current device (the array's name is a host-side stand-in,
// not a device address).
#include <mutex>

#include "host_common.h"
#include "primitives.h"

namespace dgmr {
namespace {

// sorts of up to 2^25 keys: 16 384 histogram blocks of 256 digits (16 MiB)
// plus the scan of those 4M counts (about 9 KiB)
constexpr int64_t kNoAllocMaxItems = DGMR_SORT_NOALLOC_MAX_ITEMS;
constexpr size_t kScratchBytes = (size_t(16) << 20) + (size_t(64) << 10);

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// the device address of `scratch` on the current device, looked up once per device
int scratch_address(void** out)
{
    int dev = -1;
    DGMR_HIP_CHECK(hipGetDevice(&dev));
    DGMR_ARG_CHECK(dev >= 0 && dev < kMaxDevices, "dgmr_sort_pairs_noalloc: device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        DGMR_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        DGMR_ARG_CHECK(p != nullptr, "dgmr_sort_pairs_noalloc: no device address for the scratch array");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return DGMR_OK;
}

} // namespace
} // namespace dgmr

extern "C" {

int dgmr_sort_pairs_noalloc(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp, int64_t n, int32_t begin_bit,
    int32_t end_bit, dgmr_stream stream)
{
    using namespace dgmr;
    DGMR_ARG_CHECK(n >= 0, "dgmr_sort_pairs_noalloc: n must be >= 0");
    DGMR_ARG_CHECK(n <= kNoAllocMaxItems, "dgmr_sort_pairs_noalloc: n exceeds DGMR_SORT_NOALLOC_MAX_ITEMS (2^25 keys); use dgmr_sort_pairs "
                                          "with a workspace");
    DGMR_ARG_CHECK(begin_bit >= 0 && end_bit <= 32 && begin_bit <= end_bit,
        "dgmr_sort_pairs_noalloc: bits must satisfy 0 <= begin <= end <= 32");
    if (n == 0 || begin_bit == end_bit)
        return DGMR_OK;
    DGMR_ARG_CHECK(keys && vals && keys_tmp && vals_tmp, "dgmr_sort_pairs_noalloc: null buffer");
    DGMR_ARG_CHECK(radix_scratch_bytes(n) <= kScratchBytes, "dgmr_sort_pairs_noalloc: scratch array too small for n");
    void* ws = nullptr;
    const int rc = scratch_address(&ws);
    if (rc != DGMR_OK)
        return rc;
    // every pass writes all the histogram entries it scans, so nothing in the
    // array needs clearing between calls
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const SortBuffers buf { keys, vals, keys_tmp, vals_tmp };
    DGMR_HIP_CHECK(radix_sort_pairs(buf, n, nullptr, begin_bit, end_bit, ws, st));
    if (radix_passes(begin_bit, end_bit) % 2 == 1) {
        DGMR_HIP_CHECK(hipMemcpyAsync(keys, keys_tmp, size_t(n) * 4, hipMemcpyDeviceToDevice, st));
        DGMR_HIP_CHECK(hipMemcpyAsync(vals, vals_tmp, size_t(n) * 4, hipMemcpyDeviceToDevice, st));
    }
    return DGMR_OK;
}

} // extern "C"
