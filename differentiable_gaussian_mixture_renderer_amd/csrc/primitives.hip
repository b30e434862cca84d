// SPDX-License-Identifier: GPL-3.0-or-later
// Scan and radix sort for gfx950 (wave64).  See primitives.h.
#include "primitives.h"

#include "device_math.h"

namespace dgmr {
namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;                       // per thread
constexpr int kScanTile = kScanThreads * kScanItems; // 2048 items per block
// scans of 128 small tiles or more take 8 192-item tiles: a quarter of the
// tiles shortens the look-back (a 1M-item scan has 123 tiles, not 489)
constexpr int kScanItemsLarge = 32;
constexpr int64_t kScanLargeFrom = int64_t(128) * kScanTile;

// inclusive scan of one value per thread across a 256-thread block
__device__ __forceinline__ uint32_t block_inclusive_scan(uint32_t v, uint32_t* s_waves, uint32_t* block_total)
{
    const unsigned lane = lane_id();
    const unsigned wave = threadIdx.x / 64;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(v, d, 64);
        if (lane >= unsigned(d))
            v += o;
    }
    if (lane == 63)
        s_waves[wave] = v;
    __syncthreads();
    uint32_t prefix = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / 64; ++w) {
        const uint32_t c = s_waves[w];
        if (unsigned(w) < wave)
            prefix += c;
        total += c;
    }
    __syncthreads();
    *block_total = total;
    return v + prefix;
}

// scans the one tile of a scan of at most 2048 items
__global__ __launch_bounds__(kScanThreads) void k_scan_tile(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total_out, int inclusive)
{
    __shared__ uint32_t s_waves[kScanThreads / 64];
    const int64_t base = int64_t(blockIdx.x) * kScanTile + int64_t(threadIdx.x) * kScanItems;
    uint32_t v[kScanItems];
    uint32_t acc = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        v[j] = (base + j < n) ? in[base + j] : 0u;
        acc += v[j];
    }
    uint32_t total;
    const uint32_t incl = block_inclusive_scan(acc, s_waves, &total);
    uint32_t run = incl - acc;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const uint32_t x = v[j];
        if (base + j < n)
            out[base + j] = inclusive ? run + x : run;
        run += x;
    }
    if (total_out && threadIdx.x == 0)
        *total_out = total;
}

// Device-wide scans run in one pass with decoupled look-back (Merrill and
// Garland, "Single-pass Parallel Prefix Scan with Decoupled Look-back",
// 2016): each block takes the next tile index from a ticket counter, so the
// tiles before it belong to blocks that are already running; it publishes its
// tile sum as an aggregate, then learns the sum of everything before it from
// its predecessors' published words and publishes its inclusive prefix.  The
// status words are read and written with device-scope atomics (the XCDs' L2
// caches are not coherent with each other), and each word carries its whole
// state, so no other memory needs ordering.
constexpr uint32_t kLookAggregate = 1u;
constexpr uint32_t kLookPrefix = 2u;
// a predecessor publishes within microseconds of starting; the bound only
// keeps a protocol error from turning into a hang (about a second of polling)
constexpr int kLookMaxSpins = 1 << 20;

__device__ __forceinline__ void look_publish(uint64_t* slot, uint32_t flag, uint32_t value)
{
    __hip_atomic_store(slot, (uint64_t(flag) << 32) | value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sum of the tiles before tile v, called by all 64 lanes of one wave: lane l
// reads the word of tile j - l, 64 predecessors a round, and the round ends
// the walk when it holds an inclusive prefix (the nearest one counts, with the
// aggregates between it and tile v).
__device__ uint32_t look_back(const uint64_t* status, int64_t v)
{
    const unsigned lane = lane_id();
    uint32_t prefix = 0;
    for (int64_t j = v - 1; j >= 0; j -= 64) {
        const int64_t idx = j - int64_t(lane);
        uint64_t s = uint64_t(kLookPrefix) << 32; // before tile 0: a prefix of 0
        if (idx >= 0) {
            for (int spin = 0; spin < kLookMaxSpins; ++spin) {
                s = __hip_atomic_load(status + idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if ((s >> 32) != 0)
                    break;
                __builtin_amdgcn_s_sleep(1);
            }
        }
        const uint64_t pre = __ballot(uint32_t(s >> 32) == kLookPrefix);
        const unsigned nearest = pre ? unsigned(__ffsll((unsigned long long)pre) - 1) : 63u;
        uint32_t part = lane <= nearest ? uint32_t(s) : 0u;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1)
            part += __shfl_xor(part, d, 64);
        prefix += part;
        if (pre)
            break;
    }
    return prefix;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_clear(ScanState r) { scan_clear(r); }

// ITEMS consecutive items of one thread, with 16-byte loads when they are
// all in range and the array is 16-byte aligned
template <int ITEMS>
__device__ __forceinline__ uint32_t load_items(const uint32_t* in, int64_t base, int64_t n, bool vec, uint32_t (&x)[ITEMS])
{
    uint32_t acc = 0;
    if (vec && base + ITEMS <= n) {
        const uint4* src = reinterpret_cast<const uint4*>(in + base);
#pragma unroll
        for (int j = 0; j < ITEMS / 4; ++j) {
            const uint4 q = src[j];
            x[4 * j] = q.x;
            x[4 * j + 1] = q.y;
            x[4 * j + 2] = q.z;
            x[4 * j + 3] = q.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < ITEMS; ++j)
            x[j] = (base + j < n) ? in[base + j] : 0u;
    }
#pragma unroll
    for (int j = 0; j < ITEMS; ++j)
        acc += x[j];
    return acc;
}

template <int ITEMS>
__device__ __forceinline__ void store_items(uint32_t* out, int64_t base, int64_t n, bool vec, const uint32_t (&x)[ITEMS], uint32_t run,
    bool inclusive)
{
    uint32_t y[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        y[j] = inclusive ? run + x[j] : run;
        run += x[j];
    }
    if (vec && base + ITEMS <= n) {
        uint4* dst = reinterpret_cast<uint4*>(out + base);
#pragma unroll
        for (int j = 0; j < ITEMS / 4; ++j)
            dst[j] = make_uint4(y[4 * j], y[4 * j + 1], y[4 * j + 2], y[4 * j + 3]);
    } else {
#pragma unroll
        for (int j = 0; j < ITEMS; ++j)
            if (base + j < n)
                out[base + j] = y[j];
    }
}

// One tile of kScanThreads * ITEMS items of each of ARRAYS arrays per block,
// in ticket order.  Array k's status words are r.status[k * tiles + tile]:
// both aggregates go out before either look-back, so the two walks overlap.
template <int ITEMS, int ARRAYS>
__global__ __launch_bounds__(kScanThreads) void k_scan_onepass(ScanArray a0, ScanArray a1, int64_t n, ScanState r, int vec)
{
    constexpr int64_t kTile = int64_t(kScanThreads) * ITEMS;
    __shared__ uint32_t s_waves[kScanThreads / 64];
    __shared__ uint32_t s_tile, s_prefix[ARRAYS];
    if (threadIdx.x == 0)
        s_tile = atomicAdd(r.ticket, 1u);
    __syncthreads();
    const int64_t v = s_tile, tiles = gridDim.x;
    const int64_t base = v * kTile + int64_t(threadIdx.x) * ITEMS;
    const ScanArray arr[2] = { a0, a1 };
    uint32_t x[ARRAYS][ITEMS], acc[ARRAYS], incl[ARRAYS], total[ARRAYS];
#pragma unroll
    for (int k = 0; k < ARRAYS; ++k)
        acc[k] = load_items<ITEMS>(arr[k].in, base, n, vec != 0, x[k]);
#pragma unroll
    for (int k = 0; k < ARRAYS; ++k)
        incl[k] = block_inclusive_scan(acc[k], s_waves, &total[k]);
    if (threadIdx.x < 64) {
        if (threadIdx.x == 0) {
#pragma unroll
            for (int k = 0; k < ARRAYS; ++k)
                look_publish(r.status + k * tiles + v, v == 0 ? kLookPrefix : kLookAggregate, total[k]);
        }
#pragma unroll
        for (int k = 0; k < ARRAYS; ++k) {
            const uint32_t before = v == 0 ? 0u : look_back(r.status + k * tiles, v);
            if (threadIdx.x == 0) {
                if (v > 0)
                    look_publish(r.status + k * tiles + v, kLookPrefix, before + total[k]);
                s_prefix[k] = before;
                if (arr[k].total && v == tiles - 1)
                    *arr[k].total = before + total[k];
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < ARRAYS; ++k)
        store_items<ITEMS>(arr[k].out, base, n, vec != 0, x[k], incl[k] - acc[k] + s_prefix[k], arr[k].inclusive);
}

int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

} // namespace

static int64_t scan_tile_items(int64_t n) { return n >= kScanLargeFrom ? int64_t(kScanThreads) * kScanItemsLarge : kScanTile; }

size_t scan_scratch_bytes(int64_t n, int arrays)
{
    // the ticket (padded to 256 bytes), then one status word per tile and array
    const int64_t tiles = ceil_div(std::max<int64_t>(n, 1), scan_tile_items(n));
    return 256 + ((size_t(tiles) * arrays * sizeof(uint64_t) + 255) / 256) * 256;
}

ScanState scan_state(void* scratch, int64_t n, int arrays)
{
    ScanState r;
    r.ticket = static_cast<uint32_t*>(scratch);
    r.status = reinterpret_cast<uint64_t*>(static_cast<char*>(scratch) + 256);
    r.n_status = n > kScanTile ? ceil_div(n, scan_tile_items(n)) * arrays : 0;
    return r;
}

hipError_t scan_u32_arrays(const ScanArray* arrays, int count, int64_t n, void* scratch, hipStream_t stream, bool cleared)
{
    if (count < 1 || count > 2)
        return hipErrorInvalidValue;
    if (n <= 0) {
        for (int k = 0; k < count; ++k)
            if (arrays[k].total) {
                const hipError_t e = hipMemsetAsync(arrays[k].total, 0, sizeof(uint32_t), stream);
                if (e != hipSuccess)
                    return e;
            }
        return hipSuccess;
    }
    if (n <= kScanTile) {
        for (int k = 0; k < count; ++k)
            k_scan_tile<<<1, kScanThreads, 0, stream>>>(arrays[k].in, arrays[k].out, n, arrays[k].total, arrays[k].inclusive ? 1 : 0);
        return hipGetLastError();
    }
    const ScanState r = scan_state(scratch, n, count);
    if (!cleared)
        k_scan_clear<<<unsigned(ceil_div(r.n_status, kScanThreads)), kScanThreads, 0, stream>>>(r);
    int vec = 1;
    for (int k = 0; k < count; ++k)
        vec &= (reinterpret_cast<uintptr_t>(arrays[k].in) % 16 == 0 && reinterpret_cast<uintptr_t>(arrays[k].out) % 16 == 0) ? 1 : 0;
    const ScanArray a0 = arrays[0], a1 = count > 1 ? arrays[1] : arrays[0];
    const unsigned tiles = unsigned(ceil_div(n, scan_tile_items(n)));
    const bool large = n >= kScanLargeFrom;
    if (large && count == 1)
        k_scan_onepass<kScanItemsLarge, 1><<<tiles, kScanThreads, 0, stream>>>(a0, a1, n, r, vec);
    else if (large)
        k_scan_onepass<kScanItemsLarge, 2><<<tiles, kScanThreads, 0, stream>>>(a0, a1, n, r, vec);
    else if (count == 1)
        k_scan_onepass<kScanItems, 1><<<tiles, kScanThreads, 0, stream>>>(a0, a1, n, r, vec);
    else
        k_scan_onepass<kScanItems, 2><<<tiles, kScanThreads, 0, stream>>>(a0, a1, n, r, vec);
    return hipGetLastError();
}

hipError_t scan_u32(const uint32_t* in, uint32_t* out, uint32_t* total, int64_t n, bool inclusive, void* scratch, hipStream_t stream,
    bool cleared)
{
    const ScanArray a { in, out, total, inclusive };
    return scan_u32_arrays(&a, 1, n, scratch, stream, cleared);
}

// ------------------------------------------------------------------ radix sort
namespace {

constexpr int kSortThreads = 256;
// items per thread: 8 (2 048 per block) below kSmallSort keys, so small
// sorts still give every CU several workgroups; 16 (4 096) below kLargeSort;
// 32 above, where the longer per-digit runs of an 8 192-item block write more
// whole lines (433M keys: 6.6 -> 5.4 ms) and the grid is large enough to fill
// the chip at the lower occupancy (15M keys are faster with 16: 0.155 ms,
// against 0.171 with 8 and more with 32)
constexpr int kTinyItems = 8;
constexpr int kSmallItems = 16;
constexpr int kLargeItems = 32;
constexpr int64_t kSmallSort = int64_t(4) << 20;
constexpr int64_t kLargeSort = int64_t(48) << 20;
constexpr int kRadixMax = 256;

__device__ __forceinline__ int64_t sort_count(int64_t n_max, const uint32_t* n_dev)
{
    return n_dev ? min(int64_t(*n_dev), n_max) : n_max;
}

// per-block histogram of the digit (key >> shift) & (2^bits - 1), stored
// digit-major: hist[digit * n_blocks + block]; also clears the state of the
// scan over hist that follows
template <int ITEMS, class KeyT>
__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const KeyT* __restrict__ keys, int64_t n_max, const uint32_t* n_dev,
    int shift, int bits, int n_blocks, uint32_t* __restrict__ hist, ScanState scan, int vec)
{
    constexpr int kSortItems = ITEMS, kSortTile = kSortThreads * ITEMS;
    __shared__ uint32_t s_hist[4][kRadixMax];
    scan_clear(scan);
    const int64_t n = sort_count(n_max, n_dev);
    const int64_t base = int64_t(blockIdx.x) * kSortTile;
    const uint32_t mask = (1u << bits) - 1u;
    for (int i = threadIdx.x; i < 4 * kRadixMax; i += kSortThreads)
        (&s_hist[0][0])[i] = 0;
    __syncthreads();
    if (base < n) {
        const unsigned wave = threadIdx.x / 64;
        // each thread counts ITEMS consecutive keys (the order of counting
        // does not matter), read with 16-byte loads when they are all in
        // range; every load is in flight before the LDS counts
        const int64_t first = base + int64_t(threadIdx.x) * kSortItems;
        uint32_t k[kSortItems];
        if (vec && first + kSortItems <= n) {
            constexpr int kPerVec = 16 / int(sizeof(KeyT));
            const uint4* src = reinterpret_cast<const uint4*>(keys + first);
#pragma unroll
            for (int v = 0; v < kSortItems / kPerVec; ++v) {
                const uint4 q = src[v];
                const uint32_t w[4] = { q.x, q.y, q.z, q.w };
#pragma unroll
                for (int e = 0; e < kPerVec; ++e)
                    k[v * kPerVec + e] = sizeof(KeyT) == 2 ? ((w[e / 2] >> (16 * (e % 2))) & 0xffffu) : w[e];
            }
#pragma unroll
            for (int j = 0; j < kSortItems; ++j)
                atomicAdd(&s_hist[wave][(k[j] >> shift) & mask], 1u);
        } else {
#pragma unroll
            for (int j = 0; j < kSortItems; ++j)
                k[j] = first + j < n ? uint32_t(keys[first + j]) : 0u;
#pragma unroll
            for (int j = 0; j < kSortItems; ++j)
                if (first + j < n)
                    atomicAdd(&s_hist[wave][(k[j] >> shift) & mask], 1u);
        }
    }
    __syncthreads();
    if (threadIdx.x <= mask) {
        const uint32_t c = s_hist[0][threadIdx.x] + s_hist[1][threadIdx.x] + s_hist[2][threadIdx.x] + s_hist[3][threadIdx.x];
        hist[int64_t(threadIdx.x) * n_blocks + blockIdx.x] = c;
    }
}

// Stable scatter.  Wave w of the block owns a contiguous quarter of the
// block's items and walks it in rounds of 64: lanes holding the same digit
// find each other with one ballot per digit bit, the lowest of them advances
// the wave's running count of that digit, and every item keeps its rank
// inside (wave, digit).  One block scan over (digit, wave) then places every
// item in LDS in digit order (stable: digit, wave, round, lane = input
// order), and the LDS copy leaves the block one contiguous run per digit.
// Four barriers per block.  Tile b is workgroup b: the hardware deals
// consecutive workgroups to different XCDs, so the tiles in flight are
// neighbours and their runs of a digit land close together in DRAM (giving
// each XCD a contiguous range of tiles instead made a 434M-key pass three
// times slower).  RADIX (256, or 128 for digits of at most 7
// bits) sizes the per-digit tables: 128 keeps a 16-bit-key block of 8 192
// items within a third of the CU's LDS, so three blocks fit instead of two.
// BITS: the digit width when known at compile time (the bit loop then has no
// run-time tests: with a run-time width each round's eight tests cost VALU
// work to turn the uniform condition into a lane mask and back); 0 reads bits.
template <int ITEMS, class KeyT, int RADIX, int BITS>
__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(const KeyT* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    KeyT* __restrict__ keys_out, uint32_t* __restrict__ vals_out, int64_t n_max, const uint32_t* n_dev, int shift, int bits_arg, int n_blocks,
    const uint32_t* __restrict__ offsets, KeyOut ko)
{
    const int bits = BITS > 0 ? BITS : bits_arg;
    constexpr int kSortTile = kSortThreads * ITEMS;
    constexpr int kWaves = kSortThreads / 64;
    constexpr int kRounds = kSortTile / kSortThreads; // rounds of 64 items per wave
    __shared__ KeyT s_keys[kSortTile];
    __shared__ uint32_t s_vals[kSortTile];
    __shared__ uint32_t s_count[kWaves][RADIX]; // per-wave digit counts, then per-wave starts in LDS order
    __shared__ uint32_t s_local[RADIX];         // block-local start of each digit
    __shared__ uint32_t s_global[RADIX];        // global start of each digit for this block
    __shared__ uint32_t s_scan[kWaves];
    const int64_t n = sort_count(n_max, n_dev);
    const int blk = int(blockIdx.x);
    const int64_t base = int64_t(blk) * kSortTile;
    if (base >= n)
        return;
    const int count = int(min<int64_t>(kSortTile, n - base));
    const unsigned lane = lane_id();
    const unsigned wave = threadIdx.x / 64;
    const uint64_t lt_mask = (uint64_t(1) << lane) - 1;
    const uint32_t mask = (1u << bits) - 1u;
    for (int i = threadIdx.x; i < kWaves * RADIX; i += kSortThreads)
        (&s_count[0][0])[i] = 0;
    __syncthreads();
    uint32_t keys[kRounds], vals[kRounds], rank[sizeof(KeyT) == 2 ? 1 : kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        const int i = int(wave) * (kRounds * 64) + r * 64 + int(lane);
        const bool valid = i < count;
        keys[r] = valid ? uint32_t(keys_in[base + i]) : 0u;
        vals[r] = valid ? vals_in[base + i] : 0u;
    }
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        const int i = int(wave) * (kRounds * 64) + r * 64 + int(lane);
        const bool valid = i < count;
        const uint32_t digit = (keys[r] >> shift) & mask;
        const uint64_t valid_lanes = __ballot(valid);
        // the lanes whose digit equals this lane's, bit by bit: with s = 0 or
        // ~0 the sign extension of this lane's bit b, the lanes agreeing on
        // bit b are ~(ballot(bit b) ^ s), one three-input bit operation per
        // half of the mask (half the instructions of selecting m or ~m)
        uint32_t peers_lo = uint32_t(valid_lanes), peers_hi = uint32_t(valid_lanes >> 32);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            if (BITS > 0 ? b < BITS : b < bits) {
                const uint32_t s = uint32_t(int32_t(digit << (31 - b)) >> 31);
                const uint64_t m = __ballot(s != 0u);
                peers_lo &= ~(uint32_t(m) ^ s);
                peers_hi &= ~(uint32_t(m >> 32) ^ s);
            }
        }
        const uint64_t peers = (uint64_t(peers_hi) << 32) | peers_lo;
        // the wave's LDS operations run in order: every lane reads the running
        // count before the lowest lane of its digit group writes it back
        const uint32_t before = s_count[wave][digit];
        const uint32_t rk = before + uint32_t(__popcll(peers & lt_mask));
        // a 16-bit key and its rank (below 2^16) share one register.  The
        // empty asm makes the value opaque, so the rank is formed here and
        // the placement recomputes its counter address: otherwise the
        // compiler keeps every round's counter value and address live until
        // the placement (about 64 registers at 32 rounds, one wave fewer)
        if constexpr (sizeof(KeyT) == 2) {
            keys[r] |= rk << 16;
            asm volatile("" : "+v"(keys[r]));
        } else {
            rank[r] = rk;
            asm volatile("" : "+v"(rank[r]));
        }
        if (valid && (peers & lt_mask) == 0)
            s_count[wave][digit] = before + uint32_t(__popcll(peers));
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    {
        // thread d: starts of digit d per wave, block-exclusive over digits
        const uint32_t d = threadIdx.x;
        const bool has = d < uint32_t(RADIX);
        uint32_t c[kWaves], tot = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            c[w] = has ? s_count[w][d] : 0u;
            tot += c[w];
        }
        uint32_t v = tot;
#pragma unroll
        for (int k = 1; k < 64; k <<= 1) {
            const uint32_t o = __shfl_up(v, k, 64);
            if (lane >= unsigned(k))
                v += o;
        }
        if (lane == 63)
            s_scan[wave] = v;
        __syncthreads();
        uint32_t start = v - tot;
        for (unsigned w = 0; w < wave; ++w)
            start += s_scan[w];
        if (has) {
            s_local[d] = start;
            if (d <= mask)
                s_global[d] = offsets[int64_t(d) * n_blocks + blk];
#pragma unroll
            for (int w = 0; w < kWaves; ++w) {
                s_count[w][d] = start;
                start += c[w];
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        const int i = int(wave) * (kRounds * 64) + r * 64 + int(lane);
        if (i < count) {
            const uint32_t key = sizeof(KeyT) == 2 ? (keys[r] & 0xffffu) : keys[r];
            const uint32_t rk = sizeof(KeyT) == 2 ? (keys[r] >> 16) : rank[r];
            const uint32_t pos = s_count[wave][(key >> shift) & mask] + rk;
            s_keys[pos] = KeyT(key);
            s_vals[pos] = vals[r];
        }
    }
    __syncthreads();
    const auto write_item = [&](int q) {
        const KeyT key = s_keys[q];
        const uint32_t d = (uint32_t(key) >> shift) & mask;
        const uint32_t dst = s_global[d] + (uint32_t(q) - s_local[d]);
        const uint32_t val = s_vals[q];
        if (ko.n_per_view == 0)
            keys_out[dst] = key;
        else // the view of the id replaces the key (culled_view for a culled key 0xffffffff when set)
            keys_out[dst] = KeyT((ko.culled_view != 0xffffffffu && uint32_t(key) == 0xffffffffu) ? ko.culled_view : val / ko.n_per_view);
        vals_out[dst] = val;
    };
    if (count == kSortTile) {
        // a full block: a fixed trip count, unrolled by 8, so the LDS reads
        // of several items are in flight together
#pragma unroll 8
        for (int r = 0; r < kRounds; ++r)
            write_item(r * kSortThreads + int(threadIdx.x));
    } else {
        for (int q = threadIdx.x; q < count; q += kSortThreads)
            write_item(q);
    }
}

// ------------------------------------------------------------- compaction
constexpr int kCompactThreads = 256;
constexpr int kCompactItems = 16;
constexpr int kCompactTile = kCompactThreads * kCompactItems;

__global__ __launch_bounds__(kCompactThreads) void k_compact_count(const uint32_t* __restrict__ keys, int64_t n, uint32_t skip_key,
    uint32_t* __restrict__ block_counts)
{
    __shared__ uint32_t s_wave[kCompactThreads / 64];
    const int64_t base = int64_t(blockIdx.x) * kCompactTile;
    uint32_t c = 0;
#pragma unroll
    for (int j = 0; j < kCompactItems; ++j) { // every load in flight at once
        const int64_t i = base + int64_t(j) * kCompactThreads + threadIdx.x;
        c += (i < n && keys[i] != skip_key) ? 1u : 0u;
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1)
        c += __shfl_xor(c, d, 64);
    if (lane_id() == 0)
        s_wave[threadIdx.x / 64] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kCompactThreads / 64; ++w)
            t += s_wave[w];
        block_counts[blockIdx.x] = t;
    }
}

// Each wave owns a contiguous quarter of the block's items and walks it in
// rounds of 64; kept items take ranks by ballot prefix, so the output keeps
// the input order.  vals null: each item's value is its input index.
__global__ __launch_bounds__(kCompactThreads) void k_compact_scatter(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals,
    int64_t n, uint32_t skip_key, const uint32_t* __restrict__ block_offsets, uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out)
{
    constexpr int kWaves = kCompactThreads / 64;
    constexpr int kRounds = kCompactTile / kCompactThreads;
    __shared__ uint32_t s_wave[kWaves];
    const int64_t base = int64_t(blockIdx.x) * kCompactTile;
    const unsigned lane = lane_id(), wave = threadIdx.x / 64;
    const uint64_t lt = (uint64_t(1) << lane) - 1;
    uint32_t k[kRounds], v[kRounds];
    bool keep[kRounds];
    uint32_t wave_count = 0;
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        const int64_t i = base + int64_t(wave) * (kRounds * 64) + r * 64 + lane;
        k[r] = i < n ? keys[i] : skip_key;
        v[r] = i < n ? (vals ? vals[i] : uint32_t(i)) : 0u;
        keep[r] = i < n && k[r] != skip_key;
        wave_count += uint32_t(__popcll(__ballot(keep[r])));
    }
    if (lane == 0)
        s_wave[wave] = wave_count;
    __syncthreads();
    uint32_t pos = block_offsets[blockIdx.x];
    for (unsigned w = 0; w < wave; ++w)
        pos += s_wave[w];
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        const uint64_t m = __ballot(keep[r]);
        if (keep[r]) {
            const uint32_t o = pos + uint32_t(__popcll(m & lt));
            keys_out[o] = k[r];
            vals_out[o] = v[r];
        }
        pos += uint32_t(__popcll(m));
    }
}

} // namespace

int64_t compact_blocks(int64_t n) { return std::max<int64_t>(1, ceil_div(n, kCompactTile)); }

hipError_t compact_pairs(const uint32_t* keys_in, const uint32_t* vals_in, int64_t n, uint32_t skip_key, uint32_t* keys_out,
    uint32_t* vals_out, uint32_t* count, uint32_t* block_counts, uint32_t* block_offsets, void* scratch, hipStream_t stream)
{
    const int64_t blocks = compact_blocks(n);
    k_compact_count<<<unsigned(blocks), kCompactThreads, 0, stream>>>(keys_in, n, skip_key, block_counts);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    e = scan_u32(block_counts, block_offsets, count, blocks, false, scratch, stream);
    if (e != hipSuccess)
        return e;
    k_compact_scatter<<<unsigned(blocks), kCompactThreads, 0, stream>>>(keys_in, vals_in, n, skip_key, block_offsets, keys_out, vals_out);
    return hipGetLastError();
}

int radix_passes(int begin_bit, int end_bit)
{
    return end_bit <= begin_bit ? 0 : (end_bit - begin_bit + 7) / 8;
}

// the passes split the bits evenly (13 bits: 7 + 6, not 8 + 5): fewer
// digits per pass mean longer runs per digit in each block's output
static int digit_bits(int begin_bit, int end_bit)
{
    const int passes = radix_passes(begin_bit, end_bit);
    return passes == 0 ? 0 : (end_bit - begin_bit + passes - 1) / passes;
}


size_t radix_scratch_bytes(int64_t n_max)
{
    // sized for the smallest tile, so it covers any sort of up to n_max keys
    const int64_t blocks = std::max<int64_t>(1, ceil_div(n_max, int64_t(kSortThreads) * kTinyItems));
    const size_t hist = size_t(blocks) * kRadixMax * sizeof(uint32_t);
    return ((hist + 255) / 256) * 256 + scan_scratch_bytes(blocks * kRadixMax);
}

template <int ITEMS, class KeyT>
static void radix_pass(int blocks, const KeyT* kin, const uint32_t* vin, KeyT* kout, uint32_t* vout, int64_t n_max, const uint32_t* n_dev,
    int shift, int bits, uint32_t* hist, void* scan_scratch, KeyOut ko, hipStream_t stream, hipError_t* err)
{
    const int64_t n_hist = int64_t(blocks) << bits;
    // the histogram grid clears the scan's state (it has a thread per status word and more)
    const ScanState scan = scan_state(scan_scratch, n_hist);
    if (scan.n_status > int64_t(blocks) * kSortThreads) {
        *err = hipErrorInvalidValue;
        return;
    }
    const int vec = reinterpret_cast<uintptr_t>(kin) % 16 == 0 ? 1 : 0;
    k_radix_hist<ITEMS, KeyT><<<blocks, kSortThreads, 0, stream>>>(kin, n_max, n_dev, shift, bits, blocks, hist, scan, vec);
    *err = scan_u32(hist, hist, nullptr, n_hist, false, scan_scratch, stream, true);
    if (*err != hipSuccess)
        return;
    // the digit widths of the binnings (8-bit depth passes, 7 + 6 for tile
    // ids) are compiled in; other widths read theirs at run time
    if (bits == 8)
        k_radix_scatter<ITEMS, KeyT, 256, 8><<<blocks, kSortThreads, 0, stream>>>(kin, vin, kout, vout, n_max, n_dev, shift, bits, blocks, hist, ko);
    else if (bits == 7)
        k_radix_scatter<ITEMS, KeyT, 128, 7><<<blocks, kSortThreads, 0, stream>>>(kin, vin, kout, vout, n_max, n_dev, shift, bits, blocks, hist, ko);
    else if (bits == 6)
        k_radix_scatter<ITEMS, KeyT, 128, 6><<<blocks, kSortThreads, 0, stream>>>(kin, vin, kout, vout, n_max, n_dev, shift, bits, blocks, hist, ko);
    else if (bits <= 7)
        k_radix_scatter<ITEMS, KeyT, 128, 0><<<blocks, kSortThreads, 0, stream>>>(kin, vin, kout, vout, n_max, n_dev, shift, bits, blocks, hist, ko);
    else
        k_radix_scatter<ITEMS, KeyT, 256, 0><<<blocks, kSortThreads, 0, stream>>>(kin, vin, kout, vout, n_max, n_dev, shift, bits, blocks, hist, ko);
    *err = hipGetLastError();
}

template <class KeyT>
static hipError_t radix_sort_impl(const SortBuffersT<KeyT>& buf, int64_t n_max, const uint32_t* n_dev, int begin_bit, int end_bit, void* scratch,
    hipStream_t stream, KeyOut last = KeyOut {})
{
    const int passes = radix_passes(begin_bit, end_bit);
    if (n_max <= 0 || passes == 0)
        return hipSuccess;
    const int width = digit_bits(begin_bit, end_bit);
    // items per block: 2 048 below 4M keys (1M keys in 4 096-item blocks left
    // the 256 CUs with one workgroup each), 4 096, and 8 192 from 48M keys up
    // (16-bit keys too: 434M of them sort in 5.13 ms against 5.36 with 4 096)
    const int items = n_max < kSmallSort ? kTinyItems : (n_max < kLargeSort ? kSmallItems : kLargeItems);
    const int blocks = int(std::max<int64_t>(1, ceil_div(n_max, int64_t(kSortThreads) * items)));
    uint32_t* hist = static_cast<uint32_t*>(scratch);
    void* scan_scratch = static_cast<char*>(scratch) + ((size_t(blocks) * kRadixMax * sizeof(uint32_t) + 255) / 256) * 256;
    for (int p = 0; p < passes; ++p) {
        const int shift = begin_bit + width * p;
        const int bits = std::min(width, end_bit - shift);
        const KeyT* kin = (p % 2 == 0) ? buf.keys_a : buf.keys_b;
        const uint32_t* vin = (p % 2 == 0) ? buf.vals_a : buf.vals_b;
        KeyT* kout = (p % 2 == 0) ? buf.keys_b : buf.keys_a;
        uint32_t* vout = (p % 2 == 0) ? buf.vals_b : buf.vals_a;
        const KeyOut ko = p + 1 == passes ? last : KeyOut {};
        hipError_t e = hipSuccess;
        if (items == kTinyItems)
            radix_pass<kTinyItems>(blocks, kin, vin, kout, vout, n_max, n_dev, shift, bits, hist, scan_scratch, ko, stream, &e);
        else if (items == kSmallItems)
            radix_pass<kSmallItems>(blocks, kin, vin, kout, vout, n_max, n_dev, shift, bits, hist, scan_scratch, ko, stream, &e);
        else
            radix_pass<kLargeItems>(blocks, kin, vin, kout, vout, n_max, n_dev, shift, bits, hist, scan_scratch, ko, stream, &e);
        if (e != hipSuccess)
            return e;
    }
    return hipSuccess;
}

hipError_t radix_sort_pairs(const SortBuffers& buf, int64_t n_max, const uint32_t* n_dev, int begin_bit, int end_bit, void* scratch,
    hipStream_t stream, KeyOut last_pass_keys)
{
    return radix_sort_impl(buf, n_max, n_dev, begin_bit, end_bit, scratch, stream, last_pass_keys);
}

hipError_t radix_sort_pairs(const SortBuffers16& buf, int64_t n_max, const uint32_t* n_dev, int begin_bit, int end_bit, void* scratch,
    hipStream_t stream)
{
    return end_bit > 16 ? hipErrorInvalidValue : radix_sort_impl(buf, n_max, n_dev, begin_bit, end_bit, scratch, stream);
}

} // namespace dgmr
