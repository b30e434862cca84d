// SPDX-License-Identifier: GPL-3.0-or-later
// Tile splatter for gfx950: forward (reference: src/dgmr/splat_forward.cu:30-338)
// and the backward the reference does not ship.
//
// Pipeline (one stream, one 4-byte read-back for the pair count):
//   1. k_splat_preprocess   per Gaussian: EWA splat (Ots, 0.3 px^2 filter), rect,
//                           tiles touched, depth key, SH colour, conic.
//   2. radix sort of the Gaussians by depth bits (stable, ties by index)
//   3. k_gather_owners      tiles touched and tile rectangle in depth order,
//                           exclusive scan -> pair offsets
//   4. k_emit_bounds, k_emit_pairs   pair-balanced emission of (tile id, Gaussian id)
//   5. radix sort of the pairs by tile id only (stable)
//      Steps 2-5 yield exactly the order of the reference's single sort on
//      (tile << 32 | depth bits): within a tile, pairs stay in (depth bits,
//      Gaussian id) order, which is the reference's stable order.
//   6. k_tile_ranges        [first, last) per tile (binary search per tile)
//   7. k_blend_forward      per tile, LDS-staged batches of Gaussians
// Backward: k_blend_backward (front-to-back replay, transmittance recomputed,
// wave-reduced atomics) then k_splat_preprocess_backward (per Gaussian).
#include <algorithm>
#include <cstring>

#include "device_grad.h"
#include "device_math.h"
#include "host_common.h"
#include "primitives.h"
#include "splat_device.h"

namespace dgmr {

namespace {

constexpr int kPreThreads = 256;
// depth sorts of at least this many keys compact the visible entries first
constexpr int64_t kCompactMin = int64_t(4) << 20;

__global__ __launch_bounds__(kPreThreads) void k_splat_preprocess(int64_t n, SplatParams p, DevCamera<float> cam, SplatOut out)
{
    // SH rows are read straight from global memory: the 12 row loads of a wave
    // cover one contiguous span that stays in cache, and keeping LDS free
    // keeps the occupancy that hides the gather latency
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        preprocess_gaussian(i, i, p, p.sh + i * p.sh_coeffs * 3, cam, out);
}

} // namespace

// Emission of (tile, Gaussian) pairs balanced over pairs: block b writes pairs
// [b * kEmitPairs, (b + 1) * kEmitPairs) of the depth-ordered pair list.  Its
// first owner and an upper bound of its last come from k_emit_bounds; the
// owners in between (culled Gaussians carry the largest depth key, so the
// owners of a block's pairs are contiguous and each owns at least one pair)
// are staged in LDS with their tile rectangle, and every thread writes
// kEmitPer consecutive pairs.  A window wider than the LDS stage (more than
// kEmitWindow owners in one block's pairs, or zero-pair Gaussians among
// visible ones) takes a per-pair search instead.
// views: per-view (grid_x, grid_y, first global tile, -) when ids span several
// views of n_per_view Gaussians each; null for a single view.
constexpr int kEmitThreads = 256;
// pairs per thread: 8, or 16 from kEmitLarge pairs up (the per-thread search
// and setup spread over more pairs: 434M pairs 1.13 -> 1.06 ms; at 15M, with
// half the blocks, 16 was slower, 0.042 against 0.032 ms)
constexpr int kEmitPerSmall = 8, kEmitPerLarge = 16;
constexpr int64_t kEmitLarge = int64_t(48) << 20;
constexpr int kEmitWindow = 2048; // owners a block stages in LDS (more take the per-pair search)

// First owners of the emission blocks: bounds[b] is the Gaussian (position
// in depth order) that owns pair b * block_pairs, for every block that emits,
// and bounds[n_blocks] owns the last emitted pair.  Position k owns pairs
// [offsets[k], offsets[k + 1]) and writes the entries whose pair falls in its
// range, so each entry is written exactly once and no block searches.
__global__ void k_emit_bounds(int64_t n, const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ n_pairs, uint32_t capacity,
    uint32_t block_pairs, uint32_t* __restrict__ bounds)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n)
        return;
    const uint64_t total = *n_pairs;
    const uint64_t R = min(total, uint64_t(capacity)); // pairs past the capacity are not emitted
    const uint64_t lo = offsets[k];
    const uint64_t hi = min(k + 1 < n ? uint64_t(offsets[k + 1]) : total, R);
    if (lo >= hi)
        return;
    for (uint64_t b = (lo + block_pairs - 1) / block_pairs; b * block_pairs < hi; ++b)
        bounds[b] = uint32_t(k);
    if (R - 1 < hi)
        bounds[(R + block_pairs - 1) / block_pairs] = uint32_t(k);
}

// tile rectangle of Gaussian id: first tile and (grid_x << 16 | width); 0 shape when it has no tiles
template <class S>
__device__ __forceinline__ void owner_rect(uint32_t id, const int32_t* __restrict__ rects, const uint32_t* __restrict__ tiles,
    const S* __restrict__ xy, int grid_x, int grid_y, int tile_w, int tile_h, const int4* __restrict__ views, int64_t n_per_view,
    uint32_t* start, uint32_t* shape)
{
    *start = 0;
    *shape = 0;
    if (tiles[id] == 0)
        return;
    int gx = grid_x, gy = grid_y, tbase = 0;
    if (views) {
        const int4 v = views[id / uint32_t(n_per_view)]; // ids < 2^32: a 32-bit division
        gx = v.x;
        gy = v.y;
        tbase = v.z;
    }
    int a, b, c, d;
    get_rect(xy[2 * id], xy[2 * id + 1], rects[2 * id], rects[2 * id + 1], gx, gy, tile_w, tile_h, &a, &b, &c, &d);
    // sort keys are view-local tile ids (the view comes back through the id)
    (void)tbase;
    *start = uint32_t(b) * uint32_t(gx) + uint32_t(a);
    *shape = (uint32_t(gx) << 16) | uint32_t(c - a);
}

// Per position j of the depth order: the pair count of Gaussian ids[j] (for
// the scan) and its tile rectangle (for the emission).  The random reads of
// per-Gaussian data happen here, in a streaming kernel, so the emission reads
// its owners' rectangles as contiguous runs; with the preprocess's packed bin
// rectangle that is one 8-byte read per Gaussian.
template <class S, bool RECTS>
__global__ __launch_bounds__(256) void k_gather_owners(int64_t n, const uint32_t* __restrict__ n_sorted, const uint32_t* __restrict__ ids,
    const uint2* __restrict__ bin_rects,
    const int32_t* __restrict__ rects, const uint32_t* __restrict__ tiles, const S* __restrict__ xy, int grid_x, int grid_y, int tile_w,
    int tile_h, const int4* __restrict__ views, int64_t n_per_view, uint32_t* __restrict__ counts, uint2* __restrict__ owner_rects,
    ScanState scan)
{
    scan_clear(scan); // for the scan of counts that follows
    const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= n)
        return;
    if (n_sorted && j >= int64_t(*n_sorted)) {
        // past the compacted (visible) entries: no pairs
        counts[j] = 0;
        owner_rects[j] = make_uint2(0u, 0u);
        return;
    }
    const uint32_t id = ids[j];
    uint32_t start, shape, count;
    if constexpr (RECTS) {
        const uint2 r = bin_rects[id];
        const uint32_t cols = r.y & 0xffffu, rows = r.y >> 16;
        const uint32_t gx = views ? uint32_t(views[id / uint32_t(n_per_view)].x) : uint32_t(grid_x);
        start = r.x;
        shape = cols * rows == 0 ? 0u : (gx << 16) | cols;
        count = cols * rows;
    } else {
        owner_rect(id, rects, tiles, xy, grid_x, grid_y, tile_w, tile_h, views, n_per_view, &start, &shape);
        count = tiles[id];
    }
    counts[j] = count;
    owner_rects[j] = make_uint2(start, shape);
}

// kEmitPer consecutive pairs at a 16-byte aligned position: 32-bit keys and
// the values four to a store, 16-bit keys packed eight to a store
template <int kEmitPer>
__device__ __forceinline__ void store_run(uint32_t* keys, uint32_t* vals, const uint32_t (&kv)[kEmitPer], const uint32_t (&vv)[kEmitPer])
{
    static_assert(kEmitPer % 8 == 0, "pairs per thread: a multiple of 8");
    uint4* kd = reinterpret_cast<uint4*>(keys);
    uint4* vd = reinterpret_cast<uint4*>(vals);
#pragma unroll
    for (int j = 0; j < kEmitPer / 4; ++j) {
        kd[j] = make_uint4(kv[4 * j], kv[4 * j + 1], kv[4 * j + 2], kv[4 * j + 3]);
        vd[j] = make_uint4(vv[4 * j], vv[4 * j + 1], vv[4 * j + 2], vv[4 * j + 3]);
    }
}
template <int kEmitPer>
__device__ __forceinline__ void store_run(uint16_t* keys, uint32_t* vals, const uint32_t (&kv)[kEmitPer], const uint32_t (&vv)[kEmitPer])
{
    uint4* kd = reinterpret_cast<uint4*>(keys);
    uint4* vd = reinterpret_cast<uint4*>(vals);
#pragma unroll
    for (int j = 0; j < kEmitPer / 8; ++j)
        kd[j] = make_uint4(kv[8 * j] | (kv[8 * j + 1] << 16), kv[8 * j + 2] | (kv[8 * j + 3] << 16), kv[8 * j + 4] | (kv[8 * j + 5] << 16),
            kv[8 * j + 6] | (kv[8 * j + 7] << 16));
#pragma unroll
    for (int j = 0; j < kEmitPer / 4; ++j)
        vd[j] = make_uint4(vv[4 * j], vv[4 * j + 1], vv[4 * j + 2], vv[4 * j + 3]);
}

// Tight culling: per position j of the depth order, the pair count of
// Gaussian ids[j] (tiles its reach ellipse meets, row by row) and its record,
// copied so the emission reads its owners' records as contiguous runs.
// Entries the preprocess culled have no record: they are those past n_sorted
// (compacted sorts) or whose final sort key is culled_key.
__global__ __launch_bounds__(256) void k_gather_reach(int64_t n, const uint32_t* __restrict__ n_sorted, const uint32_t* __restrict__ keys,
    uint32_t culled_key, const uint32_t* __restrict__ ids, const ReachRec* __restrict__ reach, int tile_w, int tile_h,
    uint32_t* __restrict__ counts, ReachRec* __restrict__ owners, ScanState scan)
{
    scan_clear(scan); // for the scan of counts that follows
    const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= n)
        return;
    uint4* dst = reinterpret_cast<uint4*>(owners + j);
    if ((n_sorted && j >= int64_t(*n_sorted)) || keys[j] == culled_key) {
        counts[j] = 0;
        dst[0] = make_uint4(0u, 0u, 0u, 0u);
        return;
    }
    const uint4* src = reinterpret_cast<const uint4*>(reach + ids[j]);
    const uint4 w0 = src[0], w1 = src[1];
    uint32_t count = 0;
    if (w0.y != 0u) {
        const ReachRec r { w0.x, w0.y, w0.z, __uint_as_float(w0.w), __uint_as_float(w1.x), __uint_as_float(w1.y), __uint_as_float(w1.z),
            __uint_as_float(w1.w) };
        const ReachEllipse e = reach_ellipse(r);
        const int x0 = int(r.origin & 0xffffu), y0 = int(r.origin >> 16), cols = int(r.dims & 0xffffu), rows = int(r.dims >> 16);
        for (int q = 0; q < rows; ++q) {
            const int2 c = reach_columns(e, y0 + q, x0, cols, tile_w, tile_h);
            if (c.x <= c.y)
                count += uint32_t(c.y - c.x + 1);
        }
    }
    counts[j] = count;
    dst[0] = w0;
    dst[1] = w1;
}

// the emission cursor of one owner: its tile row r (from the rectangle's
// first), that row's reached columns c, and the column offset within them
struct ReachCursor {
    ReachEllipse e;
    int x0, y0, cols, rows, gx;
    int r;
    int2 c;
};

__device__ __forceinline__ void cursor_load(ReachCursor& cu, const ReachRec* __restrict__ owners, uint32_t k, int tile_w, int tile_h)
{
    const uint4* src = reinterpret_cast<const uint4*>(owners + k);
    const uint4 w0 = src[0], w1 = src[1];
    const ReachRec r { w0.x, w0.y, w0.z, __uint_as_float(w0.w), __uint_as_float(w1.x), __uint_as_float(w1.y), __uint_as_float(w1.z),
        __uint_as_float(w1.w) };
    cu.e = reach_ellipse(r);
    cu.x0 = int(r.origin & 0xffffu);
    cu.y0 = int(r.origin >> 16);
    cu.cols = int(r.dims & 0xffffu);
    cu.rows = int(r.dims >> 16);
    cu.gx = int(r.grid_x);
    cu.r = 0;
    cu.c = reach_columns(cu.e, cu.y0, cu.x0, cu.cols, tile_w, tile_h);
}

// moves the cursor to the row holding column offset *local (from the current
// row) and returns the tile; rows past the rectangle are never entered
__device__ __forceinline__ uint32_t cursor_tile(ReachCursor& cu, uint32_t* local, int tile_w, int tile_h)
{
    while (cu.r < cu.rows) {
        const uint32_t w = cu.c.x <= cu.c.y ? uint32_t(cu.c.y - cu.c.x + 1) : 0u;
        if (*local < w)
            break;
        *local -= w;
        ++cu.r;
        if (cu.r < cu.rows)
            cu.c = reach_columns(cu.e, cu.y0 + cu.r, cu.x0, cu.cols, tile_w, tile_h);
    }
    return uint32_t(cu.y0 + cu.r) * uint32_t(cu.gx) + uint32_t(cu.c.x) + *local;
}

// Tight-culling emission: the same block layout as k_emit_pairs; each thread
// writes kEmitPer consecutive pairs, walking its owners' reached rows.
template <class KeyT, int kEmitPer>
__global__ __launch_bounds__(kEmitThreads) void k_emit_reach(const uint32_t* __restrict__ n_pairs, uint32_t capacity,
    const uint32_t* __restrict__ sorted_ids, const uint32_t* __restrict__ bounds, const uint32_t* __restrict__ offsets,
    const ReachRec* __restrict__ owners, int tile_w, int tile_h, KeyT* __restrict__ keys_out, uint32_t* __restrict__ vals_out)
{
    constexpr int kEmitPairs = kEmitThreads * kEmitPer;
    __shared__ uint32_t s_off[kEmitWindow];
    const uint32_t R = min(*n_pairs, capacity);
    const uint32_t first = blockIdx.x * uint32_t(kEmitPairs);
    if (first >= R)
        return;
    const uint32_t end = min(R, first + uint32_t(kEmitPairs));
    const uint32_t g_lo = bounds[blockIdx.x], g_hi = bounds[blockIdx.x + 1];
    const uint32_t window = g_hi - g_lo + 1;
    const bool staged = window <= uint32_t(kEmitWindow);
    if (staged) {
        for (uint32_t k = threadIdx.x; k < window; k += kEmitThreads)
            s_off[k] = offsets[g_lo + k];
        __syncthreads();
    }
    // start of owner g_lo + k's pairs; past the window: no end within this block
    auto off = [&](uint32_t k) -> uint32_t {
        if (k >= window)
            return 0xffffffffu;
        return staged ? s_off[k] : offsets[g_lo + k];
    };
    const uint32_t p0 = first + threadIdx.x * kEmitPer;
    if (p0 >= end)
        return;
    uint32_t lo = 0, hi = window - 1; // owner of p0: last k with off(k) <= p0
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (off(mid) <= p0)
            lo = mid;
        else
            hi = mid - 1;
    }
    uint32_t k = lo;
    uint32_t next = off(k + 1);
    uint32_t local = p0 - off(k);
    ReachCursor cu;
    cursor_load(cu, owners, g_lo + k, tile_w, tile_h);
    uint32_t id = sorted_ids[g_lo + k];
    uint32_t kv[kEmitPer], vv[kEmitPer];
#pragma unroll
    for (int q = 0; q < kEmitPer; ++q) {
        const uint32_t p = p0 + q;
        if (p >= end) {
            kv[q] = vv[q] = 0u;
            continue;
        }
        if (p >= next) {
            do {
                ++k;
                next = off(k + 1);
            } while (p >= next);
            cursor_load(cu, owners, g_lo + k, tile_w, tile_h);
            id = sorted_ids[g_lo + k];
            local = 0;
        }
        kv[q] = cursor_tile(cu, &local, tile_w, tile_h);
        vv[q] = id;
        ++local;
    }
    const bool aligned = ((reinterpret_cast<uintptr_t>(keys_out) | reinterpret_cast<uintptr_t>(vals_out)) & 15u) == 0;
    if (aligned && p0 + kEmitPer <= end) {
        store_run(keys_out + p0, vals_out + p0, kv, vv);
    } else {
#pragma unroll
        for (int q = 0; q < kEmitPer; ++q)
            if (p0 + q < end) {
                keys_out[p0 + q] = KeyT(kv[q]);
                vals_out[p0 + q] = vv[q];
            }
    }
}

template <class KeyT, int kEmitPer>
__global__ __launch_bounds__(kEmitThreads) void k_emit_pairs(int64_t n, const uint32_t* __restrict__ n_pairs, uint32_t capacity,
    const uint32_t* __restrict__ sorted_ids, const uint32_t* __restrict__ bounds, const uint32_t* __restrict__ offsets,
    const uint2* __restrict__ owner_rects, KeyT* __restrict__ keys_out, uint32_t* __restrict__ vals_out)
{
    constexpr int kEmitPairs = kEmitThreads * kEmitPer;
    __shared__ uint32_t s_off[kEmitWindow], s_id[kEmitWindow], s_start[kEmitWindow], s_shape[kEmitWindow];
    const uint32_t R = min(*n_pairs, capacity); // pairs past the capacity are not written
    const uint32_t first = blockIdx.x * uint32_t(kEmitPairs);
    if (first >= R)
        return;
    const uint32_t end = min(R, first + uint32_t(kEmitPairs));
    // owners of the block's first pair and of the next block's first pair (or
    // of the last pair): the owners of this block's pairs lie in between
    const uint32_t g_lo = bounds[blockIdx.x], g_hi = bounds[blockIdx.x + 1];
    const uint32_t window = g_hi - g_lo + 1;
    if (window <= uint32_t(kEmitWindow)) {
        for (uint32_t k = threadIdx.x; k < window; k += kEmitThreads) {
            const uint2 r = owner_rects[g_lo + k];
            s_off[k] = offsets[g_lo + k];
            s_id[k] = sorted_ids[g_lo + k];
            s_start[k] = r.x;
            s_shape[k] = r.y;
        }
        __syncthreads();
        const uint32_t p0 = first + threadIdx.x * kEmitPer;
        if (p0 >= end)
            return;
        // owner of p0: last k with s_off[k] <= p0
        uint32_t lo = 0, hi = window - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (s_off[mid] <= p0)
                lo = mid;
            else
                hi = mid - 1;
        }
        uint32_t k = lo;
        uint32_t next = k + 1 < window ? s_off[k + 1] : 0xffffffffu;
        uint32_t shape = s_shape[k], local = p0 - s_off[k];
        uint32_t w = shape & 0xffffu;
        uint32_t ty = local / w, tx = local - ty * w;
        uint32_t kv[kEmitPer], vv[kEmitPer];
#pragma unroll
        for (int q = 0; q < kEmitPer; ++q) {
            const uint32_t p = p0 + q;
            if (p >= next) {
                do {
                    ++k;
                    next = k + 1 < window ? s_off[k + 1] : 0xffffffffu;
                } while (p >= next);
                shape = s_shape[k];
                w = shape & 0xffffu;
                tx = 0;
                ty = 0;
            }
            kv[q] = s_start[k] + ty * (shape >> 16) + tx;
            vv[q] = s_id[k];
            if (++tx == w) {
                tx = 0;
                ++ty;
            }
        }
        const bool aligned = ((reinterpret_cast<uintptr_t>(keys_out) | reinterpret_cast<uintptr_t>(vals_out)) & 15u) == 0;
        if (aligned && p0 + kEmitPer <= end) {
            // 16-byte aligned: p0 is a multiple of kEmitPer
            store_run(keys_out + p0, vals_out + p0, kv, vv);
        } else {
#pragma unroll
            for (int q = 0; q < kEmitPer; ++q)
                if (p0 + q < end) {
                    keys_out[p0 + q] = KeyT(kv[q]);
                    vals_out[p0 + q] = vv[q];
                }
        }
    } else {
        for (uint32_t p = first + threadIdx.x; p < end; p += kEmitThreads) {
            uint32_t lo = g_lo, hi = g_hi;
            while (lo < hi) {
                const uint32_t mid = lo + ((hi - lo + 1) >> 1);
                if (offsets[mid] <= p)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            const uint32_t id = sorted_ids[lo];
            const uint2 r = owner_rects[lo];
            const uint32_t start = r.x, shape = r.y;
            const uint32_t local = p - offsets[lo];
            const uint32_t w = shape & 0xffffu;
            const uint32_t ty = local / w, tx = local - ty * w;
            keys_out[p] = KeyT(start + ty * (shape >> 16) + tx);
            vals_out[p] = id;
        }
    }
}

// [first, last) of every tile by binary search over the sorted pairs: one
// thread per tile, every tile written (an empty tile gets (0, 0)).
// The sorted order is (view-local tile, view, depth); a pair's view is
// val / n_per_view, and val >= v * n_per_view exactly when its view is >= v.
// So both ends of tile t of view v are partition points of a predicate that
// is monotone along the array: "key > t, or key == t and val >= bound".  With
// one view the bound is 0 and the two ends are the lower bounds of t and t + 1.
// The two searches of a thread run interleaved.
template <class KeyT>
__global__ __launch_bounds__(256) void k_tile_ranges(const KeyT* __restrict__ keys, const uint32_t* __restrict__ vals,
    const uint32_t* __restrict__ n_pairs, uint32_t capacity, const int4* __restrict__ views, int n_views, int64_t n_per_view, int64_t n_tiles,
    uint32_t* __restrict__ ranges)
{
    const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= n_tiles)
        return;
    const uint32_t R = min(*n_pairs, capacity);
    uint32_t t0 = uint32_t(g), t1 = uint32_t(g) + 1;
    uint64_t b0 = 0, b1 = 0;
    if (views) {
        int lo = 0, hi = n_views - 1; // last view whose first tile is <= g
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (int64_t(views[mid].z) <= g)
                lo = mid;
            else
                hi = mid - 1;
        }
        t0 = t1 = uint32_t(g - views[lo].z);
        b0 = uint64_t(lo) * uint64_t(n_per_view);
        b1 = b0 + uint64_t(n_per_view);
    }
    uint32_t lo0 = 0, hi0 = R, lo1 = 0, hi1 = R;
    while (lo0 < hi0 || lo1 < hi1) {
        const uint32_t m0 = lo0 + (hi0 - lo0) / 2, m1 = lo1 + (hi1 - lo1) / 2;
        const bool a0 = lo0 < hi0, a1 = lo1 < hi1;
        const uint32_t k0 = a0 ? uint32_t(keys[m0]) : 0u, k1 = a1 ? uint32_t(keys[m1]) : 0u;
        const uint32_t v0 = a0 && views ? vals[m0] : 0u, v1 = a1 && views ? vals[m1] : 0u;
        const bool after0 = k0 > t0 || (k0 == t0 && uint64_t(v0) >= b0);
        const bool after1 = k1 > t1 || (k1 == t1 && uint64_t(v1) >= b1);
        if (a0) {
            if (after0)
                hi0 = m0;
            else
                lo0 = m0 + 1;
        }
        if (a1) {
            if (after1)
                hi1 = m1;
            else
                lo1 = m1 + 1;
        }
    }
    // an empty tile holds (0, 0), as the reference's zero-filled ranges
    ranges[2 * g] = lo0 < lo1 ? lo0 : 0u;
    ranges[2 * g + 1] = lo0 < lo1 ? lo1 : 0u;
}

// kBlendWaves independent waves per workgroup, one tile each, LDS private per wave
template <int TW, int TH, int B>
__global__ __launch_bounds__(64 * kBlendWaves) void k_blend_forward(const uint32_t* __restrict__ ranges, const uint32_t* __restrict__ point_list,
    const float2* __restrict__ xy, const float4* __restrict__ conic_opacity, const float* __restrict__ rgb, int width, int height, int grid_x,
    int n_tiles, float bg0, float bg1, float bg2, float* __restrict__ framebuffer, float* __restrict__ final_T)
{
    __shared__ BlendSmem<B> sm_all[kBlendWaves];
    const int wave = int(uniform(threadIdx.x / 64));
    const int tile = xcd_tile(blockIdx.x, gridDim.x) * kBlendWaves + wave;
    if (tile >= n_tiles)
        return;
    BlendSmem<B>& sm = sm_all[wave];
    const uint32_t first = uniform(ranges[2 * tile]), last = uniform(ranges[2 * tile + 1]);
    blend_tile_forward<TW, TH, B>(sm, first, last, (tile % grid_x) * TW, (tile / grid_x) * TH, width, height, bg0, bg1,
        bg2, point_list, xy, conic_opacity, rgb, framebuffer, final_T);
}

template <int TW, int TH, int B>
__global__ __launch_bounds__(64 * kBlendWaves) void k_blend_backward(const uint32_t* __restrict__ ranges, const uint32_t* __restrict__ point_list,
    const float2* __restrict__ xy, const float4* __restrict__ conic_opacity, const float* __restrict__ rgb, int width, int height, int grid_x,
    int n_tiles, float bg0, float bg1, float bg2, const float* __restrict__ framebuffer, const float* __restrict__ final_T,
    const float* __restrict__ grad_fb, float* __restrict__ screen_grads, WavePriority prio)
{
    // kBwdSplit waves per tile, each on TH / kBwdSplit rows: the tile's list
    // is walked once per part, and the longest tile's work is spread over
    // kBwdSplit waves with less per-lane state
    constexpr int SPLIT = (TW * TH / kBwdSplit >= 64) ? kBwdSplit : 1;
    __shared__ BlendSmemBwd<B> sm_all[kBlendWaves];
    const int wave = int(uniform(threadIdx.x / 64));
    const int unit = xcd_tile(blockIdx.x, gridDim.x) * kBlendWaves + wave;
    const int tile = unit / SPLIT, part = unit % SPLIT;
    if (tile >= n_tiles)
        return;
    BlendSmemBwd<B>& sm = sm_all[wave];
    const uint32_t first = uniform(ranges[2 * tile]), last = uniform(ranges[2 * tile + 1]);
    set_wave_priority(last - first, prio);
    blend_tile_backward<TW, TH / SPLIT, B>(sm, first, last, (tile % grid_x) * TW, (tile / grid_x) * TH + part * (TH / SPLIT), width, height,
        bg0, bg1, bg2, point_list, xy, conic_opacity, rgb, framebuffer, final_T, grad_fb, screen_grads);
}

namespace {

// SH gradient rows (3 * sh_coeffs floats per Gaussian, 192 B at degree 3)
// are written by the whole wave: each lane stages its Gaussian's factors
// (16 basis values, 3 colour gradients) in LDS, then the wave writes its 64
// rows as one contiguous span, 256 B per store instruction.  Per-lane row
// stores would leave lines partly written while 64 rows of every resident
// wave compete for L2, and each partial eviction is a separate HBM write.
constexpr int kShStage = 19; // odd row stride: the 64 lanes' rows hit distinct banks

__global__ __launch_bounds__(kPreThreads) void k_splat_preprocess_backward(int64_t n, SplatParams p, DevCamera<float> cam,
    const uint32_t* __restrict__ tiles, const float4* __restrict__ conic_opacity, const uint8_t* __restrict__ clamped,
    const float* __restrict__ screen_grads, float* __restrict__ g_means, float* __restrict__ g_scales, float* __restrict__ g_rotations,
    float* __restrict__ g_opacity, float* __restrict__ g_sh)
{
    __shared__ float s_stage[kPreThreads * kShStage];
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int lane = int(threadIdx.x & 63);
    const int kc = 3 * p.sh_coeffs;
    float* row = s_stage + threadIdx.x * kShStage;
    if (i < n) {
        float g_mu[3] = { 0.f, 0.f, 0.f }, g_s[3] = { 0.f, 0.f, 0.f }, g_q[4] = { 0.f, 0.f, 0.f, 0.f }, g_w = 0.f;
        if (tiles[i] != 0) {
            preprocess_backward_gaussian<ShGrad::Export>(i, p, p.sh + i * kc, cam, conic_opacity[i], clamped + 3 * i, screen_grads + i * 9,
                g_mu, g_s, g_q, &g_w, row);
        } else {
            for (int k = 0; k < kShStage; ++k)
                row[k] = 0.f;
        }
        activation_backward(p, i, g_s, g_q, &g_w);
        for (int a = 0; a < 3; ++a) {
            g_means[3 * i + a] = g_mu[a];
            g_scales[3 * i + a] = g_s[a];
        }
        for (int a = 0; a < 4; ++a)
            g_rotations[4 * i + a] = g_q[a];
        g_opacity[i] = g_w;
    }
    wave_lds_sync();
    const int64_t first = i - lane; // the wave's first Gaussian
    if (first >= n)
        return;
    const int count = int(min<int64_t>(64, n - first)) * kc;
    const float* rows = s_stage + (threadIdx.x - lane) * kShStage;
    float* dst = g_sh + first * kc;
    // f = g * kc + 3 * k + ch; the quotients come from float products, exact
    // here since f < 64 * 48 keeps every quotient 1/(2 kc) away from an integer
    const float inv_kc = 1.f / float(kc);
    for (int f = lane; f < count; f += 64) {
        const int g = int((float(f) + 0.5f) * inv_kc);
        const int e = f - g * kc;
        const int k = int((float(e) + 0.5f) * (1.f / 3.f));
        const float* r = rows + g * kShStage;
        dst[f] = r[k] * r[16 + (e - 3 * k)];
    }
}

SplatParams make_params(const dgmr_gaussians* g)
{
    return SplatParams { static_cast<const float*>(g->means), static_cast<const float*>(g->scales), static_cast<const float*>(g->rotations),
        static_cast<const float*>(g->opacity), static_cast<const float*>(g->sh), g->sh_coeffs, g->sh_degree, float(g->scale_multiplier),
        g->activation };
}

} // namespace

SplatParams make_splat_params(const dgmr_gaussians* g) { return make_params(g); }

hipError_t launch_splat_preprocess(const dgmr_gaussians* g, const DevCamera<float>& cam, const SplatOut& out, hipStream_t stream)
{
    k_splat_preprocess<<<unsigned(ceil_div64(g->n, kPreThreads)), kPreThreads, 0, stream>>>(g->n, make_params(g), cam, out);
    return hipGetLastError();
}

namespace {
// one wave: out[l] = wave_sum8 of in[q * 64 + l] (q < 8), out[64 + l] = wave_sum_to_last of in[512 + l]
__global__ __launch_bounds__(64) void k_selftest_wave_sum(const float* __restrict__ in, float* __restrict__ out)
{
    const int l = int(threadIdx.x);
    out[l] = wave_sum8(in[l], in[64 + l], in[128 + l], in[192 + l], in[256 + l], in[320 + l], in[384 + l], in[448 + l]);
    out[64 + l] = wave_sum_to_last(in[512 + l]);
}

} // namespace

// ------------------------------------------------------------------ host side
struct SplatWorkspace {
    uint32_t* dkeys_a;
    uint32_t* dvals_a;
    uint32_t* dkeys_b;
    uint32_t* dvals_b;
    uint32_t* counts;
    uint32_t* emit_bounds;
    uint2* owner_rects;   // without tight culling
    uint2* bin_rects;
    ReachRec* owner_reach; // with tight culling (the same memory as owner_rects / bin_rects)
    ReachRec* reach;
    uint32_t* pair_offsets;
    uint32_t* pkeys_a;
    uint32_t* pvals_a;
    uint32_t* pkeys_b;
    uint32_t* counter;
    void* sort_scratch;
    void* scan_scratch;
};

static SplatWorkspace carve_splat(Carver& cv, int64_t n, int64_t cap, bool cull)
{
    SplatWorkspace w;
    const int64_t nn = std::max<int64_t>(n, 1), cc = std::max<int64_t>(cap, 1);
    const int64_t rec = cull ? int64_t(sizeof(ReachRec)) : int64_t(sizeof(uint2)); // bytes per Gaussian record
    w.dkeys_a = cv.take<uint32_t>(nn);
    w.dvals_a = cv.take<uint32_t>(nn);
    w.reach = reinterpret_cast<ReachRec*>(cv.take<char>(nn * rec));
    w.bin_rects = reinterpret_cast<uint2*>(w.reach);
    w.dkeys_b = cv.take<uint32_t>(nn);
    w.dvals_b = cv.take<uint32_t>(nn);
    w.counts = cv.take<uint32_t>(nn);
    w.pair_offsets = cv.take<uint32_t>(nn);
    w.pkeys_a = cv.take<uint32_t>(cc);
    w.pvals_a = cv.take<uint32_t>(cc);
    w.pkeys_b = cv.take<uint32_t>(cc);
    w.counter = cv.take<uint32_t>(64);
    w.emit_bounds = cv.take<uint32_t>(cc / (kEmitThreads * kEmitPerSmall) + 2);
    w.owner_reach = reinterpret_cast<ReachRec*>(cv.take<char>(nn * rec));
    w.owner_rects = reinterpret_cast<uint2*>(w.owner_reach);
    w.sort_scratch = cv.take<char>(radix_scratch_bytes(std::max(nn, cc)));
    w.scan_scratch = cv.take<char>(scan_scratch_bytes(nn, 2));
    return w;
}

size_t splat_workspace_bytes(int64_t n, int64_t cap, bool cull)
{
    Carver cv(nullptr, 0);
    carve_splat(cv, n, cap, cull);
    return cv.used;
}

template <int TW, int TH>
static void launch_blend_fwd(int batch, int n_tiles, hipStream_t st, const uint32_t* ranges, const uint32_t* pl, const float2* xy,
    const float4* co, const float* rgb, int W, int H, int gx, const double* bg, float* fb, float* T)
{
    const dim3 grid(unsigned(ceil_div64(n_tiles, kBlendWaves))), block(64 * kBlendWaves);
    const float b0 = float(bg[0]), b1 = float(bg[1]), b2 = float(bg[2]);
    if (batch == 64)
        k_blend_forward<TW, TH, 64><<<grid, block, 0, st>>>(ranges, pl, xy, co, rgb, W, H, gx, n_tiles, b0, b1, b2, fb, T);
    else if (batch == 128)
        k_blend_forward<TW, TH, 128><<<grid, block, 0, st>>>(ranges, pl, xy, co, rgb, W, H, gx, n_tiles, b0, b1, b2, fb, T);
    else
        k_blend_forward<TW, TH, 256><<<grid, block, 0, st>>>(ranges, pl, xy, co, rgb, W, H, gx, n_tiles, b0, b1, b2, fb, T);
}

template <int TW, int TH>
static void launch_blend_bwd(int batch, int n_tiles, hipStream_t st, const uint32_t* ranges, const uint32_t* pl, const float2* xy,
    const float4* co, const float* rgb, int W, int H, int gx, const double* bg, const float* fb, const float* T, const float* gfb, float* sg,
    WavePriority prio)
{
    constexpr int SPLIT = (TW * TH / kBwdSplit >= 64) ? kBwdSplit : 1;
    const dim3 grid(unsigned(ceil_div64(int64_t(n_tiles) * SPLIT, kBlendWaves))), block(64 * kBlendWaves);
    const float b0 = float(bg[0]), b1 = float(bg[1]), b2 = float(bg[2]);
    if (batch == 64)
        k_blend_backward<TW, TH, 64><<<grid, block, 0, st>>>(ranges, pl, xy, co, rgb, W, H, gx, n_tiles, b0, b1, b2, fb, T, gfb, sg, prio);
    else if (batch == 128)
        k_blend_backward<TW, TH, 128><<<grid, block, 0, st>>>(ranges, pl, xy, co, rgb, W, H, gx, n_tiles, b0, b1, b2, fb, T, gfb, sg, prio);
    else
        k_blend_backward<TW, TH, 256><<<grid, block, 0, st>>>(ranges, pl, xy, co, rgb, W, H, gx, n_tiles, b0, b1, b2, fb, T, gfb, sg, prio);
}

void splat_blend_forward(const dgmr_tiling& t, int n_tiles, hipStream_t st, const uint32_t* ranges, const uint32_t* pl, const float* xy,
    const float* co, const float* rgb, int W, int H, int gx, const double* bg, float* fb, float* T)
{
    const auto* xy2 = reinterpret_cast<const float2*>(xy);
    const auto* co4 = reinterpret_cast<const float4*>(co);
    if (t.tile_w == 16)
        launch_blend_fwd<16, 16>(t.batch, n_tiles, st, ranges, pl, xy2, co4, rgb, W, H, gx, bg, fb, T);
    else if (t.tile_w == 8)
        launch_blend_fwd<8, 8>(t.batch, n_tiles, st, ranges, pl, xy2, co4, rgb, W, H, gx, bg, fb, T);
    else
        launch_blend_fwd<32, 8>(t.batch, n_tiles, st, ranges, pl, xy2, co4, rgb, W, H, gx, bg, fb, T);
}

void splat_blend_backward(const dgmr_tiling& t, int n_tiles, int64_t n_pairs, hipStream_t st, const uint32_t* ranges, const uint32_t* pl,
    const float* xy, const float* co, const float* rgb, int W, int H, int gx, const double* bg, const float* fb, const float* T, const float* gfb,
    float* sg)
{
    const auto* xy2 = reinterpret_cast<const float2*>(xy);
    const auto* co4 = reinterpret_cast<const float4*>(co);
    // priority thresholds at 0.55, 0.93 and 1.37 times the mean list length
    // (1M Gaussians at 1080p: blend backward 0.210 -> 0.200 ms)
    const double mean = double(n_pairs) / double(std::max(n_tiles, 1));
    const auto at = [&](double f) { return uint32_t(std::min(f * mean, double(UINT32_MAX))); };
    const WavePriority prio { at(0.55), at(0.93), at(1.37) };
    if (t.tile_w == 16)
        launch_blend_bwd<16, 16>(t.batch, n_tiles, st, ranges, pl, xy2, co4, rgb, W, H, gx, bg, fb, T, gfb, sg, prio);
    else if (t.tile_w == 8)
        launch_blend_bwd<8, 8>(t.batch, n_tiles, st, ranges, pl, xy2, co4, rgb, W, H, gx, bg, fb, T, gfb, sg, prio);
    else
        launch_blend_bwd<32, 8>(t.batch, n_tiles, st, ranges, pl, xy2, co4, rgb, W, H, gx, bg, fb, T, gfb, sg, prio);
}

struct EmitArgs {
    int64_t n;
    unsigned blocks_n;
    uint32_t cap;
    int64_t n_tiles;
    int tile_bits;
    const uint32_t* sorted_ids;
    const int4* views;
    int64_t n_per_view;
    int n_views;
    int tile_w, tile_h;
    bool cull;
};

// emission, pair sort on the tile bits, tile ranges; pair keys of type KeyT
template <class KeyT>
static int emit_sort_ranges(const EmitArgs& a, const SplatWorkspace& w, dgmr_binning* bins, PendingCount* pending, hipStream_t st)
{
    // emit into the buffer pair that leaves the sorted ids in bins->point_list
    const int passes = radix_passes(0, a.tile_bits);
    KeyT* ka = reinterpret_cast<KeyT*>(w.pkeys_a);
    KeyT* kb = reinterpret_cast<KeyT*>(w.pkeys_b);
    SortBuffersT<KeyT> pb;
    if (passes % 2 == 0)
        pb = SortBuffersT<KeyT> { ka, bins->point_list, kb, w.pvals_a };
    else
        pb = SortBuffersT<KeyT> { ka, w.pvals_a, kb, bins->point_list };
    const int per = int64_t(a.cap) >= kEmitLarge ? kEmitPerLarge : kEmitPerSmall;
    const uint32_t block_pairs = uint32_t(kEmitThreads * per);
    const unsigned emit_blocks = unsigned(ceil_div64(a.cap, block_pairs));
    {
        StageScope sc("bin.emit", st);
        pending->add_stage(sc.slot, 4.0 + sizeof(KeyT), 28.0 * double(a.n));
        k_emit_bounds<<<a.blocks_n, 256, 0, st>>>(a.n, w.pair_offsets, w.counter, a.cap, block_pairs, w.emit_bounds);
        DGMR_HIP_CHECK(hipGetLastError());
        if (a.cull && per == kEmitPerLarge)
            k_emit_reach<KeyT, kEmitPerLarge><<<emit_blocks, kEmitThreads, 0, st>>>(w.counter, a.cap, a.sorted_ids, w.emit_bounds,
                w.pair_offsets, w.owner_reach, a.tile_w, a.tile_h, pb.keys_a, pb.vals_a);
        else if (a.cull)
            k_emit_reach<KeyT, kEmitPerSmall><<<emit_blocks, kEmitThreads, 0, st>>>(w.counter, a.cap, a.sorted_ids, w.emit_bounds,
                w.pair_offsets, w.owner_reach, a.tile_w, a.tile_h, pb.keys_a, pb.vals_a);
        else if (per == kEmitPerLarge)
            k_emit_pairs<KeyT, kEmitPerLarge><<<emit_blocks, kEmitThreads, 0, st>>>(a.n, w.counter, a.cap, a.sorted_ids, w.emit_bounds,
                w.pair_offsets, w.owner_rects, pb.keys_a, pb.vals_a);
        else
            k_emit_pairs<KeyT, kEmitPerSmall><<<emit_blocks, kEmitThreads, 0, st>>>(a.n, w.counter, a.cap, a.sorted_ids, w.emit_bounds,
                w.pair_offsets, w.owner_rects, pb.keys_a, pb.vals_a);
        DGMR_HIP_CHECK(hipGetLastError());
    }
    {
        StageScope sc("bin.pair_sort", st);
        pending->add_stage(sc.slot, (8.0 + 3.0 * sizeof(KeyT)) * passes, 0.0);
        DGMR_HIP_CHECK(radix_sort_pairs(pb, a.cap, w.counter, 0, a.tile_bits, w.sort_scratch, st));
    }
    const KeyT* sorted_keys = (passes % 2 == 0) ? pb.keys_a : pb.keys_b;
    const uint32_t* sorted_vals = (passes % 2 == 0) ? pb.vals_a : pb.vals_b;
    {
        StageScope sc("bin.ranges", st);
        pending->add_stage(sc.slot, 0.0, 8.0 * double(a.n_tiles));
        k_tile_ranges<KeyT><<<unsigned(ceil_div64(a.n_tiles, 256)), 256, 0, st>>>(sorted_keys, sorted_vals, w.counter, a.cap, a.views,
            a.n_views, a.n_per_view, a.n_tiles, bins->ranges);
        DGMR_HIP_CHECK(hipGetLastError());
    }
    return DGMR_OK;
}

// binning shared by the splatter and the marcher: depth sort, scan, emission,
// pair sort, ranges.  depth_keys/ids are the workspace's A buffers, filled by
// the preprocess.  Returns the pair count through *n_pairs (host).
int bin_gaussians(int64_t n, const int32_t* rects, const uint32_t* tiles, const void* xy, bool xy_double, bool bin_rects_written, bool cull,
    int grid_x, int grid_y, int tile_w, int tile_h, dgmr_binning* bins, Carver& cv, PendingCount* pending, hipStream_t st,
    const BatchBinning* batch)
{
    const int4* views = batch ? batch->views : nullptr;
    const int64_t n_per_view = batch ? batch->n_per_view : 0;
    SplatWorkspace w = carve_splat(cv, n, bins->pair_capacity, cull);
    // stable sort of Gaussian ids by depth bits.  Large sorts first keep only
    // the visible entries (culled ones carry the key 0xffffffff and would sort
    // last anyway), in order, so the sort and everything after it sees fewer
    // keys; n_sorted (device) is then their count.
    uint32_t *K = w.dkeys_a, *V = w.dvals_a, *K2 = w.dkeys_b, *V2 = w.dvals_b;
    const uint32_t* n_sorted = nullptr;
    if (n >= kCompactMin) {
        StageScope sc("bin.compact", st, 16.0 * double(n));
        uint32_t* count = w.counter + 1;
        // the preprocess left the ids unwritten: the compaction numbers the entries
        DGMR_HIP_CHECK(compact_pairs(K, nullptr, n, 0xffffffffu, K2, V2, count, w.counts, w.pair_offsets, w.scan_scratch, st));
        std::swap(K, K2);
        std::swap(V, V2);
        n_sorted = count;
    }
    const bool multi_view = batch && batch->n_views > 1;
    // with tight culling the culled entries sort last in the view pass, under the key n_views
    const uint32_t culled_view = cull && multi_view ? uint32_t(batch->n_views) : 0xffffffffu;
    {
        StageScope sc("bin.depth_sort", st, 16.0 * double(n) * radix_passes(0, 32));
        // several views: the last depth pass writes each entry's view as its key
        const KeyOut view_keys = multi_view ? KeyOut { uint32_t(n_per_view), culled_view } : KeyOut {};
        DGMR_HIP_CHECK(radix_sort_pairs(SortBuffers { K, V, K2, V2 }, n, n_sorted, 0, 32, w.sort_scratch, st, view_keys));
        if (radix_passes(0, 32) % 2 == 1) {
            std::swap(K, K2);
            std::swap(V, V2);
        }
    }
    if (multi_view) {
        // stable pass on the view index: order becomes (view, depth bits, id)
        StageScope sc("bin.view_sort", st, 16.0 * double(n));
        const int vbits = int(higher_msb(cull ? uint32_t(batch->n_views) : uint32_t(batch->n_views - 1)));
        DGMR_HIP_CHECK(radix_sort_pairs(SortBuffers { K, V, K2, V2 }, n, n_sorted, 0, vbits, w.sort_scratch, st));
        if (radix_passes(0, vbits) % 2 == 1) {
            std::swap(K, K2);
            std::swap(V, V2);
        }
    }
    const uint32_t* sorted_ids = V;
    const unsigned blocks_n = unsigned(ceil_div64(n, 256));
    {
        StageScope sc("bin.scan", st, 44.0 * double(n));
        // the pair offsets of the depth order and, when asked for, the
        // reference cache's inclusive scan of tiles in Gaussian order: one
        // launch, its state cleared by the gather (one thread per item)
        const ScanArray scans[2] = { { w.counts, w.pair_offsets, w.counter, false }, { tiles, bins->point_offsets, nullptr, true } };
        const int n_scans = bins->point_offsets ? 2 : 1;
        const ScanState scan = scan_state(w.scan_scratch, n, n_scans);
        if (cull)
            k_gather_reach<<<blocks_n, 256, 0, st>>>(n, n_sorted, K, batch && batch->n_views > 1 ? uint32_t(batch->n_views) : 0xffffffffu,
                sorted_ids, w.reach, tile_w, tile_h, w.counts, w.owner_reach, scan);
        else if (bin_rects_written)
            k_gather_owners<float, true><<<blocks_n, 256, 0, st>>>(n, n_sorted, sorted_ids, w.bin_rects, rects, tiles, static_cast<const float*>(xy),
                grid_x, grid_y, tile_w, tile_h, views, n_per_view, w.counts, w.owner_rects, scan);
        else if (xy_double)
            k_gather_owners<double, false><<<blocks_n, 256, 0, st>>>(n, n_sorted, sorted_ids, w.bin_rects, rects, tiles,
                static_cast<const double*>(xy), grid_x, grid_y, tile_w, tile_h, views, n_per_view, w.counts, w.owner_rects, scan);
        else
            k_gather_owners<float, false><<<blocks_n, 256, 0, st>>>(n, n_sorted, sorted_ids, w.bin_rects, rects, tiles,
                static_cast<const float*>(xy), grid_x, grid_y, tile_w, tile_h, views, n_per_view, w.counts, w.owner_rects, scan);
        DGMR_HIP_CHECK(hipGetLastError());
        DGMR_HIP_CHECK(scan_u32_arrays(scans, n_scans, n, w.scan_scratch, st, true));
    }
    // R stays on the device: the copy to the host is awaited by the caller
    const int rc_count = enqueue_count(w.counter, st, pending);
    if (rc_count != DGMR_OK)
        return rc_count;
    const int64_t n_tiles = batch ? batch->total_tiles : int64_t(grid_x) * grid_y;
    const int64_t cap = std::min<int64_t>(bins->pair_capacity, int64_t(UINT32_MAX));
    if (cap <= 0) {
        DGMR_HIP_CHECK(hipMemsetAsync(bins->ranges, 0, size_t(n_tiles) * 2 * sizeof(uint32_t), st));
        return DGMR_OK;
    }
    // view-local tile ids as keys: 16 bits when they fit (a quarter less
    // traffic in emission and sort), else 32
    const int tile_bits = int(higher_msb(uint32_t(batch ? batch->max_view_tiles : n_tiles)));
    EmitArgs ea { n, blocks_n, uint32_t(cap), n_tiles, tile_bits, sorted_ids, views, n_per_view, batch ? int(batch->n_views) : 0, tile_w, tile_h,
        cull };
    if (tile_bits <= 16)
        return emit_sort_ranges<uint16_t>(ea, w, bins, pending, st);
    return emit_sort_ranges<uint32_t>(ea, w, bins, pending, st);
}

// depth keys / ids buffers the preprocess writes and its per-Gaussian
// binning records (the first three slices of carve_splat)
void binning_key_buffers(void* workspace, int64_t n, bool cull, uint32_t** keys, uint32_t** ids, void** records)
{
    // these slices come first, so they do not depend on the pair capacity
    Carver cv(workspace, ~size_t(0));
    const SplatWorkspace w = carve_splat(cv, n, 1, cull);
    *keys = w.dkeys_a;
    // sorts that compact first take the entry index as the id
    *ids = n >= kCompactMin ? nullptr : w.dvals_a;
    if (records)
        *records = w.reach;
}

} // namespace dgmr

using namespace dgmr;

extern "C" {

int dgmr_workspace_size(int64_t n, int32_t width, int32_t height, const dgmr_tiling* tiling, int64_t pair_capacity, size_t* bytes)
{
    DGMR_ARG_CHECK(bytes && n >= 0 && valid_image(width, height) && valid_tiling(tiling) && pair_capacity >= 0, "dgmr_workspace_size: bad argument");
    *bytes = splat_workspace_bytes(n, pair_capacity, tiling->cull != 0);
    return DGMR_OK;
}

int dgmr_splat_forward(const dgmr_gaussians* g, const dgmr_camera* cam, const dgmr_tiling* tiling, dgmr_splat_geometry* geom,
    dgmr_binning* bins, float* framebuffer, float* final_transmittance, void* workspace, size_t workspace_bytes, dgmr_stats* stats,
    dgmr_stream stream)
{
    DGMR_ARG_CHECK(g && cam && geom && bins && framebuffer && final_transmittance && workspace, "dgmr_splat_forward: null argument");
    DGMR_ARG_CHECK(g->dtype == DGMR_F32, "dgmr_splat_forward: the splatter computes in float32");
    DGMR_ARG_CHECK(valid_tiling(tiling), "dgmr_splat_forward: tiling must be 8x8, 16x16 or 32x8 with batch 64/128/256");
    DGMR_ARG_CHECK(g->n >= 0 && g->n < (int64_t(1) << 31), "dgmr_splat_forward: bad Gaussian count");
    DGMR_ARG_CHECK(g->sh_coeffs >= 1 && g->sh_coeffs <= 16 && g->sh_degree >= 0 && g->sh_degree <= 3, "dgmr_splat_forward: bad SH layout");
    DGMR_ARG_CHECK(valid_image(cam->width, cam->height), "dgmr_splat_forward: bad image size");
    DGMR_ARG_CHECK(bins->point_list && bins->ranges, "dgmr_splat_forward: null binning buffers");
    const int64_t n = g->n;
    const bool cull = tiling->cull != 0;
    DGMR_ARG_CHECK(workspace_bytes >= splat_workspace_bytes(n, bins->pair_capacity, cull), "dgmr_splat_forward: workspace too small");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const DevCamera<float> dc = make_dev_camera<float>(*cam, *tiling);
    const int64_t n_tiles = int64_t(dc.grid_x) * dc.grid_y;
    const int64_t plane = int64_t(cam->width) * cam->height;
    if (stats) {
        stats->n_pairs = 0;
        stats->n_visible = 0;
    }
    if (n == 0) {
        DGMR_HIP_CHECK(hipMemsetAsync(bins->ranges, 0, size_t(n_tiles) * 2 * sizeof(uint32_t), st));
        splat_blend_forward(*tiling, int(n_tiles), st, bins->ranges, bins->point_list, geom->xy, geom->conic_opacity, geom->rgb, cam->width,
            cam->height, dc.grid_x, cam->background, framebuffer, final_transmittance);
        DGMR_HIP_CHECK(hipGetLastError());
        return DGMR_OK;
    }
    uint32_t *dkeys, *dids;
    void* recs;
    binning_key_buffers(workspace, n, cull, &dkeys, &dids, &recs);
    {
    StageScope sc("splat.preprocess", st, double(n) * (44.0 + 12.0 * g->sh_coeffs + 75.0));
    k_splat_preprocess<<<unsigned(ceil_div64(n, kPreThreads)), kPreThreads, 0, st>>>(n, make_params(g), dc,
        SplatOut { geom->rects, geom->tiles_touched, geom->xy, reinterpret_cast<float4*>(geom->conic_opacity), geom->rgb, geom->rgb_clamped,
            geom->depths, dkeys, dids, cull ? nullptr : static_cast<uint2*>(recs), cull ? static_cast<ReachRec*>(recs) : nullptr });
    DGMR_HIP_CHECK(hipGetLastError());
    }
    Carver cv(workspace, workspace_bytes);
    PendingCount pending;
    const int rc = bin_gaussians(n, geom->rects, geom->tiles_touched, geom->xy, false, true, cull, dc.grid_x, dc.grid_y, dc.tile_w, dc.tile_h, bins, cv,
        &pending, st, nullptr);
    if (rc != DGMR_OK)
        return rc;
    int blend_slot = -1;
    {
        StageScope sc("splat.blend_forward", st);
        blend_slot = sc.slot;
        splat_blend_forward(*tiling, int(n_tiles), st, bins->ranges, bins->point_list, geom->xy, geom->conic_opacity, geom->rgb, cam->width,
            cam->height, dc.grid_x, cam->background, framebuffer, final_transmittance);
        DGMR_HIP_CHECK(hipGetLastError());
    }
    pending.add_stage(blend_slot, 40.0, 16.0 * double(plane) + 8.0 * double(n_tiles));
    // every launch is queued: now wait for the pair count
    const int64_t R = pending.wait();
    if (R < 0)
        return DGMR_ERR_DEVICE;
    if (stats)
        stats->n_pairs = R;
    if (R > bins->pair_capacity) {
        set_error("pair capacity too small");
        return DGMR_ERR_CAPACITY;
    }
    return DGMR_OK;
}

int dgmr_splat_backward(const dgmr_gaussians* g, const dgmr_camera* cam, const dgmr_tiling* tiling, const dgmr_splat_geometry* geom,
    const dgmr_binning* bins, int64_t n_pairs, const float* framebuffer, const float* final_transmittance, const float* grad_framebuffer,
    float* screen_grads, float* grad_means, float* grad_scales, float* grad_rotations, float* grad_opacity, float* grad_sh, dgmr_stream stream)
{
    DGMR_ARG_CHECK(g && cam && geom && bins && framebuffer && final_transmittance && grad_framebuffer && screen_grads, "dgmr_splat_backward: null argument");
    DGMR_ARG_CHECK(grad_means && grad_scales && grad_rotations && grad_opacity && grad_sh, "dgmr_splat_backward: null gradient output");
    DGMR_ARG_CHECK(g->dtype == DGMR_F32, "dgmr_splat_backward: the splatter computes in float32");
    DGMR_ARG_CHECK(valid_tiling(tiling), "dgmr_splat_backward: bad tiling");
    DGMR_ARG_CHECK(g->sh_coeffs >= 1 && g->sh_coeffs <= 16 && g->sh_degree >= 0 && g->sh_degree <= 3, "dgmr_splat_backward: bad SH layout");
    const int64_t n = g->n;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (n == 0)
        return DGMR_OK;
    const DevCamera<float> dc = make_dev_camera<float>(*cam, *tiling);
    const int64_t n_tiles = int64_t(dc.grid_x) * dc.grid_y;
    DGMR_HIP_CHECK(hipMemsetAsync(screen_grads, 0, size_t(n) * 9 * sizeof(float), st));
    if (n_pairs > 0) {
        StageScope sc("splat.blend_backward", st, 40.0 * double(n_pairs) + 28.0 * double(int64_t(cam->width) * cam->height) + 8.0 * double(n_tiles));
        splat_blend_backward(*tiling, int(n_tiles), n_pairs, st, bins->ranges, bins->point_list, geom->xy, geom->conic_opacity, geom->rgb, cam->width,
            cam->height, dc.grid_x, cam->background, framebuffer, final_transmittance, grad_framebuffer, screen_grads);
        DGMR_HIP_CHECK(hipGetLastError());
    }
    StageScope sc("splat.preprocess_backward", st, double(n) * (2.0 * (44.0 + 12.0 * g->sh_coeffs) + 59.0));
    k_splat_preprocess_backward<<<unsigned(ceil_div64(n, kPreThreads)), kPreThreads, 0, st>>>(n, make_params(g), dc, geom->tiles_touched,
        reinterpret_cast<const float4*>(geom->conic_opacity), geom->rgb_clamped, screen_grads, grad_means, grad_scales, grad_rotations,
        grad_opacity, grad_sh);
    DGMR_HIP_CHECK(hipGetLastError());
    return DGMR_OK;
}

int dgmr_sort_workspace_size(int64_t n, size_t* bytes)
{
    DGMR_ARG_CHECK(bytes && n >= 0 && n <= int64_t(UINT32_MAX), "dgmr_sort_workspace_size: n out of range");
    *bytes = radix_scratch_bytes(std::max<int64_t>(n, 1));
    return DGMR_OK;
}

int dgmr_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp, int64_t n, int32_t begin_bit, int32_t end_bit,
    void* ws, size_t ws_bytes, dgmr_stream stream)
{
    DGMR_ARG_CHECK(n >= 0 && n <= int64_t(UINT32_MAX), "dgmr_sort_pairs: n out of range");
    DGMR_ARG_CHECK(begin_bit >= 0 && end_bit <= 32 && begin_bit <= end_bit, "dgmr_sort_pairs: bits must satisfy 0 <= begin <= end <= 32");
    if (n == 0 || begin_bit == end_bit)
        return DGMR_OK;
    DGMR_ARG_CHECK(keys && vals && keys_tmp && vals_tmp && ws, "dgmr_sort_pairs: null buffer");
    DGMR_ARG_CHECK(ws_bytes >= radix_scratch_bytes(n), "dgmr_sort_pairs: workspace too small");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const SortBuffers buf { keys, vals, keys_tmp, vals_tmp };
    DGMR_HIP_CHECK(radix_sort_pairs(buf, n, nullptr, begin_bit, end_bit, ws, st));
    if (radix_passes(begin_bit, end_bit) % 2 == 1) {
        DGMR_HIP_CHECK(hipMemcpyAsync(keys, keys_tmp, size_t(n) * 4, hipMemcpyDeviceToDevice, st));
        DGMR_HIP_CHECK(hipMemcpyAsync(vals, vals_tmp, size_t(n) * 4, hipMemcpyDeviceToDevice, st));
    }
    return DGMR_OK;
}

int dgmr_scan_workspace_size(int64_t n, size_t* bytes)
{
    DGMR_ARG_CHECK(bytes && n >= 0, "dgmr_scan_workspace_size: n out of range");
    *bytes = scan_scratch_bytes(n);
    return DGMR_OK;
}

int dgmr_scan_u32(const uint32_t* in, uint32_t* out, uint32_t* total, int64_t n, int32_t inclusive, void* ws, size_t ws_bytes,
    dgmr_stream stream)
{
    DGMR_ARG_CHECK(n >= 0, "dgmr_scan_u32: n out of range");
    DGMR_ARG_CHECK(n == 0 || (in && out && ws), "dgmr_scan_u32: null buffer");
    DGMR_ARG_CHECK(ws_bytes >= scan_scratch_bytes(n), "dgmr_scan_u32: workspace too small");
    DGMR_HIP_CHECK(scan_u32(in, out, total, n, inclusive != 0, ws, reinterpret_cast<hipStream_t>(stream)));
    return DGMR_OK;
}

int dgmr_selftest_wave_sum(const float* in, float* out, dgmr_stream stream)
{
    DGMR_ARG_CHECK(in && out, "dgmr_selftest_wave_sum: null argument");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    k_selftest_wave_sum<<<1, 64, 0, st>>>(in, out);
    DGMR_HIP_CHECK(hipGetLastError());
    return DGMR_OK;
}

} // extern "C"
