// SPDX-License-Identifier: GPL-3.0-or-later
// Batched multi-view splatting: many cameras of one scene with uneven
// resolutions in one preprocess launch, one binning pass and one persistent
// blend launch (forward and backward).
//
// Layout: the Gaussians take slots in the Morton order of their centres
// (slot s holds Gaussian order[s]; in id order when dgmr_tiling.slot_order is
// 1), and per-view geometry is stored at entry v * n + s, so the entry ids in the pair list index every array directly;
// view v's tiles occupy a contiguous range of the global tile index space.
// The blend runs as a persistent grid whose workgroups take the next tile of
// any view from a device counter, so small and large views share the chip
// without per-view launches or syncs.
#include <algorithm>
#include <numeric>
#include <type_traits>
#include <vector>

#include "host_common.h"
#include "primitives.h"
#include "splat_device.h"

namespace dgmr {
namespace {

struct ViewBlend {
    int width, height, grid_x, tile_base;
    float bg[3];
    float* framebuffer;
    float* final_T;
    const float* grad_fb;
    const float* l1_target; // non-null: dL/dimage = sign(image - target) * l1_scale instead of grad_fb
    float l1_scale;
};

// ------------------------------------------------------------ spatial order
// The preprocess kernels below give each thread one Gaussian and loop over the
// views, and a wave runs a view's projection (or its adjoint) when any of its
// 64 Gaussians is visible there.  In id order a wave's Gaussians are spread
// over the scene and nearly every wave runs every view with about 0.6 of its
// lanes busy (64 cameras around a 1M-Gaussian scene).  Slots therefore follow
// the Morton order of the centres, 8 bits per axis over the scene's bounding
// box: a wave's Gaussians are neighbours and are culled together.  Any
// permutation is correct; this one only decides which thread does what.
constexpr int kOrderBits = 24;
// the radix sort ping-pongs between two buffers: an odd pass count leaves the
// sorted ids in the second one, which is the order array
static_assert(((kOrderBits + 7) / 8) % 2 == 1, "the sorted ids must land in the second sort buffer");

// float bits whose unsigned order is the float order (NaN above +inf)
__device__ __forceinline__ uint32_t ordered_bits(float x)
{
    const uint32_t b = __float_as_uint(x);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float from_ordered_bits(uint32_t u)
{
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

// smallest and largest centre per axis (as ordered bits) of each block's
// share: partial[6 b + a] and partial[6 b + 3 + a]
constexpr int kBoundsBlocks = 256;
__device__ __forceinline__ void bounds_reduce(uint32_t (&lo)[3], uint32_t (&hi)[3])
{
    __shared__ uint32_t s_box[4][6];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            lo[a] = min(lo[a], uint32_t(__shfl_xor(int(lo[a]), off)));
            hi[a] = max(hi[a], uint32_t(__shfl_xor(int(hi[a]), off)));
        }
    }
    const int wave = int(threadIdx.x) / 64;
    if (lane_id() == 0) {
        for (int a = 0; a < 3; ++a) {
            s_box[wave][a] = lo[a];
            s_box[wave][3 + a] = hi[a];
        }
    }
    __syncthreads();
    // thread 0 ends with the block's bounds
    if (threadIdx.x == 0) {
        for (int w = 1; w < int(blockDim.x) / 64; ++w) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                lo[a] = min(lo[a], s_box[w][a]);
                hi[a] = max(hi[a], s_box[w][3 + a]);
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_bounds(int64_t n, const float* __restrict__ means, uint32_t* __restrict__ partial)
{
    uint32_t lo[3] = { 0xffffffffu, 0xffffffffu, 0xffffffffu }, hi[3] = { 0u, 0u, 0u };
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const uint32_t u = ordered_bits(means[3 * i + a]);
            lo[a] = min(lo[a], u);
            hi[a] = max(hi[a], u);
        }
    }
    bounds_reduce(lo, hi);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            partial[6 * blockIdx.x + a] = lo[a];
            partial[6 * blockIdx.x + 3 + a] = hi[a];
        }
    }
}

// box[0..2], box[3..5]: the blocks' partial bounds combined (one block)
__global__ __launch_bounds__(256) void k_bounds_final(int blocks, const uint32_t* __restrict__ partial, uint32_t* __restrict__ box)
{
    uint32_t lo[3] = { 0xffffffffu, 0xffffffffu, 0xffffffffu }, hi[3] = { 0u, 0u, 0u };
    for (int b = int(threadIdx.x); b < blocks; b += blockDim.x) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            lo[a] = min(lo[a], partial[6 * b + a]);
            hi[a] = max(hi[a], partial[6 * b + 3 + a]);
        }
    }
    bounds_reduce(lo, hi);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            box[a] = lo[a];
            box[3 + a] = hi[a];
        }
    }
}

// bit k of an 8-bit value to bit 3k
__device__ __forceinline__ uint32_t spread_bits8(uint32_t x)
{
    x = (x | (x << 8)) & 0x0000f00fu;
    x = (x | (x << 4)) & 0x000c30c3u;
    x = (x | (x << 2)) & 0x00249249u;
    return x;
}

// sort keys (Morton code of the centre in the bounding box) and ids 0..n-1;
// non-finite centres or bounds clamp to cell 0 (the order stays a permutation)
__global__ __launch_bounds__(256) void k_morton_keys(int64_t n, const float* __restrict__ means, const uint32_t* __restrict__ box,
    uint32_t* __restrict__ keys, uint32_t* __restrict__ ids)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    uint32_t key = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float lo = from_ordered_bits(box[a]), hi = from_ordered_bits(box[3 + a]);
        const float scale = 256.f / fmaxf(hi - lo, 1e-30f);
        const float c = fminf(fmaxf((means[3 * i + a] - lo) * scale, 0.f), 255.f);
        key |= spread_bits8(uint32_t(c)) << a;
    }
    keys[i] = key;
    ids[i] = uint32_t(i);
}

// The per-Gaussian parameters (centre, scales, rotation, opacity) copied
// into slot order (row s = Gaussian order[s]), so the preprocess kernels read
// them as contiguous rows; the backward reads the forward's copy.  The SH
// rows, 192 bytes each at degree 3, are read in place where they are staged.
__global__ __launch_bounds__(256) void k_gather_params(int64_t n, SplatParams p, const uint32_t* __restrict__ order, float* __restrict__ means,
    float* __restrict__ scales, float* __restrict__ rotations, float* __restrict__ opacity)
{
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= n)
        return;
    const int64_t i = order[s];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        means[3 * s + a] = p.means[3 * i + a];
        scales[3 * s + a] = p.scales[3 * i + a];
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
        rotations[4 * s + a] = p.rotations[4 * i + a];
    opacity[s] = p.opacity[i];
}

// slot s holds Gaussian s (slot_order 1)
__global__ __launch_bounds__(256) void k_iota(int64_t n, uint32_t* __restrict__ ids)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        ids[i] = uint32_t(i);
}

// thread per slot s (Gaussian order[s]; p in slot order but for its SH rows,
// read in place): the Gaussian's parameters are read once and projected into
// every view
__global__ __launch_bounds__(256) void k_batch_preprocess(int64_t n, int n_views, SplatParams p, const DevCamera<float>* __restrict__ cams,
    const uint32_t* __restrict__ order, SplatOut out)
{
    __shared__ float s_sh[256 * kShStride];
    __shared__ uint32_t s_ids[256];
    const int64_t first = int64_t(blockIdx.x) * blockDim.x;
    const int count = int(min<int64_t>(blockDim.x, n - first));
    if (int(threadIdx.x) < count)
        s_ids[threadIdx.x] = order[first + threadIdx.x];
    __syncthreads();
    stage_sh(s_sh, p.sh, s_ids, count, p.sh_coeffs);
    __syncthreads();
    const int64_t s = first + threadIdx.x;
    if (s >= n)
        return;
    // the view-independent inputs once, then each view
    GaussIn g;
    gauss_mean(p, s, g);
    gauss_rest(p, s, g);
    for (int v = 0; v < n_views; ++v)
        preprocess_view<true, false>(g, s, int64_t(v) * n + s, p, s_sh + threadIdx.x * kShStride, cams[v], out);
}

__device__ __forceinline__ int find_view(const ViewBlend* views, int n_views, int tile)
{
    int lo = 0, hi = n_views - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (views[mid].tile_base <= tile)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Persistent blend over the global tile queue: kBlendWaves independent waves
// per workgroup, each popping tiles for itself.  The queue counter starts at
// 0; a wave exits once the counter has passed the last tile.  Queue position
// k is tile k, or tile_order[k] when given.
template <int TW, int TH, int B, bool BACKWARD>
__global__ __launch_bounds__(64 * kBlendWaves) void k_batch_blend(const uint32_t* __restrict__ ranges, const uint32_t* __restrict__ point_list,
    const float2* __restrict__ xy, const float4* __restrict__ conic_opacity, const float* __restrict__ rgb, const ViewBlend* __restrict__ views,
    int n_views, int total_tiles, uint32_t* __restrict__ queue, float* __restrict__ screen_grads, const uint32_t* __restrict__ tile_order)
{
    using Smem = typename std::conditional<BACKWARD, BlendSmemBwd<B>, BlendSmem<B>>::type;
    __shared__ Smem sm_all[kBlendWaves];
    Smem& sm = sm_all[threadIdx.x / 64];
    while (true) {
        uint32_t t = 0;
        if (lane_id() == 0)
            t = atomicAdd(queue, 1u);
        // lane 0's value, read into a scalar register: the view record and
        // its pointers then stay in scalar registers too
        int tile = int(__builtin_amdgcn_readfirstlane(t));
        if (tile >= total_tiles)
            break;
        if (tile_order)
            tile = int(tile_order[tile]);
        const ViewBlend& vb = views[find_view(views, n_views, tile)];
        const int local = tile - vb.tile_base;
        const int px0 = (local % vb.grid_x) * TW, py0 = (local / vb.grid_x) * TH;
        if constexpr (BACKWARD) {
            // a view with neither a gradient image nor an L1 target has a zero
            // dL/dimage (or no gradient group): its tiles add nothing, so they
            // are skipped (the view record is the same for the whole wave)
            if (!vb.grad_fb && !vb.l1_target)
                continue;
            blend_tile_backward<TW, TH, B>(sm, uniform(ranges[2 * tile]), uniform(ranges[2 * tile + 1]), px0, py0, vb.width, vb.height, vb.bg[0],
                vb.bg[1], vb.bg[2], point_list, xy, conic_opacity, rgb, vb.framebuffer, vb.final_T, vb.grad_fb, screen_grads, vb.l1_target,
                vb.l1_scale);
        } else {
            blend_tile_forward<TW, TH, B>(sm, uniform(ranges[2 * tile]), uniform(ranges[2 * tile + 1]), px0, py0, vb.width, vb.height, vb.bg[0], vb.bg[1],
                vb.bg[2], point_list, xy, conic_opacity, rgb, vb.framebuffer, vb.final_T);
        }
        wave_lds_sync();
    }
}

// Tile order of the blend backward.  It adds each (tile, Gaussian) sum into
// the entry's screen gradients with float atomics, and tiles in flight at the
// same time that touch the same cache lines (a shared Gaussian, or Gaussians
// in neighbouring slots, which the spatial order makes neighbours on screen)
// serialise on them.  The queue therefore visits the views round-robin, the
// largest first, and steps through each view's tiles with a stride near
// 0.618 of its tile count (coprime with it), so tiles in flight together
// belong to different views or lie far apart.  by_size[j] = (first tile,
// tiles, stride) of the j-th largest view; segment q = (first queue
// position, first round, views) of the rounds in which the first `views`
// views still have tiles.
__global__ __launch_bounds__(256) void k_tile_order(int total, const int4* __restrict__ by_size, const int4* __restrict__ segs, int n_segs,
    uint32_t* __restrict__ order)
{
    const int k = int(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= total)
        return;
    int lo = 0, hi = n_segs - 1; // the last segment starting at or before k
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (segs[mid].x <= k)
            lo = mid;
        else
            hi = mid - 1;
    }
    const int4 q = segs[lo];
    const int d = k - q.x;
    const int round = q.y + d / q.z;
    const int4 v = by_size[d % q.z];
    order[k] = uint32_t(v.x + int((int64_t(round) * v.z) % v.y));
}

// clears the screen gradients (9 floats per entry) as one streaming pass of
// 16-byte stores over the whole slice: the blend backward adds into the
// visible entries' values and only those are read.  Clearing only the
// visible entries left culled holes in the written lines and took twice as
// long (0.96 ms over 64M entries).
__global__ __launch_bounds__(256) void k_zero_screen_grads(int64_t n_floats, float* __restrict__ screen_grads)
{
    const int64_t n4 = n_floats / 4;
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    float4* d = reinterpret_cast<float4*>(screen_grads);
    for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n4; k += stride)
        d[k] = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int64_t k = 4 * n4 + int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n_floats; k += stride)
        screen_grads[k] = 0.f;
}

// a block's per-Gaussian gradient rows staged in LDS: centre 3, scales 3,
// rotation 4, opacity 1; columns [col, col + W) of row r go to
// dst[ids[r] * W + k], consecutive lanes on consecutive floats
constexpr int kSmallRow = 11;
template <int W>
__device__ __forceinline__ void unstage_rows(float* __restrict__ dst, const float* s_rows, int col, const uint32_t* ids, int count)
{
    for (int e = threadIdx.x; e < count * W; e += blockDim.x) {
        const int r = e / W, k = e - r * W;
        dst[int64_t(ids[r]) * W + k] = s_rows[r * kSmallRow + col + k];
    }
}

// thread per slot s (Gaussian order[s]; p in slot order but for its SH rows,
// read in place), summing its
// adjoint over the views of each gradient group: group[v] is -1 (no
// gradient) or a group id whose views are consecutive; group g's sums go to
// block g of the outputs ([n_groups][n][..], in Gaussian order).  n_views here is the number of
// views up to the last one with a group: the views after it are neither
// loaded nor visited.
__global__ __launch_bounds__(256) void k_batch_preprocess_backward(int64_t n, int n_views, SplatParams p, const DevCamera<float>* __restrict__ cams,
    const uint32_t* __restrict__ order, bool id_order, const int32_t* __restrict__ group, const uint32_t* __restrict__ tiles,
    const float4* __restrict__ conic_opacity, const uint8_t* __restrict__ clamped, const float* __restrict__ screen_grads,
    float* __restrict__ g_means, float* __restrict__ g_scales, float* __restrict__ g_rotations, float* __restrict__ g_opacity,
    float* __restrict__ g_sh)
{
    // the block's SH coefficients, staged once; the SH gradient rows of a
    // group leave through s_out, half a block at a time (75 KB of LDS keeps
    // two blocks per CU, which the registers allow anyway)
    __shared__ float s_sh[256 * kShStride];
    __shared__ float s_out[128 * kShStride];
    __shared__ uint32_t s_ids[256];
    const int64_t first = int64_t(blockIdx.x) * blockDim.x;
    const int count = int(min<int64_t>(blockDim.x, n - first));
    if (int(threadIdx.x) < count)
        s_ids[threadIdx.x] = order[first + threadIdx.x];
    __syncthreads();
    stage_sh(s_sh, p.sh, s_ids, count, p.sh_coeffs);
    __syncthreads();
    // the view loop and its group boundaries are uniform over the block, so
    // every thread takes part in the block-wide copies at a boundary
    const int64_t i = first + threadIdx.x; // the slot: row of p and of the per-view arrays
    const bool valid = i < n;
    const int64_t gid = valid ? int64_t(s_ids[threadIdx.x]) : 0; // the Gaussian: row of the outputs
    const float* row = s_sh + threadIdx.x * kShStride;
    const int kc = 3 * p.sh_coeffs;
    // per group: the mean, opacity and SH gradients, and the view sums of
    // dL/dcov3 and dL/d(2 pi sa sb) that become scale and rotation gradients
    float g_mu[3], g_cov3[6], g_num, g_w, g_shv[48];
    // the Gaussian's own values are the same for every view: read once, and
    // its 3D covariance and largest scales computed once
    float mu[3] = { 0.f, 0.f, 0.f }, s[3] = { 1.f, 1.f, 1.f }, q[4] = { 1.f, 0.f, 0.f, 0.f }, w = 0.f;
    if (valid) {
#pragma unroll
        for (int a = 0; a < 3; ++a)
            mu[a] = p.means[3 * i + a];
        gaussian_params(p, i, s, q, &w);
    }
    const OtsPre<float> pre = ots_pre(s, q);
    // per-(view, Gaussian) inputs of view v + 1 are loaded while view v is
    // evaluated; the loads stay in bounds for every view and are unconditional
    struct ViewIn {
        uint32_t tiles;
        float4 A;
        float sg[9];
        uint8_t cl[3];
    };
    const auto load_view = [&](int v, ViewIn& x) {
        const int64_t o = int64_t(v) * n + i;
        x.tiles = tiles[o];
        x.A = conic_opacity[o];
#pragma unroll
        for (int k = 0; k < 9; ++k)
            x.sg[k] = screen_grads[o * 9 + k];
#pragma unroll
        for (int k = 0; k < 3; ++k)
            x.cl[k] = clamped[3 * o + k];
    };
    ViewIn nx {};
    if (valid && n_views > 0)
        load_view(0, nx);
    int cur = -1;
    for (int v = 0; v <= n_views; ++v) {
        const int gv = v < n_views ? group[v] : -1;
        ViewIn in = nx;
        if (valid && v + 1 < n_views)
            load_view(v + 1, nx);
        if (gv != cur) {
            if (cur >= 0) {
                // views of group cur done: write its block.  In Morton slot
                // order the slots' rows are scattered over the outputs, so they
                // leave through LDS and consecutive lanes write consecutive
                // floats of each row: a store instruction then touches a third
                // as many cache lines as one row per lane.  In id order each
                // thread writes its own rows, which the wave's lanes coalesce.
                float g_s[3] = { 0.f, 0.f, 0.f }, g_q[4] = { 0.f, 0.f, 0.f, 0.f };
                if (valid) {
                    preprocess_backward_finish(s, q, pre, p, g_cov3, g_num, g_s, g_q);
                    activation_backward(p, i, g_s, g_q, &g_w);
                }
                if (id_order) {
                    if (valid) {
                        const int64_t o = int64_t(cur) * n + i;
                        for (int a = 0; a < 3; ++a) {
                            g_means[3 * o + a] = g_mu[a];
                            g_scales[3 * o + a] = g_s[a];
                        }
                        for (int a = 0; a < 4; ++a)
                            g_rotations[4 * o + a] = g_q[a];
                        g_opacity[o] = g_w;
                    }
                } else {
                    __syncthreads(); // s_out free
                    if (valid) {
                        float* r = s_out + threadIdx.x * kSmallRow;
                        for (int a = 0; a < 3; ++a) {
                            r[a] = g_mu[a];
                            r[3 + a] = g_s[a];
                        }
                        for (int a = 0; a < 4; ++a)
                            r[6 + a] = g_q[a];
                        r[10] = g_w;
                    }
                    __syncthreads();
                    unstage_rows<3>(g_means + int64_t(cur) * n * 3, s_out, 0, s_ids, count);
                    unstage_rows<3>(g_scales + int64_t(cur) * n * 3, s_out, 3, s_ids, count);
                    unstage_rows<4>(g_rotations + int64_t(cur) * n * 4, s_out, 6, s_ids, count);
                    unstage_rows<1>(g_opacity + int64_t(cur) * n, s_out, 10, s_ids, count);
                }
                float* dst = g_sh + int64_t(cur) * n * kc;
                for (int half = 0; half < 2; ++half) {
                    const int lo = half * 128;
                    __syncthreads(); // s_out free
                    if (valid && int(threadIdx.x) - lo >= 0 && int(threadIdx.x) - lo < 128) {
                        float* out_row = s_out + (int(threadIdx.x) - lo) * kShStride;
#pragma unroll
                        for (int k = 0; k < 48; ++k)
                            if (k < kc)
                                out_row[k] = g_shv[k];
                    }
                    __syncthreads();
                    if (count > lo)
                        unstage_sh(dst, s_out, s_ids + lo, min(128, count - lo), p.sh_coeffs);
                }
            }
            cur = gv;
            for (int a = 0; a < 3; ++a)
                g_mu[a] = 0.f;
            for (int a = 0; a < 6; ++a)
                g_cov3[a] = 0.f;
            g_num = 0.f;
            g_w = 0.f;
#pragma unroll
            for (int k = 0; k < 48; ++k)
                g_shv[k] = 0.f;
        }
        if (gv < 0 || !valid)
            continue;
        if (in.tiles != 0)
            preprocess_backward_core(mu, pre, w, p, row, cams[v], in.A, in.cl, in.sg, g_mu, g_cov3, &g_num, &g_w, g_shv);
    }
}

// Workspace layout shared by forward and backward (the forward's geometry and
// binning must survive until the backward).
struct BatchLayout {
    DevCamera<float>* cams;
    int32_t* groups;
    ViewBlend* views;
    int4* grids;
    uint32_t* queue;
    uint32_t* tiles;
    float* xy;
    float4* conic_opacity;
    float* rgb;
    uint8_t* clamped;
    uint32_t* point_list;
    uint32_t* ranges;
    float* screen_grads;
    // the spatial order (slot -> Gaussian), its sort's buffers, and the
    // parameters in slot order
    uint32_t* order;
    uint32_t* order_keys;
    uint32_t* order_ids;
    uint32_t* order_keys2;
    uint32_t* order_box; // 6 bounds, then kBoundsBlocks x 6 partial bounds
    void* order_scratch;
    float* slot_means;
    float* slot_scales;
    float* slot_rotations;
    float* slot_opacity;
    // the blend backward's tile order and the tables it is built from
    uint32_t* tile_order;
    int4* tile_sizes;
    int4* tile_segs;
    void* binning;
    size_t binning_bytes;
    int64_t total_tiles;
};

BatchLayout carve_batch(void* ws, int64_t n, const dgmr_batch_view* views, int n_views, const dgmr_tiling& t, int64_t cap, size_t* used)
{
    Carver cv(ws, ~size_t(0));
    BatchLayout L;
    int64_t total_tiles = 0;
    for (int v = 0; v < n_views; ++v)
        total_tiles += int64_t((views[v].cam.width + t.tile_w - 1) / t.tile_w) * ((views[v].cam.height + t.tile_h - 1) / t.tile_h);
    const int64_t m = std::max<int64_t>(1, n * n_views);
    L.total_tiles = total_tiles;
    L.cams = cv.take<DevCamera<float>>(n_views);
    L.groups = cv.take<int32_t>(n_views);
    L.views = cv.take<ViewBlend>(n_views);
    L.grids = cv.take<int4>(n_views);
    L.queue = cv.take<uint32_t>(64);
    L.tiles = cv.take<uint32_t>(m);
    L.xy = cv.take<float>(2 * m);
    L.conic_opacity = cv.take<float4>(m);
    L.rgb = cv.take<float>(3 * m);
    L.clamped = cv.take<uint8_t>(3 * m);
    L.point_list = cv.take<uint32_t>(std::max<int64_t>(cap, 1));
    L.ranges = cv.take<uint32_t>(2 * std::max<int64_t>(total_tiles, 1));
    L.screen_grads = cv.take<float>(9 * m);
    const int64_t n1 = std::max<int64_t>(n, 1);
    L.order = cv.take<uint32_t>(n1);
    L.order_keys = cv.take<uint32_t>(n1);
    L.order_ids = cv.take<uint32_t>(n1);
    L.order_keys2 = cv.take<uint32_t>(n1);
    L.order_box = cv.take<uint32_t>(64 + 6 * kBoundsBlocks);
    L.order_scratch = cv.take<char>(radix_scratch_bytes(n1));
    L.slot_means = cv.take<float>(3 * n1);
    L.slot_scales = cv.take<float>(3 * n1);
    L.slot_rotations = cv.take<float>(4 * n1);
    L.slot_opacity = cv.take<float>(n1);
    L.tile_order = cv.take<uint32_t>(std::max<int64_t>(total_tiles, 1));
    L.tile_sizes = cv.take<int4>(n_views);
    L.tile_segs = cv.take<int4>(n_views);
    L.binning_bytes = splat_workspace_bytes(m, cap, t.cull != 0);
    L.binning = cv.take<char>(L.binning_bytes);
    *used = cv.used;
    return L;
}

// the Gaussians' parameters in slot order, as the forward copied them, and
// their SH rows in place (in id order: the parameters themselves)
SplatParams slot_params(const BatchLayout& L, const dgmr_gaussians* g, const dgmr_tiling& t)
{
    if (t.slot_order == 1)
        return make_splat_params(g);
    return SplatParams { L.slot_means, L.slot_scales, L.slot_rotations, L.slot_opacity, static_cast<const float*>(g->sh), g->sh_coeffs,
        g->sh_degree, float(g->scale_multiplier), g->activation };
}

template <int TW, int TH>
void launch_batch_blend(bool backward, int batch, int grid, hipStream_t st, const BatchLayout& L, int n_views)
{
    const auto* xy2 = reinterpret_cast<const float2*>(L.xy);
    const int tt = int(L.total_tiles);
#define DGMR_BATCH_BLEND(B)                                                                                                            \
    if (backward)                                                                                                                      \
        k_batch_blend<TW, TH, B, true><<<grid, 64 * kBlendWaves, 0, st>>>(L.ranges, L.point_list, xy2, L.conic_opacity, L.rgb, L.views, n_views, tt,  \
            L.queue, L.screen_grads, L.tile_order);                                                                                    \
    else                                                                                                                               \
        k_batch_blend<TW, TH, B, false><<<grid, 64 * kBlendWaves, 0, st>>>(L.ranges, L.point_list, xy2, L.conic_opacity, L.rgb, L.views, n_views, tt, \
            L.queue, L.screen_grads, nullptr);
    if (batch == 64) {
        DGMR_BATCH_BLEND(64)
    } else if (batch == 128) {
        DGMR_BATCH_BLEND(128)
    } else {
        DGMR_BATCH_BLEND(256)
    }
#undef DGMR_BATCH_BLEND
}

void batch_blend(bool backward, const dgmr_tiling& t, hipStream_t st, const BatchLayout& L, int n_views)
{
    // persistent grid: 32 waves per CU, never more waves than there are tiles
    const int grid = int(std::min<int64_t>(ceil_div64(L.total_tiles, kBlendWaves), 256 * 32 / kBlendWaves));
    if (t.tile_w == 16)
        launch_batch_blend<16, 16>(backward, t.batch, grid, st, L, n_views);
    else if (t.tile_w == 8)
        launch_batch_blend<8, 8>(backward, t.batch, grid, st, L, n_views);
    else
        launch_batch_blend<32, 8>(backward, t.batch, grid, st, L, n_views);
}

int check_batch_args(const dgmr_gaussians* g, const dgmr_batch_view* views, int32_t n_views, const dgmr_tiling* tiling)
{
    DGMR_ARG_CHECK(g && views && n_views > 0 && n_views <= 4096, "dgmr batch: bad views");
    DGMR_ARG_CHECK(g->dtype == DGMR_F32, "dgmr batch: the splatter computes in float32");
    DGMR_ARG_CHECK(valid_tiling(tiling), "dgmr batch: bad tiling");
    DGMR_ARG_CHECK(g->n >= 0 && g->n * int64_t(n_views) < (int64_t(1) << 32), "dgmr batch: n * views must stay below 2^32");
    DGMR_ARG_CHECK(g->sh_coeffs >= 1 && g->sh_coeffs <= 16 && g->sh_degree >= 0 && g->sh_degree <= 3, "dgmr batch: bad SH layout");
    for (int v = 0; v < n_views; ++v)
        DGMR_ARG_CHECK(valid_image(views[v].cam.width, views[v].cam.height) && views[v].framebuffer && views[v].final_transmittance,
            "dgmr batch: bad view");
    return DGMR_OK;
}

} // namespace
} // namespace dgmr

using namespace dgmr;

extern "C" {

int dgmr_batch_workspace_size(int64_t n, const dgmr_batch_view* views, int32_t n_views, const dgmr_tiling* tiling, int64_t pair_capacity,
    size_t* bytes)
{
    DGMR_ARG_CHECK(views && n_views > 0 && valid_tiling(tiling) && bytes && pair_capacity >= 0 && n >= 0, "dgmr_batch_workspace_size: bad argument");
    carve_batch(nullptr, n, views, n_views, *tiling, pair_capacity, bytes);
    return DGMR_OK;
}

int dgmr_splat_batch_forward(const dgmr_gaussians* g, const dgmr_batch_view* views, int32_t n_views, const dgmr_tiling* tiling, void* workspace,
    size_t workspace_bytes, int64_t pair_capacity, dgmr_stats* stats, dgmr_stream stream)
{
    const int rc0 = check_batch_args(g, views, n_views, tiling);
    if (rc0 != DGMR_OK)
        return rc0;
    DGMR_ARG_CHECK(workspace, "dgmr_splat_batch_forward: null workspace");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int64_t n = g->n;
    size_t need = 0;
    BatchLayout L = carve_batch(workspace, n, views, n_views, *tiling, pair_capacity, &need);
    DGMR_ARG_CHECK(workspace_bytes >= need, "dgmr_splat_batch_forward: workspace too small");
    std::vector<DevCamera<float>> cams(n_views);
    std::vector<ViewBlend> vb(n_views);
    std::vector<int4> grids(n_views);
    int64_t base = 0;
    for (int v = 0; v < n_views; ++v) {
        cams[v] = make_dev_camera<float>(views[v].cam, *tiling);
        vb[v] = ViewBlend { views[v].cam.width, views[v].cam.height, cams[v].grid_x, int(base),
            { float(views[v].cam.background[0]), float(views[v].cam.background[1]), float(views[v].cam.background[2]) }, views[v].framebuffer,
            views[v].final_transmittance, views[v].grad_framebuffer, views[v].l1_target, views[v].l1_scale };
        grids[v] = make_int4(cams[v].grid_x, cams[v].grid_y, int(base), 0);
        base += int64_t(cams[v].grid_x) * cams[v].grid_y;
    }
    DGMR_HIP_CHECK(hipMemcpyAsync(L.cams, cams.data(), sizeof(DevCamera<float>) * n_views, hipMemcpyHostToDevice, st));
    DGMR_HIP_CHECK(hipMemcpyAsync(L.views, vb.data(), sizeof(ViewBlend) * n_views, hipMemcpyHostToDevice, st));
    DGMR_HIP_CHECK(hipMemcpyAsync(L.grids, grids.data(), sizeof(int4) * n_views, hipMemcpyHostToDevice, st));
    if (stats)
        stats->n_pairs = 0;
    const int64_t m = n * n_views;
    const bool cull = tiling->cull != 0;
    uint32_t *dkeys, *dids;
    void* recs;
    binning_key_buffers(L.binning, m, cull, &dkeys, &dids, &recs);
    if (n > 0 && tiling->slot_order == 1) {
        k_iota<<<unsigned(ceil_div64(n, 256)), 256, 0, st>>>(n, L.order); // id order
        DGMR_HIP_CHECK(hipGetLastError());
    } else if (n > 0) {
        // the spatial order of the slots: bounding box, Morton keys, 24-bit sort
        StageScope sc("batch.order", st, 52.0 * double(n));
        const float* means = static_cast<const float*>(g->means);
        const int bounds_blocks = int(std::min<int64_t>(ceil_div64(n, 256), kBoundsBlocks));
        k_bounds<<<unsigned(bounds_blocks), 256, 0, st>>>(n, means, L.order_box + 64);
        k_bounds_final<<<1, 256, 0, st>>>(bounds_blocks, L.order_box + 64, L.order_box);
        k_morton_keys<<<unsigned(ceil_div64(n, 256)), 256, 0, st>>>(n, means, L.order_box, L.order_keys, L.order_ids);
        DGMR_HIP_CHECK(hipGetLastError());
        DGMR_HIP_CHECK(radix_sort_pairs(SortBuffers { L.order_keys, L.order_ids, L.order_keys2, L.order }, n, nullptr, 0, kOrderBits,
            L.order_scratch, st));
        k_gather_params<<<unsigned(ceil_div64(n, 256)), 256, 0, st>>>(n, make_splat_params(g), L.order, L.slot_means, L.slot_scales,
            L.slot_rotations, L.slot_opacity);
        DGMR_HIP_CHECK(hipGetLastError());
    }
    if (n > 0) {
        StageScope sc("batch.preprocess", st, double(n) * (44.0 + 12.0 * g->sh_coeffs) + double(m) * 55.0);
        k_batch_preprocess<<<unsigned(ceil_div64(n, 256)), 256, 0, st>>>(n, n_views, slot_params(L, g, *tiling), L.cams, L.order,
            SplatOut { nullptr, L.tiles, L.xy, L.conic_opacity, L.rgb, L.clamped, nullptr, dkeys, dids,
                cull ? nullptr : static_cast<uint2*>(recs), cull ? static_cast<ReachRec*>(recs) : nullptr });
        DGMR_HIP_CHECK(hipGetLastError());
    }
    dgmr_binning bins { nullptr, L.point_list, L.ranges, pair_capacity };
    Carver cv(L.binning, L.binning_bytes);
    int64_t max_view_tiles = 1;
    for (int v = 0; v < n_views; ++v) {
        const DevCamera<float> c = make_dev_camera<float>(views[v].cam, *tiling);
        max_view_tiles = std::max<int64_t>(max_view_tiles, int64_t(c.grid_x) * c.grid_y);
    }
    const BatchBinning bb { L.grids, std::max<int64_t>(n, 1), n_views, L.total_tiles, max_view_tiles };
    PendingCount pending;
    if (m > 0) {
        const int rc = bin_gaussians(m, nullptr, L.tiles, L.xy, false, true, cull, 0, 0, tiling->tile_w, tiling->tile_h, &bins, cv, &pending, st, &bb);
        if (rc != DGMR_OK)
            return rc;
    } else {
        DGMR_HIP_CHECK(hipMemsetAsync(L.ranges, 0, size_t(L.total_tiles) * 2 * sizeof(uint32_t), st));
    }
    {
        StageScope sc("batch.blend_forward", st);
        pending.add_stage(sc.slot, 40.0, 0.0);
        DGMR_HIP_CHECK(hipMemsetAsync(L.queue, 0, sizeof(uint32_t), st));
        batch_blend(false, *tiling, st, L, n_views);
        DGMR_HIP_CHECK(hipGetLastError());
    }
    const int64_t R = pending.wait();
    if (R < 0)
        return DGMR_ERR_DEVICE;
    if (stats)
        stats->n_pairs = R;
    if (R > pair_capacity) {
        set_error("pair capacity too small");
        return DGMR_ERR_CAPACITY;
    }
    return DGMR_OK;
}

int dgmr_splat_batch_backward(const dgmr_gaussians* g, const dgmr_batch_view* views, int32_t n_views, const dgmr_tiling* tiling, void* workspace,
    size_t workspace_bytes, int64_t pair_capacity, int64_t n_pairs, const int32_t* view_group, int32_t n_groups, float* grad_means,
    float* grad_scales, float* grad_rotations, float* grad_opacity, float* grad_sh, dgmr_stream stream)
{
    const int rc0 = check_batch_args(g, views, n_views, tiling);
    if (rc0 != DGMR_OK)
        return rc0;
    DGMR_ARG_CHECK(workspace && grad_means && grad_scales && grad_rotations && grad_opacity && grad_sh, "dgmr_splat_batch_backward: null argument");
    // gradient groups: ids in [-1, n_groups), each group's views consecutive
    std::vector<int32_t> grp(n_views, 0);
    bool every_group_has_views = true;
    if (view_group) {
        DGMR_ARG_CHECK(n_groups >= 1, "dgmr_splat_batch_backward: n_groups must be >= 1 with view_group");
        std::vector<char> seen(n_groups, 0);
        for (int v = 0; v < n_views; ++v) {
            const int32_t gv = view_group[v];
            DGMR_ARG_CHECK(gv >= -1 && gv < n_groups, "dgmr_splat_batch_backward: view_group out of range");
            if (gv >= 0 && (v == 0 || view_group[v - 1] != gv)) {
                DGMR_ARG_CHECK(!seen[gv], "dgmr_splat_batch_backward: a group's views must be consecutive");
                seen[gv] = 1;
            }
            grp[v] = gv;
        }
        for (int gi = 0; gi < n_groups; ++gi)
            every_group_has_views = every_group_has_views && seen[gi];
    } else {
        n_groups = 1;
    }
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int64_t n = g->n;
    size_t need = 0;
    BatchLayout L = carve_batch(workspace, n, views, n_views, *tiling, pair_capacity, &need);
    DGMR_ARG_CHECK(workspace_bytes >= need, "dgmr_splat_batch_backward: workspace too small");
    if (n == 0)
        return DGMR_OK;
    // the grad pointers may differ from the forward call: refresh the view table
    std::vector<ViewBlend> vb(n_views);
    int64_t base = 0;
    for (int v = 0; v < n_views; ++v) {
        const DevCamera<float> c = make_dev_camera<float>(views[v].cam, *tiling);
        vb[v] = ViewBlend { views[v].cam.width, views[v].cam.height, c.grid_x, int(base),
            { float(views[v].cam.background[0]), float(views[v].cam.background[1]), float(views[v].cam.background[2]) }, views[v].framebuffer,
            views[v].final_transmittance, views[v].grad_framebuffer, views[v].l1_target, views[v].l1_scale };
        base += int64_t(c.grid_x) * c.grid_y;
    }
    DGMR_HIP_CHECK(hipMemcpyAsync(L.views, vb.data(), sizeof(ViewBlend) * n_views, hipMemcpyHostToDevice, st));
    DGMR_HIP_CHECK(hipMemcpyAsync(L.groups, grp.data(), sizeof(int32_t) * n_views, hipMemcpyHostToDevice, st));
    // views after the last one with a gradient group take no part in the
    // backward: their screen gradients are neither cleared nor read
    int n_grad_views = 0;
    for (int v = 0; v < n_views; ++v)
        if (grp[v] >= 0)
            n_grad_views = v + 1;
    const int64_t m = n * n_grad_views;
    {
        StageScope sc("batch.clear", st, 36.0 * double(m));
        const int64_t blocks = std::min<int64_t>(ceil_div64(ceil_div64(9 * m, 4), 256), 8192);
        k_zero_screen_grads<<<unsigned(std::max<int64_t>(blocks, 1)), 256, 0, st>>>(9 * m, L.screen_grads);
        DGMR_HIP_CHECK(hipGetLastError());
    }
    if (n_pairs > 0) {
        StageScope sc("batch.blend_backward", st, 40.0 * double(n_pairs));
        {
            // views by tile count, largest first, and the segments of rounds
            // with the same number of views (k_tile_order)
            std::vector<int4> sizes(n_views);
            for (int v = 0; v < n_views; ++v) {
                const int nt = (v + 1 < n_views ? vb[v + 1].tile_base : int(L.total_tiles)) - vb[v].tile_base;
                int64_t stride = std::max<int64_t>(1, int64_t(double(nt) * 0.6180339887498949 + 0.5));
                while (std::gcd(stride, int64_t(nt)) != 1)
                    ++stride;
                sizes[v] = make_int4(vb[v].tile_base, nt, int(stride), 0);
            }
            std::stable_sort(sizes.begin(), sizes.end(), [](const int4& a, const int4& b) { return a.y > b.y; });
            std::vector<int4> segs;
            int start = 0;
            for (int c = n_views; c >= 1; --c) {
                const int r0 = c < n_views ? sizes[c].y : 0, r1 = sizes[c - 1].y;
                if (r1 > r0) {
                    segs.push_back(make_int4(start, r0, c, 0));
                    start += c * (r1 - r0);
                }
            }
            DGMR_ARG_CHECK(start == L.total_tiles && !segs.empty(), "dgmr_splat_batch_backward: inconsistent view tiles");
            DGMR_HIP_CHECK(hipMemcpyAsync(L.tile_sizes, sizes.data(), sizeof(int4) * n_views, hipMemcpyHostToDevice, st));
            DGMR_HIP_CHECK(hipMemcpyAsync(L.tile_segs, segs.data(), sizeof(int4) * segs.size(), hipMemcpyHostToDevice, st));
            k_tile_order<<<unsigned(ceil_div64(L.total_tiles, 256)), 256, 0, st>>>(int(L.total_tiles), L.tile_sizes, L.tile_segs, int(segs.size()),
                L.tile_order);
            DGMR_HIP_CHECK(hipGetLastError());
        }
        DGMR_HIP_CHECK(hipMemsetAsync(L.queue, 0, sizeof(uint32_t), st));
        batch_blend(true, *tiling, st, L, n_views);
        DGMR_HIP_CHECK(hipGetLastError());
    }
    StageScope sc("batch.preprocess_backward", st, double(n) * 2.0 * (44.0 + 12.0 * g->sh_coeffs) + double(m) * 40.0);
    if (!every_group_has_views) {
        // a group with no view is written by no thread: clear the outputs first
        const size_t per = size_t(n) * n_groups;
        DGMR_HIP_CHECK(hipMemsetAsync(grad_means, 0, per * 3 * sizeof(float), st));
        DGMR_HIP_CHECK(hipMemsetAsync(grad_scales, 0, per * 3 * sizeof(float), st));
        DGMR_HIP_CHECK(hipMemsetAsync(grad_rotations, 0, per * 4 * sizeof(float), st));
        DGMR_HIP_CHECK(hipMemsetAsync(grad_opacity, 0, per * sizeof(float), st));
        DGMR_HIP_CHECK(hipMemsetAsync(grad_sh, 0, per * 3 * g->sh_coeffs * sizeof(float), st));
    }
    k_batch_preprocess_backward<<<unsigned(ceil_div64(n, 256)), 256, 0, st>>>(n, n_grad_views, slot_params(L, g, *tiling), L.cams, L.order, tiling->slot_order == 1, L.groups, L.tiles,
        L.conic_opacity, L.clamped, L.screen_grads, grad_means, grad_scales, grad_rotations, grad_opacity, grad_sh);
    DGMR_HIP_CHECK(hipGetLastError());
    return DGMR_OK;
}

} // extern "C"
