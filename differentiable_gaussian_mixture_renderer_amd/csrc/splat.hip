// SPDX-License-Identifier: GPL-3.0-or-later
// Tile splatter for gfx950: forward (reference: src/dgmr/splat_forward.cu:30-338)
// and the backward the reference does not ship.
//
// Pipeline (one stream, one 4-byte read-back for the pair count):
//   1. k_splat_preprocess   per Gaussian: EWA splat (Ots, 0.3 px^2 filter), rect,
//                           tiles touched, depth key, SH colour, conic.
//   2. radix sort of the Gaussians by depth bits (stable, ties by index)
//   3. gather tiles touched in depth order + exclusive scan -> pair offsets
//   4. k_emit_pairs         wave-balanced emission of (tile id, Gaussian id)
//   5. radix sort of the pairs by tile id only (stable)
//      Steps 2-5 yield exactly the order of the reference's single sort on
//      (tile << 32 | depth bits): within a tile, pairs stay in (depth bits,
//      Gaussian id) order, which is the reference's stable order.
//   6. k_tile_ranges        [first, last) per tile
//   7. k_blend_forward      per tile, LDS-staged batches of Gaussians
// Backward: k_blend_backward (front-to-back replay, transmittance recomputed,
// wave-reduced atomics) then k_splat_preprocess_backward (per Gaussian).
#include <algorithm>
#include <cstring>

#include "device_grad.h"
#include "device_math.h"
#include "host_common.h"
#include "primitives.h"

namespace dgmr {

namespace {

constexpr int kPreThreads = 256;

__global__ __launch_bounds__(kPreThreads) void k_splat_preprocess(int64_t n, const float* __restrict__ means, const float* __restrict__ scales,
    const float* __restrict__ rotations, const float* __restrict__ opacity, const float* __restrict__ sh, int sh_coeffs, int sh_degree,
    float scale_multiplier, DevCamera<float> cam, int32_t* __restrict__ rects, uint32_t* __restrict__ tiles, float* __restrict__ xy_out,
    float4* __restrict__ conic_opacity, float* __restrict__ rgb_out, uint8_t* __restrict__ clamped_out, float* __restrict__ depths,
    uint32_t* __restrict__ depth_keys, uint32_t* __restrict__ ids)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    tiles[i] = 0;
    rects[2 * i] = 0;
    rects[2 * i + 1] = 0;
    xy_out[2 * i] = 0.f;
    xy_out[2 * i + 1] = 0.f;
    depth_keys[i] = 0xffffffffu;
    ids[i] = uint32_t(i);
    const float mu[3] = { means[3 * i], means[3 * i + 1], means[3 * i + 2] };
    float t[3];
    transform_point(cam.view, mu, t, 3);
    if (t[2] < 0.2f)
        return;
    const float s[3] = { scales[3 * i] * scale_multiplier, scales[3 * i + 1] * scale_multiplier, scales[3 * i + 2] * scale_multiplier };
    const float q[4] = { rotations[4 * i], rotations[4 * i + 1], rotations[4 * i + 2], rotations[4 * i + 3] };
    const float w = opacity[i];
    SplatIn<float> in;
    splat_internals(mu, s, q, cam, &in);
    const float cov_f[3] = { in.cov2[0] + 0.3f, in.cov2[1], in.cov2[2] + 0.3f };
    float xy[2];
    project_to_screen(mu, cam, xy);
    const int ext_x = guarded_int(ceilf(3.f * __fsqrt_rn(cov_f[0])));
    const int ext_y = guarded_int(ceilf(3.f * __fsqrt_rn(cov_f[2])));
    rects[2 * i] = ext_x;
    rects[2 * i + 1] = ext_y;
    int x0, y0, x1, y1;
    get_rect(xy[0], xy[1], ext_x, ext_y, cam.grid_x, cam.grid_y, cam.tile_w, cam.tile_h, &x0, &y0, &x1, &y1);
    const uint32_t touched = uint32_t(x1 - x0) * uint32_t(y1 - y0);
    if (touched == 0)
        return;
    tiles[i] = touched;
    xy_out[2 * i] = xy[0];
    xy_out[2 * i + 1] = xy[1];
    const float dv[3] = { cam.pos[0] - mu[0], cam.pos[1] - mu[1], cam.pos[2] - mu[2] };
    const float depth = length3(dv);
    depths[i] = depth;
    depth_keys[i] = __float_as_uint(depth);
    // colour from SH along normalize(mu - cam) (math.h:233-271)
    const float v[3] = { -dv[0], -dv[1], -dv[2] };
    const float inv_len = __fdiv_rn(1.f, __fsqrt_rn(dot3(v, v)));
    const float dir[3] = { v[0] * inv_len, v[1] * inv_len, v[2] * inv_len };
    float Y[16];
    sh_basis(sh_degree, dir, Y);
    const int kmax = min(sh_coeffs, (sh_degree + 1) * (sh_degree + 1));
    const float* c = sh + int64_t(i) * sh_coeffs * 3;
    float r0 = 0.f, r1 = 0.f, r2 = 0.f;
    for (int k = 0; k < kmax; ++k) {
        r0 = fmaf(Y[k], c[3 * k + 0], r0);
        r1 = fmaf(Y[k], c[3 * k + 1], r1);
        r2 = fmaf(Y[k], c[3 * k + 2], r2);
    }
    r0 += 0.5f;
    r1 += 0.5f;
    r2 += 0.5f;
    rgb_out[3 * i + 0] = fmaxf(r0, 0.f);
    rgb_out[3 * i + 1] = fmaxf(r1, 0.f);
    rgb_out[3 * i + 2] = fmaxf(r2, 0.f);
    clamped_out[3 * i + 0] = r0 >= 0.f;
    clamped_out[3 * i + 1] = r1 >= 0.f;
    clamped_out[3 * i + 2] = r2 >= 0.f;
    // screen-space weight, Ots formulation (math.h:389-397)
    int i0, i1;
    larger2_idx(s, &i0, &i1);
    const float i2prime = float(2 * kPi) * s[i0] * s[i1];
    const float det2 = cov_f[0] * cov_f[2] - cov_f[1] * cov_f[1];
    const float i2 = float(2 * kPi) * __fsqrt_rn(det2);
    const float weight = w * det3(in.J) * i2prime / i2;
    conic_opacity[i] = make_float4(cov_f[2] / det2, -cov_f[1] / det2, cov_f[0] / det2, weight);
}

__global__ void k_gather_counts(int64_t n, const uint32_t* __restrict__ ids, const uint32_t* __restrict__ tiles, uint32_t* __restrict__ counts)
{
    const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j < n)
        counts[j] = tiles[ids[j]];
}

} // namespace

// Emission of (tile, Gaussian) pairs with per-wave load balancing: a wave
// owns 64 consecutive depth-sorted Gaussians and writes their pairs 64 at a
// time, each lane locating its Gaussian by binary search over the wave's
// offsets in LDS, so one large Gaussian no longer serialises a thread.
__global__ __launch_bounds__(256) void k_emit_pairs(int64_t n, const uint32_t* __restrict__ sorted_ids, const uint32_t* __restrict__ offsets,
    const int32_t* __restrict__ rects, const uint32_t* __restrict__ tiles, const float* __restrict__ xy, int grid_x, int grid_y, int tile_w,
    int tile_h, uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out)
{
    __shared__ uint32_t s_off[4][64], s_x0[4][64], s_y0[4][64], s_w[4][64], s_id[4][64];
    const unsigned lane = lane_id();
    const unsigned wave = threadIdx.x / 64;
    const int64_t j = (int64_t(blockIdx.x) * 4 + wave) * 64 + lane;
    uint32_t off = 0xffffffffu, cnt = 0, x0 = 0, y0 = 0, w = 1, id = 0;
    if (j < n) {
        id = sorted_ids[j];
        off = offsets[j];
        cnt = tiles[id];
        if (cnt) {
            int a, b, c, d;
            get_rect(xy[2 * id], xy[2 * id + 1], rects[2 * id], rects[2 * id + 1], grid_x, grid_y, tile_w, tile_h, &a, &b, &c, &d);
            x0 = uint32_t(a);
            y0 = uint32_t(b);
            w = uint32_t(c - a);
        }
    }
    s_off[wave][lane] = off;
    s_x0[wave][lane] = x0;
    s_y0[wave][lane] = y0;
    s_w[wave][lane] = w;
    s_id[wave][lane] = id;
    uint32_t end = (j < n) ? off + cnt : 0u;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1)
        end = max(end, __shfl_xor(end, d, 64));
    const uint32_t base = __shfl(off, 0, 64);
    __syncthreads();
    if (base == 0xffffffffu)
        return;
    const uint32_t total = end - base;
    for (uint32_t p = lane; p < total; p += 64) {
        const uint32_t target = base + p;
        // last lane whose offset is <= target
        unsigned lo = 0, hi = 63;
        while (lo < hi) {
            const unsigned mid = (lo + hi + 1) >> 1;
            if (s_off[wave][mid] <= target)
                lo = mid;
            else
                hi = mid - 1;
        }
        const uint32_t local = target - s_off[wave][lo];
        const uint32_t ww = s_w[wave][lo];
        uint32_t ty = uint32_t(__fdiv_rn(float(local), float(ww)));
        int32_t tx = int32_t(local) - int32_t(ty * ww);
        if (tx < 0) {
            --ty;
            tx += int32_t(ww);
        } else if (tx >= int32_t(ww)) {
            ++ty;
            tx -= int32_t(ww);
        }
        keys_out[target] = (s_y0[wave][lo] + ty) * uint32_t(grid_x) + s_x0[wave][lo] + uint32_t(tx);
        vals_out[target] = s_id[wave][lo];
    }
}

__global__ void k_tile_ranges(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ n_pairs, uint32_t* __restrict__ ranges)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t R = *n_pairs;
    if (k >= R)
        return;
    const uint32_t tile = keys[k];
    if (k == 0 || keys[k - 1] != tile)
        ranges[2 * tile] = k;
    if (k + 1 == R || keys[k + 1] != tile)
        ranges[2 * tile + 1] = k + 1;
}

// logical tile of a block: workgroups are dealt round-robin over the 8 XCDs,
// so give each XCD one contiguous band of tiles (shared Gaussians stay in
// that XCD's L2)
__device__ __forceinline__ int xcd_tile(int b, int n_tiles)
{
    const int q = n_tiles / 8, r = n_tiles % 8;
    const int x = b % 8, slot = b / 8;
    return x * q + min(x, r) + slot;
}

template <int TW, int TH, int B>
__global__ __launch_bounds__(TW* TH) void k_blend_forward(const uint32_t* __restrict__ ranges, const uint32_t* __restrict__ point_list,
    const float2* __restrict__ xy, const float4* __restrict__ conic_opacity, const float* __restrict__ rgb, int width, int height, int grid_x,
    int n_tiles, float bg0, float bg1, float bg2, float* __restrict__ framebuffer, float* __restrict__ final_T)
{
    constexpr int NT = TW * TH;
    __shared__ float4 s_geo[B];  // x, y, -a/2, -b
    __shared__ float4 s_col[B];  // -c/2, opacity, r, g
    __shared__ float s_blue[B];
    const int tile = xcd_tile(blockIdx.x, n_tiles);
    const int tx = tile % grid_x, ty = tile / grid_x;
    const int px = tx * TW + int(threadIdx.x) % TW;
    const int py = ty * TH + int(threadIdx.x) / TW;
    const bool inside = px < width && py < height;
    const float pxf = float(px), pyf = float(py);
    const uint32_t first = ranges[2 * tile], last = ranges[2 * tile + 1];
    float T = 1.f, C0 = 0.f, C1 = 0.f, C2 = 0.f;
    bool done = !inside;
    for (uint32_t base = first; base < last; base += B) {
        if (__syncthreads_count(!done) == 0)
            break;
        for (int i = threadIdx.x; i < B; i += NT) {
            const uint32_t k = base + uint32_t(i);
            if (k < last) {
                const uint32_t id = point_list[k];
                const float2 p = xy[id];
                const float4 co = conic_opacity[id];
                s_geo[i] = make_float4(p.x, p.y, -0.5f * co.x, -co.y);
                s_col[i] = make_float4(-0.5f * co.z, co.w, rgb[3 * id], rgb[3 * id + 1]);
                s_blue[i] = rgb[3 * id + 2];
            }
        }
        __syncthreads();
        if (!done) {
            const int cnt = int(min<uint32_t>(B, last - base));
            for (int j = 0; j < cnt; ++j) {
                const float4 g = s_geo[j];
                const float4 c = s_col[j];
                const float dx = pxf - g.x, dy = pyf - g.y;
                const float power = fmaf(g.z * dx, dx, fmaf(c.x * dy, dy, g.w * dx * dy));
                const float alpha = fminf(0.99f, c.y * __expf(power));
                if (alpha < 1.0f / 255.0f)
                    continue;
                const float test_T = T * (1.f - alpha);
                if (test_T < 0.0001f) {
                    done = true;
                    break;
                }
                const float wgt = alpha * T;
                C0 = fmaf(c.z, wgt, C0);
                C1 = fmaf(c.w, wgt, C1);
                C2 = fmaf(s_blue[j], wgt, C2);
                T = test_T;
            }
        }
    }
    if (!inside)
        return;
    const int64_t p = int64_t(py) * width + px;
    const int64_t plane = int64_t(width) * height;
    framebuffer[p] = fmaf(T, bg0, C0);
    framebuffer[plane + p] = fmaf(T, bg1, C1);
    framebuffer[2 * plane + p] = fmaf(T, bg2, C2);
    final_T[p] = T;
}

// Blend adjoint, front to back.  The remaining colour is carried unnormalised
// (C_final minus what has been composited so far), which is the reference
// prototype's remaining colour (grad_inria_splatter_render.cu:100-156) times
// the current transmittance, without the repeated division by (1 - alpha).
// Per-Gaussian gradients are summed over the 64 pixels of a wave with DPP
// before one lane issues the global atomics.
template <int TW, int TH, int B>
__global__ __launch_bounds__(TW* TH) void k_blend_backward(const uint32_t* __restrict__ ranges, const uint32_t* __restrict__ point_list,
    const float2* __restrict__ xy, const float4* __restrict__ conic_opacity, const float* __restrict__ rgb, int width, int height, int grid_x,
    int n_tiles, float bg0, float bg1, float bg2, const float* __restrict__ framebuffer, const float* __restrict__ final_T,
    const float* __restrict__ grad_fb, float* __restrict__ screen_grads)
{
    constexpr int NT = TW * TH;
    __shared__ float4 s_geo[B];  // x, y, a, b
    __shared__ float4 s_col[B];  // c, opacity, r, g
    __shared__ float s_blue[B];
    __shared__ uint32_t s_id[B];
    const int tile = xcd_tile(blockIdx.x, n_tiles);
    const int tx = tile % grid_x, ty = tile / grid_x;
    const int px = tx * TW + int(threadIdx.x) % TW;
    const int py = ty * TH + int(threadIdx.x) / TW;
    const bool inside = px < width && py < height;
    const float pxf = float(px), pyf = float(py);
    const uint32_t first = ranges[2 * tile], last = ranges[2 * tile + 1];
    float gC0 = 0.f, gC1 = 0.f, gC2 = 0.f, Tf = 1.f, S0 = 0.f, S1 = 0.f, S2 = 0.f;
    if (inside) {
        const int64_t p = int64_t(py) * width + px;
        const int64_t plane = int64_t(width) * height;
        gC0 = grad_fb[p];
        gC1 = grad_fb[plane + p];
        gC2 = grad_fb[2 * plane + p];
        Tf = final_T[p];
        S0 = framebuffer[p] - Tf * bg0;
        S1 = framebuffer[plane + p] - Tf * bg1;
        S2 = framebuffer[2 * plane + p] - Tf * bg2;
    }
    const float gT_Tf = (gC0 * bg0 + gC1 * bg1 + gC2 * bg2) * Tf;
    float T = 1.f;
    bool done = !inside || (gC0 == 0.f && gC1 == 0.f && gC2 == 0.f);
    for (uint32_t base = first; base < last; base += B) {
        if (__syncthreads_count(!done) == 0)
            break;
        for (int i = threadIdx.x; i < B; i += NT) {
            const uint32_t k = base + uint32_t(i);
            if (k < last) {
                const uint32_t id = point_list[k];
                const float2 p = xy[id];
                const float4 co = conic_opacity[id];
                s_geo[i] = make_float4(p.x, p.y, co.x, co.y);
                s_col[i] = make_float4(co.z, co.w, rgb[3 * id], rgb[3 * id + 1]);
                s_blue[i] = rgb[3 * id + 2];
                s_id[i] = id;
            }
        }
        __syncthreads();
        const int cnt = int(min<uint32_t>(B, last - base));
        for (int j = 0; j < cnt; ++j) {
            if (__ballot(!done) == 0)
                break; // the whole wave has finished
            const float4 g = s_geo[j];
            const float4 c = s_col[j];
            const float blue = s_blue[j];
            float gx = 0.f, gy = 0.f, ga = 0.f, gb = 0.f, gc = 0.f, go = 0.f, gr = 0.f, gg = 0.f, gbl = 0.f;
            bool contrib = false;
            if (!done) {
                const float dx = pxf - g.x, dy = pyf - g.y;
                const float power = fmaf(-0.5f * g.z * dx, dx, fmaf(-0.5f * c.x * dy, dy, -g.w * dx * dy));
                const float G = __expf(power);
                const float raw = c.y * G;
                const float alpha = fminf(0.99f, raw);
                if (alpha >= 1.0f / 255.0f) {
                    const float tk = 1.f - alpha;
                    const float test_T = T * tk;
                    if (test_T < 0.0001f) {
                        done = true;
                    } else {
                        contrib = true;
                        const float wgt = alpha * T;
                        S0 = fmaf(-c.z, wgt, S0);
                        S1 = fmaf(-c.w, wgt, S1);
                        S2 = fmaf(-blue, wgt, S2);
                        gr = gC0 * wgt;
                        gg = gC1 * wgt;
                        gbl = gC2 * wgt;
                        const float g_alpha = (gC0 * c.z + gC1 * c.w + gC2 * blue) * T - (gC0 * S0 + gC1 * S1 + gC2 * S2 + gT_Tf) / tk;
                        const float g_raw = raw >= 0.99f ? 0.f : g_alpha;
                        go = G * g_raw;
                        const float gp = c.y * g_raw * G;
                        ga = gp * (-0.5f * dx * dx);
                        gb = gp * (-dx * dy);
                        gc = gp * (-0.5f * dy * dy);
                        gx = gp * (g.z * dx + g.w * dy);
                        gy = gp * (c.x * dy + g.w * dx);
                        T = test_T;
                    }
                }
            }
            if (__ballot(contrib) != 0) {
                gx = wave_sum_to_last(gx);
                gy = wave_sum_to_last(gy);
                ga = wave_sum_to_last(ga);
                gb = wave_sum_to_last(gb);
                gc = wave_sum_to_last(gc);
                go = wave_sum_to_last(go);
                gr = wave_sum_to_last(gr);
                gg = wave_sum_to_last(gg);
                gbl = wave_sum_to_last(gbl);
                if (lane_id() == 63) {
                    float* dst = screen_grads + int64_t(s_id[j]) * 9;
                    unsafeAtomicAdd(dst + 0, gx);
                    unsafeAtomicAdd(dst + 1, gy);
                    unsafeAtomicAdd(dst + 2, ga);
                    unsafeAtomicAdd(dst + 3, gb);
                    unsafeAtomicAdd(dst + 4, gc);
                    unsafeAtomicAdd(dst + 5, go);
                    unsafeAtomicAdd(dst + 6, gr);
                    unsafeAtomicAdd(dst + 7, gg);
                    unsafeAtomicAdd(dst + 8, gbl);
                }
            }
        }
    }
}

namespace {

// Per-Gaussian adjoint of the preprocess: conic -> covariance (2x2 inverse
// adjoint), grad::splat<Ots>, grad::sh_to_colour, normalize.
__global__ __launch_bounds__(kPreThreads) void k_splat_preprocess_backward(int64_t n, const float* __restrict__ means,
    const float* __restrict__ scales, const float* __restrict__ rotations, const float* __restrict__ opacity, const float* __restrict__ sh,
    int sh_coeffs, int sh_degree, float scale_multiplier, DevCamera<float> cam, const uint32_t* __restrict__ tiles,
    const float4* __restrict__ conic_opacity, const uint8_t* __restrict__ clamped, const float* __restrict__ screen_grads,
    float* __restrict__ g_means, float* __restrict__ g_scales, float* __restrict__ g_rotations, float* __restrict__ g_opacity,
    float* __restrict__ g_sh)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    float* gsh = g_sh + i * sh_coeffs * 3;
    if (tiles[i] == 0) {
        for (int a = 0; a < 3; ++a)
            g_means[3 * i + a] = g_scales[3 * i + a] = 0.f;
        for (int a = 0; a < 4; ++a)
            g_rotations[4 * i + a] = 0.f;
        g_opacity[i] = 0.f;
        for (int k = 0; k < sh_coeffs * 3; ++k)
            gsh[k] = 0.f;
        return;
    }
    const float* sg = screen_grads + i * 9;
    const float mu[3] = { means[3 * i], means[3 * i + 1], means[3 * i + 2] };
    const float m = scale_multiplier;
    const float s[3] = { scales[3 * i] * m, scales[3 * i + 1] * m, scales[3 * i + 2] * m };
    const float q[4] = { rotations[4 * i], rotations[4 * i + 1], rotations[4 * i + 2], rotations[4 * i + 3] };
    // conic = cov^-1: dL/dcov = -A G A
    const float4 A = conic_opacity[i];
    const float G00 = sg[2], G01 = 0.5f * sg[3], G11 = sg[4];
    const float t00 = A.x * G00 + A.y * G01, t01 = A.x * G01 + A.y * G11;
    const float t10 = A.y * G00 + A.z * G01, t11 = A.y * G01 + A.z * G11;
    const float r00 = t00 * A.x + t01 * A.y, r01 = t00 * A.y + t01 * A.z;
    const float r10 = t10 * A.x + t11 * A.y, r11 = t10 * A.y + t11 * A.z;
    const float g_cov2[3] = { -r00, -(r01 + r10), -r11 };
    float g_w, g_mu[3], g_s[3], g_q[4];
    splat_ots_grad(opacity[i], mu, s, q, cam, 0.3f, sg[5], sg, g_cov2, &g_w, g_mu, g_s, g_q);
    // colour
    const float v[3] = { mu[0] - cam.pos[0], mu[1] - cam.pos[1], mu[2] - cam.pos[2] };
    const float len = sqrtf(dot3(v, v));
    const float dir[3] = { v[0] / len, v[1] / len, v[2] / len };
    const float gr[3] = { clamped[3 * i] ? sg[6] : 0.f, clamped[3 * i + 1] ? sg[7] : 0.f, clamped[3 * i + 2] ? sg[8] : 0.f };
    float Y[16], dY[16][3];
    sh_basis(sh_degree, dir, Y);
    sh_basis_grad(sh_degree, dir, dY);
    const float* c = sh + i * sh_coeffs * 3;
    const int kmax = min(sh_coeffs, (sh_degree + 1) * (sh_degree + 1));
    float g_dir[3] = { 0.f, 0.f, 0.f };
    for (int k = 0; k < sh_coeffs; ++k) {
        if (k < kmax) {
            const float c0 = c[3 * k], c1 = c[3 * k + 1], c2 = c[3 * k + 2];
            gsh[3 * k + 0] = Y[k] * gr[0];
            gsh[3 * k + 1] = Y[k] * gr[1];
            gsh[3 * k + 2] = Y[k] * gr[2];
            const float dot_c = c0 * gr[0] + c1 * gr[1] + c2 * gr[2];
            g_dir[0] = fmaf(dY[k][0], dot_c, g_dir[0]);
            g_dir[1] = fmaf(dY[k][1], dot_c, g_dir[1]);
            g_dir[2] = fmaf(dY[k][2], dot_c, g_dir[2]);
        } else {
            gsh[3 * k + 0] = gsh[3 * k + 1] = gsh[3 * k + 2] = 0.f;
        }
    }
    const float dd = dot3(g_dir, dir);
    for (int a = 0; a < 3; ++a) {
        g_means[3 * i + a] = g_mu[a] + (g_dir[a] - dd * dir[a]) / len;
        g_scales[3 * i + a] = g_s[a] * m;
    }
    for (int a = 0; a < 4; ++a)
        g_rotations[4 * i + a] = g_q[a];
    g_opacity[i] = g_w;
}

} // namespace

// ------------------------------------------------------------------ host side
struct SplatWorkspace {
    uint32_t* dkeys_a;
    uint32_t* dvals_a;
    uint32_t* dkeys_b;
    uint32_t* dvals_b;
    uint32_t* counts;
    uint32_t* pair_offsets;
    uint32_t* pkeys_a;
    uint32_t* pvals_a;
    uint32_t* pkeys_b;
    uint32_t* counter;
    void* sort_scratch;
    void* scan_scratch;
};

static SplatWorkspace carve_splat(Carver& cv, int64_t n, int64_t cap)
{
    SplatWorkspace w;
    const int64_t nn = std::max<int64_t>(n, 1), cc = std::max<int64_t>(cap, 1);
    w.dkeys_a = cv.take<uint32_t>(nn);
    w.dvals_a = cv.take<uint32_t>(nn);
    w.dkeys_b = cv.take<uint32_t>(nn);
    w.dvals_b = cv.take<uint32_t>(nn);
    w.counts = cv.take<uint32_t>(nn);
    w.pair_offsets = cv.take<uint32_t>(nn);
    w.pkeys_a = cv.take<uint32_t>(cc);
    w.pvals_a = cv.take<uint32_t>(cc);
    w.pkeys_b = cv.take<uint32_t>(cc);
    w.counter = cv.take<uint32_t>(64);
    w.sort_scratch = cv.take<char>(radix_scratch_bytes(std::max(nn, cc)));
    w.scan_scratch = cv.take<char>(scan_scratch_bytes(nn));
    return w;
}

size_t splat_workspace_bytes(int64_t n, int64_t cap)
{
    Carver cv(nullptr, 0);
    carve_splat(cv, n, cap);
    return cv.used;
}

template <int TW, int TH>
static void launch_blend_fwd(int batch, int n_tiles, hipStream_t st, const uint32_t* ranges, const uint32_t* pl, const float2* xy,
    const float4* co, const float* rgb, int W, int H, int gx, const double* bg, float* fb, float* T)
{
    const dim3 grid(n_tiles), block(TW * TH);
    const float b0 = float(bg[0]), b1 = float(bg[1]), b2 = float(bg[2]);
    if (batch == 128)
        k_blend_forward<TW, TH, 128><<<grid, block, 0, st>>>(ranges, pl, xy, co, rgb, W, H, gx, n_tiles, b0, b1, b2, fb, T);
    else if (batch == 256)
        k_blend_forward<TW, TH, 256><<<grid, block, 0, st>>>(ranges, pl, xy, co, rgb, W, H, gx, n_tiles, b0, b1, b2, fb, T);
    else
        k_blend_forward<TW, TH, 512><<<grid, block, 0, st>>>(ranges, pl, xy, co, rgb, W, H, gx, n_tiles, b0, b1, b2, fb, T);
}

template <int TW, int TH>
static void launch_blend_bwd(int batch, int n_tiles, hipStream_t st, const uint32_t* ranges, const uint32_t* pl, const float2* xy,
    const float4* co, const float* rgb, int W, int H, int gx, const double* bg, const float* fb, const float* T, const float* gfb, float* sg)
{
    const dim3 grid(n_tiles), block(TW * TH);
    const float b0 = float(bg[0]), b1 = float(bg[1]), b2 = float(bg[2]);
    if (batch == 128)
        k_blend_backward<TW, TH, 128><<<grid, block, 0, st>>>(ranges, pl, xy, co, rgb, W, H, gx, n_tiles, b0, b1, b2, fb, T, gfb, sg);
    else if (batch == 256)
        k_blend_backward<TW, TH, 256><<<grid, block, 0, st>>>(ranges, pl, xy, co, rgb, W, H, gx, n_tiles, b0, b1, b2, fb, T, gfb, sg);
    else
        k_blend_backward<TW, TH, 512><<<grid, block, 0, st>>>(ranges, pl, xy, co, rgb, W, H, gx, n_tiles, b0, b1, b2, fb, T, gfb, sg);
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

void splat_blend_backward(const dgmr_tiling& t, int n_tiles, hipStream_t st, const uint32_t* ranges, const uint32_t* pl, const float* xy,
    const float* co, const float* rgb, int W, int H, int gx, const double* bg, const float* fb, const float* T, const float* gfb, float* sg)
{
    const auto* xy2 = reinterpret_cast<const float2*>(xy);
    const auto* co4 = reinterpret_cast<const float4*>(co);
    if (t.tile_w == 16)
        launch_blend_bwd<16, 16>(t.batch, n_tiles, st, ranges, pl, xy2, co4, rgb, W, H, gx, bg, fb, T, gfb, sg);
    else if (t.tile_w == 8)
        launch_blend_bwd<8, 8>(t.batch, n_tiles, st, ranges, pl, xy2, co4, rgb, W, H, gx, bg, fb, T, gfb, sg);
    else
        launch_blend_bwd<32, 8>(t.batch, n_tiles, st, ranges, pl, xy2, co4, rgb, W, H, gx, bg, fb, T, gfb, sg);
}

// binning shared by the splatter and the marcher: depth sort, scan, emission,
// pair sort, ranges.  depth_keys/ids are the workspace's A buffers, filled by
// the preprocess.  Returns the pair count through *n_pairs (host).
int bin_gaussians(int64_t n, const int32_t* rects, const uint32_t* tiles, const float* xy_f32, int grid_x, int grid_y, int tile_w,
    int tile_h, dgmr_binning* bins, Carver& cv, int64_t* n_pairs, hipStream_t st)
{
    SplatWorkspace w = carve_splat(cv, n, bins->pair_capacity);
    // stable sort of Gaussian ids by depth bits
    SortBuffers db { w.dkeys_a, w.dvals_a, w.dkeys_b, w.dvals_b };
    DGMR_HIP_CHECK(radix_sort_pairs(db, n, nullptr, 0, 32, w.sort_scratch, st));
    const uint32_t* sorted_ids = (radix_passes(0, 32) % 2 == 0) ? w.dvals_a : w.dvals_b;
    const unsigned blocks_n = unsigned(ceil_div64(n, 256));
    k_gather_counts<<<blocks_n, 256, 0, st>>>(n, sorted_ids, tiles, w.counts);
    DGMR_HIP_CHECK(hipGetLastError());
    DGMR_HIP_CHECK(scan_u32(w.counts, w.pair_offsets, w.counter, n, false, w.scan_scratch, st));
    // the reference's cache keeps the inclusive scan in Gaussian order
    if (bins->point_offsets)
        DGMR_HIP_CHECK(scan_u32(tiles, bins->point_offsets, nullptr, n, true, w.scan_scratch, st));
    uint32_t R = 0;
    DGMR_HIP_CHECK(hipMemcpyAsync(&R, w.counter, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    DGMR_HIP_CHECK(hipStreamSynchronize(st));
    *n_pairs = int64_t(R);
    const int64_t n_tiles = int64_t(grid_x) * grid_y;
    DGMR_HIP_CHECK(hipMemsetAsync(bins->ranges, 0, size_t(n_tiles) * 2 * sizeof(uint32_t), st));
    if (R == 0)
        return DGMR_OK;
    if (int64_t(R) > bins->pair_capacity) {
        set_error("pair capacity too small");
        return DGMR_ERR_CAPACITY;
    }
    // emit into the buffer pair that leaves the sorted ids in bins->point_list
    const int tile_bits = int(higher_msb(uint32_t(n_tiles)));
    const int passes = radix_passes(0, tile_bits);
    SortBuffers pb;
    if (passes % 2 == 0)
        pb = SortBuffers { w.pkeys_a, bins->point_list, w.pkeys_b, w.pvals_a };
    else
        pb = SortBuffers { w.pkeys_a, w.pvals_a, w.pkeys_b, bins->point_list };
    const unsigned emit_blocks = unsigned(ceil_div64(n, 256));
    k_emit_pairs<<<emit_blocks, 256, 0, st>>>(n, sorted_ids, w.pair_offsets, rects, tiles, xy_f32, grid_x, grid_y, tile_w, tile_h, pb.keys_a,
        pb.vals_a);
    DGMR_HIP_CHECK(hipGetLastError());
    DGMR_HIP_CHECK(radix_sort_pairs(pb, int64_t(R), nullptr, 0, tile_bits, w.sort_scratch, st));
    const uint32_t* sorted_keys = (passes % 2 == 0) ? pb.keys_a : pb.keys_b;
    k_tile_ranges<<<unsigned(ceil_div64(R, 256)), 256, 0, st>>>(sorted_keys, w.counter, bins->ranges);
    DGMR_HIP_CHECK(hipGetLastError());
    return DGMR_OK;
}

size_t binning_workspace_bytes(int64_t n, int64_t cap) { return splat_workspace_bytes(n, cap); }

// depth keys / ids buffers the preprocess writes (the first two slices of the workspace)
void binning_key_buffers(void* workspace, int64_t n, uint32_t** keys, uint32_t** ids)
{
    Carver cv(workspace, ~size_t(0));
    const int64_t nn = std::max<int64_t>(n, 1);
    *keys = cv.take<uint32_t>(nn);
    *ids = cv.take<uint32_t>(nn);
}

} // namespace dgmr

using namespace dgmr;

extern "C" {

int dgmr_workspace_size(int64_t n, int32_t width, int32_t height, const dgmr_tiling* tiling, int64_t pair_capacity, size_t* bytes)
{
    DGMR_ARG_CHECK(bytes && n >= 0 && width > 0 && height > 0 && valid_tiling(tiling) && pair_capacity >= 0, "dgmr_workspace_size: bad argument");
    *bytes = splat_workspace_bytes(n, pair_capacity);
    return DGMR_OK;
}

int dgmr_splat_forward(const dgmr_gaussians* g, const dgmr_camera* cam, const dgmr_tiling* tiling, dgmr_splat_geometry* geom,
    dgmr_binning* bins, float* framebuffer, float* final_transmittance, void* workspace, size_t workspace_bytes, dgmr_stats* stats,
    dgmr_stream stream)
{
    DGMR_ARG_CHECK(g && cam && geom && bins && framebuffer && final_transmittance && workspace, "dgmr_splat_forward: null argument");
    DGMR_ARG_CHECK(g->dtype == DGMR_F32, "dgmr_splat_forward: the splatter computes in float32");
    DGMR_ARG_CHECK(valid_tiling(tiling), "dgmr_splat_forward: tiling must be 8x8, 16x16 or 32x8 with batch 128/256/512");
    DGMR_ARG_CHECK(g->n >= 0 && g->n < (int64_t(1) << 31), "dgmr_splat_forward: bad Gaussian count");
    DGMR_ARG_CHECK(g->sh_coeffs >= 1 && g->sh_coeffs <= 16 && g->sh_degree >= 0 && g->sh_degree <= 3, "dgmr_splat_forward: bad SH layout");
    DGMR_ARG_CHECK(cam->width > 0 && cam->height > 0, "dgmr_splat_forward: bad image size");
    DGMR_ARG_CHECK(bins->point_list && bins->ranges, "dgmr_splat_forward: null binning buffers");
    const int64_t n = g->n;
    DGMR_ARG_CHECK(workspace_bytes >= splat_workspace_bytes(n, bins->pair_capacity), "dgmr_splat_forward: workspace too small");
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
    binning_key_buffers(workspace, n, &dkeys, &dids);
    k_splat_preprocess<<<unsigned(ceil_div64(n, kPreThreads)), kPreThreads, 0, st>>>(n, static_cast<const float*>(g->means),
        static_cast<const float*>(g->scales), static_cast<const float*>(g->rotations), static_cast<const float*>(g->opacity),
        static_cast<const float*>(g->sh), g->sh_coeffs, g->sh_degree, float(g->scale_multiplier), dc, geom->rects, geom->tiles_touched,
        geom->xy, reinterpret_cast<float4*>(geom->conic_opacity), geom->rgb, geom->rgb_clamped, geom->depths, dkeys, dids);
    DGMR_HIP_CHECK(hipGetLastError());
    Carver cv(workspace, workspace_bytes);
    int64_t R = 0;
    const int rc = bin_gaussians(n, geom->rects, geom->tiles_touched, geom->xy, dc.grid_x, dc.grid_y, dc.tile_w, dc.tile_h, bins, cv, &R, st);
    if (stats)
        stats->n_pairs = R;
    if (rc != DGMR_OK)
        return rc;
    (void)plane;
    splat_blend_forward(*tiling, int(n_tiles), st, bins->ranges, bins->point_list, geom->xy, geom->conic_opacity, geom->rgb, cam->width,
        cam->height, dc.grid_x, cam->background, framebuffer, final_transmittance);
    DGMR_HIP_CHECK(hipGetLastError());
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
        splat_blend_backward(*tiling, int(n_tiles), st, bins->ranges, bins->point_list, geom->xy, geom->conic_opacity, geom->rgb, cam->width,
            cam->height, dc.grid_x, cam->background, framebuffer, final_transmittance, grad_framebuffer, screen_grads);
        DGMR_HIP_CHECK(hipGetLastError());
    }
    k_splat_preprocess_backward<<<unsigned(ceil_div64(n, kPreThreads)), kPreThreads, 0, st>>>(n, static_cast<const float*>(g->means),
        static_cast<const float*>(g->scales), static_cast<const float*>(g->rotations), static_cast<const float*>(g->opacity),
        static_cast<const float*>(g->sh), g->sh_coeffs, g->sh_degree, float(g->scale_multiplier), dc, geom->tiles_touched,
        reinterpret_cast<const float4*>(geom->conic_opacity), geom->rgb_clamped, screen_grads, grad_means, grad_scales, grad_rotations,
        grad_opacity, grad_sh);
    DGMR_HIP_CHECK(hipGetLastError());
    return DGMR_OK;
}

} // extern "C"
