// SPDX-License-Identifier: GPL-3.0-or-later
// Device functions of the tile splatter shared by the single-view kernels and
// the persistent multi-view kernels (reference: src/dgmr/splat_forward.cu).
#pragma once

#include "device_grad.h"
#include "device_math.h"

namespace dgmr {

struct SplatParams {
    const float* means;
    const float* scales;
    const float* rotations;
    const float* opacity;
    const float* sh;
    int sh_coeffs, sh_degree;
    float scale_multiplier;
};

// per-Gaussian outputs of the preprocess; index o = view * n + i
struct SplatOut {
    int32_t* rects;
    uint32_t* tiles;
    float* xy;
    float4* conic_opacity;
    float* rgb;
    uint8_t* clamped;
    float* depths;
    uint32_t* depth_keys;
    uint32_t* ids;
};

// SH coefficients of a block of Gaussians staged in LDS: the block's
// contiguous [n_block, K, 3] slice is read with coalesced loads and each
// Gaussian's row sits at a stride of 3K + 1 floats (odd, so the 64 lanes of a
// wave reading their own rows hit distinct LDS banks).
constexpr int kShStride = 49;

__device__ __forceinline__ void stage_sh(float* s_sh, const float* sh, int64_t first, int count, int sh_coeffs)
{
    const int row = sh_coeffs * 3;
    const int total = count * row;
    const float* src = sh + first * row;
    for (int e = threadIdx.x; e < total; e += blockDim.x)
        s_sh[(e / row) * kShStride + (e % row)] = src[e];
}

__device__ __forceinline__ void unstage_sh(float* dst, const float* s_sh, int64_t first, int count, int sh_coeffs)
{
    const int row = sh_coeffs * 3;
    const int total = count * row;
    float* out = dst + first * row;
    for (int e = threadIdx.x; e < total; e += blockDim.x)
        out[e] = s_sh[(e / row) * kShStride + (e % row)];
}

// preprocess of Gaussian i for one camera (splat_forward.cu:71-118): cull
// view-z < 0.2, EWA splat (Ots, 0.3 px^2 filter), rect ceil(3 sigma), tiles
// touched, depth, SH colour, conic.  Writes slot o.  c: this Gaussian's SH row.
__device__ __forceinline__ void preprocess_gaussian(int64_t i, int64_t o, const SplatParams& p, const float* c, const DevCamera<float>& cam,
    const SplatOut& out)
{
    out.tiles[o] = 0;
    out.rects[2 * o] = 0;
    out.rects[2 * o + 1] = 0;
    out.xy[2 * o] = 0.f;
    out.xy[2 * o + 1] = 0.f;
    out.depth_keys[o] = 0xffffffffu;
    out.ids[o] = uint32_t(o);
    const float mu[3] = { p.means[3 * i], p.means[3 * i + 1], p.means[3 * i + 2] };
    float t[3];
    transform_point(cam.view, mu, t, 3);
    if (t[2] < 0.2f)
        return;
    const float m = p.scale_multiplier;
    const float s[3] = { p.scales[3 * i] * m, p.scales[3 * i + 1] * m, p.scales[3 * i + 2] * m };
    const float q[4] = { p.rotations[4 * i], p.rotations[4 * i + 1], p.rotations[4 * i + 2], p.rotations[4 * i + 3] };
    const float w = p.opacity[i];
    SplatIn<float> in;
    splat_internals(mu, s, q, cam, &in);
    const float cov_f[3] = { in.cov2[0] + 0.3f, in.cov2[1], in.cov2[2] + 0.3f };
    float xy[2];
    project_to_screen(mu, cam, xy);
    const int ext_x = guarded_int(ceilf(3.f * sqrt_rn(cov_f[0])));
    const int ext_y = guarded_int(ceilf(3.f * sqrt_rn(cov_f[2])));
    out.rects[2 * o] = ext_x;
    out.rects[2 * o + 1] = ext_y;
    int x0, y0, x1, y1;
    get_rect(xy[0], xy[1], ext_x, ext_y, cam.grid_x, cam.grid_y, cam.tile_w, cam.tile_h, &x0, &y0, &x1, &y1);
    const uint32_t touched = uint32_t(x1 - x0) * uint32_t(y1 - y0);
    if (touched == 0)
        return;
    out.tiles[o] = touched;
    out.xy[2 * o] = xy[0];
    out.xy[2 * o + 1] = xy[1];
    const float dv[3] = { cam.pos[0] - mu[0], cam.pos[1] - mu[1], cam.pos[2] - mu[2] };
    const float depth = length3(dv);
    out.depths[o] = depth;
    out.depth_keys[o] = __float_as_uint(depth);
    // colour from SH along normalize(mu - cam) (math.h:233-271)
    const float v[3] = { -dv[0], -dv[1], -dv[2] };
    const float inv_len = __fdiv_rn(1.f, sqrt_rn(dot3(v, v)));
    const float dir[3] = { v[0] * inv_len, v[1] * inv_len, v[2] * inv_len };
    float Y[16];
    sh_basis(p.sh_degree, dir, Y);
    const int kmax = min(p.sh_coeffs, (p.sh_degree + 1) * (p.sh_degree + 1));
    float r0 = 0.f, r1 = 0.f, r2 = 0.f;
    for (int k = 0; k < kmax; ++k) {
        r0 = fmaf(Y[k], c[3 * k + 0], r0);
        r1 = fmaf(Y[k], c[3 * k + 1], r1);
        r2 = fmaf(Y[k], c[3 * k + 2], r2);
    }
    r0 += 0.5f;
    r1 += 0.5f;
    r2 += 0.5f;
    out.rgb[3 * o + 0] = fmaxf(r0, 0.f);
    out.rgb[3 * o + 1] = fmaxf(r1, 0.f);
    out.rgb[3 * o + 2] = fmaxf(r2, 0.f);
    out.clamped[3 * o + 0] = r0 >= 0.f;
    out.clamped[3 * o + 1] = r1 >= 0.f;
    out.clamped[3 * o + 2] = r2 >= 0.f;
    // screen-space weight, Ots formulation (math.h:389-397)
    int i0, i1;
    larger2_idx(s, &i0, &i1);
    const float i2prime = float(2 * kPi) * s[i0] * s[i1];
    const float det2 = cov_f[0] * cov_f[2] - cov_f[1] * cov_f[1];
    const float i2 = float(2 * kPi) * sqrt_rn(det2);
    const float weight = w * det3(in.J) * i2prime / i2;
    out.conic_opacity[o] = make_float4(cov_f[2] / det2, -cov_f[1] / det2, cov_f[0] / det2, weight);
}

template <int B>
struct BlendSmem {
    float4 geo[B];  // x, y, conic a, conic b (forward: pre-scaled -a/2, -b)
    float4 col[B];  // conic c (forward: -c/2), opacity, r, g
    float blue[B];
    uint32_t id[B];
};

// One tile of the front-to-back blend (splat_forward.cu:242-335): pixels at
// integer coordinates, alpha = min(0.99, o G), skip alpha < 1/255, stop before
// T (1 - alpha) < 1e-4.  Must be called by all threads of the block.
template <int TW, int TH, int B>
__device__ __forceinline__ void blend_tile_forward(BlendSmem<B>& sm, uint32_t first, uint32_t last, int px0, int py0, int width, int height,
    float bg0, float bg1, float bg2, const uint32_t* __restrict__ point_list, const float2* __restrict__ xy,
    const float4* __restrict__ conic_opacity, const float* __restrict__ rgb, float* __restrict__ framebuffer, float* __restrict__ final_T)
{
    constexpr int NT = TW * TH;
    const int px = px0 + int(threadIdx.x) % TW;
    const int py = py0 + int(threadIdx.x) / TW;
    const bool inside = px < width && py < height;
    const float pxf = float(px), pyf = float(py);
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
                sm.geo[i] = make_float4(p.x, p.y, -0.5f * co.x, -co.y);
                sm.col[i] = make_float4(-0.5f * co.z, co.w, rgb[3 * id], rgb[3 * id + 1]);
                sm.blue[i] = rgb[3 * id + 2];
            }
        }
        __syncthreads();
        if (!done) {
            const int cnt = int(min<uint32_t>(B, last - base));
            for (int j = 0; j < cnt; ++j) {
                const float4 g = sm.geo[j];
                const float4 c = sm.col[j];
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
                C2 = fmaf(sm.blue[j], wgt, C2);
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

template <int B>
struct BlendSmemBwd {
    float4 geo[B];  // x, y, conic a, conic b
    float4 col[B];  // conic c, opacity, r, g
    float blue[B];
    uint32_t id[B];
    float acc[9][B]; // per-batch gradient sums: xy 2, conic 3, opacity 1, rgb 3
};

// Blend adjoint of one tile, front to back.  The remaining colour is carried
// unnormalised (C_final minus what has been composited so far), which is the
// reference prototype's remaining colour (grad_inria_splatter_render.cu:100-156)
// times the current transmittance, without its repeated division by (1 - alpha).
// Per-Gaussian gradients (xy 2, conic 3, opacity 1, rgb 3) are summed over the
// 64 pixels of a wave with DPP, over the waves of the tile with LDS atomics,
// and leave the tile as one global atomic per value and Gaussian.
template <int TW, int TH, int B>
__device__ __forceinline__ void blend_tile_backward(BlendSmemBwd<B>& sm, uint32_t first, uint32_t last, int px0, int py0, int width,
    int height, float bg0, float bg1, float bg2, const uint32_t* __restrict__ point_list, const float2* __restrict__ xy,
    const float4* __restrict__ conic_opacity, const float* __restrict__ rgb, const float* __restrict__ framebuffer,
    const float* __restrict__ final_T, const float* __restrict__ grad_fb, float* __restrict__ screen_grads)
{
    constexpr int NT = TW * TH;
    const int px = px0 + int(threadIdx.x) % TW;
    const int py = py0 + int(threadIdx.x) / TW;
    const bool inside = px < width && py < height;
    const float pxf = float(px), pyf = float(py);
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
        const int cnt = int(min<uint32_t>(B, last - base));
        for (int i = threadIdx.x; i < B; i += NT) {
            const uint32_t k = base + uint32_t(i);
            if (k < last) {
                const uint32_t id = point_list[k];
                const float2 p = xy[id];
                const float4 co = conic_opacity[id];
                sm.geo[i] = make_float4(p.x, p.y, co.x, co.y);
                sm.col[i] = make_float4(co.z, co.w, rgb[3 * id], rgb[3 * id + 1]);
                sm.blue[i] = rgb[3 * id + 2];
                sm.id[i] = id;
            }
#pragma unroll
            for (int q = 0; q < 9; ++q)
                sm.acc[q][i] = 0.f;
        }
        __syncthreads();
        for (int j = 0; j < cnt; ++j) {
            if (__ballot(!done) == 0)
                break; // the whole wave has finished
            const float4 g = sm.geo[j];
            const float4 c = sm.col[j];
            const float blue = sm.blue[j];
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
                        // 1 - alpha >= 0.01, so the hardware reciprocal (1 ulp) is ample
                        const float inv_tk = __builtin_amdgcn_rcpf(tk);
                        const float g_alpha = (gC0 * c.z + gC1 * c.w + gC2 * blue) * T - (gC0 * S0 + gC1 * S1 + gC2 * S2 + gT_Tf) * inv_tk;
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
                    atomicAdd(&sm.acc[0][j], gx);
                    atomicAdd(&sm.acc[1][j], gy);
                    atomicAdd(&sm.acc[2][j], ga);
                    atomicAdd(&sm.acc[3][j], gb);
                    atomicAdd(&sm.acc[4][j], gc);
                    atomicAdd(&sm.acc[5][j], go);
                    atomicAdd(&sm.acc[6][j], gr);
                    atomicAdd(&sm.acc[7][j], gg);
                    atomicAdd(&sm.acc[8][j], gbl);
                }
            }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < cnt; i += NT) {
            float* dst = screen_grads + int64_t(sm.id[i]) * 9;
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                const float v = sm.acc[q][i];
                if (v != 0.f)
                    unsafeAtomicAdd(dst + q, v);
            }
        }
    }
}

// Per-Gaussian adjoint of the preprocess for one camera: conic -> covariance
// (2x2 inverse adjoint), grad::splat<Ots> (grad/math.h:333-527),
// grad::sh_to_colour (grad/math.h:203-282) and normalize.  Adds into the
// accumulators (g_sh has 48 slots; only k < sh_coeffs are meaningful).
// c: this Gaussian's SH row.
__device__ __forceinline__ void preprocess_backward_gaussian(int64_t i, const SplatParams& p, const float* c, const DevCamera<float>& cam,
    const float4 A, const uint8_t* clamped3, const float* sg, float* g_mu_acc, float* g_s_acc, float* g_q_acc, float* g_w_acc, float* g_sh_acc)
{
    const float mu[3] = { p.means[3 * i], p.means[3 * i + 1], p.means[3 * i + 2] };
    const float m = p.scale_multiplier;
    const float s[3] = { p.scales[3 * i] * m, p.scales[3 * i + 1] * m, p.scales[3 * i + 2] * m };
    const float q[4] = { p.rotations[4 * i], p.rotations[4 * i + 1], p.rotations[4 * i + 2], p.rotations[4 * i + 3] };
    // conic = cov^-1: dL/dcov = -A G A
    const float G00 = sg[2], G01 = 0.5f * sg[3], G11 = sg[4];
    const float t00 = A.x * G00 + A.y * G01, t01 = A.x * G01 + A.y * G11;
    const float t10 = A.y * G00 + A.z * G01, t11 = A.y * G01 + A.z * G11;
    const float r00 = t00 * A.x + t01 * A.y, r01 = t00 * A.y + t01 * A.z;
    const float r10 = t10 * A.x + t11 * A.y, r11 = t10 * A.y + t11 * A.z;
    const float g_cov2[3] = { -r00, -(r01 + r10), -r11 };
    float g_w, g_mu[3], g_s[3], g_q[4];
    splat_ots_grad(p.opacity[i], mu, s, q, cam, 0.3f, sg[5], sg, g_cov2, &g_w, g_mu, g_s, g_q);
    const float v[3] = { mu[0] - cam.pos[0], mu[1] - cam.pos[1], mu[2] - cam.pos[2] };
    const float len = sqrtf(dot3(v, v));
    const float dir[3] = { v[0] / len, v[1] / len, v[2] / len };
    const float gr[3] = { clamped3[0] ? sg[6] : 0.f, clamped3[1] ? sg[7] : 0.f, clamped3[2] ? sg[8] : 0.f };
    float Y[16], dY[16][3];
    sh_basis(p.sh_degree, dir, Y);
    sh_basis_grad(p.sh_degree, dir, dY);
    const int kmax = min(p.sh_coeffs, (p.sh_degree + 1) * (p.sh_degree + 1));
    float g_dir[3] = { 0.f, 0.f, 0.f };
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        if (k < kmax) {
            const float c0 = c[3 * k], c1 = c[3 * k + 1], c2 = c[3 * k + 2];
            g_sh_acc[3 * k + 0] += Y[k] * gr[0];
            g_sh_acc[3 * k + 1] += Y[k] * gr[1];
            g_sh_acc[3 * k + 2] += Y[k] * gr[2];
            const float dot_c = c0 * gr[0] + c1 * gr[1] + c2 * gr[2];
            g_dir[0] = fmaf(dY[k][0], dot_c, g_dir[0]);
            g_dir[1] = fmaf(dY[k][1], dot_c, g_dir[1]);
            g_dir[2] = fmaf(dY[k][2], dot_c, g_dir[2]);
        }
    }
    const float dd = dot3(g_dir, dir);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        g_mu_acc[a] += g_mu[a] + (g_dir[a] - dd * dir[a]) / len;
        g_s_acc[a] += g_s[a] * m;
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
        g_q_acc[a] += g_q[a];
    *g_w_acc += g_w;
}

// logical tile of a block: workgroups are dealt round-robin over the 8 XCDs,
// so give each XCD one contiguous band of tiles (shared Gaussians stay in that
// XCD's L2)
__device__ __forceinline__ int xcd_tile(int b, int n_tiles)
{
    const int q = n_tiles / 8, r = n_tiles % 8;
    const int x = b % 8, slot = b / 8;
    return x * q + min(x, r) + slot;
}

} // namespace dgmr
