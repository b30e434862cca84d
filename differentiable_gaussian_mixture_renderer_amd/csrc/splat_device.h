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

// ------------------------------------------------------------------ blending
// One wave (64 lanes) blends one tile.  Lane l owns the NP = TW*TH/64 pixels
// l, l + 64, ... of the tile in row-major order, so each Gaussian read from
// LDS feeds NP independent pixel chains.  The tile's Gaussians go through LDS
// in batches of B (B/64 per lane); the next batch is fetched into registers
// while the current one is blended.
template <int B>
struct BlendSmem {
    float4 geo[B];  // x, y, -a/2, -b (conic a, b, c)
    float4 col[B];  // -c/2, opacity, r, g
    float blue[B];
};

template <int B>
struct BlendFetch {
    static constexpr int NS = B / 64;
    uint32_t id[NS];
    float2 xy[NS];
    float4 co[NS];
    float r[NS], g[NS], b[NS];

    __device__ __forceinline__ void load(uint32_t base, uint32_t last, const uint32_t* __restrict__ point_list, const float2* __restrict__ xy_in,
        const float4* __restrict__ conic_opacity, const float* __restrict__ rgb)
    {
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const uint32_t k = base + uint32_t(lane_id() + 64 * s);
            if (k < last) {
                const uint32_t i = point_list[k];
                id[s] = i;
                xy[s] = xy_in[i];
                co[s] = conic_opacity[i];
                r[s] = rgb[3 * i];
                g[s] = rgb[3 * i + 1];
                b[s] = rgb[3 * i + 2];
            }
        }
    }
};

template <int TW, int TH, int B>
__device__ __forceinline__ void blend_tile_forward(BlendSmem<B>& sm, uint32_t first, uint32_t last, int px0, int py0, int width, int height,
    float bg0, float bg1, float bg2, const uint32_t* __restrict__ point_list, const float2* __restrict__ xy,
    const float4* __restrict__ conic_opacity, const float* __restrict__ rgb, float* __restrict__ framebuffer, float* __restrict__ final_T)
{
    constexpr int NP = TW * TH / 64;
    const int lane = int(lane_id());
    float pxf[NP], pyf[NP], T[NP], C0[NP], C1[NP], C2[NP];
    bool live[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const int q = lane + 64 * k;
        const int px = px0 + q % TW, py = py0 + q / TW;
        pxf[k] = float(px);
        pyf[k] = float(py);
        live[k] = px < width && py < height;
        T[k] = 1.f;
        C0[k] = C1[k] = C2[k] = 0.f;
    }
    BlendFetch<B> f;
    if (first < last)
        f.load(first, last, point_list, xy, conic_opacity, rgb);
    for (uint32_t base = first; base < last; base += B) {
        bool any = false;
#pragma unroll
        for (int k = 0; k < NP; ++k)
            any |= live[k];
        if (__ballot(any) == 0)
            break;
        __syncthreads(); // the previous batch has been read
#pragma unroll
        for (int s = 0; s < BlendFetch<B>::NS; ++s) {
            const int i = lane + 64 * s;
            sm.geo[i] = make_float4(f.xy[s].x, f.xy[s].y, -0.5f * f.co[s].x, -f.co[s].y);
            sm.col[i] = make_float4(-0.5f * f.co[s].z, f.co[s].w, f.r[s], f.g[s]);
            sm.blue[i] = f.b[s];
        }
        __syncthreads();
        if (base + B < last)
            f.load(base + B, last, point_list, xy, conic_opacity, rgb);
        const int cnt = int(min<uint32_t>(B, last - base));
        for (int j = 0; j < cnt; ++j) {
            const float4 g = sm.geo[j];
            const float4 c = sm.col[j];
            const float blue = sm.blue[j];
            bool still = false;
#pragma unroll
            for (int k = 0; k < NP; ++k) {
                if (live[k]) {
                    const float dx = pxf[k] - g.x, dy = pyf[k] - g.y;
                    const float power = fmaf(g.z * dx, dx, fmaf(c.x * dy, dy, g.w * dx * dy));
                    const float alpha = fminf(0.99f, c.y * __expf(power));
                    if (alpha >= 1.0f / 255.0f) {
                        const float test_T = T[k] * (1.f - alpha);
                        if (test_T < 0.0001f) {
                            live[k] = false;
                        } else {
                            const float wgt = alpha * T[k];
                            C0[k] = fmaf(c.z, wgt, C0[k]);
                            C1[k] = fmaf(c.w, wgt, C1[k]);
                            C2[k] = fmaf(blue, wgt, C2[k]);
                            T[k] = test_T;
                        }
                    }
                }
                still |= live[k];
            }
            if (__ballot(still) == 0)
                break;
        }
    }
    const int64_t plane = int64_t(width) * height;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const int q = lane + 64 * k;
        const int px = px0 + q % TW, py = py0 + q / TW;
        if (px < width && py < height) {
            const int64_t p = int64_t(py) * width + px;
            framebuffer[p] = fmaf(T[k], bg0, C0[k]);
            framebuffer[plane + p] = fmaf(T[k], bg1, C1[k]);
            framebuffer[2 * plane + p] = fmaf(T[k], bg2, C2[k]);
            final_T[p] = T[k];
        }
    }
}

template <int B>
struct BlendSmemBwd {
    float4 geo[B];   // x, y, conic a, conic b
    float4 col[B];   // conic c, opacity, r, g
    float blue[B];
    float acc[9][B]; // per-batch gradient sums of the tile: xy 2, conic 3, opacity 1, rgb 3
};

// Blend adjoint of one tile, front to back.  The remaining colour is carried
// unnormalised (C_final minus what has been composited so far), which is the
// reference prototype's remaining colour (grad_inria_splatter_render.cu:100-156)
// times the current transmittance, without its repeated division by (1 - alpha).
// Per-Gaussian gradients (xy 2, conic 3, opacity 1, rgb 3) are summed over a
// lane's pixels in registers, over the wave with DPP, kept per batch in LDS
// and leave the tile as one global atomic per nonzero value and Gaussian.
template <int TW, int TH, int B>
__device__ __forceinline__ void blend_tile_backward(BlendSmemBwd<B>& sm, uint32_t first, uint32_t last, int px0, int py0, int width,
    int height, float bg0, float bg1, float bg2, const uint32_t* __restrict__ point_list, const float2* __restrict__ xy,
    const float4* __restrict__ conic_opacity, const float* __restrict__ rgb, const float* __restrict__ framebuffer,
    const float* __restrict__ final_T, const float* __restrict__ grad_fb, float* __restrict__ screen_grads)
{
    constexpr int NP = TW * TH / 64;
    constexpr int NS = B / 64;
    const int lane = int(lane_id());
    const int64_t plane = int64_t(width) * height;
    float pxf[NP], pyf[NP], gC0[NP], gC1[NP], gC2[NP], S0[NP], S1[NP], S2[NP], gT_Tf[NP], T[NP];
    bool live[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const int q = lane + 64 * k;
        const int px = px0 + q % TW, py = py0 + q / TW;
        pxf[k] = float(px);
        pyf[k] = float(py);
        gC0[k] = gC1[k] = gC2[k] = 0.f;
        S0[k] = S1[k] = S2[k] = 0.f;
        float Tf = 1.f;
        if (px < width && py < height) {
            const int64_t p = int64_t(py) * width + px;
            gC0[k] = grad_fb[p];
            gC1[k] = grad_fb[plane + p];
            gC2[k] = grad_fb[2 * plane + p];
            Tf = final_T[p];
            S0[k] = framebuffer[p] - Tf * bg0;
            S1[k] = framebuffer[plane + p] - Tf * bg1;
            S2[k] = framebuffer[2 * plane + p] - Tf * bg2;
        }
        gT_Tf[k] = (gC0[k] * bg0 + gC1[k] * bg1 + gC2[k] * bg2) * Tf;
        T[k] = 1.f;
        live[k] = px < width && py < height && !(gC0[k] == 0.f && gC1[k] == 0.f && gC2[k] == 0.f);
    }
    BlendFetch<B> f;
    if (first < last)
        f.load(first, last, point_list, xy, conic_opacity, rgb);
    for (uint32_t base = first; base < last; base += B) {
        bool any = false;
#pragma unroll
        for (int k = 0; k < NP; ++k)
            any |= live[k];
        if (__ballot(any) == 0)
            break;
        __syncthreads(); // the previous batch has been read and flushed
        uint32_t ids[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const int i = lane + 64 * s;
            ids[s] = f.id[s];
            sm.geo[i] = make_float4(f.xy[s].x, f.xy[s].y, f.co[s].x, f.co[s].y);
            sm.col[i] = make_float4(f.co[s].z, f.co[s].w, f.r[s], f.g[s]);
            sm.blue[i] = f.b[s];
#pragma unroll
            for (int q = 0; q < 9; ++q)
                sm.acc[q][i] = 0.f;
        }
        __syncthreads();
        if (base + B < last)
            f.load(base + B, last, point_list, xy, conic_opacity, rgb);
        const int cnt = int(min<uint32_t>(B, last - base));
        for (int j = 0; j < cnt; ++j) {
            const float4 g = sm.geo[j];
            const float4 c = sm.col[j];
            const float blue = sm.blue[j];
            float sx = 0.f, sy = 0.f, sa = 0.f, sb = 0.f, sc = 0.f, so = 0.f, sr = 0.f, sg = 0.f, sbl = 0.f;
            bool contrib = false, still = false;
#pragma unroll
            for (int k = 0; k < NP; ++k) {
                if (live[k]) {
                    const float dx = pxf[k] - g.x, dy = pyf[k] - g.y;
                    const float power = fmaf(-0.5f * g.z * dx, dx, fmaf(-0.5f * c.x * dy, dy, -g.w * dx * dy));
                    const float G = __expf(power);
                    const float raw = c.y * G;
                    const float alpha = fminf(0.99f, raw);
                    if (alpha >= 1.0f / 255.0f) {
                        const float tk = 1.f - alpha;
                        const float test_T = T[k] * tk;
                        if (test_T < 0.0001f) {
                            live[k] = false;
                        } else {
                            contrib = true;
                            const float wgt = alpha * T[k];
                            S0[k] = fmaf(-c.z, wgt, S0[k]);
                            S1[k] = fmaf(-c.w, wgt, S1[k]);
                            S2[k] = fmaf(-blue, wgt, S2[k]);
                            sr += gC0[k] * wgt;
                            sg += gC1[k] * wgt;
                            sbl += gC2[k] * wgt;
                            // 1 - alpha >= 0.01, so the hardware reciprocal (1 ulp) is ample
                            const float inv_tk = __builtin_amdgcn_rcpf(tk);
                            const float g_alpha = (gC0[k] * c.z + gC1[k] * c.w + gC2[k] * blue) * T[k]
                                - (gC0[k] * S0[k] + gC1[k] * S1[k] + gC2[k] * S2[k] + gT_Tf[k]) * inv_tk;
                            const float g_raw = raw >= 0.99f ? 0.f : g_alpha;
                            so += G * g_raw;
                            const float gp = c.y * g_raw * G;
                            sa += gp * (-0.5f * dx * dx);
                            sb += gp * (-dx * dy);
                            sc += gp * (-0.5f * dy * dy);
                            sx += gp * (g.z * dx + g.w * dy);
                            sy += gp * (c.x * dy + g.w * dx);
                            T[k] = test_T;
                        }
                    }
                }
                still |= live[k];
            }
            if (__ballot(contrib) != 0) {
                sx = wave_sum_to_last(sx);
                sy = wave_sum_to_last(sy);
                sa = wave_sum_to_last(sa);
                sb = wave_sum_to_last(sb);
                sc = wave_sum_to_last(sc);
                so = wave_sum_to_last(so);
                sr = wave_sum_to_last(sr);
                sg = wave_sum_to_last(sg);
                sbl = wave_sum_to_last(sbl);
                if (lane == 63) {
                    sm.acc[0][j] = sx;
                    sm.acc[1][j] = sy;
                    sm.acc[2][j] = sa;
                    sm.acc[3][j] = sb;
                    sm.acc[4][j] = sc;
                    sm.acc[5][j] = so;
                    sm.acc[6][j] = sr;
                    sm.acc[7][j] = sg;
                    sm.acc[8][j] = sbl;
                }
            }
            if (__ballot(still) == 0)
                break;
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const int i = lane + 64 * s;
            if (i < cnt) {
                float* dst = screen_grads + int64_t(ids[s]) * 9;
#pragma unroll
                for (int q = 0; q < 9; ++q) {
                    const float v = sm.acc[q][i];
                    if (v != 0.f)
                        unsafeAtomicAdd(dst + q, v);
                }
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
