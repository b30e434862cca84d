// SPDX-License-Identifier: GPL-3.0-or-later
// Device functions of the tile splatter shared by the single-view kernels and
// the persistent multi-view kernels (reference: src/dgmr/splat_forward.cu).
#pragma once

#include "device_grad.h"
#include "device_math.h"
#include "tile_reach.h"

namespace dgmr {

struct SplatParams {
    const float* means;
    const float* scales;
    const float* rotations;
    const float* opacity;
    const float* sh;
    int sh_coeffs, sh_degree;
    float scale_multiplier;
    int activation; // DGMR_ACT_*: 1 = log-scales, unnormalised quaternions, opacity logits
};

// scales (times the multiplier), rotation and opacity of Gaussian i as the
// renderer uses them; with activation the stored values are log-scales, an
// unnormalised quaternion and an opacity logit (exp, q/|q|, sigmoid as
// torch.exp, torch.nn.functional.normalize and torch.sigmoid compute them)
__device__ __forceinline__ void gaussian_params(const SplatParams& p, int64_t i, float* s, float* q, float* w)
{
    const float m = p.scale_multiplier;
    if (p.activation) {
#pragma unroll
        for (int a = 0; a < 3; ++a)
            s[a] = expf(p.scales[3 * i + a]) * m;
        const float r[4] = { p.rotations[4 * i], p.rotations[4 * i + 1], p.rotations[4 * i + 2], p.rotations[4 * i + 3] };
        const float nrm = fmaxf(sqrtf(r[0] * r[0] + r[1] * r[1] + r[2] * r[2] + r[3] * r[3]), 1e-12f);
#pragma unroll
        for (int a = 0; a < 4; ++a)
            q[a] = r[a] / nrm;
        *w = 1.f / (1.f + expf(-p.opacity[i]));
    } else {
#pragma unroll
        for (int a = 0; a < 3; ++a)
            s[a] = p.scales[3 * i + a] * m;
#pragma unroll
        for (int a = 0; a < 4; ++a)
            q[a] = p.rotations[4 * i + a];
        *w = p.opacity[i];
    }
}

// turns gradients w.r.t. the used scale (before the multiplier), rotation and
// opacity into gradients w.r.t. the stored log-scale, quaternion and logit
__device__ __forceinline__ void activation_backward(const SplatParams& p, int64_t i, float* g_s, float* g_q, float* g_w)
{
    if (!p.activation)
        return;
#pragma unroll
    for (int a = 0; a < 3; ++a)
        g_s[a] *= expf(p.scales[3 * i + a]);
    const float r[4] = { p.rotations[4 * i], p.rotations[4 * i + 1], p.rotations[4 * i + 2], p.rotations[4 * i + 3] };
    const float nrm = fmaxf(sqrtf(r[0] * r[0] + r[1] * r[1] + r[2] * r[2] + r[3] * r[3]), 1e-12f);
    float qn[4], d = 0.f;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        qn[a] = r[a] / nrm;
        d += g_q[a] * qn[a];
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
        g_q[a] = (g_q[a] - d * qn[a]) / nrm;
    const float o = 1.f / (1.f + expf(-p.opacity[i]));
    *g_w *= o * (1.f - o);
}

// ------------------------------------------------------------ tight culling
// With tight culling a Gaussian is binned only into the tiles of its
// rectangle that hold a pixel where its alpha can reach 1/255.  The blend
// skips every (pixel, Gaussian) with alpha < 1/255 in both directions, so the
// dropped tiles change no image, transmittance or gradient.  alpha >= 1/255
// needs w exp(-Q/2) >= 1/255 with Q = a u^2 + 2 b u v + c v^2 (u, v: pixel
// minus centre, (a, b, c) the conic), i.e. Q <= t = 2 ln(255 w).  The region
// binned is the ellipse Q <= t' = t (1 + 1e-5 k) + 1e-3, where
// k = (a + c)^2 / (ac - b^2) bounds the blend's relative rounding of Q, and a
// tile counts when the ellipse meets its pixel box grown by
// delta = 1/16 + 2^-20 max(|mx|, |my|) pixels (the rounding of the offsets to
// the centre).  Both allowances only ever add tiles.
struct __align__(16) ReachRec {
    uint32_t origin; // (first tile row << 16) | first tile column of the rectangle
    uint32_t dims;   // (rows << 16) | columns; 0 when nothing is binned
    uint32_t grid_x; // tile columns of the Gaussian's view
    float mx, my;    // centre in pixels
    float A, B, C;   // (a, b, c) / t'; A < 0: the whole rectangle (degenerate conic)
};
static_assert(sizeof(ReachRec) == 32, "ReachRec is two 16-byte words");

struct ReachEllipse {
    float mx, my, A, inv_A, B, det, vmax, vr, delta;
    bool whole;
};

// Both the counting and the emission of pairs evaluate the ellipse through
// these two functions, with contraction off and the hardware reciprocal and
// square root (an ulp or two: far inside delta), so they agree bit for bit.
__device__ __forceinline__ ReachEllipse reach_ellipse(const ReachRec& r)
{
#pragma clang fp contract(off)
    ReachEllipse e;
    e.mx = r.mx;
    e.my = r.my;
    e.A = r.A;
    e.inv_A = __builtin_amdgcn_rcpf(r.A);
    e.B = r.B;
    e.whole = !(r.A > 0.f);
    e.det = r.A * r.C - r.B * r.B;
    e.vmax = __builtin_amdgcn_sqrtf(r.A * __builtin_amdgcn_rcpf(e.det));                // |v| of the top and bottom
    e.vr = -r.B * __builtin_amdgcn_rcpf(__builtin_amdgcn_sqrtf(r.C * e.det));           // v of the rightmost point (leftmost: -vr)
    e.delta = 0.0625f + fmaxf(fabsf(r.mx), fabsf(r.my)) * 0x1p-20f;
    return e;
}

// tile columns [x, y] of tile row ty (within the rectangle's columns
// [x0, x0 + cols)) whose pixel boxes the ellipse meets; x > y when none.
// Across the row's band of pixel rows the reached x range is an interval:
// its right end is the concave u_R(v) = (-B v + sqrt(A - det v^2)) / A at the
// band's clamp of the rightmost point, its left end likewise.
__device__ __forceinline__ int2 reach_columns(const ReachEllipse& e, int ty, int x0, int cols, int tw, int th)
{
#pragma clang fp contract(off)
    if (e.whole)
        return make_int2(x0, x0 + cols - 1);
    const float vb = float(ty * th) - e.my;
    const float lo = fmaxf(vb - e.delta, -e.vmax), hi = fminf((vb + float(th - 1)) + e.delta, e.vmax);
    if (!(lo <= hi))
        return make_int2(1, 0);
    const float vR = fminf(fmaxf(e.vr, lo), hi), vL = fminf(fmaxf(-e.vr, lo), hi);
    const float sR = __builtin_amdgcn_sqrtf(fmaxf(e.A - e.det * (vR * vR), 0.f));
    const float sL = __builtin_amdgcn_sqrtf(fmaxf(e.A - e.det * (vL * vL), 0.f));
    const float xR = e.mx + (sR - e.B * vR) * e.inv_A;
    const float xL = e.mx - (sL + e.B * vL) * e.inv_A;
    // column tx holds pixels tx * tw .. tx * tw + tw - 1 (tw a power of two: the scaling is exact)
    const float inv_tw = __builtin_amdgcn_rcpf(float(tw));
    const float c0 = fmaxf(ceilf(((xL - e.delta) - float(tw - 1)) * inv_tw), float(x0));
    const float c1 = fminf(floorf((xR + e.delta) * inv_tw), float(x0 + cols - 1));
    if (!(c0 <= c1))
        return make_int2(1, 0);
    return make_int2(int(c0), int(c1));
}

// the record of a visible Gaussian: rectangle x0..x0+cols, y0..y0+rows; conic (a, b, c), weight w
__device__ __forceinline__ ReachRec make_reach(int x0, int y0, int cols, int rows, int grid_x, float mx, float my, float a, float b, float c,
    float w)
{
    ReachRec r { (uint32_t(y0) << 16) | uint32_t(x0), (uint32_t(rows) << 16) | uint32_t(cols), uint32_t(grid_x), mx, my, 0.f, 0.f, 0.f };
    // 2 ln(255 w) from the hardware log2 (error ~1e-7, inside the 1e-3 allowance)
    const float t = 2.f * 0.6931471805599453f * __builtin_amdgcn_logf(255.f * w);
    if (!(t > 0.f)) {
        r.dims = 0; // alpha < 1/255 everywhere
        return r;
    }
    const float det = a * c - b * b;
    const float k = ((a + c) * (a + c)) * __builtin_amdgcn_rcpf(det);
    const float tp = (t + t * 1e-5f * k) + 1e-3f;
    if (!(det > 0.f) || !(tp < 3.0e38f) || !(a > 0.f) || !(c > 0.f)) {
        r.A = -1.f;
        return r;
    }
    const float inv_tp = __builtin_amdgcn_rcpf(tp);
    r.A = a * inv_tp;
    r.B = b * inv_tp;
    r.C = c * inv_tp;
    if (!(r.A > 0.f) || !(r.C > 0.f) || !(r.A * r.C - r.B * r.B > 0.f))
        r.A = -1.f;
    return r;
}

// per-Gaussian outputs of the preprocess; index o = view * n + i
// rects, depths and ids may be null (not written): the batched path reads
// neither extents nor depths, and ids are not needed when the binning's
// compaction numbers the entries itself
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
    uint2* bin_rects; // (first tile, rows << 16 | columns) for the binning
    ReachRec* reach;  // tight culling: written instead of bin_rects (null otherwise)
};

// host side of the splatter preprocess, shared with the rasteriser
SplatParams make_splat_params(const dgmr_gaussians* g);
hipError_t launch_splat_preprocess(const dgmr_gaussians* g, const DevCamera<float>& cam, const SplatOut& out, hipStream_t stream);

// SH coefficients of a block of Gaussians staged in LDS: row r of the block
// is Gaussian ids[r], whose [K, 3] coefficients are contiguous and read by
// consecutive lanes, eight loads in flight per thread (the rows may be
// scattered over the scene).  Each row sits in LDS at a stride of 3K + 1
// floats (odd, so the 64 lanes of a wave reading their own rows hit distinct
// LDS banks).  The block has 256 threads.
constexpr int kShStride = 49;
constexpr int kStageInFlight = 8;

__device__ __forceinline__ void stage_sh(float* s_sh, const float* sh, const uint32_t* ids, int count, int sh_coeffs)
{
    const int row = sh_coeffs * 3;
    const int total = count * row;
    for (int base = 0; base < total; base += kStageInFlight * 256) {
        float v[kStageInFlight];
#pragma unroll
        for (int u = 0; u < kStageInFlight; ++u) {
            const int e = base + u * 256 + int(threadIdx.x);
            if (e < total) {
                const int r = e / row;
                v[u] = sh[int64_t(ids[r]) * row + (e - r * row)];
            }
        }
#pragma unroll
        for (int u = 0; u < kStageInFlight; ++u) {
            const int e = base + u * 256 + int(threadIdx.x);
            if (e < total) {
                const int r = e / row;
                s_sh[r * kShStride + (e - r * row)] = v[u];
            }
        }
    }
}

// the staged rows written back to Gaussians ids[0..count): each row's [K, 3]
// coefficients are contiguous, written by consecutive lanes
__device__ __forceinline__ void unstage_sh(float* dst, const float* s_sh, const uint32_t* __restrict__ ids, int count, int sh_coeffs)
{
    const int row = sh_coeffs * 3;
    const int total = count * row;
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
        const int r = e / row, k = e - r * row;
        dst[int64_t(ids[r]) * row + k] = s_sh[r * kShStride + k];
    }
}

// preprocess of Gaussian i for one camera (splat_forward.cu:71-118): cull
// view-z < 0.2, EWA splat (Ots, 0.3 px^2 filter), rect ceil(3 sigma), tiles
// touched, depth, SH colour, conic.  Writes slot o.  c: this Gaussian's SH row.
// A Gaussian's SH row of coefficients, element f = 3k + channel for the
// first kmax coefficients.  A row in global memory (ROW_GLOBAL) is read in
// one burst of loads issued before any is used: a loop that loads and
// accumulates a coefficient at a time waits for each load in turn, 16 round
// trips to memory per Gaussian.  A row staged in LDS is read where it is
// used, which keeps the 48 values out of the registers.
template <bool ROW_GLOBAL>
struct ShRow {
    float v[ROW_GLOBAL ? 48 : 1];
    const float* c;
    __device__ __forceinline__ ShRow(const float* row, int kmax) : c(row)
    {
        if constexpr (ROW_GLOBAL) {
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const bool on = k < kmax;
#pragma unroll
                for (int ch = 0; ch < 3; ++ch)
                    v[3 * k + ch] = on ? row[3 * k + ch] : 0.f;
            }
        }
    }
    __device__ __forceinline__ float operator[](int f) const
    {
        if constexpr (ROW_GLOBAL)
            return v[f];
        else
            return c[f];
    }
};

// a Gaussian's view-independent preprocess inputs: centre, activated scales,
// rotation and opacity, 3D covariance and its two largest scales
struct GaussIn {
    float mu[3], s[3], q[4], w, cov3[6], sa, sb;
};

__device__ __forceinline__ void gauss_mean(const SplatParams& p, int64_t i, GaussIn& g)
{
    g.mu[0] = p.means[3 * i];
    g.mu[1] = p.means[3 * i + 1];
    g.mu[2] = p.means[3 * i + 2];
}

__device__ __forceinline__ void gauss_rest(const SplatParams& p, int64_t i, GaussIn& g)
{
    gaussian_params(p, i, g.s, g.q, &g.w);
    compute_cov(g.s, g.q, g.cov3);
    int i0, i1;
    larger2_values(g.s, &g.sa, &g.sb, &i0, &i1);
}

template <bool REST_READY, bool ROW_GLOBAL>
__device__ __forceinline__ void preprocess_view(GaussIn& g, int64_t i, int64_t o, const SplatParams& p, const float* c, const DevCamera<float>& cam,
    const SplatOut& out);

// one view: g.mu set, the rest computed here after the view-space cull; c in
// global memory
__device__ __forceinline__ void preprocess_gaussian(int64_t i, int64_t o, const SplatParams& p, const float* c, const DevCamera<float>& cam,
    const SplatOut& out)
{
    GaussIn g;
    gauss_mean(p, i, g);
    preprocess_view<false, true>(g, i, o, p, c, cam, out);
}

// REST_READY: g fully set (a Gaussian seen from many views fills it once);
// ROW_GLOBAL: c is in global memory, not staged in LDS
template <bool REST_READY, bool ROW_GLOBAL>
__device__ __forceinline__ void preprocess_view(GaussIn& g, int64_t i, int64_t o, const SplatParams& p, const float* c, const DevCamera<float>& cam,
    const SplatOut& out)
{
    if (out.ids)
        out.ids[o] = uint32_t(o);
    // a culled entry: no tiles, centre 0, the culled depth key (each output is
    // written once, here or on the visible path)
    const auto culled = [&]() {
        out.tiles[o] = 0;
        out.xy[2 * o] = 0.f;
        out.xy[2 * o + 1] = 0.f;
        out.depth_keys[o] = 0xffffffffu;
        if (!out.reach) // reach records are read only for entries with a depth key
            out.bin_rects[o] = make_uint2(0u, 0u);
    };
    const float* mu = g.mu;
    float t[3];
    transform_point(cam.view, mu, t, 3);
    if (t[2] < 0.2f) {
        if (out.rects) {
            out.rects[2 * o] = 0;
            out.rects[2 * o + 1] = 0;
        }
        culled();
        return;
    }
    if (!REST_READY)
        gauss_rest(p, i, g);
    const float w = g.w;
    SplatIn<float> in;
    splat_projection(mu, cam, &in);
#pragma unroll
    for (int k = 0; k < 6; ++k)
        in.cov3[k] = g.cov3[k];
    affine_cut(in.cov3, in.T, in.cov2);
    const float cov_f[3] = { in.cov2[0] + 0.3f, in.cov2[1], in.cov2[2] + 0.3f };
    float xy[2];
    project_to_screen(mu, cam, xy);
    const int ext_x = guarded_int(ceilf(3.f * sqrt_rn(cov_f[0])));
    const int ext_y = guarded_int(ceilf(3.f * sqrt_rn(cov_f[2])));
    if (out.rects) {
        out.rects[2 * o] = ext_x;
        out.rects[2 * o + 1] = ext_y;
    }
    int x0, y0, x1, y1;
    get_rect(xy[0], xy[1], ext_x, ext_y, cam.grid_x, cam.grid_y, cam.tile_w, cam.tile_h, &x0, &y0, &x1, &y1);
    const uint32_t touched = uint32_t(x1 - x0) * uint32_t(y1 - y0);
    if (touched == 0) {
        culled(); // the extents stay in rects, as the reference leaves them
        return;
    }
    out.tiles[o] = touched;
    if (!out.reach)
        out.bin_rects[o] = make_uint2(uint32_t(y0) * uint32_t(cam.grid_x) + uint32_t(x0), (uint32_t(y1 - y0) << 16) | uint32_t(x1 - x0));
    out.xy[2 * o] = xy[0];
    out.xy[2 * o + 1] = xy[1];
    const float dv[3] = { cam.pos[0] - mu[0], cam.pos[1] - mu[1], cam.pos[2] - mu[2] };
    const float depth = length3(dv);
    if (out.depths)
        out.depths[o] = depth;
    out.depth_keys[o] = __float_as_uint(depth);
    // colour from SH along normalize(mu - cam) (math.h:233-271)
    // the colour, conic and weight feed only the blend: hardware reciprocal
    // and square root (an ulp); centre, depth and extents above, which the
    // binning depends on, stay correctly rounded
    const float v[3] = { -dv[0], -dv[1], -dv[2] };
    const float inv_len = fast_div(1.f, fast_sqrt(dot3(v, v)));
    const float dir[3] = { v[0] * inv_len, v[1] * inv_len, v[2] * inv_len };
    float Y[16];
    sh_basis(p.sh_degree, dir, Y);
    const int kmax = min(p.sh_coeffs, (p.sh_degree + 1) * (p.sh_degree + 1));
    const ShRow<ROW_GLOBAL> cf(c, kmax);
    float r0 = 0.f, r1 = 0.f, r2 = 0.f;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        if (k < kmax) {
            r0 = fmaf(Y[k], cf[3 * k + 0], r0);
            r1 = fmaf(Y[k], cf[3 * k + 1], r1);
            r2 = fmaf(Y[k], cf[3 * k + 2], r2);
        }
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
    const float i2prime = float(2 * kPi) * g.sa * g.sb;
    const float det2 = cov_f[0] * cov_f[2] - cov_f[1] * cov_f[1];
    const float inv_i2 = fast_div(1.f, float(2 * kPi) * fast_sqrt(det2));
    const float weight = w * det3(in.J) * i2prime * inv_i2;
    const float inv_det2 = fast_div(1.f, det2);
    const float4 co = make_float4(cov_f[2] * inv_det2, -cov_f[1] * inv_det2, cov_f[0] * inv_det2, weight);
    out.conic_opacity[o] = co;
    if (out.reach) {
        const ReachRec r = make_reach(x0, y0, x1 - x0, y1 - y0, cam.grid_x, xy[0], xy[1], co.x, co.y, co.z, co.w);
        uint4* d = reinterpret_cast<uint4*>(out.reach + o);
        d[0] = make_uint4(r.origin, r.dims, r.grid_x, __float_as_uint(r.mx));
        d[1] = make_uint4(__float_as_uint(r.my), __float_as_uint(r.A), __float_as_uint(r.B), __float_as_uint(r.C));
    }
}

// ------------------------------------------------------------------ blending
// One wave (64 lanes) blends one tile.  Lane l owns the NP = TW*TH/64 pixels
// l, l + 64, ... of the tile in row-major order: one pixel column, rows
// RS = 64/TW apart, so the x offset to a Gaussian is shared by the lane's
// pixels.  The tile's Gaussians go through LDS in batches of B (B/64 per
// lane); the next batch is fetched into registers while the current one is
// blended.  The conic is staged pre-scaled by -log2(e)/2 (diagonal) and
// -log2(e) (off-diagonal), so G = exp2(A dx^2 + B dx dy + C dy^2).
constexpr float kLog2e = 1.4426950408889634f;
constexpr int kBlendWaves = 4; // tile waves per workgroup

// conic and opacity co = (a, b, c, o) at screen centre xy (tile_reach.h)
__device__ __forceinline__ bool reaches_box(float2 xy, float4 co, float bx0, float bx1, float by0, float by1)
{
    return reaches_box(xy.x, xy.y, co.x, co.y, co.z, co.w, bx0, bx1, by0, by1);
}

template <int B>
struct BlendSmem {
    float4 geo[B];  // x, y, A, B
    float4 col[B];  // C, opacity, r, g
    float blue[B];
};

// Two-stage fetch of a tile's list: the ids of batch b+2 and the Gaussian data
// of batch b+1 are in flight while batch b is blended.  Every load is issued
// unconditionally at an index clamped to the list (the list is non-empty), so
// no value has to wait for a branch; slots past the end are never read.
template <int B>
struct BlendFetch {
    static constexpr int NS = B / 64;
    uint32_t id_next[NS]; // ids of the batch after the one whose data is in flight
    uint32_t id[NS];
    float2 xy[NS];
    float4 co[NS];
    float r[NS], g[NS], b[NS];

    __device__ __forceinline__ static uint32_t clamp_k(uint32_t base, int s, uint32_t last)
    {
        return min(base + uint32_t(lane_id() + 64 * s), last - 1);
    }

    __device__ __forceinline__ void load_ids(uint32_t base, uint32_t last, const uint32_t* __restrict__ point_list)
    {
#pragma unroll
        for (int s = 0; s < NS; ++s)
            id_next[s] = point_list[clamp_k(base, s, last)];
    }

    // data of the batch whose ids are in id_next
    __device__ __forceinline__ void load_data(const float2* __restrict__ xy_in, const float4* __restrict__ conic_opacity,
        const float* __restrict__ rgb)
    {
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const uint32_t i = id_next[s];
            id[s] = i;
            xy[s] = xy_in[i];
            co[s] = conic_opacity[i];
            r[s] = rgb[3 * i];
            g[s] = rgb[3 * i + 1];
            b[s] = rgb[3 * i + 2];
        }
    }

    // slot s of this lane into LDS slot i, conic pre-scaled
    __device__ __forceinline__ void stage_at(float4* geo, float4* col, float* blue, int s, int i) const
    {
        geo[i] = make_float4(xy[s].x, xy[s].y, (-0.5f * kLog2e) * co[s].x, -kLog2e * co[s].y);
        col[i] = make_float4((-0.5f * kLog2e) * co[s].z, co[s].w, r[s], g[s]);
        blue[i] = b[s];
    }

    // Stages, in list order, only the batch's Gaussians whose alpha can reach
    // 1/255 at a pixel of the tile box (reaches_box); the others would blend
    // nowhere in the tile, so skipping them changes no value.  gid (may be
    // null) receives the staged Gaussians' ids.  Returns the staged count.
    __device__ __forceinline__ int stage_reaching(float4* geo, float4* col, float* blue, uint32_t* gid, uint32_t base, uint32_t last,
        float bx0, float bx1, float by0, float by1) const
    {
        const uint64_t below = (uint64_t(1) << lane_id()) - 1;
        int cnt = 0;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const bool keep = base + lane_id() + 64u * s < last && reaches_box(xy[s], co[s], bx0, bx1, by0, by1);
            const uint64_t m = __builtin_amdgcn_ballot_w64(keep);
            if (keep) {
                const int i = cnt + __builtin_popcountll(m & below);
                stage_at(geo, col, blue, s, i);
                if (gid)
                    gid[i] = id[s];
            }
            cnt += __builtin_popcountll(m);
        }
        return cnt;
    }
};

// LDS of one wave is private to it: LDS operations of a wave execute in
// order, so a compiler barrier orders its staging writes before its reads
__device__ __forceinline__ void wave_lds_sync() { __builtin_amdgcn_wave_barrier(); }

// true when the predicate holds on some lane of the wave: the ballot of the
// condition itself, so the lane mask is used as it is (a ballot of an int
// re-materialises the mask through a VGPR first)
__device__ __forceinline__ bool wave_any(bool pred) { return __builtin_amdgcn_ballot_w64(pred) != 0; }

// lane masks of float comparisons (LLVM predicates), combined with scalar
// instructions where a wave-uniform decision needs them
enum : int { kCmpOGT = 2, kCmpOGE = 3, kCmpOLT = 4, kCmpUGE = 11 };
template <int PRED>
__device__ __forceinline__ uint64_t lanes_where(float a, float b) { return __builtin_amdgcn_fcmpf(a, b, PRED); }

// per-Gaussian terms shared by a lane's pixels (dx is common to them)
struct LaneTerms {
    float dx, adx2, bdx;
    __device__ __forceinline__ LaneTerms(float pxf, const float4& g)
    {
        dx = pxf - g.x;
        adx2 = g.z * dx * dx;
        bdx = g.w * dx;
    }
};

// G of one pixel: exp2 of the pre-scaled quadratic form; the forward and the
// backward evaluate it with this one sequence of operations
__device__ __forceinline__ float gauss_at(const LaneTerms& t, float c_scaled, float dy)
{
    return __builtin_amdgcn_exp2f(fmaf(fmaf(c_scaled, dy, t.bdx), dy, t.adx2));
}

template <int TW, int TH, int B>
__device__ __forceinline__ void blend_tile_forward(BlendSmem<B>& sm, uint32_t first, uint32_t last, int px0, int py0, int width, int height,
    float bg0, float bg1, float bg2, const uint32_t* __restrict__ point_list, const float2* __restrict__ xy,
    const float4* __restrict__ conic_opacity, const float* __restrict__ rgb, float* __restrict__ framebuffer, float* __restrict__ final_T)
{
    constexpr int NP = TW * TH / 64;
    constexpr int RS = 64 / TW;
    const int lane = int(lane_id());
    const int px = px0 + lane % TW;
    const float pxf = float(px);
    const float pyf0 = float(py0 + lane / TW);
    const float bx0 = float(px0), bx1 = float(px0 + TW - 1), by0 = float(py0), by1 = float(py0 + TH - 1);
    // a pixel that has stopped carries its transmittance negated (a live
    // pixel's T is at least 1e-4), so liveness costs no separate register
    float T[NP], C0[NP], C1[NP], C2[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const int py = py0 + lane / TW + RS * k;
        T[k] = (px < width && py < height) ? 1.f : -1.f;
        C0[k] = C1[k] = C2[k] = 0.f;
    }
    if (first < last) {
        BlendFetch<B> f;
        f.load_ids(first, last, point_list);
        f.load_data(xy, conic_opacity, rgb);
        f.load_ids(first + B, last, point_list);
        for (uint32_t base = first; base < last; base += B) {
            bool any = false;
#pragma unroll
            for (int k = 0; k < NP; ++k)
                any |= T[k] > 0.f;
            if (!wave_any(any))
                break;
            wave_lds_sync(); // the previous batch has been read
            const int cnt = f.stage_reaching(sm.geo, sm.col, sm.blue, nullptr, base, last, bx0, bx1, by0, by1);
            wave_lds_sync();
            // unconditional (clamped) so the loaded registers need no merge after a branch
            f.load_data(xy, conic_opacity, rgb);
            f.load_ids(base + 2 * B, last, point_list);
            for (int j = 0; j < cnt; ++j) {
                const float4 g = sm.geo[j];
                const float4 c = sm.col[j];
                const float blue = sm.blue[j];
                const LaneTerms lt(pxf, g);
                // branch-free: every pixel evaluates, selects apply the blend rule
                bool still = false;
#pragma unroll
                for (int k = 0; k < NP; ++k) {
                    const float alpha = fminf(0.99f, c.y * gauss_at(lt, c.x, pyf0 + float(RS * k) - g.y));
                    const float test_T = T[k] * (1.f - alpha);
                    const bool blend = T[k] > 0.f && alpha >= 1.0f / 255.0f;
                    const bool stop = blend && test_T < 0.0001f;
                    const bool use = blend && !stop;
                    // one select on the weight: fma(c, 0, C) is C exactly (colours are finite, >= 0)
                    const float wgt = use ? alpha * T[k] : 0.f;
                    C0[k] = fmaf(c.z, wgt, C0[k]);
                    C1[k] = fmaf(c.w, wgt, C1[k]);
                    C2[k] = fmaf(blue, wgt, C2[k]);
                    T[k] = use ? test_T : (stop ? -T[k] : T[k]);
                    still |= T[k] > 0.f;
                }
                if (!wave_any(still))
                    break;
            }
        }
    }
    const int64_t plane = int64_t(width) * height;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const int py = py0 + lane / TW + RS * k;
        if (px < width && py < height) {
            const int64_t p = int64_t(py) * width + px;
            const float Tk = fabsf(T[k]);
            framebuffer[p] = fmaf(Tk, bg0, C0[k]);
            framebuffer[plane + p] = fmaf(Tk, bg1, C1[k]);
            framebuffer[2 * plane + p] = fmaf(Tk, bg2, C2[k]);
            final_T[p] = Tk;
        }
    }
}

// waves per tile in the plain blend backward
#ifndef DGMR_BWD_SPLIT
#define DGMR_BWD_SPLIT 1
#endif
constexpr int kBwdSplit = DGMR_BWD_SPLIT;

// Issue priority of a backward blend wave from its tile's list length
// (thresholds t1 < t2 < t3, set from the frame's mean list length): waves on
// longer lists issue first on their SIMD, so they finish with the rest of the
// launch instead of running on alone at its end.  Results do not depend on it.
struct WavePriority {
    uint32_t t1, t2, t3;
};

__device__ __forceinline__ void set_wave_priority(uint32_t len, const WavePriority& p)
{
    if (len > p.t3)
        __builtin_amdgcn_s_setprio(3);
    else if (len > p.t2)
        __builtin_amdgcn_s_setprio(2);
    else if (len > p.t1)
        __builtin_amdgcn_s_setprio(1);
}

template <int B>
struct BlendSmemBwd {
    float4 geo[B];   // x, y, A, B
    float4 col[B];   // C, opacity, r, g
    float blue[B];
    // per-batch gradient sums of the tile, Gaussian-major: acc[9 i + q] is
    // value q (xy 2, conic 3, opacity 1, rgb 3) of batch Gaussian i, so the
    // nine lanes storing one Gaussian's sums and the flush's lanes both touch
    // consecutive words (value-major rows 64 words apart put a Gaussian's
    // nine values in one bank)
    float acc[9 * B];
    uint32_t gid[B]; // the batch's Gaussian ids, for the flush
};

// Blend adjoint of one tile, front to back.  The remaining colour is carried
// unnormalised (C_final minus what has been composited so far), which is the
// reference prototype's remaining colour (grad_inria_splatter_render.cu:100-156)
// times the current transmittance, without its repeated division by (1 - alpha).
// Per Gaussian, a lane sums over its pixels the moments sum(gp), sum(gp dy),
// sum(gp dy^2) of dL/dpower and turns them into the conic and position
// gradients with its shared dx; the nine per-Gaussian sums (xy 2, conic 3,
// opacity 1, rgb 3) are reduced over the wave, kept per batch in LDS and leave
// the tile as one global atomic per nonzero value and Gaussian.
template <int TW, int TH, int B>
__device__ __forceinline__ void blend_tile_backward(BlendSmemBwd<B>& sm, uint32_t first, uint32_t last, int px0, int py0, int width,
    int height, float bg0, float bg1, float bg2, const uint32_t* __restrict__ point_list, const float2* __restrict__ xy,
    const float4* __restrict__ conic_opacity, const float* __restrict__ rgb, const float* __restrict__ framebuffer,
    const float* __restrict__ final_T, const float* __restrict__ grad_fb, float* __restrict__ screen_grads,
    const float* __restrict__ l1_target = nullptr, float l1_scale = 0.f)
{
    constexpr int NP = TW * TH / 64;
    constexpr int RS = 64 / TW;
    constexpr int NS = B / 64;
    // d(power)/d(conic) in the pre-scaled units: conic = scaled * kUnscale
    constexpr float kUnscaleDiag = -2.f / kLog2e, kUnscaleOff = -1.f / kLog2e;
    const int lane = int(lane_id());
    const int64_t plane = int64_t(width) * height;
    const int px = px0 + lane % TW;
    const float pxf = float(px);
    const float pyf0 = float(py0 + lane / TW);
    const float bx0 = float(px0), bx1 = float(px0 + TW - 1), by0 = float(py0), by1 = float(py0 + TH - 1);
    // D = <dL/dC, S> with S the framebuffer minus what has been composited so
    // far, i.e. the remaining colour plus T_final * background: D carries the
    // remaining-colour term and dL/dT_final * T_final of the alpha adjoint in
    // one scalar per pixel
    // T[k] = 0 marks a pixel that takes no further part (outside the image,
    // no gradient, or stopped); a live pixel's T never falls below 1e-4
    float gC0[NP], gC1[NP], gC2[NP], D[NP], T[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const int py = py0 + lane / TW + RS * k;
        gC0[k] = gC1[k] = gC2[k] = 0.f;
        D[k] = 0.f;
        const bool inside = px < width && py < height;
        if (inside) {
            const int64_t p = int64_t(py) * width + px;
            const float f0 = framebuffer[p], f1 = framebuffer[plane + p], f2 = framebuffer[2 * plane + p];
            if (l1_target) {
                // fused L1 loss: dL/dC = sign(C - target) * scale, as the L1 kernel writes it
                gC0[k] = l1_sign(f0 - l1_target[p]) * l1_scale;
                gC1[k] = l1_sign(f1 - l1_target[plane + p]) * l1_scale;
                gC2[k] = l1_sign(f2 - l1_target[2 * plane + p]) * l1_scale;
            } else {
                gC0[k] = grad_fb[p];
                gC1[k] = grad_fb[plane + p];
                gC2[k] = grad_fb[2 * plane + p];
            }
            D[k] = gC0[k] * f0 + gC1[k] * f1 + gC2[k] * f2;
        }
        T[k] = (inside && !(gC0[k] == 0.f && gC1[k] == 0.f && gC2[k] == 0.f)) ? 1.f : 0.f;
    }
    if (first >= last)
        return;
    BlendFetch<B> f;
    f.load_ids(first, last, point_list);
    f.load_data(xy, conic_opacity, rgb);
    f.load_ids(first + B, last, point_list);
    // lane 8m stores wave_sum8's total of value kSum8Order[m]; lane 63 the ninth
    const int slot8 = (lane & 7) == 0 ? kSum8Order[lane >> 3] : (lane == 63 ? 8 : -1);
    // a batch's sums leave for global memory at the top of the next batch,
    // after its data has arrived, so no wait on the loads covers the atomics
    int cnt = 0;
    // The batch's sums leave as float atomics in Gaussian-major order: lane l
    // of flush step t adds value q = f % 9 of Gaussian f / 9, f = 64 t + l, so
    // one instruction touches about seven Gaussians' 36-byte rows and lanes
    // on the same cache line travel together (value-major order would send
    // every lane of an instruction to a different line).
    const auto flush = [&]() {
        const int total = cnt * 9;
        for (int f = lane; f < total; f += 64) {
            const int i = f / 9, q = f - 9 * i;
            const float v = sm.acc[f];
            if (v != 0.f)
                unsafeAtomicAdd(screen_grads + int64_t(sm.gid[i]) * 9 + q, v);
        }
    };
    for (uint32_t base = first; base < last; base += B) {
        bool any = false;
#pragma unroll
        for (int k = 0; k < NP; ++k)
            any |= T[k] > 0.f;
        if (!wave_any(any))
            break;
        wave_lds_sync(); // the previous batch has been read
        flush();
        wave_lds_sync(); // the flush has read the previous batch's ids and sums
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const int i = lane + 64 * s;
#pragma unroll
            for (int q = 0; q < 9; ++q)
                sm.acc[q * B + i] = 0.f; // 9 B consecutive words over the lanes and steps
        }
        cnt = f.stage_reaching(sm.geo, sm.col, sm.blue, sm.gid, base, last, bx0, bx1, by0, by1);
        wave_lds_sync();
        // unconditional (clamped) so the loaded registers need no merge after a branch
        f.load_data(xy, conic_opacity, rgb);
        f.load_ids(base + 2 * B, last, point_list);
        for (int j = 0; j < cnt; ++j) {
            const float4 g = sm.geo[j];
            const float4 c = sm.col[j];
            const float blue = sm.blue[j];
            const LaneTerms lt(pxf, g);
            // cheap part for every pixel: the blend decisions
            float Gk[NP], alphak[NP];
            bool use[NP], stop[NP];
            uint64_t use_lanes[NP]; // use[k] as a lane mask, for the wave-uniform tests
            uint64_t any_use = 0;
#pragma unroll
            for (int k = 0; k < NP; ++k) {
                Gk[k] = gauss_at(lt, c.x, pyf0 + float(RS * k) - g.y);
                alphak[k] = fminf(0.99f, c.y * Gk[k]);
                const float test_T = T[k] * (1.f - alphak[k]);
                const bool blend = T[k] > 0.f && alphak[k] >= 1.0f / 255.0f;
                stop[k] = blend && test_T < 0.0001f;
                use[k] = blend && !stop[k];
                use_lanes[k] = lanes_where<kCmpOGT>(T[k], 0.f) & lanes_where<kCmpOGE>(alphak[k], 1.0f / 255.0f)
                    & lanes_where<kCmpUGE>(test_T, 0.0001f);
                any_use |= use_lanes[k];
            }
            if (any_use != 0) {
                float m0 = 0.f, m1 = 0.f, m2 = 0.f, so = 0.f, sr = 0.f, sg = 0.f, sbl = 0.f;
                // selects keep the sums exact for pixels that do not blend this Gaussian
#pragma unroll
                for (int k = 0; k < NP; ++k) {
                    if (use_lanes[k] == 0)
                        continue; // no pixel of this row group blends the Gaussian
                    // the same operations as the cheap part, so the same values
                    const float dy = pyf0 + float(RS * k) - g.y;
                    const float raw = c.y * Gk[k];
                    const float tk = 1.f - alphak[k];
                    const float test_T = T[k] * tk;
                    // one select on the weight: fma(x, 0, s) is s exactly for finite x
                    const float wgt = use[k] ? alphak[k] * T[k] : 0.f;
                    // <dL/dC, rgb> and dL/dalpha as fused multiply-adds (gradient
                    // values only: no blend decision reads them)
                    const float grgb = fmaf(gC2[k], blue, fmaf(gC1[k], c.w, gC0[k] * c.z));
                    D[k] = fmaf(-grgb, wgt, D[k]);
                    sr = fmaf(gC0[k], wgt, sr);
                    sg = fmaf(gC1[k], wgt, sg);
                    sbl = fmaf(gC2[k], wgt, sbl);
                    // 1 - alpha >= 0.01, so the hardware reciprocal (1 ulp) is ample
                    const float inv_tk = __builtin_amdgcn_rcpf(tk);
                    const float g_alpha = fmaf(grgb, T[k], -(D[k] * inv_tk));
                    const float g_raw = (use[k] && raw < 0.99f) ? g_alpha : 0.f;
                    so = fmaf(Gk[k], g_raw, so);
                    const float gp = raw * g_raw; // dL/dpower
                    m0 += gp;
                    m1 = fmaf(gp, dy, m1);
                    m2 = fmaf(gp * dy, dy, m2);
                    T[k] = use[k] ? test_T : T[k];
                }
                // power = -a/2 dx^2 - b dx dy - c/2 dy^2 with dx = px - x, dy = py - y
                const float a = g.z * kUnscaleDiag, b = g.w * kUnscaleOff, cc = c.x * kUnscaleDiag;
                const float dx = lt.dx;
                const float ga = -0.5f * dx * dx * m0;
                const float gb = -dx * m1;
                const float gc = -0.5f * m2;
                const float gx = fmaf(a * dx, m0, b * m1);
                const float gy = fmaf(cc, m1, b * dx * m0);
                const float t8 = wave_sum8(gx, gy, ga, gb, gc, so, sr, sg);
                const float t9 = wave_sum_to_last(sbl);
                if (slot8 >= 0)
                    sm.acc[9 * j + slot8] = slot8 == 8 ? t9 : t8;
            }
            // a pixel stops before the Gaussian that would take T under 1e-4
            bool still = false;
#pragma unroll
            for (int k = 0; k < NP; ++k) {
                T[k] = stop[k] ? 0.f : T[k];
                still |= T[k] > 0.f;
            }
            if (!wave_any(still))
                break;
        }
    }
    wave_lds_sync();
    flush();
}

// Per-Gaussian adjoint of the preprocess for one camera: conic -> covariance
// (2x2 inverse adjoint), grad::splat<Ots> (grad/math.h:333-527),
// grad::sh_to_colour (grad/math.h:203-282) and normalize.  Adds into the
// accumulators (g_sh has 48 slots; only k < sh_coeffs are meaningful).
// c: this Gaussian's SH row.
// SH gradient (the outer product Y_k * dL/drgb_ch), by mode:
//   ShGrad::Add     adds it to g_sh (48 accumulators);
//   ShGrad::Export  writes its factors, g_sh[k] = Y_k (k < 16, zero beyond the
//                   active degree) and g_sh[16 + ch] = dL/drgb_ch, for a
//                   wave-cooperative coalesced write of the products.
enum class ShGrad { Add, Export };
// the per-Gaussian values (mean, scales times the multiplier, rotation,
// opacity) passed in, for callers that evaluate one Gaussian for many views
// One view's adjoint of the preprocess of a Gaussian (pre = ots_pre of its
// scales and rotation): adds into g_mu_acc, g_w_acc, the SH gradient, and the
// view sums g_cov3_acc, g_num_acc that preprocess_backward_finish turns into
// scale and rotation gradients once per Gaussian.
template <ShGrad MODE = ShGrad::Add, bool ROW_GLOBAL = false>
__device__ __forceinline__ void preprocess_backward_core(const float* mu, const OtsPre<float>& pre, float w, const SplatParams& p,
    const float* c, const DevCamera<float>& cam, const float4 A, const uint8_t* clamped3, const float* sg, float* g_mu_acc, float* g_cov3_acc,
    float* g_num_acc, float* g_w_acc, float* g_sh);

// scale (times the scale multiplier) and rotation gradients from the view sums
__device__ __forceinline__ void preprocess_backward_finish(const float* s, const float* q, const OtsPre<float>& pre, const SplatParams& p,
    const float* g_cov3, float g_num, float* g_s, float* g_q)
{
#pragma clang fp contract(fast)
    splat_ots_grad_finish(s, q, pre, g_cov3, g_num, g_s, g_q);
#pragma unroll
    for (int a = 0; a < 3; ++a)
        g_s[a] *= p.scale_multiplier;
}

template <ShGrad MODE = ShGrad::Add>
__device__ __forceinline__ void preprocess_backward_gaussian(int64_t i, const SplatParams& p, const float* c, const DevCamera<float>& cam,
    const float4 A, const uint8_t* clamped3, const float* sg, float* g_mu_acc, float* g_s_acc, float* g_q_acc, float* g_w_acc, float* g_sh)
{
    const float mu[3] = { p.means[3 * i], p.means[3 * i + 1], p.means[3 * i + 2] };
    float s[3], q[4], w;
    gaussian_params(p, i, s, q, &w);
    const OtsPre<float> pre = ots_pre(s, q);
    float g_cov3[6] = { 0.f, 0.f, 0.f, 0.f, 0.f, 0.f }, g_num = 0.f;
    preprocess_backward_core<MODE, false>(mu, pre, w, p, c, cam, A, clamped3, sg, g_mu_acc, g_cov3, &g_num, g_w_acc, g_sh);
    float g_s[3], g_q[4];
    preprocess_backward_finish(s, q, pre, p, g_cov3, g_num, g_s, g_q);
#pragma unroll
    for (int a = 0; a < 3; ++a)
        g_s_acc[a] += g_s[a];
#pragma unroll
    for (int a = 0; a < 4; ++a)
        g_q_acc[a] += g_q[a];
}

template <ShGrad MODE, bool ROW_GLOBAL>
__device__ __forceinline__ void preprocess_backward_core(const float* mu, const OtsPre<float>& pre, float w, const SplatParams& p,
    const float* c, const DevCamera<float>& cam, const float4 A, const uint8_t* clamped_in, const float* sg_in, float* g_mu_acc, float* g_cov3_acc,
    float* g_num_acc, float* g_w_acc, float* g_sh)
{
#pragma clang fp contract(fast)
    // the inputs are read into registers before any arithmetic (and an SH
    // row in global memory before the basis is evaluated): read where they
    // are used, each load would be waited for in turn
    float sg[9];
#pragma unroll
    for (int a = 0; a < 9; ++a)
        sg[a] = sg_in[a];
    const bool clamped3[3] = { clamped_in[0] != 0, clamped_in[1] != 0, clamped_in[2] != 0 };
    // conic = cov^-1: dL/dcov = -A G A
    const float G00 = sg[2], G01 = 0.5f * sg[3], G11 = sg[4];
    const float t00 = A.x * G00 + A.y * G01, t01 = A.x * G01 + A.y * G11;
    const float t10 = A.y * G00 + A.z * G01, t11 = A.y * G01 + A.z * G11;
    const float r00 = t00 * A.x + t01 * A.y, r01 = t00 * A.y + t01 * A.z;
    const float r10 = t10 * A.x + t11 * A.y, r11 = t10 * A.y + t11 * A.z;
    const float g_cov2[3] = { -r00, -(r01 + r10), -r11 };
    float g_w, g_mu[3];
    splat_ots_grad_view(w, mu, pre, cam, 0.3f, sg[5], sg, g_cov2, &g_w, g_mu, g_cov3_acc, g_num_acc);
    const float v[3] = { mu[0] - cam.pos[0], mu[1] - cam.pos[1], mu[2] - cam.pos[2] };
    const float inv_len = fast_div(1.f, fast_sqrt(dot3(v, v)));
    const float dir[3] = { v[0] * inv_len, v[1] * inv_len, v[2] * inv_len };
    const float gr[3] = { clamped3[0] ? sg[6] : 0.f, clamped3[1] ? sg[7] : 0.f, clamped3[2] ? sg[8] : 0.f };
    const int kmax = min(p.sh_coeffs, (p.sh_degree + 1) * (p.sh_degree + 1));
    const ShRow<ROW_GLOBAL> cf(c, kmax);
    float Y[16];
    sh_basis(p.sh_degree, dir, Y);
    // per coefficient <c_k, dL/drgb>, the weight of dY_k/d dir
    float wdot[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        wdot[k] = 0.f;
        if (k < kmax) {
            const float c0 = cf[3 * k], c1 = cf[3 * k + 1], c2 = cf[3 * k + 2];
            if constexpr (MODE == ShGrad::Export) {
                g_sh[k] = Y[k];
            } else {
                g_sh[3 * k + 0] += Y[k] * gr[0];
                g_sh[3 * k + 1] += Y[k] * gr[1];
                g_sh[3 * k + 2] += Y[k] * gr[2];
            }
            wdot[k] = c0 * gr[0] + c1 * gr[1] + c2 * gr[2];
        }
    }
    float g_dir[3] = { 0.f, 0.f, 0.f };
    sh_dir_grad(p.sh_degree, dir, wdot, g_dir);
    if constexpr (MODE == ShGrad::Export) {
#pragma unroll
        for (int k = 0; k < 16; ++k)
            if (k >= kmax)
                g_sh[k] = 0.f; // beyond the active degree
        g_sh[16] = gr[0];
        g_sh[17] = gr[1];
        g_sh[18] = gr[2];
    }
    const float dd = dot3(g_dir, dir);
#pragma unroll
    for (int a = 0; a < 3; ++a)
        g_mu_acc[a] += g_mu[a] + (g_dir[a] - dd * dir[a]) * inv_len;
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
