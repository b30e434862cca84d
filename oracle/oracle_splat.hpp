// SPDX-License-Identifier: GPL-3.0-or-later
// dgmr oracle -- TEST INFRASTRUCTURE ONLY (see oracle_math.hpp).
//
// Plain-loop restatement of the tile splatter, src/dgmr/splat_forward.cu:30-338,
// plus the splatter backward the reference does not ship, composed from its
// adjoints: the blend adjoint of unittests/grad_inria_splatter_render.cu:67-156
// (min(0.99, o G) branch, adapted to the kernel's stop-before termination at
// splat_forward.cu:311-315) and grad::splat<Ots> (grad/math.h:333-527),
// grad::sh_to_colour (grad/math.h:203-282) and the 2x2 inverse adjoint.
#pragma once

#include <algorithm>
#ifdef _OPENMP
#include <parallel/algorithm>
#endif
#include <cstdint>
#include <vector>

#include "oracle_math.hpp"

namespace orc {

template <class S>
struct Scene {
    int64_t n = 0;
    const S* means = nullptr;     // [n,3]
    const S* scales = nullptr;    // [n,3]
    const S* rotations = nullptr; // [n,4] (w,x,y,z)
    const S* opacity = nullptr;   // [n]
    const S* sh = nullptr;        // [n,16,3]
    int sh_degree = 0;
    S scale_multiplier = 1;
};

// per-Gaussian output of the preprocess pass, splat_forward.cu:71-118
template <class S>
struct SplatGeometry {
    std::vector<int32_t> rect;      // [n,2] iso-extent in pixels
    std::vector<uint32_t> tiles;    // [n]   tiles touched (0 = culled)
    std::vector<S> xy;              // [n,2]
    std::vector<S> conic;           // [n,3] inverse filtered 2D covariance
    std::vector<S> weight2d;        // [n]   screen-space opacity
    std::vector<S> rgb;             // [n,3]
    std::vector<uint8_t> clamped;   // [n,3] 1 where the SH colour was >= 0
    std::vector<S> depth;           // [n]
    std::vector<float> key_depth;   // [n]   the float whose bits go into the sort key
};

// binning of splat_forward.cu:120-240: scan, duplicateWithKeys, stable radix sort, ranges
struct Binning {
    int grid_x = 0, grid_y = 0, tile_w = 16, tile_h = 16;
    std::vector<uint32_t> offsets;    // [n]   inclusive scan of tiles touched
    std::vector<uint64_t> keys;       // [R]   sorted keys (tile << 32 | depth bits)
    std::vector<uint32_t> point_list; // [R]   sorted Gaussian ids
    std::vector<uint32_t> ranges;     // [T,2] [first, last) per tile, zero for empty tiles
};

// `decide`, when given, supplies the discrete decisions (culling, rects, tiles,
// depth keys) from a float run, so a double evaluation bins exactly as float.
template <class S>
inline void splat_preprocess(const Scene<S>& sc, const Camera<S>& cam, int tile_w, int tile_h, SplatGeometry<S>* g,
    const SplatGeometry<float>* decide = nullptr)
{
    const int64_t n = sc.n;
    const int grid_x = (cam.width + tile_w - 1) / tile_w, grid_y = (cam.height + tile_h - 1) / tile_h;
    g->rect.assign(2 * n, 0);
    g->tiles.assign(n, 0);
    g->xy.assign(2 * n, 0);
    g->conic.assign(3 * n, 0);
    g->weight2d.assign(n, 0);
    g->rgb.assign(3 * n, 0);
    g->clamped.assign(3 * n, 0);
    g->depth.assign(n, 0);
    g->key_depth.assign(n, 0);
    // Gaussians are independent: the thread count does not change any value
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        const S* mu = sc.means + 3 * i;
        S t[3];
        transform_point(cam.view, mu, t, 3);
        if (decide ? decide->tiles[i] == 0 : t[2] < S(0.2))
            continue;
        const S s[3] = { sc.scales[3 * i] * sc.scale_multiplier, sc.scales[3 * i + 1] * sc.scale_multiplier, sc.scales[3 * i + 2] * sc.scale_multiplier };
        const Splat2d<S> sp = splat(Formulation::Ots, sc.opacity[i], mu, s, sc.rotations + 4 * i, cam, S(0.3));
        const S d[3] = { cam.pos[0] - mu[0], cam.pos[1] - mu[1], cam.pos[2] - mu[2] };
        g->depth[i] = length3(d);
        if (decide) {
            g->rect[2 * i] = decide->rect[2 * i];
            g->rect[2 * i + 1] = decide->rect[2 * i + 1];
            g->tiles[i] = decide->tiles[i];
            g->key_depth[i] = decide->key_depth[i];
        } else {
            const int ext_x = guarded_int(std::ceil(S(3) * std::sqrt(sp.cov[0])));
            const int ext_y = guarded_int(std::ceil(S(3) * std::sqrt(sp.cov[2])));
            g->rect[2 * i] = ext_x;
            g->rect[2 * i + 1] = ext_y;
            int x0, y0, x1, y1;
            get_rect(sp.xy[0], sp.xy[1], ext_x, ext_y, grid_x, grid_y, tile_w, tile_h, &x0, &y0, &x1, &y1);
            const uint32_t touched = uint32_t(x1 - x0) * uint32_t(y1 - y0);
            if (touched == 0)
                continue;
            g->tiles[i] = touched;
            g->key_depth[i] = float(g->depth[i]);
        }
        g->xy[2 * i] = sp.xy[0];
        g->xy[2 * i + 1] = sp.xy[1];
        const S v[3] = { mu[0] - cam.pos[0], mu[1] - cam.pos[1], mu[2] - cam.pos[2] };
        const S inv_len = S(1) / std::sqrt(dot3(v, v));
        const S dir[3] = { v[0] * inv_len, v[1] * inv_len, v[2] * inv_len };
        bool ok[3];
        sh_to_colour(sc.sh + 48 * i, sc.sh_degree, dir, &g->rgb[3 * i], ok);
        for (int c = 0; c < 3; ++c)
            g->clamped[3 * i + c] = ok[c];
        const S det = sp.cov[0] * sp.cov[2] - sp.cov[1] * sp.cov[1];
        g->conic[3 * i] = sp.cov[2] / det;
        g->conic[3 * i + 1] = -sp.cov[1] / det;
        g->conic[3 * i + 2] = sp.cov[0] / det;
        g->weight2d[i] = sp.weight;
    }
}

// Binning from integer rects, centroids and float depth keys.  The reference
// emits (tile << 32 | bits(depth)) per covered tile in Gaussian order and
// stable-sorts on the key; std::stable_sort reproduces that exactly.
template <class S>
inline void bin(const std::vector<int32_t>& rect, const std::vector<uint32_t>& tiles, const std::vector<S>& xy,
    const std::vector<float>& key_depth, int width, int height, int tile_w, int tile_h, Binning* b)
{
    const int64_t n = int64_t(tiles.size());
    b->tile_w = tile_w;
    b->tile_h = tile_h;
    b->grid_x = (width + tile_w - 1) / tile_w;
    b->grid_y = (height + tile_h - 1) / tile_h;
    b->offsets.assign(n, 0);
    uint64_t acc = 0;
    for (int64_t i = 0; i < n; ++i) {
        acc += tiles[i];
        b->offsets[i] = uint32_t(acc);
    }
    // emitted in Gaussian order, row-major within a rectangle; slot offsets
    // come from the scan, so the order does not depend on the thread count
    std::vector<std::pair<uint64_t, uint32_t>> kv(acc);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t i = 0; i < n; ++i) {
        if (tiles[i] == 0)
            continue;
        int x0, y0, x1, y1;
        get_rect(xy[2 * i], xy[2 * i + 1], rect[2 * i], rect[2 * i + 1], b->grid_x, b->grid_y, tile_w, tile_h, &x0, &y0, &x1, &y1);
        uint32_t bits;
        std::memcpy(&bits, &key_depth[i], 4);
        size_t k = b->offsets[i] - tiles[i];
        for (int y = y0; y < y1; ++y)
            for (int x = x0; x < x1; ++x)
                kv[k++] = { (uint64_t(uint32_t(y * b->grid_x + x)) << 32) | bits, uint32_t(i) };
    }
    // a stable sort has one result; the parallel one reaches it with all threads
    const auto by_key = [](const auto& a, const auto& c) { return a.first < c.first; };
#ifdef _OPENMP
    __gnu_parallel::stable_sort(kv.begin(), kv.end(), by_key);
#else
    std::stable_sort(kv.begin(), kv.end(), by_key);
#endif
    b->keys.resize(kv.size());
    b->point_list.resize(kv.size());
    for (size_t k = 0; k < kv.size(); ++k) {
        b->keys[k] = kv[k].first;
        b->point_list[k] = kv[k].second;
    }
    const size_t T = size_t(b->grid_x) * b->grid_y;
    b->ranges.assign(2 * T, 0);
    for (size_t k = 0; k < kv.size(); ++k) {
        const uint32_t tile = uint32_t(kv[k].first >> 32);
        if (k == 0 || uint32_t(kv[k - 1].first >> 32) != tile)
            b->ranges[2 * tile] = uint32_t(k);
        if (k + 1 == kv.size() || uint32_t(kv[k + 1].first >> 32) != tile)
            b->ranges[2 * tile + 1] = uint32_t(k + 1);
    }
}

// thresholds of the blend loop, splat_forward.cu:307-315
struct BlendThresholds {
    double rel_alpha = 2e-5;   // |alpha - 1/255| within this relative band
    double rel_clamp = 2e-5;   // |o G - 0.99|
    double rel_T = 1e-3;       // |T (1 - alpha) - 1e-4|
};

// Per-pixel front-to-back blend of splat_forward.cu:242-335 over the crop
// [x0, x1) x [y0, y1).  fb is planar [3,H,W], T_final [H,W].  `ambiguous` marks
// pixels where a float-vs-double difference could flip a discrete decision.
template <class S>
inline void splat_blend(const SplatGeometry<S>& g, const Binning& b, int width, int height, const S* bg,
    int x0, int y0, int x1, int y1, S* fb, S* T_final, uint8_t* ambiguous, uint64_t* n_evals = nullptr)
{
    const BlendThresholds th;
    uint64_t evals = 0;
    // pixels are independent: the thread count does not change any value
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : evals)
    for (int py = y0; py < y1; ++py) {
        for (int px = x0; px < x1; ++px) {
            const int tile = (py / b.tile_h) * b.grid_x + (px / b.tile_w);
            const uint32_t first = b.ranges[2 * tile], last = b.ranges[2 * tile + 1];
            S T = 1;
            S C[3] = { 0, 0, 0 };
            bool amb = false;
            for (uint32_t k = first; k < last; ++k) {
                const uint32_t id = b.point_list[k];
                const S dx = S(px) - g.xy[2 * id], dy = S(py) - g.xy[2 * id + 1];
                const S* c = &g.conic[3 * id];
                const S power = S(-0.5) * (c[0] * dx * dx + c[2] * dy * dy) - c[1] * dx * dy;
                const S G = std::exp(power);
                const S raw = g.weight2d[id] * G;
                const S alpha = std::min(S(0.99), raw);
                ++evals;
                if (std::abs(double(alpha) - 1.0 / 255.0) <= th.rel_alpha / 255.0)
                    amb = true;
                if (std::abs(double(raw) - 0.99) <= th.rel_clamp * 0.99)
                    amb = true;
                if (alpha < S(1.0f / 255.0f))
                    continue;
                const S test_T = T * (S(1) - alpha);
                if (std::abs(double(test_T) - 1e-4) <= th.rel_T * 1e-4)
                    amb = true;
                if (test_T < S(0.0001f))
                    break;
                for (int ch = 0; ch < 3; ++ch)
                    C[ch] += g.rgb[3 * id + ch] * alpha * T;
                T = test_T;
            }
            const size_t p = size_t(py) * width + px;
            for (int ch = 0; ch < 3; ++ch)
                fb[size_t(ch) * width * height + p] = C[ch] + T * bg[ch];
            T_final[p] = T;
            if (ambiguous)
                ambiguous[p] = amb ? 1 : 0;
        }
    }
    if (n_evals)
        *n_evals += evals;
}

// gradients of the screen-space quantities, accumulated over pixels
template <class S>
struct Splat2dGrads {
    std::vector<S> xy;      // [n,2]
    std::vector<S> conic;   // [n,3] stored-value gradients
    std::vector<S> weight;  // [n]
    std::vector<S> rgb;     // [n,3]
};

// Blend adjoint over the crop.  Front to back as the reference's prototype
// (grad_inria_splatter_render.cu:67-156, #else branch): the remaining colour is
// rebuilt by dividing by each transparency (clamped to 255), and the Gaussian
// that would bring T below 1e-4 is skipped as in the forward kernel.
template <class S>
inline void splat_blend_backward(const SplatGeometry<S>& g, const Binning& b, int width, int height, const S* bg,
    int x0, int y0, int x1, int y1, const S* fb, const S* T_final, const S* grad_fb, Splat2dGrads<S>* gr)
{
    const size_t n = g.tiles.size();
    gr->xy.assign(2 * n, 0);
    gr->conic.assign(3 * n, 0);
    gr->weight.assign(n, 0);
    gr->rgb.assign(3 * n, 0);
    const size_t plane = size_t(width) * height;
    for (int py = y0; py < y1; ++py) {
        for (int px = x0; px < x1; ++px) {
            const size_t p = size_t(py) * width + px;
            const S gC[3] = { grad_fb[p], grad_fb[plane + p], grad_fb[2 * plane + p] };
            if (gC[0] == 0 && gC[1] == 0 && gC[2] == 0)
                continue;
            const S Tf = T_final[p];
            S cur[3];
            for (int ch = 0; ch < 3; ++ch)
                cur[ch] = fb[ch * plane + p] - Tf * bg[ch];
            const S gT = (gC[0] * bg[0] + gC[1] * bg[1]) + gC[2] * bg[2];
            const int tile = (py / b.tile_h) * b.grid_x + (px / b.tile_w);
            S T = 1;
            for (uint32_t k = b.ranges[2 * tile]; k < b.ranges[2 * tile + 1]; ++k) {
                const uint32_t id = b.point_list[k];
                const S dx = S(px) - g.xy[2 * id], dy = S(py) - g.xy[2 * id + 1];
                const S* c = &g.conic[3 * id];
                const S power = S(-0.5) * (c[0] * dx * dx + c[2] * dy * dy) - c[1] * dx * dy;
                const S G = std::exp(power);
                const S o = g.weight2d[id];
                const S raw = o * G;
                const S alpha = std::min(S(0.99), raw);
                if (alpha < S(1.0f / 255.0f))
                    continue;
                const S tk = S(1) - alpha;
                if (T * tk < S(0.0001f))
                    break;
                const S inv_tk = std::min(S(255), S(1) / tk);
                const S* rgb = &g.rgb[3 * id];
                for (int ch = 0; ch < 3; ++ch)
                    cur[ch] = (cur[ch] - rgb[ch] * alpha) * inv_tk;
                const S g_tk = Tf * inv_tk * gT + ((gC[0] * cur[0] + gC[1] * cur[1]) + gC[2] * cur[2]) * T;
                const S g_alpha = ((gC[0] * rgb[0] + gC[1] * rgb[1]) + gC[2] * rgb[2]) * T - g_tk;
                for (int ch = 0; ch < 3; ++ch)
                    gr->rgb[3 * id + ch] += gC[ch] * T * alpha;
                const S g_raw = raw >= S(0.99) ? S(0) : g_alpha;
                const S g_G = o * g_raw;
                gr->weight[id] += G * g_raw;
                // power = -1/2 (a dx^2 + c dy^2) - b dx dy, d = pixel - centroid
                const S g_power = g_G * G;
                gr->conic[3 * id + 0] += g_power * (S(-0.5) * dx * dx);
                gr->conic[3 * id + 1] += g_power * (-dx * dy);
                gr->conic[3 * id + 2] += g_power * (S(-0.5) * dy * dy);
                gr->xy[2 * id + 0] += g_power * (c[0] * dx + c[1] * dy);
                gr->xy[2 * id + 1] += g_power * (c[2] * dy + c[1] * dx);
                T = T * tk;
            }
        }
    }
}

// The blend adjoint with image rows split over threads, each summing into its
// own gradient arrays, reduced at the end.  Same math as splat_blend_backward;
// the sums differ from it by float reassociation only.  Used for host timing.
template <class S>
inline void splat_blend_backward_threads(const SplatGeometry<S>& g, const Binning& b, int width, int height, const S* bg,
    int x0, int y0, int x1, int y1, const S* fb, const S* T_final, const S* grad_fb, Splat2dGrads<S>* gr, int threads)
{
    std::vector<Splat2dGrads<S>> part(threads);
#pragma omp parallel for num_threads(threads) schedule(static, 1)
    for (int t = 0; t < threads; ++t) {
        const int rows = y1 - y0;
        const int a = y0 + int(int64_t(rows) * t / threads), c = y0 + int(int64_t(rows) * (t + 1) / threads);
        splat_blend_backward(g, b, width, height, bg, x0, a, x1, c, fb, T_final, grad_fb, &part[t]);
    }
    *gr = std::move(part[0]);
    for (int t = 1; t < threads; ++t) {
        for (size_t k = 0; k < gr->xy.size(); ++k)
            gr->xy[k] += part[t].xy[k];
        for (size_t k = 0; k < gr->conic.size(); ++k)
            gr->conic[k] += part[t].conic[k];
        for (size_t k = 0; k < gr->weight.size(); ++k)
            gr->weight[k] += part[t].weight[k];
        for (size_t k = 0; k < gr->rgb.size(); ++k)
            gr->rgb[k] += part[t].rgb[k];
    }
}

template <class S>
struct ParamGrads {
    std::vector<S> means;     // [n,3]
    std::vector<S> scales;    // [n,3]
    std::vector<S> rotations; // [n,4]
    std::vector<S> opacity;   // [n]
    std::vector<S> sh;        // [n,16,3]
};

// per-Gaussian adjoint of the preprocess: conic -> cov2d, grad::splat<Ots>,
// grad::sh_to_colour, normalize.  Culled Gaussians get zero gradients.
template <class S>
inline void splat_preprocess_backward(const Scene<S>& sc, const Camera<S>& cam, const SplatGeometry<S>& g,
    const Splat2dGrads<S>& gr, ParamGrads<S>* out)
{
    const int64_t n = sc.n;
    out->means.assign(3 * n, 0);
    out->scales.assign(3 * n, 0);
    out->rotations.assign(4 * n, 0);
    out->opacity.assign(n, 0);
    out->sh.assign(48 * n, 0);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        if (g.tiles[i] == 0)
            continue;
        const S* mu = sc.means + 3 * i;
        const S m = sc.scale_multiplier;
        const S s[3] = { sc.scales[3 * i] * m, sc.scales[3 * i + 1] * m, sc.scales[3 * i + 2] * m };
        // conic = inverse(cov): dL/dcov = -A G A with A = conic
        const S* A = &g.conic[3 * i];
        const S* gA = &gr.conic[3 * i];
        const S G00 = gA[0], G01 = gA[1] / 2, G11 = gA[2];
        const S t00 = A[0] * G00 + A[1] * G01, t01 = A[0] * G01 + A[1] * G11;
        const S t10 = A[1] * G00 + A[2] * G01, t11 = A[1] * G01 + A[2] * G11;
        const S r00 = t00 * A[0] + t01 * A[1], r01 = t00 * A[1] + t01 * A[2];
        const S r10 = t10 * A[0] + t11 * A[1], r11 = t10 * A[1] + t11 * A[2];
        const S g_cov2[3] = { -r00, -(r01 + r10), -r11 };
        S g_w, g_mu[3], g_s[3], g_q[4];
        splat_grad(Formulation::Ots, sc.opacity[i], mu, s, sc.rotations + 4 * i, cam, S(0.3),
            gr.weight[i], &gr.xy[2 * i], g_cov2, &g_w, g_mu, g_s, g_q);
        // colour: rgb = SH(normalize(mu - cam))
        const S v[3] = { mu[0] - cam.pos[0], mu[1] - cam.pos[1], mu[2] - cam.pos[2] };
        const S len = std::sqrt(dot3(v, v));
        const S dir[3] = { v[0] / len, v[1] / len, v[2] / len };
        bool ok[3] = { g.clamped[3 * i] != 0, g.clamped[3 * i + 1] != 0, g.clamped[3 * i + 2] != 0 };
        S g_dir[3];
        sh_to_colour_grad(sc.sh + 48 * i, sc.sh_degree, dir, &gr.rgb[3 * i], ok, &out->sh[48 * i], g_dir);
        const S dd = dot3(g_dir, dir);
        for (int a = 0; a < 3; ++a)
            g_mu[a] += (g_dir[a] - dd * dir[a]) / len;
        for (int a = 0; a < 3; ++a) {
            out->means[3 * i + a] = g_mu[a];
            out->scales[3 * i + a] = g_s[a] * m;
        }
        for (int a = 0; a < 4; ++a)
            out->rotations[4 * i + a] = g_q[a];
        out->opacity[i] = g_w;
    }
}

} // namespace orc
