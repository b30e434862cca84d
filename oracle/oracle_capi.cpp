// SPDX-License-Identifier: GPL-3.0-or-later
// dgmr oracle -- TEST INFRASTRUCTURE ONLY.  extern "C" surface of the oracle
// for ctypes (tests/, __graft_entry__.smoke(), bench.py cpu_baseline).  Every
// array crosses as a plain pointer in the scalar type chosen at creation
// (float32 when dbl == 0, float64 otherwise).
#include <chrono>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <cstdint>
#include <cstdlib>
#include <memory>
#include <vector>

#include "oracle_march.hpp"
#include "oracle_raster.hpp"
#include "oracle_splat.hpp"

using namespace orc;

namespace {

template <class S>
struct SceneStore {
    std::vector<S> means, scales, rotations, opacity, sh;
    Scene<S> scene;
    Camera<S> cam;
    void load(int64_t n, const void* m, const void* s, const void* r, const void* o, const void* h, int deg, double mult,
        const void* view, const void* proj, const void* pos, double tanx, double tany, int W, int H, bool src_double)
    {
        const auto cp = [&](const void* src, size_t cnt, std::vector<S>* dst) {
            dst->resize(cnt);
            for (size_t i = 0; i < cnt; ++i)
                (*dst)[i] = src_double ? S(static_cast<const double*>(src)[i]) : S(static_cast<const float*>(src)[i]);
        };
        cp(m, 3 * n, &means);
        cp(s, 3 * n, &scales);
        cp(r, 4 * n, &rotations);
        cp(o, n, &opacity);
        cp(h, 48 * n, &sh);
        std::vector<S> v, p, c;
        cp(view, 16, &v);
        cp(proj, 16, &p);
        cp(pos, 3, &c);
        scene.n = n;
        scene.means = means.data();
        scene.scales = scales.data();
        scene.rotations = rotations.data();
        scene.opacity = opacity.data();
        scene.sh = sh.data();
        scene.sh_degree = deg;
        scene.scale_multiplier = S(mult);
        const S tx = src_double ? S(tanx) : S(float(tanx));
        const S ty = src_double ? S(tany) : S(float(tany));
        cam = make_camera<S>(v.data(), p.data(), c.data(), tx, ty, W, H);
    }
};

template <class S>
void store_out(const std::vector<S>& v, void* dst)
{
    if (dst)
        std::memcpy(dst, v.data(), v.size() * sizeof(S));
}

struct SplatHandle {
    bool dbl = false;
    int tile_w = 16, tile_h = 16;
    SceneStore<float> sf;
    SceneStore<double> sd;
    SplatGeometry<float> gf;
    SplatGeometry<double> gd;
    Binning bins;
};

struct MarchHandle {
    bool dbl = false;
    int tile_w = 16, tile_h = 16;
    SceneStore<float> sf;
    SceneStore<double> sd;
    MarchGeometry<float> gf;
    MarchGeometry<double> gd;
    Binning bins;
    double inv_proj[16];
};

template <class S>
void crop_guard(int W, int H, int* x0, int* y0, int* x1, int* y1)
{
    *x0 = std::max(0, *x0);
    *y0 = std::max(0, *y0);
    *x1 = std::min(W, *x1);
    *y1 = std::min(H, *y1);
}

} // namespace

extern "C" {

// ------------------------------------------------------------------- splatter
// dbl: evaluate in double; src_double: input arrays are double; float_binning:
// take culling/rects/tiles/keys from a float run (always true for dbl == 0).
void* orc_splat_create(int dbl, int src_double, int float_binning, int64_t n, const void* means, const void* scales,
    const void* rotations, const void* opacity, const void* sh16, int sh_degree, double scale_multiplier,
    const void* view, const void* proj, const void* cam_pos, double tan_fovx, double tan_fovy, int width, int height,
    int tile_w, int tile_h)
{
    auto* h = new SplatHandle;
    h->dbl = dbl != 0;
    h->tile_w = tile_w;
    h->tile_h = tile_h;
    h->sf.load(n, means, scales, rotations, opacity, sh16, sh_degree, scale_multiplier, view, proj, cam_pos, tan_fovx, tan_fovy, width, height, src_double != 0);
    splat_preprocess(h->sf.scene, h->sf.cam, tile_w, tile_h, &h->gf);
    if (h->dbl) {
        h->sd.load(n, means, scales, rotations, opacity, sh16, sh_degree, scale_multiplier, view, proj, cam_pos, tan_fovx, tan_fovy, width, height, src_double != 0);
        splat_preprocess(h->sd.scene, h->sd.cam, tile_w, tile_h, &h->gd, float_binning ? &h->gf : nullptr);
        if (float_binning)
            bin(h->gf.rect, h->gf.tiles, h->gf.xy, h->gf.key_depth, width, height, tile_w, tile_h, &h->bins);
        else
            bin(h->gd.rect, h->gd.tiles, h->gd.xy, h->gd.key_depth, width, height, tile_w, tile_h, &h->bins);
    } else {
        bin(h->gf.rect, h->gf.tiles, h->gf.xy, h->gf.key_depth, width, height, tile_w, tile_h, &h->bins);
    }
    return h;
}

void orc_splat_free(void* p) { delete static_cast<SplatHandle*>(p); }

int64_t orc_splat_num_pairs(void* p) { return int64_t(static_cast<SplatHandle*>(p)->bins.point_list.size()); }

void orc_splat_get_binning(void* p, uint32_t* offsets, uint64_t* keys, uint32_t* point_list, uint32_t* ranges)
{
    auto* h = static_cast<SplatHandle*>(p);
    store_out(h->bins.offsets, offsets);
    store_out(h->bins.keys, keys);
    store_out(h->bins.point_list, point_list);
    store_out(h->bins.ranges, ranges);
}

// arrays in the handle's scalar type
void orc_splat_get_geometry(void* p, int32_t* rect, uint32_t* tiles, void* xy, void* conic, void* weight2d, void* rgb,
    uint8_t* clamped, void* depth)
{
    auto* h = static_cast<SplatHandle*>(p);
    const auto go = [&](const auto& g) {
        store_out(g.rect, rect);
        store_out(g.tiles, tiles);
        store_out(g.xy, xy);
        store_out(g.conic, conic);
        store_out(g.weight2d, weight2d);
        store_out(g.rgb, rgb);
        store_out(g.clamped, clamped);
        store_out(g.depth, depth);
    };
    if (h->dbl)
        go(h->gd);
    else
        go(h->gf);
}

void orc_splat_render(void* p, const void* bg, int x0, int y0, int x1, int y1, void* fb, void* T_final, uint8_t* ambiguous,
    uint64_t* n_evals)
{
    auto* h = static_cast<SplatHandle*>(p);
    if (h->dbl) {
        crop_guard<double>(h->sd.cam.width, h->sd.cam.height, &x0, &y0, &x1, &y1);
        splat_blend(h->gd, h->bins, h->sd.cam.width, h->sd.cam.height, static_cast<const double*>(bg), x0, y0, x1, y1,
            static_cast<double*>(fb), static_cast<double*>(T_final), ambiguous, n_evals);
    } else {
        crop_guard<float>(h->sf.cam.width, h->sf.cam.height, &x0, &y0, &x1, &y1);
        splat_blend(h->gf, h->bins, h->sf.cam.width, h->sf.cam.height, static_cast<const float*>(bg), x0, y0, x1, y1,
            static_cast<float*>(fb), static_cast<float*>(T_final), ambiguous, n_evals);
    }
}

// Work profile of the blend (float restatement), per tile: `traversed` = list
// entries visited before every pixel of the tile has stopped (or the list
// ends), `blended` = entries that reach alpha >= 1/255 at some live pixel,
// `reachable` (may be null) = entries that reach alpha >= 1/255 at some pixel
// of the tile, ignoring termination.  Used to size launches; not a parity
// reference.
void orc_splat_tile_work(void* p, uint32_t* traversed, uint32_t* blended, uint32_t* reachable)
{
    auto* h = static_cast<SplatHandle*>(p);
    const auto& g = h->gf;
    const auto& b = h->bins;
    const int W = h->sf.cam.width, H = h->sf.cam.height;
    const int n_tiles = b.grid_x * b.grid_y;
#pragma omp parallel for schedule(dynamic, 4)
    for (int tile = 0; tile < n_tiles; ++tile) {
        const uint32_t first = b.ranges[2 * tile], last = b.ranges[2 * tile + 1];
        const int px0 = (tile % b.grid_x) * b.tile_w, py0 = (tile / b.grid_x) * b.tile_h;
        const int np = b.tile_w * b.tile_h;
        std::vector<float> T(np, 1.f);
        std::vector<uint8_t> live(np, 0);
        int n_live = 0;
        for (int q = 0; q < np; ++q) {
            const int px = px0 + q % b.tile_w, py = py0 + q / b.tile_w;
            live[q] = px < W && py < H;
            n_live += live[q];
        }
        uint32_t trav = 0, blend = 0;
        for (uint32_t k = first; k < last && n_live > 0; ++k) {
            ++trav;
            const uint32_t id = b.point_list[k];
            bool any = false;
            for (int q = 0; q < np; ++q) {
                if (!live[q])
                    continue;
                const float dx = float(px0 + q % b.tile_w) - g.xy[2 * id], dy = float(py0 + q / b.tile_w) - g.xy[2 * id + 1];
                const float* c = &g.conic[3 * id];
                const float power = -0.5f * (c[0] * dx * dx + c[2] * dy * dy) - c[1] * dx * dy;
                const float alpha = std::min(0.99f, g.weight2d[id] * std::exp(power));
                if (alpha < 1.0f / 255.0f)
                    continue;
                any = true;
                const float test_T = T[q] * (1.f - alpha);
                if (test_T < 0.0001f) {
                    live[q] = 0;
                    --n_live;
                    continue;
                }
                T[q] = test_T;
            }
            blend += any;
        }
        traversed[tile] = trav;
        blended[tile] = blend;
        if (reachable) {
            uint32_t reach = 0;
            for (uint32_t k = first; k < last; ++k) {
                const uint32_t id = b.point_list[k];
                bool any = false;
                for (int q = 0; q < np && !any; ++q) {
                    const int px = px0 + q % b.tile_w, py = py0 + q / b.tile_w;
                    if (px >= W || py >= H)
                        continue;
                    const float dx = float(px) - g.xy[2 * id], dy = float(py) - g.xy[2 * id + 1];
                    const float* c = &g.conic[3 * id];
                    const float power = -0.5f * (c[0] * dx * dx + c[2] * dy * dy) - c[1] * dx * dy;
                    any = std::min(0.99f, g.weight2d[id] * std::exp(power)) >= 1.0f / 255.0f;
                }
                reach += any;
            }
            reachable[tile] = reach;
        }
    }
}

// grads of the screen-space quantities (optional, may be null) and of the parameters
void orc_splat_backward(void* p, const void* bg, int x0, int y0, int x1, int y1, const void* fb, const void* T_final,
    const void* grad_fb, void* g_means, void* g_scales, void* g_rotations, void* g_opacity, void* g_sh,
    void* g2_xy, void* g2_conic, void* g2_weight, void* g2_rgb)
{
    auto* h = static_cast<SplatHandle*>(p);
    const auto go = [&](auto& st, auto& g, auto* tag) {
        using S = std::remove_pointer_t<decltype(tag)>;
        crop_guard<S>(st.cam.width, st.cam.height, &x0, &y0, &x1, &y1);
        Splat2dGrads<S> gr;
        splat_blend_backward(g, h->bins, st.cam.width, st.cam.height, static_cast<const S*>(bg), x0, y0, x1, y1,
            static_cast<const S*>(fb), static_cast<const S*>(T_final), static_cast<const S*>(grad_fb), &gr);
        ParamGrads<S> pg;
        splat_preprocess_backward(st.scene, st.cam, g, gr, &pg);
        store_out(pg.means, g_means);
        store_out(pg.scales, g_scales);
        store_out(pg.rotations, g_rotations);
        store_out(pg.opacity, g_opacity);
        store_out(pg.sh, g_sh);
        store_out(gr.xy, g2_xy);
        store_out(gr.conic, g2_conic);
        store_out(gr.weight, g2_weight);
        store_out(gr.rgb, g2_rgb);
    };
    if (h->dbl)
        go(h->sd, h->gd, static_cast<double*>(nullptr));
    else
        go(h->sf, h->gf, static_cast<float*>(nullptr));
}

// -------------------------------------------------------------------- marcher
void* orc_march_create(int dbl, int src_double, int float_binning, int64_t n, const void* means, const void* scales,
    const void* rotations, const void* opacity, const void* sh16, int sh_degree, double scale_multiplier,
    const void* view, const void* proj, const void* cam_pos, double tan_fovx, double tan_fovy, int width, int height,
    int tile_w, int tile_h)
{
    auto* h = new MarchHandle;
    h->dbl = dbl != 0;
    h->tile_w = tile_w;
    h->tile_h = tile_h;
    h->sf.load(n, means, scales, rotations, opacity, sh16, sh_degree, scale_multiplier, view, proj, cam_pos, tan_fovx, tan_fovy, width, height, src_double != 0);
    march_preprocess(h->sf.scene, h->sf.cam, tile_w, tile_h, &h->gf);
    double pd[16];
    if (h->dbl) {
        h->sd.load(n, means, scales, rotations, opacity, sh16, sh_degree, scale_multiplier, view, proj, cam_pos, tan_fovx, tan_fovy, width, height, src_double != 0);
        march_preprocess(h->sd.scene, h->sd.cam, tile_w, tile_h, &h->gd, float_binning ? &h->gf : nullptr);
        if (float_binning)
            bin(h->gf.rect, h->gf.tiles, h->gf.xy, h->gf.key_depth, width, height, tile_w, tile_h, &h->bins);
        else
            bin(h->gd.rect, h->gd.tiles, h->gd.xy, h->gd.key_depth, width, height, tile_w, tile_h, &h->bins);
        for (int i = 0; i < 16; ++i)
            pd[i] = h->sd.cam.proj[i];
    } else {
        bin(h->gf.rect, h->gf.tiles, h->gf.xy, h->gf.key_depth, width, height, tile_w, tile_h, &h->bins);
        for (int i = 0; i < 16; ++i)
            pd[i] = h->sf.cam.proj[i];
    }
    invert4(pd, h->inv_proj);
    return h;
}

void orc_march_free(void* p) { delete static_cast<MarchHandle*>(p); }

int64_t orc_march_num_pairs(void* p) { return int64_t(static_cast<MarchHandle*>(p)->bins.point_list.size()); }

void orc_march_get_binning(void* p, uint32_t* offsets, uint64_t* keys, uint32_t* point_list, uint32_t* ranges)
{
    auto* h = static_cast<MarchHandle*>(p);
    store_out(h->bins.offsets, offsets);
    store_out(h->bins.keys, keys);
    store_out(h->bins.point_list, point_list);
    store_out(h->bins.ranges, ranges);
}

void orc_march_get_geometry(void* p, int32_t* rect, uint32_t* tiles, void* xy, void* depth, void* rgb, uint8_t* clamped,
    void* inv_cov, void* mass)
{
    auto* h = static_cast<MarchHandle*>(p);
    const auto go = [&](const auto& g) {
        store_out(g.rect, rect);
        store_out(g.tiles, tiles);
        store_out(g.xy, xy);
        store_out(g.depth, depth);
        store_out(g.rgb, rgb);
        store_out(g.clamped, clamped);
        store_out(g.inv_cov, inv_cov);
        store_out(g.mass, mass);
    };
    if (h->dbl)
        go(h->gd);
    else
        go(h->gf);
}

void orc_march_render(void* p, const void* bg, int x0, int y0, int x1, int y1, int max_large_steps, int n_frozen,
    const double* frozen, int float_sampling, int render_mode, int debug_bin, double max_depth, void* fb, void* T_final, int32_t* n_large,
    uint8_t* ambiguous)
{
    auto* h = static_cast<MarchHandle*>(p);
    MarchOptions opt;
    opt.max_large_steps = max_large_steps;
    opt.n_frozen = n_frozen;
    opt.frozen = frozen;
    opt.mode = static_cast<RenderMode>(render_mode);
    opt.debug_bin = debug_bin;
    opt.max_depth = max_depth;
    float ipf[16];
    for (int i = 0; i < 16; ++i)
        ipf[i] = float(h->inv_proj[i]);
    const FloatSampling fsamp { &h->gf, &h->sf.scene, &h->sf.cam, ipf };
    const FloatSampling* fs = (h->dbl && float_sampling) ? &fsamp : nullptr;
    const auto go = [&](auto& st, auto& g, auto* tag) {
        using S = std::remove_pointer_t<decltype(tag)>;
        crop_guard<S>(st.cam.width, st.cam.height, &x0, &y0, &x1, &y1);
        S ip[16];
        for (int i = 0; i < 16; ++i)
            ip[i] = S(h->inv_proj[i]);
        const int W = st.cam.width, H = st.cam.height;
        const size_t plane = size_t(W) * H;
        const S* bgs = static_cast<const S*>(bg);
        S* out = static_cast<S*>(fb);
        S* Tout = static_cast<S*>(T_final);
        std::vector<int32_t> steps(plane, 0);
#pragma omp parallel for schedule(dynamic, 1)
        for (int py = y0; py < y1; ++py)
            for (int px = x0; px < x1; ++px) {
                S C[3], T;
                int nl;
                TieWatch watch;
                tie_watch = ambiguous ? &watch : nullptr;
                march_pixel(g, h->bins, st.scene, st.cam, ip, px, py, opt, C, &T, &nl, fs);
                tie_watch = nullptr;
                const size_t q = size_t(py) * W + px;
                if (ambiguous)
                    ambiguous[q] = uint8_t(watch.hit);
                for (int c = 0; c < 3; ++c)
                    out[c * plane + q] = C[c] + T * bgs[c];
                Tout[q] = T;
                steps[q] = nl;
                if (n_large)
                    n_large[q] = nl;
            }
        if (opt.mode == RenderMode::Depth) {
            // A pixel that has stopped keeps taking part in its tile's large
            // steps (vol_marcher_forward.cu:340-499 loops per block): pass 1
            // skips it, so its section list is empty, its end() +inf, and
            // the depth colour becomes +inf / max_depth; its bins stay
            // empty, so T is unchanged.  Tiles are whole in the crop only
            // when the crop is tile-aligned; steps outside the crop are not
            // known, so the rule is applied over the crop's part of a tile.
            const int tw = h->bins.tile_w, th = h->bins.tile_h;
            for (int ty = y0 / th; ty * th < y1; ++ty)
                for (int tx = x0 / tw; tx * tw < x1; ++tx) {
                    const int ya = std::max(y0, ty * th), yb = std::min(y1, (ty + 1) * th);
                    const int xa = std::max(x0, tx * tw), xb = std::min(x1, (tx + 1) * tw);
                    int most = 0;
                    for (int py = ya; py < yb; ++py)
                        for (int px = xa; px < xb; ++px)
                            most = std::max(most, steps[size_t(py) * W + px]);
                    const S far = std::numeric_limits<S>::infinity() / S(opt.max_depth);
                    for (int py = ya; py < yb; ++py)
                        for (int px = xa; px < xb; ++px) {
                            const size_t q = size_t(py) * W + px;
                            if (steps[q] < most)
                                for (int c = 0; c < 3; ++c)
                                    out[c * plane + q] = far + Tout[q] * bgs[c];
                        }
                }
        }
    };
    if (h->dbl)
        go(h->sd, h->gd, static_cast<double*>(nullptr));
    else
        go(h->sf, h->gf, static_cast<float*>(nullptr));
}

void orc_march_backward(void* p, const void* bg, int x0, int y0, int x1, int y1, int max_large_steps, int n_frozen,
    const double* frozen, int float_sampling, int sh_stored, const void* fb, const void* T_final, const void* grad_fb, void* g_means,
    void* g_scales, void* g_rotations, void* g_opacity, void* g_sh, void* r_mass, void* r_rgb, void* r_means, void* r_inv_cov)
{
    auto* h = static_cast<MarchHandle*>(p);
    MarchOptions opt;
    opt.max_large_steps = max_large_steps;
    opt.n_frozen = n_frozen;
    opt.frozen = frozen;
    float ipf[16];
    for (int i = 0; i < 16; ++i)
        ipf[i] = float(h->inv_proj[i]);
    const FloatSampling fsamp { &h->gf, &h->sf.scene, &h->sf.cam, ipf };
    const FloatSampling* fs = (h->dbl && float_sampling) ? &fsamp : nullptr;
    const auto go = [&](auto& st, auto& g, auto* tag) {
        using S = std::remove_pointer_t<decltype(tag)>;
        crop_guard<S>(st.cam.width, st.cam.height, &x0, &y0, &x1, &y1);
        S ip[16];
        for (int i = 0; i < 16; ++i)
            ip[i] = S(h->inv_proj[i]);
        const int W = st.cam.width, H = st.cam.height;
        const size_t plane = size_t(W) * H;
        const size_t n = size_t(st.scene.n);
        const S* fbs = static_cast<const S*>(fb);
        const S* Ts = static_cast<const S*>(T_final);
        const S* gs = static_cast<const S*>(grad_fb);
        const auto zeroed = [n](MarchRenderGrads<S>& a) {
            a.mass.assign(n, 0);
            a.rgb.assign(3 * n, 0);
            a.means.assign(3 * n, 0);
            a.inv_cov.assign(6 * n, 0);
        };
        // rows in fixed contiguous blocks per thread, each thread summing into
        // its own buffers, reduced in thread order: deterministic for a given
        // thread count
        const int n_threads = std::max(1, std::min(omp_get_max_threads(), y1 - y0));
        std::vector<MarchRenderGrads<S>> part(n_threads);
#pragma omp parallel for schedule(static, 1) num_threads(n_threads)
        for (int t = 0; t < n_threads; ++t) {
            zeroed(part[t]);
            const int r0 = y0 + int(int64_t(y1 - y0) * t / n_threads), r1 = y0 + int(int64_t(y1 - y0) * (t + 1) / n_threads);
            for (int py = r0; py < r1; ++py)
                for (int px = x0; px < x1; ++px) {
                    const size_t q = size_t(py) * W + px;
                    const S gC[3] = { gs[q], gs[plane + q], gs[2 * plane + q] };
                    if (gC[0] == 0 && gC[1] == 0 && gC[2] == 0)
                        continue;
                    const S fpx[3] = { fbs[q], fbs[plane + q], fbs[2 * plane + q] };
                    march_pixel_backward(g, h->bins, st.scene, st.cam, ip, px, py, opt, fpx, Ts[q], gC, static_cast<const S*>(bg),
                        &part[t], fs);
                }
        }
        MarchRenderGrads<S> gr = std::move(part[0]);
        for (int t = 1; t < n_threads; ++t) {
            for (size_t i = 0; i < n; ++i)
                gr.mass[i] += part[t].mass[i];
            for (size_t i = 0; i < 3 * n; ++i) {
                gr.rgb[i] += part[t].rgb[i];
                gr.means[i] += part[t].means[i];
            }
            for (size_t i = 0; i < 6 * n; ++i)
                gr.inv_cov[i] += part[t].inv_cov[i];
        }
        // The reference binding pads SH coefficients the caller did not pass
        // (sh_stored of 16) with ones for the backward, zeros for the forward
        // (python_bindings/vol_marcher.cu:92 vs :177); the render backward
        // replays with the forward's colours, only the preprocess backward
        // sees the ones.
        Scene<S> bsc = st.scene;
        std::vector<S> sh_ones;
        if (sh_stored > 0 && sh_stored < 16) {
            sh_ones.assign(st.scene.sh, st.scene.sh + 48 * n);
            for (size_t i = 0; i < n; ++i)
                for (int k = sh_stored; k < 16; ++k)
                    for (int c = 0; c < 3; ++c)
                        sh_ones[48 * i + 3 * k + c] = S(1);
            bsc.sh = sh_ones.data();
        }
        ParamGrads<S> pg;
        march_preprocess_backward(bsc, st.cam, g, gr, &pg);
        store_out(pg.means, g_means);
        store_out(pg.scales, g_scales);
        store_out(pg.rotations, g_rotations);
        store_out(pg.opacity, g_opacity);
        store_out(pg.sh, g_sh);
        store_out(gr.mass, r_mass);
        store_out(gr.rgb, r_rgb);
        store_out(gr.means, r_means);
        store_out(gr.inv_cov, r_inv_cov);
    };
    if (h->dbl)
        go(h->sd, h->gd, static_cast<double*>(nullptr));
    else
        go(h->sf, h->gf, static_cast<float*>(nullptr));
}

// ----------------------------------------------------- unit-level math (double)
void orc_compute_cov(const double* s, const double* q, double* cov) { compute_cov(s, q, cov); }
void orc_compute_cov_grad(const double* s, const double* q, const double* g, double* gs, double* gq) { compute_cov_grad(s, q, g, gs, gq); }

static Camera<double> cam_from(const double* view, const double* proj, const double* pos, double tanx, double tany, int W, int H)
{
    return make_camera<double>(view, proj, pos, tanx, tany, W, H);
}

// out: weight, x, y, cov[3]
void orc_splat2d(int formulation, double w, const double* mu, const double* s, const double* q, const double* view,
    const double* proj, const double* pos, double tanx, double tany, int W, int H, double filter, double* out)
{
    const auto cam = cam_from(view, proj, pos, tanx, tany, W, H);
    const auto g = splat(Formulation(formulation), w, mu, s, q, cam, filter);
    out[0] = g.weight;
    out[1] = g.xy[0];
    out[2] = g.xy[1];
    out[3] = g.cov[0];
    out[4] = g.cov[1];
    out[5] = g.cov[2];
}

// g_in: weight, x, y, cov[3]; out: g_w, g_mu[3], g_s[3], g_q[4]
void orc_splat2d_grad(int formulation, double w, const double* mu, const double* s, const double* q, const double* view,
    const double* proj, const double* pos, double tanx, double tany, int W, int H, double filter, const double* g_in, double* out)
{
    const auto cam = cam_from(view, proj, pos, tanx, tany, W, H);
    splat_grad(Formulation(formulation), w, mu, s, q, cam, filter, g_in[0], g_in + 1, g_in + 3, out, out + 1, out + 4, out + 7);
}

void orc_sh_to_colour(const double* sh, int deg, const double* dir, double* rgb, uint8_t* ok)
{
    bool b[3];
    sh_to_colour(sh, deg, dir, rgb, b);
    for (int i = 0; i < 3; ++i)
        ok[i] = b[i];
}

void orc_sh_to_colour_grad(const double* sh, int deg, const double* dir, const double* g_rgb, const uint8_t* ok, double* g_sh, double* g_dir)
{
    bool b[3] = { ok[0] != 0, ok[1] != 0, ok[2] != 0 };
    sh_to_colour_grad(sh, deg, dir, g_rgb, b, g_sh, g_dir);
}

void orc_intersect_ray(const double* mu, const double* A, const double* o, const double* d, double* out)
{
    const auto g = intersect_ray(mu, A, o, d);
    out[0] = g.weight;
    out[1] = g.centre;
    out[2] = g.var;
}

void orc_intersect_ray_grad(const double* mu, const double* A, const double* o, const double* d, const double* g, double* g_mu, double* g_A)
{
    intersect_ray_grad(mu, A, o, d, g[0], g[1], g[2], g_mu, g_A);
}

void orc_sample_gaussian(double mass, const double* rgb, const double* mu, const double* A, const double* o, const double* d,
    const double* borders, double* bins)
{
    sample_gaussian(mass, rgb, mu, A, o, d, borders, reinterpret_cast<double(*)[4]>(bins));
}

void orc_sample_gaussian_grad(double mass, const double* rgb, const double* mu, const double* A, const double* o, const double* d,
    const double* borders, const double* g_bins, double* g_mass, double* g_rgb, double* g_mu, double* g_A)
{
    sample_gaussian_grad(mass, rgb, mu, A, o, d, borders, reinterpret_cast<const double(*)[4]>(g_bins), g_mass, g_rgb, g_mu, g_A);
}

void orc_integrate_bins(double* C, double* T, const double* bins)
{
    integrate_bins(C, T, reinterpret_cast<const double(*)[4]>(bins));
}

void orc_integrate_bins_grad(double* C, double* T, double T_final, const double* bins, const double* gC, double gT, double* g_bins)
{
    integrate_bins_grad(C, T, T_final, reinterpret_cast<const double(*)[4]>(bins), gC, gT, reinterpret_cast<double(*)[4]>(g_bins));
}

float orc_portable_log_f(float x) { return portable_log(x); }
double orc_portable_log_d(double x) { return portable_log(x); }

void orc_get_rect_f(float px, float py, int ex, int ey, int gx, int gy, int tw, int th, int* out)
{
    get_rect(px, py, ex, ey, gx, gy, tw, th, out, out + 1, out + 2, out + 3);
}

uint32_t orc_higher_msb(uint32_t n) { return higher_msb(n); }

// section list driver: runs `n_puts` insertions into a list of capacity `cap`
// (4, 6, 8 or 16) starting at `start`; writes the final sections and end()
int orc_sections_run(int dbl, int cap, double start, int n_puts, const double* puts, double* out_sections, double* out_end)
{
    const auto go = [&](auto* tag, auto cap_tag) -> int {
        using S = std::remove_pointer_t<decltype(tag)>;
        constexpr unsigned CAP = decltype(cap_tag)::value;
        SectionList<S, CAP> sl{ static_cast<S>(start) };
        for (int i = 0; i < n_puts; ++i)
            sl.put({ S(puts[3 * i]), S(puts[3 * i + 1]), S(puts[3 * i + 2]) });
        for (unsigned i = 0; i < sl.size(); ++i) {
            out_sections[3 * i] = double(sl[i].start);
            out_sections[3 * i + 1] = double(sl[i].end);
            out_sections[3 * i + 2] = double(sl[i].dt);
        }
        *out_end = double(sl.end());
        return int(sl.size());
    };
    const auto dispatch = [&](auto* tag) -> int {
        switch (cap) {
        case 4: return go(tag, std::integral_constant<unsigned, 4>());
        case 6: return go(tag, std::integral_constant<unsigned, 6>());
        case 8: return go(tag, std::integral_constant<unsigned, 8>());
        case 16: return go(tag, std::integral_constant<unsigned, 16>());
        default: return -1;
        }
    };
    return dbl ? dispatch(static_cast<double*>(nullptr)) : dispatch(static_cast<float*>(nullptr));
}

void orc_sections_combine(int dbl, const double* a, const double* b, double* out)
{
    const auto go = [&](auto* tag) {
        using S = std::remove_pointer_t<decltype(tag)>;
        Section<S> p[3];
        combine_sections<S>({ S(a[0]), S(a[1]), S(a[2]) }, { S(b[0]), S(b[1]), S(b[2]) }, p);
        for (int i = 0; i < 3; ++i) {
            out[3 * i] = p[i].start;
            out[3 * i + 1] = p[i].end;
            out[3 * i + 2] = p[i].dt;
        }
    };
    if (dbl)
        go(static_cast<double*>(nullptr));
    else
        go(static_cast<float*>(nullptr));
}

// sample<n> over the list built by `n_puts` insertions (capacity cap)
int orc_sections_sample(int dbl, int cap, double start, int n_puts, const double* puts, int n_samples, double* out)
{
    const auto go = [&](auto* tag, auto cap_tag) {
        using S = std::remove_pointer_t<decltype(tag)>;
        constexpr unsigned CAP = decltype(cap_tag)::value;
        SectionList<S, CAP> sl{ static_cast<S>(start) };
        for (int i = 0; i < n_puts; ++i)
            sl.put({ S(puts[3 * i]), S(puts[3 * i + 1]), S(puts[3 * i + 2]) });
        std::vector<S> b(n_samples);
        sample_borders(sl, b.data(), unsigned(n_samples));
        for (int i = 0; i < n_samples; ++i)
            out[i] = double(b[i]);
        return 0;
    };
    const auto dispatch = [&](auto* tag) -> int {
        switch (cap) {
        case 4: return go(tag, std::integral_constant<unsigned, 4>());
        case 16: return go(tag, std::integral_constant<unsigned, 16>());
        default: return -1;
        }
    };
    return dbl ? dispatch(static_cast<double*>(nullptr)) : dispatch(static_cast<float*>(nullptr));
}

// The reference's seeded randomised section-list check (unittests/marching_steps.cpp:939-1003):
// std::srand(0), 1000 lists of capacity 8, 100 insertions each, every list validated and
// compared with the brute-force minimum step at 50 random positions.  Returns the
// number of failed checks.
int orc_sections_randomised_check(int dbl)
{
    const auto go = [&](auto* tag) -> int {
        using S = std::remove_pointer_t<decltype(tag)>;
        std::srand(0);
        const auto r = []() { return S(std::rand()) / S(RAND_MAX); };
        int failures = 0;
        for (unsigned i = 0; i < 1000; ++i) {
            const S smallest = r();
            SectionList<S, 8> arr(smallest);
            std::vector<Section<S>> entries;
            const auto real_dt = [&](S t) {
                S dt = S(1) / S(0);
                for (const auto& e : entries)
                    if (e.start <= t && t < e.end)
                        dt = std::min(dt, e.dt);
                return dt;
            };
            const auto computed_dt = [&](S t) {
                for (unsigned k = 0; k < arr.size(); ++k)
                    if (arr[k].start <= t && t < arr[k].end)
                        return arr[k].dt;
                return S(1) / S(0);
            };
            for (unsigned j = 0; j < 100; ++j) {
                const S start = r() * 10.0f;
                const S end = start + r();
                entries.push_back({ start, end, r() });
                arr.put(entries.back());
                if (arr.size()) {
                    failures += !(arr[0].start >= smallest);
                    failures += !(arr[0].start < arr[0].end);
                }
                for (unsigned k = 1; k < arr.size(); ++k) {
                    const auto& last = arr[k - 1];
                    const auto& curr = arr[k];
                    failures += !(last.start != curr.start || last.end != curr.end || last.dt != curr.dt);
                    failures += !(curr.start < curr.end);
                    failures += !(last.end <= curr.start);
                }
                if (arr.size() == 0)
                    continue;
                for (unsigned l = 0; l < 50; ++l) {
                    const S t = r() * (arr[arr.size() - 1].end - smallest - 0.00001f) + smallest;
                    failures += !(real_dt(t) == computed_dt(t));
                }
            }
        }
        return failures;
    };
    return dbl ? go(static_cast<double*>(nullptr)) : go(static_cast<float*>(nullptr));
}

// ------------------------------------------------------------ CPU baseline
// Times splat forward+backward of the float restatement on the host with
// `threads` OpenMP threads (preprocess and its backward over Gaussians, pair
// emission over Gaussians, a parallel stable sort, the blend and its adjoint
// over image rows).  Returns wall seconds.
double orc_time_splat_fwd_bwd(void* p, const float* bg, const float* grad_fb, float* fb, float* T_final, int x0, int y0, int x1, int y1,
    int threads)
{
    auto* h = static_cast<SplatHandle*>(p);
    threads = std::max(1, threads);
#ifdef _OPENMP
    const int saved = omp_get_max_threads();
    omp_set_num_threads(threads);
#endif
    const auto t0 = std::chrono::steady_clock::now();
    auto& st = h->sf;
    SplatGeometry<float> g;
    splat_preprocess(st.scene, st.cam, h->tile_w, h->tile_h, &g);
    Binning b;
    bin(g.rect, g.tiles, g.xy, g.key_depth, st.cam.width, st.cam.height, h->tile_w, h->tile_h, &b);
    splat_blend(g, b, st.cam.width, st.cam.height, bg, x0, y0, x1, y1, fb, T_final, nullptr, nullptr);
    Splat2dGrads<float> gr;
    if (threads > 1)
        splat_blend_backward_threads(g, b, st.cam.width, st.cam.height, bg, x0, y0, x1, y1, fb, T_final, grad_fb, &gr, threads);
    else
        splat_blend_backward(g, b, st.cam.width, st.cam.height, bg, x0, y0, x1, y1, fb, T_final, grad_fb, &gr);
    ParamGrads<float> pg;
    splat_preprocess_backward(st.scene, st.cam, g, gr, &pg);
    const auto t1 = std::chrono::steady_clock::now();
#ifdef _OPENMP
    omp_set_num_threads(saved);
#endif
    return std::chrono::duration<double>(t1 - t0).count();
}

// ----------------------------------------------------------------- rasteriser
// Float restatement of vol_raster_forward (16x16 tiles) for the pixels
// [x0, x1) x [y0, y1) of the planar [3, H, W] framebuffer fb (float inputs).
// n_pairs (may be null) receives the binning's pair count.
void orc_raster_render(int64_t n, const float* means, const float* scales, const float* rotations, const float* opacity, const float* sh16,
    int sh_degree, double scale_multiplier, const float* view, const float* proj, const float* cam_pos, double tan_fovx, double tan_fovy,
    int width, int height, const float* bg, int mode, int debug_bin, float max_depth, int x0, int y0, int x1, int y1, float* fb,
    int64_t* n_pairs)
{
    SceneStore<float> sf;
    sf.load(n, means, scales, rotations, opacity, sh16, sh_degree, scale_multiplier, view, proj, cam_pos, tan_fovx, tan_fovy, width, height,
        false);
    SplatGeometry<float> g;
    splat_preprocess(sf.scene, sf.cam, 16, 16, &g);
    Binning b;
    bin(g.rect, g.tiles, g.xy, g.key_depth, width, height, 16, 16, &b);
    if (n_pairs)
        *n_pairs = int64_t(b.point_list.size());
    // nothing binned: the reference returns before rendering (vol_raster_forward.cu:153-155)
    if (b.point_list.empty())
        return;
    RasterExtra<float> x;
    raster_preprocess_extra(sf.scene, sf.cam, g, &x);
    double pd[16], ipd[16];
    for (int i = 0; i < 16; ++i)
        pd[i] = sf.cam.proj[i];
    invert4(pd, ipd);
    float ip[16];
    for (int i = 0; i < 16; ++i)
        ip[i] = float(ipd[i]);
    crop_guard<float>(width, height, &x0, &y0, &x1, &y1);
    const size_t plane = size_t(width) * height;
#pragma omp parallel for schedule(dynamic, 1)
    for (int py = y0; py < y1; ++py)
        for (int px = x0; px < x1; ++px) {
            float out[3];
            raster_pixel(sf.scene, g, x, b, sf.cam, ip, px, py, bg, mode, debug_bin, max_depth, out);
            for (int c = 0; c < 3; ++c)
                fb[c * plane + size_t(py) * width + px] = out[c];
        }
}

// The rasteriser evaluated in FP64 on the float32 pipeline's placement: the
// scene (float32 arrays) converted to double, the binning and the per-pixel
// ray taken from the float32 restatement, everything else in double.
// ambiguous (optional, [H,W]) gets each pixel's RasterTieKind bits: the
// decisions that lay within float rounding of their thresholds.
void orc_raster_render64(int64_t n, const float* means, const float* scales, const float* rotations, const float* opacity, const float* sh16,
    int sh_degree, double scale_multiplier, const float* view, const float* proj, const float* cam_pos, double tan_fovx, double tan_fovy,
    int width, int height, const double* bg, int mode, int debug_bin, float max_depth, int x0, int y0, int x1, int y1, double* fb,
    uint8_t* ambiguous, int64_t* n_pairs)
{
    SceneStore<float> sf;
    sf.load(n, means, scales, rotations, opacity, sh16, sh_degree, scale_multiplier, view, proj, cam_pos, tan_fovx, tan_fovy, width, height,
        false);
    SceneStore<double> sd;
    sd.load(n, means, scales, rotations, opacity, sh16, sh_degree, scale_multiplier, view, proj, cam_pos, double(float(tan_fovx)),
        double(float(tan_fovy)), width, height, false);
    SplatGeometry<float> gf;
    splat_preprocess(sf.scene, sf.cam, 16, 16, &gf);
    SplatGeometry<double> gd;
    splat_preprocess(sd.scene, sd.cam, 16, 16, &gd, &gf);
    Binning b;
    bin(gf.rect, gf.tiles, gf.xy, gf.key_depth, width, height, 16, 16, &b);
    if (n_pairs)
        *n_pairs = int64_t(b.point_list.size());
    if (b.point_list.empty())
        return;
    RasterExtra<double> x;
    raster_preprocess_extra(sd.scene, sd.cam, gd, &x);
    double pd[16], ipd[16];
    for (int i = 0; i < 16; ++i)
        pd[i] = sf.cam.proj[i];
    invert4(pd, ipd);
    float ipf[16];
    for (int i = 0; i < 16; ++i)
        ipf[i] = float(ipd[i]);
    crop_guard<double>(width, height, &x0, &y0, &x1, &y1);
    const size_t plane = size_t(width) * height;
#pragma omp parallel for schedule(dynamic, 1)
    for (int py = y0; py < y1; ++py)
        for (int px = x0; px < x1; ++px) {
            float dirf[3];
            raster_ray(sf.cam, ipf, px, py, dirf);
            TieWatch tw;
            tie_watch = ambiguous ? &tw : nullptr;
            double out[3];
            raster_pixel(sd.scene, gd, x, b, sd.cam, ipd, px, py, bg, mode, debug_bin, max_depth, out, dirf);
            tie_watch = nullptr;
            for (int c = 0; c < 3; ++c)
                fb[c * plane + size_t(py) * width + px] = out[c];
            if (ambiguous)
                ambiguous[size_t(py) * width + px] = uint8_t(tw.hit);
        }
}

// RasterBinSizer with n_bins = nb (4 or 32) in float: the adds (mass, centre,
// sd) in order, full_after[k] = is_full() after add k, then finalise();
// begin/end of every bin and is_full() afterwards.  Returns 0, or -1 for an
// unsupported nb.
int orc_binsizer_run(int nb, float transmission_threshold, float relevance_sigma, int n_adds, const float* adds, uint8_t* full_after,
    float* begin, float* end, uint8_t* full_final)
{
    const auto go = [&](auto& z) {
        for (int k = 0; k < n_adds; ++k) {
            z.add_gaussian(adds[3 * k], adds[3 * k + 1], adds[3 * k + 2]);
            full_after[k] = z.is_full();
        }
        z.finalise();
        for (int i = 0; i < nb; ++i) {
            begin[i] = z.begin_of(i);
            end[i] = z.end_of(i);
        }
        *full_final = z.is_full();
    };
    if (nb == 4) {
        BinSizer<float, 4> z(transmission_threshold, relevance_sigma);
        go(z);
    } else if (nb == 32) {
        BinSizer<float, 32> z(transmission_threshold, relevance_sigma);
        go(z);
    } else {
        return -1;
    }
    return 0;
}

// piecewise_linear in double: create_approximation -> 21 values (d0, k0, t1,
// t2, k2 per channel, then t_right); sample; volume integration over a chain
// of groups with n_steps per group.
void orc_pl_create(const double* percent_left, const double* total, const double* eval_left, const double* eval_right, double t_right,
    double* out)
{
    const PiecewiseLinear<double> f = create_approximation(percent_left, total, eval_left, eval_right, t_right);
    for (int i = 0; i < 4; ++i) {
        out[i] = f.d0[i];
        out[4 + i] = f.k0[i];
        out[8 + i] = f.t1[i];
        out[12 + i] = f.t2[i];
        out[16 + i] = f.k2[i];
    }
    out[20] = f.t_right;
}

static PiecewiseLinear<double> pl_from(const double* in)
{
    PiecewiseLinear<double> f;
    for (int i = 0; i < 4; ++i) {
        f.d0[i] = in[i];
        f.k0[i] = in[4 + i];
        f.t1[i] = in[8 + i];
        f.t2[i] = in[12 + i];
        f.k2[i] = in[16 + i];
    }
    f.t_right = in[20];
    return f;
}

void orc_pl_sample(const double* params, double t, double* out) { pl_from(params).sample(t, out); }

void orc_pl_sample_many(const double* params, int64_t n, const double* ts, double* out)
{
    const PiecewiseLinear<double> f = pl_from(params);
    for (int64_t i = 0; i < n; ++i)
        f.sample(ts[i], out + 4 * i);
}

void orc_pl_volume_integrate(int n_groups, const double* params, int n_steps, double* out_rgb)
{
    double result[3] = { 0, 0, 0 }, transparency = 1;
    for (int k = 0; k < n_groups; ++k)
        volume_integrate_step(pl_from(params + 21 * k), n_steps, result, &transparency);
    for (int c = 0; c < 3; ++c)
        out_rgb[c] = result[c];
}

} // extern "C"
