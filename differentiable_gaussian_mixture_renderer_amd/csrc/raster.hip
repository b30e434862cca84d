// SPDX-License-Identifier: GPL-3.0-or-later
// Volumetric rasteriser for gfx950 (reference: src/dgmr/vol_raster_forward.cu:36-520),
// forward only, as the reference.
//
//   1. k_splat_preprocess (the splatter's: cull, rect, tiles, xy, depth, SH
//      colour; the reference's preprocess computes the same quantities the
//      same way) and k_raster_extra: inverse of the filtered 3D covariance and
//      the filtered Ots mass per visible Gaussian.
//   2. bin_gaussians: the splatter's binning (same keys, same order).
//   3. k_raster_render: one 16x16 tile per 256-thread workgroup, one pixel per
//      thread.  Pass 1 sizes 32 bins per pixel (math::RasterBinSizer, kept in
//      registers with fully unrolled, select-based updates so no runtime index
//      sends the arrays to scratch); pass 2 rasterises the tile's Gaussians
//      into the bins kChunk bins at a time and integrates each chunk's bins
//      piecewise-linearly before the next chunk, so only kChunk bins of
//      (rgb*mass, mass) and kChunk + 1 border evaluations are live.
#include <cmath>

#include "device_grad.h"
#include "device_math.h"
#include "host_common.h"
#include "march_device.h"
#include "splat_device.h"

namespace dgmr {
namespace {

constexpr int kRasterBins = 32;         // vol_raster::config::n_rasterisation_bins
constexpr int kRasterSteps = 256;       // vol_raster::config::n_steps_per_bin
constexpr float kRasterFilterSD = 0.55f; // vol_raster::config::filter_kernel_SD
constexpr float kRelevanceSigma = 3.f;  // vol_raster::config::gaussian_relevance_sigma
constexpr int kTile = 16;               // render_block_width/height (constants.h:45-46)
constexpr int kThreads = kTile * kTile;
constexpr int kChunk = 8;               // bins rasterised per walk of the tile's list
constexpr int kBorderStride = kRasterBins + 1; // odd: one LDS row per pixel, conflict-free

// vol_raster_forward.cu:100-136 for the Gaussians the splatter preprocess kept:
// filtered inverse covariance (xx, xy, xz, yy, yz, zz) and the mass
// weight_to_mass<Ots>(w, scales + f^2, .) = w * 2 pi * (two largest of s + f^2)
__global__ __launch_bounds__(256) void k_raster_extra(int64_t n, SplatParams p, DevCamera<float> cam, const uint32_t* __restrict__ tiles,
    float* __restrict__ inv_cov, float* __restrict__ mass)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n || tiles[i] == 0)
        return;
    float s[3], q[4], w;
    gaussian_params(p, i, s, q, &w);
    const float mu[3] = { p.means[3 * i], p.means[3 * i + 1], p.means[3 * i + 2] };
    float cov[6];
    compute_cov(s, q, cov);
    const float dv[3] = { mu[0] - cam.pos[0], mu[1] - cam.pos[1], mu[2] - cam.pos[2] };
    const float aa = (kRasterFilterSD * cam.tan_fovx * 2.f) / float(cam.width);
    const float f = length3(dv) * aa;
    const float f2 = f * f;
    const float sf[3] = { s[0] + f2, s[1] + f2, s[2] + f2 };
    float la, lb;
    int i0, i1;
    larger2_values(sf, &la, &lb, &i0, &i1);
    mass[i] = w * (float(2 * kPi) * la * lb);
    const float cov_f[6] = { cov[0] + f2, cov[1], cov[2], cov[3] + f2, cov[4], cov[5] + f2 };
    float inv[6];
    sym3_inverse(cov_f, inv);
#pragma unroll
    for (int k = 0; k < 6; ++k)
        inv_cov[6 * i + k] = inv[k];
}

__device__ __forceinline__ float larger_of(float a, float b) { return a < b ? b : a; } // std::max(a, b)

// stroke::gaussian::integrate_normalised_SD over [a, b]
__device__ __forceinline__ float integrate_normalised_sd(float centre, float sd, float a, float b)
{
    const float k = 0.70710678118654752440f;
    return 0.5f * erff((b - centre) / sd * k) - 0.5f * erff((a - centre) / sd * k);
}

__device__ __forceinline__ float cdf_sd(float centre, float inv_sd, float x)
{
    return 0.5f * (1.f + erff((x - centre) * inv_sd * 0.70710678118654752440f));
}

__device__ __forceinline__ float eval_exponential(float centre, float var, float x)
{
    const float d = x - centre;
    return expf(-0.5f * d * d / var);
}

// math::RasterBinSizer<vol_raster::config> (raster_bin_sizers.h:271-485)
struct BinSizer {
    float m[kRasterBins], b[kRasterBins], s[kRasterBins];
    float acc, maxd;

    __device__ __forceinline__ void init()
    {
#pragma unroll
        for (int i = 0; i < kRasterBins; ++i) {
            m[i] = 0.f;
            b[i] = __builtin_inff();
            s[i] = 0.f;
        }
        acc = 0.f;
        maxd = 0.f;
    }

    __device__ __forceinline__ void add(float mass, float centre, float sd)
    {
        if (mass < 0.0001f)
            return;
        constexpr int L = kRasterBins - 1;
        const float t_rel = centre + sd * kRelevanceSigma;
        const float added = mass * integrate_normalised_sd(centre, sd, 0.f, t_rel);
        maxd = larger_of(t_rel, maxd);
        acc += added;
        m[L] = added;
        b[L] = centre;
        s[L] = sd;
        // insertion from the back: swap while the left neighbour lies further
        bool moving = true;
#pragma unroll
        for (int i = L; i > 0; --i) {
            const bool sw = moving && b[i - 1] > b[i];
            const float bl = b[i - 1], br = b[i], sl = s[i - 1], sr = s[i], ml = m[i - 1], mr = m[i];
            b[i - 1] = sw ? br : bl;
            b[i] = sw ? bl : br;
            s[i - 1] = sw ? sr : sl;
            s[i] = sw ? sl : sr;
            m[i - 1] = sw ? mr : ml;
            m[i] = sw ? ml : mr;
            moving = sw;
        }
        if (!(m[L] > 0.f))
            return;
        // every slot is taken: merge the cheapest adjacent pair
        float tr[L];
        float t = 1.f;
#pragma unroll
        for (int i = 0; i < L; ++i) {
            t *= expf(-m[i]);
            tr[i] = t;
        }
        int best = -1;
        float best_val = __builtin_inff();
#pragma unroll
        for (int j = 0; j < L; ++j) {
            const float ta = j == 0 ? 1.f : tr[j - 1];
            float nm, nc, nsd;
            merge(ta, tr[j], j, &nm, &nc, &nsd);
            const float v = tr[j] * (kl(b[j], s[j], nc, nsd) + kl(b[j + 1], s[j + 1], nc, nsd));
            if (v < best_val) {
                best_val = v;
                best = j;
            }
        }
        // the reference asserts a candidate exists; with none its merge would
        // index slot -1, here slot 0 merges
        best = best < 0 ? 0 : best;
        // gather the pair (best, best + 1) without a runtime index
        float ba = b[0], bb = b[1], sa = s[0], sb = s[1], ma = m[0], mb = m[1], tb_a = 1.f, tb_b = tr[0];
#pragma unroll
        for (int j = 1; j < L; ++j) {
            const bool h = best == j;
            ba = h ? b[j] : ba;
            bb = h ? b[j + 1] : bb;
            sa = h ? s[j] : sa;
            sb = h ? s[j + 1] : sb;
            ma = h ? m[j] : ma;
            mb = h ? m[j + 1] : mb;
            tb_a = h ? tr[j - 1] : tb_a;
            tb_b = h ? tr[j] : tb_b;
        }
        float nm, nc, nsd;
        merge_values(tb_a, tb_b, ma, mb, ba, bb, sa, sb, &nm, &nc, &nsd);
#pragma unroll
        for (int j = 0; j < L; ++j) {
            const bool here = j == best, after = j > best;
            const float bn = j + 1 < kRasterBins ? b[j + 1] : b[j];
            const float sn = j + 1 < kRasterBins ? s[j + 1] : s[j];
            const float mn = j + 1 < kRasterBins ? m[j + 1] : m[j];
            b[j] = here ? nc : (after ? bn : b[j]);
            s[j] = here ? nsd : (after ? sn : s[j]);
            m[j] = here ? nm : (after ? mn : m[j]);
        }
    }

    static __device__ __forceinline__ float kl(float c1, float sd1, float c2, float sd2)
    {
        return ((c1 - c2) * (c1 - c2) + sd1 * sd1 - sd2 * sd2) / (2.f * (sd2 * sd2)) + logf(sd2 / sd1);
    }

    static __device__ __forceinline__ void merge_values(float tb_a, float tb_b, float ma, float mb, float ba, float bb, float sa, float sb,
        float* nm, float* nc, float* nsd)
    {
        const float xa = tb_a * ma;
        const float xb = tb_b * mb;
        const float pa = (xa + xb >= 0.0000001f) ? fminf_ordered(xa / (xa + xb), 0.999999f) : 0.5f;
        const float pb = 1.f - pa;
        *nc = pa * ba + pb * bb;
        *nm = ma + mb;
        const float va = sa * sa;
        const float vb = sb * sb;
        const float nv = pa * (va + (ba - *nc) * (ba - *nc)) + pb * (vb + (bb - *nc) * (bb - *nc));
        *nsd = sqrt_rn(nv);
    }

    // std::min(a, b) = b < a ? b : a
    static __device__ __forceinline__ float fminf_ordered(float a, float b) { return b < a ? b : a; }

    __device__ __forceinline__ void merge(float tb_a, float tb_b, int j, float* nm, float* nc, float* nsd) const
    {
        merge_values(tb_a, tb_b, m[j], m[j + 1], b[j], b[j + 1], s[j], s[j + 1], nm, nc, nsd);
    }

    __device__ __forceinline__ bool full(float threshold) const { return acc >= threshold; }

    // borders of never-filled slots spread evenly up to the furthest relevance end
    __device__ __forceinline__ void finalise()
    {
        float last = 0.f;
#pragma unroll
        for (int i = 0; i < kRasterBins; ++i) {
            const bool inf = __builtin_isinf(b[i]);
            const float kept = larger_of(b[i], last + 0.00001f);
            const float spread = last + (maxd - last) / float(kRasterBins - i);
            b[i] = inf ? spread : kept;
            m[i] = inf ? 0.f : m[i];
            last = b[i];
        }
    }
};

// piecewise_linear::create_approximation for one channel with the left share
// 0.5 (the reference's share is m / (m + m) of one border mass: exactly 0.5,
// and 0/0 -> NaN -> 0.5, vol_raster_forward.cu:450-453)
struct Linear {
    float d0, k0, t1, t2, k2;
};

__device__ __forceinline__ Linear fit_channel(float total, float el, float er, float t_right)
{
    const float m_l = 0.5f * total;
    const float m_r = total - m_l;
    Linear f { el, 0.f, 2.f * m_l / el, t_right - 2.f * m_r / er, 0.f };
    if (total <= 0.000001f)
        return Linear { 0.f, 0.f, 0.f, t_right, 0.f };
    if (f.t2 > f.t1) {
        f.k0 = -el * el / (2.f * m_l);
        f.k2 = er * er / (2.f * m_r);
        return f;
    }
    const float c1 = 2.f * total - (el + er) * t_right;
    const float c2 = er * m_l + el * m_r;
    const float c3 = el * er * t_right;
    const float dm = larger_of(0.f, (c1 + sqrt_rn(c1 * c1 + 4.f * t_right * (2.f * c2 - c3))) / (2.f * t_right));
    float tl, tr;
    if (m_r > m_l) {
        tr = 2.f * m_r / (dm + er);
        tl = larger_of(0.f, t_right - tr);
    } else {
        tl = 2.f * m_l / (dm + el);
        tr = larger_of(0.f, t_right - tl);
    }
    f.k0 = tl == 0.f ? 0.f : -(el - dm) / tl;
    f.t1 = tl;
    f.t2 = tl;
    f.k2 = tr == 0.f ? 0.f : (er - dm) / tr;
    return f;
}

__device__ __forceinline__ float sample_channel(const Linear& f, float t)
{
    const float left = f.d0 + f.k0 * t;
    const float right = f.d0 + f.t1 * f.k0 + (t - f.t2) * f.k2;
    return t <= f.t1 ? left : (t <= f.t2 ? 0.f : right);
}

struct RasterArgs {
    const float* means;
    const float* inv_cov;
    const float* mass;
    const float* rgb;
    const uint32_t* point_list;
    const uint32_t* ranges;
    float inv_proj[16];
    float pos[3];
    float bg[3];
    int width, height, grid_x, grid_y;
    int mode, debug_bin;
    float max_depth;
    float mass_threshold; // -log(transmission_threshold)
    const uint32_t* n_pairs; // device pair count: none leaves the framebuffer untouched
    float* framebuffer;
};

__global__ __launch_bounds__(kThreads) void k_raster_render(RasterArgs a)
{
    __shared__ float s_mu[3][kThreads];
    __shared__ float s_inv[6][kThreads];
    __shared__ float s_mass[kThreads];
    __shared__ float s_rgb[3][kThreads];
    __shared__ float s_border[kThreads * kBorderStride];
    // the reference returns before rendering when no Gaussian is binned
    // (vol_raster_forward.cu:153-155): the caller's framebuffer stays as it was
    if (*a.n_pairs == 0)
        return;
    const int tile = int(blockIdx.x);
    const int tx = tile % a.grid_x, ty = tile / a.grid_x;
    const int t_id = int(threadIdx.x);
    const int px = tx * kTile + (t_id % kTile), py = ty * kTile + (t_id / kTile);
    const bool inside = px < a.width && py < a.height;
    // ray through the integer pixel coordinate on the near plane (vol_raster_forward.cu:272-278)
    float dir[3];
    {
        const float ndc[3] = { float(px * 2) / float(a.width) - 1.f, float(py * 2) / float(a.height) - 1.f, -1.f };
        float v[4];
        transform_point(a.inv_proj, ndc, v, 4);
        float pv[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
            pv[i] = v[i] / v[3] - a.pos[i];
        const float inv_len = 1.f / sqrt_rn(dot3(pv, pv));
#pragma unroll
        for (int i = 0; i < 3; ++i)
            dir[i] = pv[i] * inv_len;
    }
    const uint32_t first = a.ranges[2 * tile], last = a.ranges[2 * tile + 1];
    const int n_rounds = int((last - first + kThreads - 1) / kThreads);

    const auto stage = [&](int round) {
        const uint32_t k = first + uint32_t(round) * kThreads + uint32_t(t_id);
        if (k < last) {
            const uint32_t id = a.point_list[k];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                s_mu[c][t_id] = a.means[3 * id + c];
                s_rgb[c][t_id] = a.rgb[3 * id + c];
            }
#pragma unroll
            for (int c = 0; c < 6; ++c)
                s_inv[c][t_id] = a.inv_cov[6 * id + c];
            s_mass[t_id] = a.mass[id];
        }
    };
    const auto ray1d = [&](int j) {
        const float mu[3] = { s_mu[0][j], s_mu[1][j], s_mu[2][j] };
        const float A[6] = { s_inv[0][j], s_inv[1][j], s_inv[2][j], s_inv[3][j], s_inv[4][j], s_inv[5][j] };
        return intersect_ray(mu, A, a.pos, dir);
    };

    // ---------------------------------------------------------------- pass 1
    BinSizer z;
    z.init();
    if (a.mode != 1) { // Bins mode renders the background only
        bool done = !inside;
        for (int r = 0; r < n_rounds; ++r) {
            if (__syncthreads_and(done))
                break;
            stage(r);
            __syncthreads();
            if (done)
                continue;
            const int count = int(min(uint32_t(kThreads), last - first - uint32_t(r) * kThreads));
            for (int j = 0; j < count; ++j) {
                const Ray1d<float> g = ray1d(j);
                const float mor = g.weight * s_mass[j];
                if (mor < 0.001f || mor > 1000.f)
                    continue;
                if (g.var <= 0.f)
                    continue;
                z.add(mor, g.centre, sqrt_rn(g.var));
                if (z.full(a.mass_threshold)) {
                    done = true;
                    break;
                }
            }
        }
        __syncthreads();
        z.finalise();
    }
    // end_of(i): border i, the last one being the furthest relevance end
    float* border = s_border + t_id * kBorderStride;
#pragma unroll
    for (int i = 0; i < kRasterBins - 1; ++i)
        border[i] = z.b[i];
    border[kRasterBins - 1] = z.maxd;
    const auto end_of = [&](int i) { return border[i]; };
    const auto begin_of = [&](int i) { return i == 0 ? 0.f : border[i - 1]; };

    float C[3] = { 0.f, 0.f, 0.f };
    float T = 1.f; // Full mode keeps T = 1: the background is added in full (vol_raster_forward.cu:509-513)
    if (a.mode == 0) {
        // ------------------------------------------------------------ pass 2
        float result[3] = { 0.f, 0.f, 0.f };
        float transparency = 1.f;
        const float max_distance = end_of(kRasterBins - 1);
        for (int k0 = 0; k0 < kRasterBins; k0 += kChunk) {
            float bm[kChunk][4], be[kChunk + 1][4];
#pragma unroll
            for (int q = 0; q < kChunk; ++q)
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    bm[q][c] = 0.f;
#pragma unroll
            for (int q = 0; q <= kChunk; ++q)
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    be[q][c] = 0.f;
            const bool last_chunk = k0 + kChunk == kRasterBins;
            for (int r = 0; r < n_rounds; ++r) {
                __syncthreads();
                stage(r);
                __syncthreads();
                if (!inside)
                    continue;
                const int count = int(min(uint32_t(kThreads), last - first - uint32_t(r) * kThreads));
                for (int j = 0; j < count; ++j) {
                    const Ray1d<float> g = ray1d(j);
                    const float c = g.centre;
                    const float var = g.var;
                    const float inv_sd = 1.f / sqrt_rn(var);
                    const float w = g.weight * s_mass[j];
                    if (var <= 0.f || __builtin_isnan(var) || __builtin_isnan(w) || w > 100000.f)
                        continue;
                    if (w * (cdf_sd(c, inv_sd, max_distance) - cdf_sd(c, inv_sd, 0.f)) <= 0.001f)
                        continue;
                    const float rgb[3] = { s_rgb[0][j], s_rgb[1][j], s_rgb[2][j] };
                    float cdf_start = cdf_sd(c, inv_sd, begin_of(k0));
#pragma unroll
                    for (int q = 0; q < kChunk; ++q) {
                        const float cdf_end = cdf_sd(c, inv_sd, end_of(k0 + q));
                        const float mq = (cdf_end - cdf_start) * w;
                        cdf_start = cdf_end;
                        if (mq < 0.0001f)
                            continue;
#pragma unroll
                        for (int ch = 0; ch < 3; ++ch)
                            bm[q][ch] += rgb[ch] * mq;
                        bm[q][3] += mq;
                        const float e = w * eval_exponential(c, var, begin_of(k0 + q));
#pragma unroll
                        for (int ch = 0; ch < 3; ++ch)
                            be[q][ch] += rgb[ch] * e;
                        be[q][3] += e;
                    }
                    // the evaluation at the chunk's right border belongs to the
                    // next bin, added when that bin's mass passes (or, after the
                    // last bin, always)
                    bool add_right = last_chunk;
                    float x_right = max_distance;
                    if (!last_chunk) {
                        const float cdf_end = cdf_sd(c, inv_sd, end_of(k0 + kChunk));
                        add_right = !((cdf_end - cdf_start) * w < 0.0001f);
                        x_right = begin_of(k0 + kChunk);
                    }
                    if (add_right) {
                        const float e = w * eval_exponential(c, var, x_right);
#pragma unroll
                        for (int ch = 0; ch < 3; ++ch)
                            be[kChunk][ch] += rgb[ch] * e;
                        be[kChunk][3] += e;
                    }
                }
            }
            // integrate the chunk's bins (piecewise_linear.h:152-178)
#pragma unroll
            for (int q = 0; q < kChunk; ++q) {
                const float t_right = end_of(k0 + q) - begin_of(k0 + q);
                if (!(t_right > 0.f))
                    continue;
                Linear f[4];
#pragma unroll
                for (int ch = 0; ch < 4; ++ch)
                    f[ch] = fit_channel(bm[q][ch], be[q][ch], be[q + 1][ch], t_right);
                const float dt = t_right / float(kRasterSteps);
                float t = dt / 2.f;
                for (int st = 0; st < kRasterSteps; ++st) {
                    const float e0 = sample_channel(f[0], t), e1 = sample_channel(f[1], t), e2 = sample_channel(f[2], t);
                    const float e3 = sample_channel(f[3], t);
                    result[0] += e0 * transparency * dt;
                    result[1] += e1 * transparency * dt;
                    result[2] += e2 * transparency * dt;
                    transparency *= larger_of(0.f, 1.f - e3 * dt);
                    t += dt;
                }
#pragma unroll
                for (int ch = 0; ch < 3; ++ch)
                    C[ch] = result[ch];
            }
        }
    } else if (a.mode == 2) {
        const int bin = int(min(unsigned(a.debug_bin), unsigned(kRasterBins - 1)));
        const float dist = end_of(bin);
        const float v = dist / a.max_depth;
        C[0] = C[1] = C[2] = v;
        if (dist == 0.f) {
            C[0] = 0.f;
            C[1] = 1.f;
            C[2] = 0.f;
        }
        if (__builtin_isnan(dist)) {
            C[0] = 1.f;
            C[1] = 0.f;
            C[2] = 0.5f;
        }
        if (dist < 0.f) {
            C[0] = 1.f;
            C[1] = 0.5f;
            C[2] = 0.f;
        }
        T = 0.f;
    }
    if (!inside)
        return;
    const int64_t plane = int64_t(a.width) * a.height;
    const int64_t q = int64_t(py) * a.width + px;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch)
        a.framebuffer[ch * plane + q] = C[ch] + T * a.bg[ch];
}

} // namespace
} // namespace dgmr

using namespace dgmr;

extern "C" {

int dgmr_raster_forward(const dgmr_gaussians* g, const dgmr_camera* cam, const dgmr_raster_options* opt, dgmr_splat_geometry* geom,
    dgmr_binning* bins, float* inv_cov, float* mass, float* framebuffer, void* workspace, size_t workspace_bytes, dgmr_stats* stats,
    dgmr_stream stream)
{
    DGMR_ARG_CHECK(g && cam && opt && geom && bins && inv_cov && mass && framebuffer && workspace, "dgmr_raster_forward: null argument");
    DGMR_ARG_CHECK(g->dtype == DGMR_F32, "dgmr_raster_forward: the rasteriser computes in float32");
    DGMR_ARG_CHECK(g->n >= 0 && g->n < (int64_t(1) << 31), "dgmr_raster_forward: bad Gaussian count");
    DGMR_ARG_CHECK(g->sh_coeffs >= 1 && g->sh_coeffs <= 16 && g->sh_degree >= 0 && g->sh_degree <= 3, "dgmr_raster_forward: bad SH layout");
    DGMR_ARG_CHECK(valid_image(cam->width, cam->height), "dgmr_raster_forward: bad image size");
    DGMR_ARG_CHECK(opt->render_mode >= 0 && opt->render_mode <= 2, "dgmr_raster_forward: render mode must be 0 (full), 1 (bins) or 2 (depth)");
    DGMR_ARG_CHECK(bins->point_list && bins->ranges, "dgmr_raster_forward: null binning buffers");
    const int64_t n = g->n;
    DGMR_ARG_CHECK(workspace_bytes >= splat_workspace_bytes(n, bins->pair_capacity), "dgmr_raster_forward: workspace too small");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const dgmr_tiling tiling { kTile, kTile, 64, 0, 0 }; // the reference's rectangles
    const DevCamera<float> dc = make_dev_camera<float>(*cam, tiling);
    const int64_t n_tiles = int64_t(dc.grid_x) * dc.grid_y;
    if (stats) {
        stats->n_pairs = 0;
        stats->n_visible = 0;
    }
    RasterArgs ra {};
    ra.means = static_cast<const float*>(g->means);
    ra.inv_cov = inv_cov;
    ra.mass = mass;
    ra.rgb = geom->rgb;
    ra.point_list = bins->point_list;
    ra.ranges = bins->ranges;
    for (int i = 0; i < 16; ++i)
        ra.inv_proj[i] = dc.inv_proj[i];
    for (int i = 0; i < 3; ++i) {
        ra.pos[i] = dc.pos[i];
        ra.bg[i] = dc.bg[i];
    }
    ra.width = cam->width;
    ra.height = cam->height;
    ra.grid_x = dc.grid_x;
    ra.grid_y = dc.grid_y;
    ra.mode = opt->render_mode;
    ra.debug_bin = opt->debug_bin;
    ra.max_depth = opt->max_depth;
    ra.mass_threshold = float(-std::log(double(0.005f))); // -log(vol_raster::config::transmission_threshold)
    ra.framebuffer = framebuffer;
    if (n == 0) {
        // nothing binned: the reference returns before rendering, the framebuffer is left as it was
        DGMR_HIP_CHECK(hipMemsetAsync(bins->ranges, 0, size_t(n_tiles) * 2 * sizeof(uint32_t), st));
        return DGMR_OK;
    }
    uint32_t *dkeys, *dids;
    uint2* brects;
    binning_key_buffers(workspace, n, false, &dkeys, &dids, reinterpret_cast<void**>(&brects));
    {
        StageScope sc("raster.preprocess", st, double(n) * (44.0 + 12.0 * g->sh_coeffs + 75.0 + 28.0));
        DGMR_HIP_CHECK(launch_splat_preprocess(g, dc,
            SplatOut { geom->rects, geom->tiles_touched, geom->xy, reinterpret_cast<float4*>(geom->conic_opacity), geom->rgb, geom->rgb_clamped,
                geom->depths, dkeys, dids, brects, nullptr },
            st));
        k_raster_extra<<<unsigned(ceil_div64(n, 256)), 256, 0, st>>>(n, make_splat_params(g), dc, geom->tiles_touched, inv_cov, mass);
        DGMR_HIP_CHECK(hipGetLastError());
    }
    Carver cv(workspace, workspace_bytes);
    PendingCount pending;
    const int rc = bin_gaussians(n, geom->rects, geom->tiles_touched, geom->xy, false, true, false, dc.grid_x, dc.grid_y, dc.tile_w, dc.tile_h, bins, cv,
        &pending, st, nullptr);
    if (rc != DGMR_OK)
        return rc;
    ra.n_pairs = pending.device;
    {
        StageScope sc("raster.render", st);
        k_raster_render<<<unsigned(n_tiles), kThreads, 0, st>>>(ra);
        DGMR_HIP_CHECK(hipGetLastError());
    }
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

} // extern "C"
