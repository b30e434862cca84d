// SPDX-License-Identifier: GPL-3.0-or-later
// dgmr oracle -- TEST INFRASTRUCTURE ONLY (see oracle_math.hpp).
//
// Plain-loop restatement of the volumetric rasteriser,
// src/dgmr/vol_raster_forward.cu:36-520: the splatter's preprocess and
// binning plus, per Gaussian, the inverse of the filtered 3D covariance and
// the filtered mass; per pixel two passes over the tile's Gaussians (bin
// sizing with math::RasterBinSizer, raster_bin_sizers.h:271-485, then
// rasterisation into the bins) and the piecewise-linear volume integration of
// piecewise_linear.h:25-180.
#pragma once

#include <cmath>
#include <limits>

#include "oracle_march.hpp"
#include "oracle_splat.hpp"

namespace orc {

// vol_raster::config, vol_raster_forward.h:59-68
struct RasterConfig {
    static constexpr int n_bins = 32;
    static constexpr int n_steps_per_bin = 256;
    static constexpr float transmission_threshold = 0.005f;
    static constexpr float relevance_sigma = 3.f;
    static constexpr float filter_kernel_sd = 0.55f;
};

// Decisions of the rasteriser that a float32 run can take the other way when
// the FP64 value lies within float rounding of the threshold (recorded through
// the marcher's TieWatch, per pixel, in FP64 runs only)
enum RasterTieKind : unsigned {
    kRasterTieMass = 1,   // a mass against its threshold: on the ray (1e-3, 1000), the sizer's 1e-4,
                          // in the bins (1e-3), per bin (1e-4), a bin's total (1e-6), a merge's weights (1e-7)
    kRasterTieOrder = 2,  // the sizer's order of centres
    kRasterTieMerge = 4,  // the sizer's cheapest merge against the runner-up
    kRasterTieFull = 8,   // the sizer's accumulated mass against -ln(0.005)
    kRasterTieShape = 16  // the piecewise-linear approximation's branch t2 > t1 (its m_r > m_l is decided
                          // exactly: the reference's left share is 0.5, so m_l = m_r in every precision)
};

struct RasterTolerance {
    static constexpr double rel = 1e-4;      // relative, for masses, costs and the fullness sum
    // relative with a floor of 1, for the centres the sizer orders: a float
    // centre is a few ulp of its distance off (no accumulation along the ray)
    static constexpr double position = 4e-6;
    static constexpr double knot = 2e-5;     // relative to the bin's width, for the approximation's knots
    static constexpr double cdf_abs = 4e-7;  // absolute error of a float cdf difference (per unit of mass on the ray)
};

inline void raster_watch(double a, double b, double rel, double abs_tol, unsigned kind)
{
    if (!tie_watch || !std::isfinite(a) || !std::isfinite(b))
        return;
    if (std::abs(a - b) <= rel * std::max(std::abs(a), std::abs(b)) + abs_tol)
        tie_watch->hit |= kind;
}

// stroke::gaussian::integrate_normalised_SD over [a, b]: Phi((b - c)/sd) - Phi((a - c)/sd)
template <class S>
inline S integrate_normalised_sd(S centre, S sd, S a, S b)
{
    const S k = S(0.70710678118654752440);
    return S(0.5) * std::erf((b - centre) / sd * k) - S(0.5) * std::erf((a - centre) / sd * k);
}

// the 1D exponential exp(-(x - c)^2 / (2 var)) without normalisation (stroke::gaussian::eval_exponential)
template <class S>
inline S eval_exponential(S centre, S var, S x)
{
    const S d = x - centre;
    return std::exp(S(-0.5) * d * d / var);
}

// math::RasterBinSizer, raster_bin_sizers.h:271-485: NB slots (border =
// centre, SD, mass) kept sorted by centre; when all slots are taken the
// adjacent pair whose merge costs the least (symmetric KL divergence weighted
// by the transmission before it) is merged.
template <class S, int NB>
struct BinSizer {
    S masses[NB];
    S borders[NB];
    S sds[NB];
    S accumulated = 0;
    S max_depth = 0;
    S mass_threshold;
    S relevance_sigma;

    BinSizer(float transmission_threshold, float relevance)
        : mass_threshold(S(float(-std::log(double(transmission_threshold))))), relevance_sigma(S(relevance))
    {
        for (int i = 0; i < NB; ++i) {
            masses[i] = 0;
            borders[i] = std::numeric_limits<S>::infinity();
            sds[i] = 0;
        }
    }
    S end_of(int i) const { return i == NB - 1 ? max_depth : borders[i]; }
    S begin_of(int i) const { return i == 0 ? S(0) : end_of(i - 1); }
    bool is_full() const
    {
        raster_watch(double(accumulated), double(mass_threshold), RasterTolerance::rel, 0.0, kRasterTieFull);
        return accumulated >= mass_threshold;
    }

    void add_gaussian(S mass, S centre, S sd)
    {
        raster_watch(double(mass), double(0.0001f), RasterTolerance::rel, 0.0, kRasterTieMass);
        if (mass < S(0.0001f))
            return;
        const S t_rel = centre + sd * relevance_sigma;
        const S added = mass * integrate_normalised_sd(centre, sd, S(0), t_rel);
        max_depth = std::max(t_rel, max_depth);
        accumulated += added;
        const int last = NB - 1;
        masses[last] = added;
        borders[last] = centre;
        sds[last] = sd;
        for (int i = last; i > 0; --i) {
            raster_watch(double(borders[i - 1]), double(borders[i]), RasterTolerance::position, RasterTolerance::position, kRasterTieOrder);
            if (borders[i - 1] > borders[i]) {
                std::swap(borders[i - 1], borders[i]);
                std::swap(sds[i - 1], sds[i]);
                std::swap(masses[i - 1], masses[i]);
                continue;
            }
            break;
        }
        S trans[NB];
        S tr = 1;
        for (int i = 0; i < last; ++i) {
            tr *= std::exp(-masses[i]);
            trans[i] = tr;
        }
        const auto before = [&](int index) { return index <= 0 ? S(1) : trans[index - 1]; };
        struct G1 {
            S mass, centre, sd;
        };
        const auto merge = [&](int a, int b) {
            const S sa = before(a) * masses[a];
            const S sb = before(b) * masses[b];
            raster_watch(double(sa + sb), double(0.0000001f), RasterTolerance::rel, 0.0, kRasterTieMass);
            const S pa = (sa + sb >= S(0.0000001f)) ? std::min(sa / (sa + sb), S(0.999999f)) : S(0.5f);
            const S pb = 1 - pa;
            const S nc = pa * borders[a] + pb * borders[b];
            const S nm = masses[a] + masses[b];
            const S va = sds[a] * sds[a];
            const S vb = sds[b] * sds[b];
            const S nv = pa * (va + (borders[a] - nc) * (borders[a] - nc)) + pb * (vb + (borders[b] - nc) * (borders[b] - nc));
            return G1 { nm, nc, std::sqrt(nv) };
        };
        const auto kl = [](S c1, S sd1, S c2, S sd2) {
            return ((c1 - c2) * (c1 - c2) + sd1 * sd1 - sd2 * sd2) / (2 * (sd2 * sd2)) + std::log(sd2 / sd1);
        };
        const auto cost = [&](int j) {
            const G1 m = merge(j, j + 1);
            return before(j + 1) * (kl(borders[j], sds[j], m.centre, m.sd) + kl(borders[j + 1], sds[j + 1], m.centre, m.sd));
        };
        if (masses[last] > 0) {
            int best = -1;
            S best_val = std::numeric_limits<S>::infinity();
            S runner_up = std::numeric_limits<S>::infinity();
            for (int j = 0; j < NB - 1; ++j) {
                const S v = cost(j);
                if (v < best_val) {
                    runner_up = best_val;
                    best_val = v;
                    best = j;
                } else if (v < runner_up) {
                    runner_up = v;
                }
            }
            raster_watch(double(best_val), double(runner_up), RasterTolerance::rel, 1e-12, kRasterTieMerge);
            // the reference asserts a candidate exists; with none (every cost
            // NaN or infinite) its merge would index slot -1, here slot 0 merges
            if (best < 0)
                best = 0;
            const G1 m = merge(best, best + 1);
            borders[best] = m.centre;
            masses[best] = m.mass;
            sds[best] = m.sd;
            for (int j = best + 1; j < NB - 1; ++j) {
                borders[j] = borders[j + 1];
                masses[j] = masses[j + 1];
                sds[j] = sds[j + 1];
            }
        }
    }

    void finalise()
    {
        S last_valid = 0;
        for (int i = 0; i < NB; ++i) {
            S& b = borders[i];
            if (!std::isinf(b)) {
                b = std::max(b, last_valid + S(0.00001f));
                last_valid = b;
            } else {
                b = last_valid + (max_depth - last_valid) / S(NB - i);
                last_valid = b;
                masses[i] = 0;
            }
        }
    }
};

// piecewise_linear::FunctionGroup (piecewise_linear.h:25-52), four channels
template <class S>
struct PiecewiseLinear {
    S d0[4], k0[4], t1[4], t2[4], k2[4];
    S t_right;
    void sample(S t, S* out) const
    {
        for (int i = 0; i < 4; ++i) {
            if (t <= t1[i])
                out[i] = d0[i] + k0[i] * t;
            else if (t <= t2[i])
                out[i] = 0;
            else
                out[i] = d0[i] + t1[i] * k0[i] + (t - t2[i]) * k2[i];
        }
    }
};

// piecewise_linear::create_approximation, piecewise_linear.h:54-150
template <class S>
inline PiecewiseLinear<S> create_approximation(const S* percent_left, const S* total, const S* eval_left, const S* eval_right, S t_right)
{
    PiecewiseLinear<S> f;
    f.t_right = t_right;
    for (int i = 0; i < 4; ++i) {
        const S m_l = percent_left[i] * total[i];
        const S m_r = total[i] - m_l;
        f.d0[i] = eval_left[i];
        f.k0[i] = 0;
        f.t1[i] = S(2) * m_l / eval_left[i];
        f.t2[i] = t_right - S(2) * m_r / eval_right[i];
        f.k2[i] = 0;
        raster_watch(double(total[i]), 0.000001, RasterTolerance::rel, 0.0, kRasterTieMass);
        raster_watch(double(f.t2[i]), double(f.t1[i]), 0.0, RasterTolerance::knot * double(t_right), kRasterTieShape);
        if (total[i] <= S(0.000001)) {
            f.d0[i] = 0;
            f.k0[i] = 0;
            f.t1[i] = 0;
            f.t2[i] = t_right;
            f.k2[i] = 0;
            continue;
        }
        if (f.t2[i] > f.t1[i]) {
            f.k0[i] = -eval_left[i] * eval_left[i] / (2 * m_l);
            f.k2[i] = eval_right[i] * eval_right[i] / (2 * m_r);
            continue;
        }
        const S m = total[i];
        const S gl = eval_left[i], gr = eval_right[i];
        const S to = t_right;
        const S c1 = 2 * m - (gl + gr) * to;
        const S c2 = gr * m_l + gl * m_r;
        const S c3 = gl * gr * to;
        const S dm = std::max(S(0), (c1 + std::sqrt(c1 * c1 + 4 * to * (2 * c2 - c3))) / (2 * to));
        // the reference returns (tl, tr) as a pair of floats whatever scalar_t is
        float tlf, trf;
        if (m_r > m_l) {
            trf = float(2 * m_r / (dm + gr));
            tlf = float(std::max(S(0), to - S(trf)));
        } else {
            tlf = float(2 * m_l / (dm + gl));
            trf = float(std::max(S(0), to - S(tlf)));
        }
        const S tl = S(tlf), tr = S(trf);
        f.k0[i] = tl == 0 ? S(0) : -(gl - dm) / tl;
        f.t1[i] = tl;
        f.t2[i] = tl;
        f.k2[i] = tr == 0 ? S(0) : (gr - dm) / tr;
    }
    return f;
}

// piecewise_linear::volume_integrate, piecewise_linear.h:152-178 (as its
// callers use it: transmittance carried across the bins, colour returned)
template <class S>
inline void volume_integrate_step(const PiecewiseLinear<S>& f, int n_steps, S* result, S* transparency)
{
    const S dt = f.t_right / S(n_steps);
    S t = dt / 2;
    for (int s = 0; s < n_steps; ++s) {
        S e[4];
        f.sample(t, e);
        for (int c = 0; c < 3; ++c)
            result[c] += e[c] * *transparency * dt;
        *transparency *= std::max(S(0), S(1) - e[3] * dt);
        t += dt;
    }
}

// per-Gaussian data of the rasteriser beyond the splatter's preprocess
template <class S>
struct RasterExtra {
    std::vector<S> inv_cov; // [n,6] inverse of the filtered 3D covariance (xx, xy, xz, yy, yz, zz)
    std::vector<S> mass;    // [n]   filtered mass, weight_to_mass<Ots>
};

// vol_raster_forward.cu:100-136 for the Gaussians the splatter preprocess kept
template <class S>
inline void raster_preprocess_extra(const Scene<S>& sc, const Camera<S>& cam, const SplatGeometry<S>& g, RasterExtra<S>* x)
{
    const int64_t n = sc.n;
    x->inv_cov.assign(6 * n, 0);
    x->mass.assign(n, 0);
    const S aa = (S(RasterConfig::filter_kernel_sd) * cam.tan_fovx * S(2)) / S(cam.width);
    for (int64_t i = 0; i < n; ++i) {
        if (g.tiles[i] == 0)
            continue;
        const S* mu = sc.means + 3 * i;
        const S s[3] = { sc.scales[3 * i] * sc.scale_multiplier, sc.scales[3 * i + 1] * sc.scale_multiplier, sc.scales[3 * i + 2] * sc.scale_multiplier };
        S cov[6];
        compute_cov(s, sc.rotations + 4 * i, cov);
        const S dv[3] = { mu[0] - cam.pos[0], mu[1] - cam.pos[1], mu[2] - cam.pos[2] };
        const S f = length3(dv) * aa;
        const S f2 = f * f;
        // weight_to_mass<Ots> of the reference's call: the scales (not the
        // variances) plus f^2, two largest, times 2 pi (math.h:311-314)
        const S sf[3] = { s[0] + f2, s[1] + f2, s[2] + f2 };
        int i0, i1;
        larger2_idx(sf, &i0, &i1);
        x->mass[i] = sc.opacity[i] * (S(2 * kPi) * sf[i0] * sf[i1]);
        const S cov_f[6] = { cov[0] + f2, cov[1], cov[2], cov[3] + f2, cov[4], cov[5] + f2 };
        sym3_inverse(cov_f, &x->inv_cov[6 * i]);
    }
}

// per-pixel ray of the rasteriser (vol_raster_forward.cu:272-278): the integer
// pixel coordinate (no +0.5) to NDC on the near plane (z = -1), through the
// inverse projection, normalised direction from the camera
template <class S>
inline void raster_ray(const Camera<S>& cam, const S* inv_proj, int px, int py, S* dir)
{
    const S ndc[3] = { S(px * 2) / S(cam.width) - S(1), S(py * 2) / S(cam.height) - S(1), S(-1) };
    S v[4];
    transform_point(inv_proj, ndc, v, 4);
    S p[3];
    for (int i = 0; i < 3; ++i)
        p[i] = v[i] / v[3] - cam.pos[i];
    const S inv_len = S(1) / std::sqrt(dot3(p, p));
    for (int i = 0; i < 3; ++i)
        dir[i] = p[i] * inv_len;
}

enum RasterMode { RasterFull = 0, RasterBins = 1, RasterDepth = 2 };

// one pixel, vol_raster_forward.cu:280-516
// ray_f (optional): the pixel's ray direction as the float32 pipeline
// computes it, used instead of this precision's (the near-plane unprojection
// leaves a float32 direction ~1e-5 off)
template <class S>
inline void raster_pixel(const Scene<S>& sc, const SplatGeometry<S>& g, const RasterExtra<S>& x, const Binning& b, const Camera<S>& cam,
    const S* inv_proj, int px, int py, const S* bg, int mode, int debug_bin, float max_depth_scale, S* out, const float* ray_f = nullptr)
{
    constexpr int NB = RasterConfig::n_bins;
    S d[3];
    if (ray_f) {
        for (int i = 0; i < 3; ++i)
            d[i] = S(ray_f[i]);
    } else {
        raster_ray(cam, inv_proj, px, py, d);
    }
    const int tile = (py / b.tile_h) * b.grid_x + (px / b.tile_w);
    const uint32_t first = b.ranges[2 * tile], last = b.ranges[2 * tile + 1];
    BinSizer<S, NB> sizer(RasterConfig::transmission_threshold, RasterConfig::relevance_sigma);
    for (uint32_t k = first; k < last; ++k) {
        const uint32_t id = b.point_list[k];
        const Gaussian1d<S> g1 = intersect_ray(sc.means + 3 * id, &x.inv_cov[6 * id], cam.pos, d);
        const S mor = g1.weight * x.mass[id];
        raster_watch(double(mor), 0.001, RasterTolerance::rel, 0.0, kRasterTieMass);
        raster_watch(double(mor), 1000.0, RasterTolerance::rel, 0.0, kRasterTieMass);
        if (mor < S(0.001) || mor > S(1000))
            continue;
        if (g1.var <= 0)
            continue;
        sizer.add_gaussian(mor, g1.centre, std::sqrt(g1.var));
        if (sizer.is_full())
            break;
    }
    sizer.finalise();
    S bin_mass[NB][4] = {};
    S bin_eval[NB + 1][4] = {};
    for (uint32_t k = first; k < last; ++k) {
        const uint32_t id = b.point_list[k];
        const Gaussian1d<S> g1 = intersect_ray(sc.means + 3 * id, &x.inv_cov[6 * id], cam.pos, d);
        const S c = g1.centre;
        const S var = g1.var;
        const S sd = std::sqrt(var);
        const S inv_sd = S(1) / sd;
        const S w = g1.weight * x.mass[id];
        raster_watch(double(w), 100000.0, RasterTolerance::rel, 0.0, kRasterTieMass);
        if (var <= 0 || std::isnan(var) || std::isnan(w) || w > S(100000))
            continue;
        const S in_bins = w * (cdf_inv_sd(c, inv_sd, sizer.end_of(NB - 1)) - cdf_inv_sd(c, inv_sd, S(0)));
        raster_watch(double(in_bins), double(0.001f), RasterTolerance::rel, double(w) * RasterTolerance::cdf_abs, kRasterTieMass);
        if (in_bins <= S(0.001f))
            continue;
        const S* rgb = &g.rgb[3 * id];
        S cdf_start = cdf_inv_sd(c, inv_sd, S(0));
        for (int k2 = 0; k2 < NB; ++k2) {
            const S cdf_end = cdf_inv_sd(c, inv_sd, sizer.end_of(k2));
            const S m = (cdf_end - cdf_start) * w;
            cdf_start = cdf_end;
            raster_watch(double(m), double(0.0001f), RasterTolerance::rel, double(w) * RasterTolerance::cdf_abs, kRasterTieMass);
            if (m < S(0.0001f))
                continue;
            for (int ch = 0; ch < 3; ++ch)
                bin_mass[k2][ch] += rgb[ch] * m;
            bin_mass[k2][3] += m;
            const S e = w * eval_exponential(c, var, sizer.begin_of(k2));
            for (int ch = 0; ch < 3; ++ch)
                bin_eval[k2][ch] += rgb[ch] * e;
            bin_eval[k2][3] += e;
        }
        const S e = w * eval_exponential(c, var, sizer.end_of(NB - 1));
        for (int ch = 0; ch < 3; ++ch)
            bin_eval[NB][ch] += rgb[ch] * e;
        bin_eval[NB][3] += e;
    }
    S C[3] = { 0, 0, 0 };
    S T = 1;
    if (mode == RasterFull) {
        S result[3] = { 0, 0, 0 };
        S transparency = 1;
        for (int k2 = 0; k2 < NB; ++k2) {
            const S t_right = sizer.end_of(k2) - sizer.begin_of(k2);
            if (t_right <= 0)
                continue;
            // the reference's left share is m / (m + m) of one border mass:
            // exactly 0.5, and 0/0 -> NaN -> 0.5 (vol_raster_forward.cu:450-453)
            const S half[4] = { S(0.5), S(0.5), S(0.5), S(0.5) };
            const PiecewiseLinear<S> f = create_approximation(half, bin_mass[k2], bin_eval[k2], bin_eval[k2 + 1], t_right);
            volume_integrate_step(f, RasterConfig::n_steps_per_bin, result, &transparency);
            for (int ch = 0; ch < 3; ++ch)
                C[ch] = result[ch];
        }
    } else if (mode == RasterDepth) {
        const int bin = std::min(unsigned(debug_bin), unsigned(NB - 1));
        const S dist = sizer.end_of(bin);
        for (int ch = 0; ch < 3; ++ch)
            C[ch] = dist / S(max_depth_scale);
        if (dist == 0) {
            C[0] = 0;
            C[1] = 1;
            C[2] = 0;
        }
        if (std::isnan(dist)) {
            C[0] = 1;
            C[1] = 0;
            C[2] = S(0.5);
        }
        if (dist < 0) {
            C[0] = 1;
            C[1] = S(0.5);
            C[2] = 0;
        }
        T = 0;
    }
    // Full mode leaves T = 1: the background is added in full (vol_raster_forward.cu:509-513)
    for (int ch = 0; ch < 3; ++ch)
        out[ch] = C[ch] + T * bg[ch];
}

} // namespace orc
