// SPDX-License-Identifier: GPL-3.0-or-later
// dgmr oracle -- TEST INFRASTRUCTURE ONLY (see oracle_math.hpp).
//
// Plain-loop restatement of the volumetric ray marcher:
//   preprocess        src/dgmr/vol_marcher_forward.cu:98-176
//   binning           vol_marcher_forward.cu:178-293 (depth cast to float for the key)
//   per-pixel march   vol_marcher_forward.cu:304-500
//   backward          src/dgmr/vol_marcher_backward.cu:115-384
// with the sampling machinery of src/dgmr/marching_steps.h:28-220 and the 1D
// ray restriction, cdf and adjoints of src/dgmr/math.h:412-468 and
// src/dgmr/grad/math.h:529-620.
#pragma once

#include <vector>

#include "oracle_math.hpp"
#include "oracle_splat.hpp"

namespace orc {

// vol_marcher.h:59-67
struct MarchConfig {
    static constexpr float filter_kernel_sd = 0.55f;
    static constexpr unsigned n_large_steps = 16;   // DensityArray capacity
    static constexpr unsigned n_small_steps = 128;  // bin borders per large step
    static constexpr unsigned n_steps_per_gaussian = 32;
    static constexpr float relevance_sigma = 3.f;
};

// ---------------------------------------------------------------- float ambiguity
// A pixel's march takes discrete decisions: section comparisons, a bin border
// snapping to a section end, mass and transmittance thresholds.  The FP64 run
// marks a pixel whose decisions compare values closer than a float32 run can
// resolve, because a float run may then take the other branch and sample
// differently from there on.  The watch is installed per pixel (thread-local,
// so OpenMP rows stay independent); with none installed the checks do nothing.
struct TieWatch {
    unsigned hit = 0;   // bit set of the decision kinds below that were near a tie
};
enum TieKind : unsigned {
    kTieSection = 1,       // section list comparisons (put / combine)
    kTieStep = 2,          // step sizes of overlapping sections
    kTieBorder = 4,        // a border snapping to a section end
    kTieMass = 8,          // mass on the ray against its thresholds
    kTieInBins = 16,       // a Gaussian's mass in the bins against 1e-4
    kTieTransmittance = 32 // transmittance against 1/255
};
inline thread_local TieWatch* tie_watch = nullptr;
inline thread_local unsigned tie_ignore = 0;   // kinds not recorded (decisions shared with the float32 run)

struct TieTolerance {
    // positions along the ray: float error of a border after up to 2 x 127
    // accumulated steps and of the 1D centre is below 1e-5 of max(|t|, 1)
    static constexpr double position = 2e-5;
    // step sizes 6 sd / 31 compared between sections (relative)
    static constexpr double step = 2e-5;
    // mass on the ray (weight x mass) against its thresholds (relative)
    static constexpr double mass = 1e-4;
    // transmittance against 1/255 (relative): the float cdf's absolute error
    // (1.5e-7 per evaluation) and skipped bins under 1e-5 move ln T by ~1e-5
    static constexpr double transmittance = 2e-4;
    // absolute error of a float mass in bins per unit of mass on the ray
    static constexpr double cdf_abs = 4e-7;
};

inline void watch_near(double a, double b, double rel, double floor, unsigned kind)
{
    if (!tie_watch || a == b || !std::isfinite(a) || !std::isfinite(b))
        return;
    const double scale = std::max({ std::abs(a), std::abs(b), floor });
    if (std::abs(a - b) <= rel * scale)
        tie_watch->hit |= kind & ~tie_ignore;
}
inline void watch_position(double a, double b, unsigned kind = kTieSection) { watch_near(a, b, TieTolerance::position, 1.0, kind); }
inline void watch_step(double a, double b) { watch_near(a, b, TieTolerance::step, 0.0, kTieStep); }
inline void watch_mass(double v, double threshold) { watch_near(v, threshold, TieTolerance::mass, 0.0, kTieMass); }
inline void watch_transmittance(double v) { watch_near(v, 1.0 / 255.0, TieTolerance::transmittance, 0.0, kTieTransmittance); }
inline void watch_abs(double v, double threshold, double tol)
{
    if (tie_watch && std::abs(v - threshold) <= tol)
        tie_watch->hit |= kTieInBins;
}

// ---------------------------------------------------------------- sampling sections
template <class S>
struct Section {
    S start, end, dt;
};

// marching_steps::DensityArray::combine, marching_steps.h:72-101: split two
// overlapping sections into (before, overlap, after); the overlap takes the
// smaller step.
template <class S>
inline void combine_sections(const Section<S>& a, const Section<S>& b, Section<S>* out)
{
    watch_position(a.start, b.start);
    watch_position(b.end, a.end);
    watch_step(a.dt, b.dt);
    if (a.start < b.start)
        out[0] = { a.start, std::min(a.end, b.start), a.dt };
    else
        out[0] = { b.start, std::min(b.end, a.start), b.dt };
    if (b.end < a.end)
        out[2] = { b.end, a.end, a.dt };
    else
        out[2] = { a.end, b.end, b.dt };
    out[1] = { out[0].end, out[2].start, std::min(a.dt, b.dt) };
}

// marching_steps::DensityArray, marching_steps.h:35-200: a sorted list of
// disjoint sections over [start, end) holding the minimum step size.  When the
// list is full, end() shrinks to the last section's end.  The insertion below
// performs the same steps in the same order as the reference (an in-place merge
// through a small ring buffer), because the exact section boundaries it leaves
// decide where the bin borders fall.
template <class S, unsigned CAP>
class SectionList {
public:
    explicit SectionList(S start) : start_(start), end_(std::numeric_limits<S>::infinity()) {}
    unsigned size() const { return size_; }
    S start() const { return start_; }
    S end() const { return end_; }
    const Section<S>& operator[](unsigned i) const { return data_[i]; }

    void put(Section<S> e)
    {
        watch_position(e.end, start_);
        watch_position(e.start, end_);
        if (e.end <= start_ || e.start >= end_)
            return;
        e.start = std::max(start_, e.start);
        e.end = std::min(end_, e.end);

        unsigned first = first_ending_after(e.start, 0);
        if (first >= size_) {
            if (first < CAP)
                data_[size_++] = e;
            else
                end_ = e.start;
            return;
        }
        unsigned stop = first_ending_after(e.end, first);
        if (stop < size_)
            watch_position(data_[stop].start, e.end);
        if (stop < size_ && data_[stop].start < e.end)
            ++stop;

        Section<S> ring[CAP];
        unsigned ring_r = 0, ring_w = 0;
        unsigned rd = first, wr = first, inserted = 0;
        while (rd < stop + inserted && wr < CAP) {
            if (rd < size_)
                ring[(ring_w++) % CAP] = data_[rd++];
            if (ring_r >= ring_w)
                break;
            const Section<S>& head = ring[ring_r % CAP];
            watch_position(head.end, e.start);
            watch_position(head.start, e.end);
            if (head.end <= e.start) {
                data_[wr++] = ring[(ring_r++) % CAP];
                continue;
            }
            if (head.start >= e.end)
                break;
            Section<S> parts[3];
            combine_sections(ring[(ring_r++) % CAP], e, parts);
            if (parts[0].dt == parts[1].dt) { // fold an equal-step prefix into the overlap
                parts[1].start = parts[0].start;
                parts[0].end = parts[0].start;
            }
            e = parts[2];
            if (parts[0].end - parts[0].start <= 0) {
                if (parts[1].dt == parts[2].dt) {
                    e.start = parts[1].start;
                    continue;
                }
                data_[wr++] = parts[1];
            } else {
                data_[wr++] = parts[0];
                if (parts[1].dt == parts[2].dt) {
                    e.start = parts[1].start;
                    continue;
                }
                ++inserted;
                ring[(--ring_r) % CAP] = parts[1];
            }
        }
        if (e.end - e.start > 0 && wr < CAP) {
            if (rd < size_)
                ring[(ring_w++) % CAP] = data_[rd++];
            data_[wr++] = e;
        }
        while (ring_w - ring_r && wr < CAP) {
            if (rd < size_)
                ring[(ring_w++) % CAP] = data_[rd++];
            data_[wr++] = ring[(ring_r++) % CAP];
        }
        size_ = wr + size_ - rd;
        if (size_ == CAP)
            end_ = std::min(end_, data_[size_ - 1].end);
    }

private:
    unsigned first_ending_after(S v, unsigned from) const
    {
        for (unsigned i = from; i < size_; ++i) {
            watch_position(data_[i].end, v);
            if (data_[i].end > v)
                return i;
        }
        return size_;
    }
    unsigned size_ = 0;
    Section<S> data_[CAP];
    S start_, end_;
};

// marching_steps::sample<n>, marching_steps.h:202-220
template <class S, unsigned CAP>
inline void sample_borders(const SectionList<S, CAP>& sec, S* borders, unsigned n)
{
    if (sec.size() == 0) {
        for (unsigned i = 0; i < n; ++i)
            borders[i] = 0;
        return;
    }
    borders[0] = sec[0].start;
    unsigned cur = 0;
    for (unsigned i = 1; i < n; ++i) {
        if (cur < sec.size() - 1)
            watch_position(borders[i - 1] + sec[cur].dt, sec[cur].end, kTieBorder);
        const S v = std::min(sec[cur].end, borders[i - 1] + sec[cur].dt);
        borders[i] = v;
        if (cur < sec.size() - 1 && v == sec[cur].end)
            ++cur;
    }
}

// ---------------------------------------------------------------- 1D restriction
template <class S>
struct Gaussian1d {
    S weight, centre, var;
};

// the 3D Gaussian N(mu, A^-1) restricted to the ray o + t d: variance
// 1/(d^T A d), centre (d^T A (mu - o)) / (d^T A d), and weight = the integral
// of the normalised 3D density along the ray (stroke::gaussian::
// intersect_with_ray_inv_C, used at math.h:436 and vol_marcher_forward.cu:378)
template <class S>
inline Gaussian1d<S> intersect_ray(const S* mu, const S* A, const S* o, const S* d)
{
    S M[3][3];
    sym3_to_mat(A, M);
    S Ad[3], m[3];
    for (int i = 0; i < 3; ++i) {
        Ad[i] = (M[i][0] * d[0] + M[i][1] * d[1]) + M[i][2] * d[2];
        m[i] = mu[i] - o[i];
    }
    const S q = dot3(d, Ad);
    const S C = S(1) / q;
    const S centre = dot3(Ad, m) * C;
    S r[3];
    for (int i = 0; i < 3; ++i)
        r[i] = (o[i] + centre * d[i]) - mu[i];
    S Ar[3];
    for (int i = 0; i < 3; ++i)
        Ar[i] = (M[i][0] * r[0] + M[i][1] * r[1]) + M[i][2] * r[2];
    const S E = dot3(r, Ar);
    const S norm3 = std::sqrt(det3(M)) / S(std::sqrt(std::pow(2 * kPi, 3.0)));
    const S norm1 = S(1) / std::sqrt(S(2 * kPi) * C);
    return { norm3 / norm1 * std::exp(S(-0.5) * E), centre, C };
}

// adjoint of intersect_ray w.r.t. mu and the stored inverse covariance
template <class S>
inline void intersect_ray_grad(const S* mu, const S* A, const S* o, const S* d, S g_w, S g_centre, S g_var, S* g_mu, S* g_A)
{
    S M[3][3];
    sym3_to_mat(A, M);
    S Ad[3], m[3];
    for (int i = 0; i < 3; ++i) {
        Ad[i] = (M[i][0] * d[0] + M[i][1] * d[1]) + M[i][2] * d[2];
        m[i] = mu[i] - o[i];
    }
    const S q = dot3(d, Ad);
    const S C = S(1) / q;
    const S k = dot3(Ad, m);
    const S centre = k * C;
    S r[3], Ar[3];
    for (int i = 0; i < 3; ++i)
        r[i] = (o[i] + centre * d[i]) - mu[i];
    for (int i = 0; i < 3; ++i)
        Ar[i] = (M[i][0] * r[0] + M[i][1] * r[1]) + M[i][2] * r[2];
    const S E = dot3(r, Ar);
    const S W = std::sqrt(det3(M)) / S(std::sqrt(std::pow(2 * kPi, 3.0))) * std::sqrt(S(2 * kPi) * C) * std::exp(S(-0.5) * E);
    for (int i = 0; i < 3; ++i)
        g_mu[i] = g_w * W * Ar[i] + g_centre * Ad[i] * C;
    S Ainv[6];
    sym3_inverse(A, Ainv);
    S Ai[3][3];
    sym3_to_mat(Ainv, Ai);
    S G[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const S ddT = d[i] * d[j];
            G[i][j] = g_w * W * S(0.5) * (Ai[i][j] - C * ddT - r[i] * r[j])
                + g_centre * (C * S(0.5) * (d[i] * m[j] + m[i] * d[j]) - k * C * C * ddT)
                - g_var * C * C * ddT;
        }
    mat_grad_to_sym3(G, g_A);
}

// stroke::gaussian::cdf_inv_SD: standard normal cdf of (x - centre) / sd
template <class S>
inline S cdf_inv_sd(S centre, S inv_sd, S x) { return S(0.5) * (S(1) + std::erf((x - centre) * inv_sd * S(0.70710678118654752440))); }

template <class S>
inline void cdf_inv_sd_grad(S centre, S inv_sd, S x, S g, S* g_centre, S* g_inv_sd)
{
    const S z = (x - centre) * inv_sd;
    const S pdf = std::exp(S(-0.5) * z * z) * S(0.39894228040143267794);
    *g_centre += g * (-inv_sd * pdf);
    *g_inv_sd += g * ((x - centre) * pdf);
}

constexpr unsigned kBins = MarchConfig::n_small_steps - 1;

// math::sample_gaussian, math.h:430-468 (bins: [127][4] = rgb * m, m)
template <class S>
inline void sample_gaussian(S mass, const S* rgb, const S* mu, const S* A, const S* o, const S* d, const S* borders, S (*bins)[4])
{
    const Gaussian1d<S> g = intersect_ray(mu, A, o, d);
    const S sd = std::sqrt(g.var);
    const S inv_sd = S(1) / sd;
    const S mor = g.weight * mass;
    if (std::isnan(g.centre))
        return;
    watch_mass(mor, 1.0f / 255.f);
    watch_mass(mor, 1000.0);
    if (mor < S(1.0f / 255.f) || mor > S(1000))
        return;
    watch_mass(sd, 0.0001);
    if (sd <= S(0.0001) || std::isnan(sd) || std::isnan(mor) || mor > S(100000))
        return;
    const S in_bins = mor * (cdf_inv_sd(g.centre, inv_sd, borders[kBins]) - cdf_inv_sd(g.centre, inv_sd, borders[0]));
    watch_abs(in_bins, 0.0001f, 0.0001 * TieTolerance::mass + double(mor) * TieTolerance::cdf_abs);
    if (in_bins < S(0.0001f))
        return;
    S c0 = cdf_inv_sd(g.centre, inv_sd, borders[0]);
    for (unsigned k = 0; k < kBins; ++k) {
        const S c1 = cdf_inv_sd(g.centre, inv_sd, borders[k + 1]);
        const S m = std::max(S(0), (c1 - c0) * mor);
        c0 = c1;
        // a bin's skip below 1e-5 is not watched: flipping it moves the
        // image by at most 1e-5 x colour x transmittance
        if (m < S(0.00001f))
            continue;
        bins[k][0] += rgb[0] * m;
        bins[k][1] += rgb[1] * m;
        bins[k][2] += rgb[2] * m;
        bins[k][3] += m;
    }
}

// grad::sample_gaussian, grad/math.h:559-620
template <class S>
inline void sample_gaussian_grad(S mass, const S* rgb, const S* mu, const S* A, const S* o, const S* d, const S* borders,
    const S (*g_bins)[4], S* g_mass, S* g_rgb, S* g_mu, S* g_A)
{
    *g_mass = 0;
    for (int i = 0; i < 3; ++i)
        g_rgb[i] = g_mu[i] = 0;
    for (int i = 0; i < 6; ++i)
        g_A[i] = 0;
    const Gaussian1d<S> g = intersect_ray(mu, A, o, d);
    const S sd = std::sqrt(g.var);
    const S inv_sd = S(1) / sd;
    const S mor = g.weight * mass;
    if (std::isnan(g.centre))
        return;
    if (mor < S(1.0f / 255.f) || mor > S(1000))
        return;
    if (sd <= S(0.0001) || std::isnan(sd) || std::isnan(mor) || mor > S(100000))
        return;
    const S in_bins = mor * (cdf_inv_sd(g.centre, inv_sd, borders[kBins]) - cdf_inv_sd(g.centre, inv_sd, borders[0]));
    if (in_bins < S(0.0001f))
        return;
    S g_centre = 0, g_inv_sd = 0, g_mor = 0;
    S c0 = cdf_inv_sd(g.centre, inv_sd, borders[0]);
    for (unsigned k = 0; k < kBins; ++k) {
        const S c1 = cdf_inv_sd(g.centre, inv_sd, borders[k + 1]);
        const S delta = c1 - c0;
        const S m = std::max(S(0), delta * mor);
        c0 = c1;
        if (m < S(0.00001f))
            continue;
        for (int i = 0; i < 3; ++i)
            g_rgb[i] += g_bins[k][i] * m;
        const S g_m = ((g_bins[k][0] * rgb[0] + g_bins[k][1] * rgb[1]) + g_bins[k][2] * rgb[2]) + g_bins[k][3];
        if (delta * mor <= 0)
            continue;
        g_mor += g_m * delta;
        const S g_delta = g_m * mor;
        cdf_inv_sd_grad(g.centre, inv_sd, borders[k], -g_delta, &g_centre, &g_inv_sd);
        cdf_inv_sd_grad(g.centre, inv_sd, borders[k + 1], g_delta, &g_centre, &g_inv_sd);
    }
    const S g_w = g_mor * mass;
    *g_mass = g_mor * g.weight;
    const S g_sd = -g_inv_sd / (sd * sd);
    const S g_var = g_sd / (S(2) * sd);
    intersect_ray_grad(mu, A, o, d, g_w, g_centre, g_var, g_mu, g_A);
}

// math::integrate_bins, math.h:412-428
template <class S>
inline void integrate_bins(S* C, S* T, const S (*bins)[4])
{
    for (unsigned k = 0; k < kBins; ++k) {
        const S tk = std::exp(-bins[k][3]);
        for (int i = 0; i < 3; ++i)
            C[i] += bins[k][i] * *T;
        *T *= tk;
        watch_transmittance(*T);
        if (*T < S(1. / 255.))
            break;
    }
}

// grad::integrate_bins, grad/math.h:529-557: C enters as the remaining colour
// (normalised by the current transparency) and leaves advanced past the bins.
template <class S>
inline void integrate_bins_grad(S* C, S* T, S T_final, const S (*bins)[4], const S* gC, S gT, S (*g_bins)[4])
{
    for (unsigned k = 0; k < kBins; ++k)
        g_bins[k][0] = g_bins[k][1] = g_bins[k][2] = g_bins[k][3] = 0;
    for (unsigned k = 0; k < kBins; ++k) {
        for (int i = 0; i < 3; ++i)
            g_bins[k][i] += *T * gC[i];
        const S tk = std::exp(-bins[k][3]);
        const S inv_tk = std::min(S(255), S(1) / tk);
        for (int i = 0; i < 3; ++i)
            C[i] = (C[i] - bins[k][i]) * inv_tk;
        const S g_tk = T_final * inv_tk * gT + ((gC[0] * C[0] + gC[1] * C[1]) + gC[2] * C[2]) * *T;
        g_bins[k][3] -= g_tk * tk;
        *T *= tk;
        if (*T < S(1. / 255.))
            break;
    }
}

// ---------------------------------------------------------------- pipeline
template <class S>
struct MarchGeometry {
    std::vector<int32_t> rect;    // [n,2]
    std::vector<uint32_t> tiles;  // [n]
    std::vector<S> xy;            // [n,2]
    std::vector<S> depth;         // [n]
    std::vector<float> key_depth; // [n]
    std::vector<S> rgb;           // [n,3]
    std::vector<uint8_t> clamped; // [n,3]
    std::vector<S> inv_cov;       // [n,6] inverse of the filtered 3D covariance
    std::vector<S> mass;          // [n]
};

template <class S>
inline S aa_distance_multiplier(const Camera<S>& cam)
{
    return (S(MarchConfig::filter_kernel_sd) * cam.tan_fovx * S(2)) / S(cam.width);
}

// vol_marcher_forward.cu:98-176
template <class S>
inline void march_preprocess(const Scene<S>& sc, const Camera<S>& cam, int tile_w, int tile_h, MarchGeometry<S>* g,
    const MarchGeometry<float>* decide = nullptr)
{
    const int64_t n = sc.n;
    const int grid_x = (cam.width + tile_w - 1) / tile_w, grid_y = (cam.height + tile_h - 1) / tile_h;
    g->rect.assign(2 * n, 0);
    g->tiles.assign(n, 0);
    g->xy.assign(2 * n, 0);
    g->depth.assign(n, 0);
    g->key_depth.assign(n, 0);
    g->rgb.assign(3 * n, 0);
    g->clamped.assign(3 * n, 0);
    g->inv_cov.assign(6 * n, 0);
    g->mass.assign(n, 0);
    const S aa = aa_distance_multiplier(cam);
    S W[3][3];
    view_rotation(cam, W);
    for (int64_t i = 0; i < n; ++i) {
        const S* mu = sc.means + 3 * i;
        S t[3];
        transform_point(cam.view, mu, t, 3);
        if (decide ? decide->tiles[i] == 0 : t[2] < S(0.2f))
            continue;
        const S w = sc.opacity[i];
        const S s[3] = { sc.scales[3 * i] * sc.scale_multiplier, sc.scales[3 * i + 1] * sc.scale_multiplier, sc.scales[3 * i + 2] * sc.scale_multiplier };
        const S* q = sc.rotations + 4 * i;
        const S dv[3] = { cam.pos[0] - mu[0], cam.pos[1] - mu[1], cam.pos[2] - mu[2] };
        const S dist = length3(dv);
        S cov[6];
        compute_cov(s, q, cov);
        const S f = dist * aa;
        const S f2 = f * f;
        const S cov_f[6] = { cov[0] + f2, cov[1], cov[2], cov[3] + f2, cov[4], cov[5] + f2 };
        S c2[3];
        affine_transform_and_cut(cov, W, c2);
        const S mass = w * integrate_exp_2d(c2);
        if (!decide && mass <= 0)
            continue;
        const Splat2d<S> sp = splat(Formulation::Opacity, w, mu, s, q, cam, S(MarchConfig::filter_kernel_sd * MarchConfig::filter_kernel_sd));
        if (decide) {
            g->rect[2 * i] = decide->rect[2 * i];
            g->rect[2 * i + 1] = decide->rect[2 * i + 1];
            g->tiles[i] = decide->tiles[i];
        } else {
            const auto iso = [](S wt, S var, S level) {
                const S v = -2 * portable_log(level / wt);
                if (v <= 0)
                    return S(0);
                return std::sqrt(v * var);
            };
            const int ext_x = guarded_int(std::ceil(iso(sp.weight, sp.cov[0], S(1.f / 255.f))));
            const int ext_y = guarded_int(std::ceil(iso(sp.weight, sp.cov[2], S(1.f / 255.f))));
            g->rect[2 * i] = ext_x;
            g->rect[2 * i + 1] = ext_y;
            int x0, y0, x1, y1;
            get_rect(sp.xy[0], sp.xy[1], ext_x, ext_y, grid_x, grid_y, tile_w, tile_h, &x0, &y0, &x1, &y1);
            const uint32_t touched = uint32_t(x1 - x0) * uint32_t(y1 - y0);
            if (touched == 0)
                continue;
            g->tiles[i] = touched;
        }
        g->xy[2 * i] = sp.xy[0];
        g->xy[2 * i + 1] = sp.xy[1];
        sym3_inverse(cov_f, &g->inv_cov[6 * i]);
        g->depth[i] = dist - std::max(s[0], std::max(s[1], s[2])) * S(MarchConfig::relevance_sigma) / S(2);
        g->key_depth[i] = decide ? decide->key_depth[i] : float(g->depth[i]);
        const S dir[3] = { (mu[0] - cam.pos[0]) / dist, (mu[1] - cam.pos[1]) / dist, (mu[2] - cam.pos[2]) / dist };
        bool ok[3];
        sh_to_colour(sc.sh + 48 * i, sc.sh_degree, dir, &g->rgb[3 * i], ok);
        for (int c = 0; c < 3; ++c)
            g->clamped[3 * i + c] = ok[c];
        g->mass[i] = mass;
    }
}

// 4x4 inverse (Gauss-Jordan, partial pivoting) in double
inline bool invert4(const double* m, double* out)
{
    double a[4][8];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 8; ++c)
            a[r][c] = c < 4 ? m[4 * r + c] : (c - 4 == r ? 1.0 : 0.0);
    for (int c = 0; c < 4; ++c) {
        int piv = c;
        for (int r = c + 1; r < 4; ++r)
            if (std::abs(a[r][c]) > std::abs(a[piv][c]))
                piv = r;
        if (a[piv][c] == 0)
            return false;
        if (piv != c)
            for (int k = 0; k < 8; ++k)
                std::swap(a[piv][k], a[c][k]);
        const double inv = 1.0 / a[c][c];
        for (int k = 0; k < 8; ++k)
            a[c][k] *= inv;
        for (int r = 0; r < 4; ++r) {
            if (r == c)
                continue;
            const double f = a[r][c];
            for (int k = 0; k < 8; ++k)
                a[r][k] -= f * a[c][k];
        }
    }
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c)
            out[4 * r + c] = a[r][c + 4];
    return true;
}

// per-pixel ray, vol_marcher_forward.cu:308-315: pixel centre -> NDC (z = 0)
// -> world through the inverse projection -> normalised direction from the camera
template <class S>
inline void pixel_ray(const Camera<S>& cam, const S* inv_proj, int px, int py, S* dir)
{
    const S pix[2] = { S(px) + S(0.5), S(py) + S(0.5) };
    const S ndc[3] = { pix[0] * S(2) / S(cam.width) - S(1), pix[1] * S(2) / S(cam.height) - S(1), S(0) };
    S v[4];
    transform_point(inv_proj, ndc, v, 4);
    S p[3];
    for (int i = 0; i < 3; ++i)
        p[i] = v[i] / v[3] - cam.pos[i];
    const S inv_len = S(1) / std::sqrt(dot3(p, p));
    for (int i = 0; i < 3; ++i)
        dir[i] = p[i] * inv_len;
}

// ForwardData::RenderMode (vol_marcher.h:48-53)
enum class RenderMode { Full = 0, Bins = 1, Depth = 2 };

struct MarchOptions {
    int max_large_steps = 0;              // 0: unbounded, as the reference
    int n_frozen = 0;                     // >0: fixed sample sections (gradient-check mode)
    const double* frozen = nullptr;       // [n_frozen][3] start, end, dt
    RenderMode mode = RenderMode::Full;   // debug_render_mode (forward only)
    int debug_bin = -1;                   // debug_render_bin; -1 (as unsigned) selects the last bin
    double max_depth = 20.0;              // ForwardData::max_depth
};

// pass 1 of one large step: build the sample sections from the tile's Gaussians
template <class S>
inline void build_sections(const MarchGeometry<S>& g, const Binning& b, int tile, const S* o, const S* d, const Scene<S>& sc,
    const MarchOptions& opt, SectionList<S, MarchConfig::n_large_steps>* sec)
{
    if (opt.n_frozen > 0) {
        for (int k = 0; k < opt.n_frozen; ++k)
            sec->put({ S(opt.frozen[3 * k]), S(opt.frozen[3 * k + 1]), S(opt.frozen[3 * k + 2]) });
        return;
    }
    for (uint32_t k = b.ranges[2 * tile]; k < b.ranges[2 * tile + 1]; ++k) {
        const uint32_t id = b.point_list[k];
        const Gaussian1d<S> g1 = intersect_ray(sc.means + 3 * id, &g.inv_cov[6 * id], o, d);
        const S sd = std::sqrt(g1.var);
        watch_position(sec->end(), g.depth[id]);
        if (sec->end() < g.depth[id])
            break;
        const S mor = g1.weight * g.mass[id];
        watch_mass(mor, 1.0f / 255.f);
        watch_mass(mor, 1000.0);
        if (mor <= S(1.0f / 255.f) || mor > S(1000))
            continue;
        if (g1.var <= 0)
            continue;
        if (std::isnan(g1.centre))
            continue;
        const S half = sd * S(MarchConfig::relevance_sigma);
        sec->put({ g1.centre - half, g1.centre + half, (sd * S(MarchConfig::relevance_sigma) * S(2)) / S(MarchConfig::n_steps_per_gaussian - 1) });
    }
}

// Ray and bin placement taken from the float32 pipeline: the pixel's ray
// direction, and the bin borders of each large step (pass 1 and
// marching_steps::sample) from the float32 geometry and section arithmetic.
// A double march that uses them samples exactly where a float32 march
// samples, so the two differ only by the rounding of what they compute along
// that ray on those bins.  Both placements are discontinuous in rounding:
//  * the ray unprojects NDC z = 0 (vol_marcher_forward.cu:308-315), a point
//    about 2 znear from the camera, and subtracts the camera position, so a
//    float32 direction carries ~1e-5 relative error by construction, which
//    moves the ray ~1e-4 at the Gaussians and their mass on it by ~1e-3;
//  * the border walk lands on section ends by design (a section is 31 steps
//    long), so a float32 and a double march that each place their own borders
//    split such ties by rounding noise and can sample a large step with one
//    bin more or less.
struct FloatSampling {
    const MarchGeometry<float>* g = nullptr;
    const Scene<float>* sc = nullptr;
    const Camera<float>* cam = nullptr;
    const float* inv_proj = nullptr;
};

template <class S>
inline unsigned large_step_borders(const MarchGeometry<S>& g, const Binning& b, const Scene<S>& sc, int tile, const S* o, const S* d,
    const MarchOptions& opt, S start, S* borders, S* end = nullptr)
{
    SectionList<S, MarchConfig::n_large_steps> sec(start);
    build_sections(g, b, tile, o, d, sc, opt, &sec);
    sample_borders(sec, borders, MarchConfig::n_small_steps);
    if (end)
        *end = sec.end();
    return sec.size();
}

template <class S>
inline void float_ray(const FloatSampling& fs, int px, int py, S* o, S* d)
{
    float df[3];
    pixel_ray(*fs.cam, fs.inv_proj, px, py, df);
    for (int i = 0; i < 3; ++i) {
        o[i] = S(fs.cam->pos[i]);
        d[i] = S(df[i]);
    }
}

template <class S>
inline unsigned large_step_borders(const FloatSampling& fs, const Binning& b, int tile, int px, int py, const MarchOptions& opt, S start,
    S* borders, S* end = nullptr)
{
    float d[3], fb[MarchConfig::n_small_steps];
    pixel_ray(*fs.cam, fs.inv_proj, px, py, d);
    // the section and border decisions are the float32 run's own: a float32
    // march that computes them in the same order takes the same ones
    tie_ignore = kTieSection | kTieStep | kTieBorder;
    float fend = 0;
    const unsigned n = large_step_borders(*fs.g, b, *fs.sc, tile, fs.cam->pos, d, opt, float(start), fb, &fend);
    tie_ignore = 0;
    for (unsigned i = 0; i < MarchConfig::n_small_steps; ++i)
        borders[i] = S(fb[i]);
    if (end)
        *end = S(fend);
    return n;
}

// one pixel of the forward march, vol_marcher_forward.cu:340-499; with `fs`
// the bin borders come from the float32 pipeline (see FloatSampling)
template <class S>
inline void march_pixel(const MarchGeometry<S>& g, const Binning& b, const Scene<S>& sc, const Camera<S>& cam,
    const S* inv_proj, int px, int py, const MarchOptions& opt, S* C, S* T_out, int* n_large, const FloatSampling* fs = nullptr)
{
    S o[3] = { cam.pos[0], cam.pos[1], cam.pos[2] };
    S d[3];
    pixel_ray(cam, inv_proj, px, py, d);
    if (fs)
        float_ray(*fs, px, py, o, d);
    const int tile = (py / b.tile_h) * b.grid_x + (px / b.tile_w);
    S T = 1;
    C[0] = C[1] = C[2] = 0;
    S start = 0;
    S distance = 0;
    int steps = 0;
    std::vector<S> borders(MarchConfig::n_small_steps);
    std::vector<S> binsv(4 * kBins);
    S (*bins)[4] = reinterpret_cast<S (*)[4]>(binsv.data());
    while (true) {
        S section_end = 0;
        const unsigned n_sections = fs ? large_step_borders(*fs, b, tile, px, py, opt, start, borders.data(), &section_end)
                                       : large_step_borders(g, b, sc, tile, o, d, opt, start, borders.data(), &section_end);
        std::fill(binsv.begin(), binsv.end(), S(0));
        for (uint32_t k = b.ranges[2 * tile]; k < b.ranges[2 * tile + 1]; ++k) {
            const uint32_t id = b.point_list[k];
            sample_gaussian(g.mass[id], &g.rgb[3 * id], sc.means + 3 * id, &g.inv_cov[6 * id], o, d, borders.data(), bins);
        }
        // the blend of vol_marcher_forward.cu:443-484 by debug_render_mode
        if (opt.mode == RenderMode::Bins) {
            const unsigned bin = std::min(unsigned(opt.debug_bin), kBins - 1);
            const S mass = ((bins[bin][0] + bins[bin][1]) + bins[bin][2]) + bins[bin][3];
            const S v = mass * S(opt.max_depth);
            C[0] = C[1] = C[2] = v;
            if (mass == 0) {
                C[0] = 0;
                C[1] = 1;
                C[2] = 0;
            }
            if (std::isnan(mass)) {
                C[0] = 1;
                C[1] = 0;
                C[2] = S(0.5);
            }
            if (mass < 0) {
                C[0] = 1;
                C[1] = 0;
                C[2] = 0;
            }
            T = 0;
        } else if (opt.mode == RenderMode::Depth) {
            for (unsigned k = 0; k < kBins; ++k)
                T *= std::exp(-bins[k][3]);
            // the section list's end: +inf unless the list filled up
            distance = std::max(section_end, distance);
            C[0] = C[1] = C[2] = distance / S(opt.max_depth);
        } else {
            integrate_bins(C, &T, bins);
        }
        ++steps;
        if (n_sections == 0 || T < S(1.f / 255.f))
            break;
        if (opt.max_large_steps > 0 && steps >= opt.max_large_steps)
            break;
        start = borders[MarchConfig::n_small_steps - 1];
    }
    *T_out = T;
    if (n_large)
        *n_large = steps;
}

template <class S>
struct MarchRenderGrads {
    std::vector<S> mass;    // [n]
    std::vector<S> rgb;     // [n,3]
    std::vector<S> means;   // [n,3] direct contribution of the ray restriction
    std::vector<S> inv_cov; // [n,6]
};

// one pixel of the backward march, vol_marcher_backward.cu:126-318 (`fs` as
// in march_pixel)
template <class S>
inline void march_pixel_backward(const MarchGeometry<S>& g, const Binning& b, const Scene<S>& sc, const Camera<S>& cam,
    const S* inv_proj, int px, int py, const MarchOptions& opt, const S* fb_px, S T_final, const S* gC, const S* bg,
    MarchRenderGrads<S>* gr, const FloatSampling* fs = nullptr)
{
    S o[3] = { cam.pos[0], cam.pos[1], cam.pos[2] };
    S d[3];
    pixel_ray(cam, inv_proj, px, py, d);
    if (fs)
        float_ray(*fs, px, py, o, d);
    const int tile = (py / b.tile_h) * b.grid_x + (px / b.tile_w);
    S C[3];
    for (int i = 0; i < 3; ++i)
        C[i] = fb_px[i] - T_final * bg[i];
    const S gT = (gC[0] * bg[0] + gC[1] * bg[1]) + gC[2] * bg[2];
    S T = 1;
    S start = 0;
    int steps = 0;
    std::vector<S> borders(MarchConfig::n_small_steps);
    std::vector<S> binsv(4 * kBins), gbinsv(4 * kBins);
    S (*bins)[4] = reinterpret_cast<S (*)[4]>(binsv.data());
    S (*gbins)[4] = reinterpret_cast<S (*)[4]>(gbinsv.data());
    while (true) {
        const unsigned n_sections = fs ? large_step_borders(*fs, b, tile, px, py, opt, start, borders.data())
                                       : large_step_borders(g, b, sc, tile, o, d, opt, start, borders.data());
        std::fill(binsv.begin(), binsv.end(), S(0));
        for (uint32_t k = b.ranges[2 * tile]; k < b.ranges[2 * tile + 1]; ++k) {
            const uint32_t id = b.point_list[k];
            sample_gaussian(g.mass[id], &g.rgb[3 * id], sc.means + 3 * id, &g.inv_cov[6 * id], o, d, borders.data(), bins);
        }
        integrate_bins_grad(C, &T, T_final, bins, gC, gT, gbins);
        for (uint32_t k = b.ranges[2 * tile]; k < b.ranges[2 * tile + 1]; ++k) {
            const uint32_t id = b.point_list[k];
            S gm, grgb[3], gmu[3], gA[6];
            sample_gaussian_grad(g.mass[id], &g.rgb[3 * id], sc.means + 3 * id, &g.inv_cov[6 * id], o, d, borders.data(), gbins, &gm, grgb, gmu, gA);
            gr->mass[id] += gm;
            for (int i = 0; i < 3; ++i) {
                gr->rgb[3 * id + i] += grgb[i];
                gr->means[3 * id + i] += gmu[i];
            }
            for (int i = 0; i < 6; ++i)
                gr->inv_cov[6 * id + i] += gA[i];
        }
        ++steps;
        if (n_sections == 0 || T < S(1.f / 255.f))
            break;
        if (opt.max_large_steps > 0 && steps >= opt.max_large_steps)
            break;
        start = borders[MarchConfig::n_small_steps - 1];
    }
}

// per-Gaussian adjoint of the preprocess, vol_marcher_backward.cu:321-384
template <class S>
inline void march_preprocess_backward(const Scene<S>& sc, const Camera<S>& cam, const MarchGeometry<S>& g,
    const MarchRenderGrads<S>& gr, ParamGrads<S>* out)
{
    const int64_t n = sc.n;
    out->means.assign(3 * n, 0);
    out->scales.assign(3 * n, 0);
    out->rotations.assign(4 * n, 0);
    out->opacity.assign(n, 0);
    out->sh.assign(48 * n, 0);
    const S aa = aa_distance_multiplier(cam);
    S W[3][3];
    view_rotation(cam, W);
    for (int64_t i = 0; i < n; ++i) {
        for (int a = 0; a < 3; ++a)
            out->means[3 * i + a] = gr.means[3 * i + a];
        if (g.tiles[i] == 0)
            continue;
        const S* mu = sc.means + 3 * i;
        const S w = sc.opacity[i];
        const S m = sc.scale_multiplier;
        const S s[3] = { sc.scales[3 * i] * m, sc.scales[3 * i + 1] * m, sc.scales[3 * i + 2] * m };
        const S* q = sc.rotations + 4 * i;
        const S delta[3] = { mu[0] - cam.pos[0], mu[1] - cam.pos[1], mu[2] - cam.pos[2] };
        const S dist = length3(delta);
        S cov[6];
        compute_cov(s, q, cov);
        const S f = dist * aa;
        S c2[3];
        affine_transform_and_cut(cov, W, c2);
        const S dir[3] = { delta[0] / dist, delta[1] / dist, delta[2] / dist };
        bool ok[3] = { g.clamped[3 * i] != 0, g.clamped[3 * i + 1] != 0, g.clamped[3 * i + 2] != 0 };
        S g_dir[3];
        sh_to_colour_grad(sc.sh + 48 * i, sc.sh_degree, dir, &gr.rgb[3 * i], ok, &out->sh[48 * i], g_dir);
        S g_mu[3];
        S g_dist = -dot3(g_dir, delta) / (dist * dist);
        for (int a = 0; a < 3; ++a)
            g_mu[a] = g_dir[a] / dist;
        // inverse of the filtered covariance
        S g_covf[6];
        sym3_inverse_grad(&g.inv_cov[6 * i], &gr.inv_cov[6 * i], g_covf);
        // mass = w * 2 pi sqrt(det(cut(W Sigma W^T)))
        const S g_mass = gr.mass[i];
        const S g_w = g_mass * integrate_exp_2d(c2);
        S g_c2[3];
        integrate_exp_2d_grad(c2, g_mass * w, g_c2);
        S g_cov[6], g_Wm[3][3];
        affine_transform_and_cut_grad(cov, W, g_c2, g_cov, g_Wm);
        for (int a = 0; a < 6; ++a)
            g_cov[a] += g_covf[a];
        const S g_f2 = g_covf[0] + g_covf[3] + g_covf[5];
        g_dist += g_f2 * S(2) * f * aa;
        S g_s[3], g_q[4];
        compute_cov_grad(s, q, g_cov, g_s, g_q);
        for (int a = 0; a < 3; ++a)
            g_mu[a] += g_dist * delta[a] / dist;
        for (int a = 0; a < 3; ++a) {
            out->means[3 * i + a] += g_mu[a];
            out->scales[3 * i + a] = g_s[a] * m;
        }
        for (int a = 0; a < 4; ++a)
            out->rotations[4 * i + a] = g_q[a];
        out->opacity[i] = g_w;
    }
}

} // namespace orc
