// SPDX-License-Identifier: GPL-3.0-or-later
// Device pieces of the volumetric ray marcher: the sample-section list, the
// border walk, the 1D restriction of a 3D Gaussian to a ray, the normal cdf and
// their adjoints.  References: src/dgmr/marching_steps.h:28-220,
// src/dgmr/math.h:412-468, src/dgmr/grad/math.h:529-620.
#pragma once

#include "device_grad.h"
#include "device_math.h"

namespace dgmr {

// vol_marcher.h:59-67
constexpr int kSections = 16;      // n_large_steps: section-list capacity
constexpr int kBorders = 128;      // n_small_steps
constexpr int kBins = kBorders - 1;
constexpr int kStepsPerGaussian = 32;
constexpr float kRelevanceSigma = 3.f;
constexpr float kFilterSD = 0.55f;

template <class S>
struct Section {
    S start, end, dt;
};

// Storage of a pixel's sections: in the thread's own array (private memory
// once indexed at run time), or in LDS, where thread t keeps field f of its
// section i at base[(3 i + f) NT] (base = the kernel's array + t, so the
// lanes of a wave touch consecutive words).
template <class S, int CAP>
struct SectionRegs {
    Section<S> d[CAP];
    __device__ __forceinline__ Section<S> get(int i) const { return d[i]; }
    __device__ __forceinline__ S start(int i) const { return d[i].start; }
    __device__ __forceinline__ S end(int i) const { return d[i].end; }
    __device__ __forceinline__ S dt(int i) const { return d[i].dt; }
    __device__ __forceinline__ void set(int i, const Section<S>& v) { d[i] = v; }
};

template <int NT>
struct SectionLds {
    float* base;
    __device__ __forceinline__ Section<float> get(int i) const { return { base[(3 * i) * NT], base[(3 * i + 1) * NT], base[(3 * i + 2) * NT] }; }
    __device__ __forceinline__ float start(int i) const { return base[(3 * i) * NT]; }
    __device__ __forceinline__ float end(int i) const { return base[(3 * i + 1) * NT]; }
    __device__ __forceinline__ float dt(int i) const { return base[(3 * i + 2) * NT]; }
    __device__ __forceinline__ void set(int i, const Section<float>& v)
    {
        base[(3 * i) * NT] = v.start;
        base[(3 * i + 1) * NT] = v.end;
        base[(3 * i + 2) * NT] = v.dt;
    }
};

// marching_steps::DensityArray (marching_steps.h:35-200): a sorted list of
// disjoint sections over [start, end) holding the minimum step; when full,
// end() shrinks to the last section's end.  put() follows the reference's
// insertion step for step (in-place merge through a small ring buffer),
// because the boundaries it leaves decide where the bin borders fall.
template <class S, int CAP, class Store = SectionRegs<S, CAP>>
struct SectionList {
    Store data;
    int size;
    S start_, end_;

    __device__ __forceinline__ void reset(S start)
    {
        size = 0;
        start_ = start;
        end_ = S(__builtin_huge_val());
    }

    __device__ __forceinline__ int first_ending_after(S v, int from) const
    {
        for (int i = from; i < size; ++i)
            if (data.end(i) > v)
                return i;
        return size;
    }

    __device__ __forceinline__ static void combine(const Section<S>& a, const Section<S>& b, Section<S>* out)
    {
        if (a.start < b.start)
            out[0] = { a.start, dmin(a.end, b.start), a.dt };
        else
            out[0] = { b.start, dmin(b.end, a.start), b.dt };
        if (b.end < a.end)
            out[2] = { b.end, a.end, a.dt };
        else
            out[2] = { a.end, b.end, b.dt };
        out[1] = { out[0].end, out[2].start, dmin(a.dt, b.dt) };
    }

    __device__ void put(Section<S> e)
    {
        if (e.end <= start_ || e.start >= end_)
            return;
        e.start = dmax(start_, e.start);
        e.end = dmin(end_, e.end);
        const int first = first_ending_after(e.start, 0);
        if (first >= size) {
            if (first < CAP)
                data.set(size++, e);
            else
                end_ = e.start;
            return;
        }
        int stop = first_ending_after(e.end, first);
        if (stop < size && data.start(stop) < e.end)
            ++stop;
        Section<S> ring[CAP];
        unsigned ring_r = 0, ring_w = 0;
        int rd = first, wr = first, inserted = 0;
        while (rd < stop + inserted && wr < CAP) {
            if (rd < size)
                ring[(ring_w++) % CAP] = data.get(rd++);
            if (ring_r >= ring_w)
                break;
            const Section<S> head = ring[ring_r % CAP];
            if (head.end <= e.start) {
                data.set(wr++, ring[(ring_r++) % CAP]);
                continue;
            }
            if (head.start >= e.end)
                break;
            Section<S> parts[3];
            combine(ring[(ring_r++) % CAP], e, parts);
            if (parts[0].dt == parts[1].dt) {
                parts[1].start = parts[0].start;
                parts[0].end = parts[0].start;
            }
            e = parts[2];
            if (parts[0].end - parts[0].start <= 0) {
                if (parts[1].dt == parts[2].dt) {
                    e.start = parts[1].start;
                    continue;
                }
                data.set(wr++, parts[1]);
            } else {
                data.set(wr++, parts[0]);
                if (parts[1].dt == parts[2].dt) {
                    e.start = parts[1].start;
                    continue;
                }
                ++inserted;
                ring[(--ring_r) % CAP] = parts[1];
            }
        }
        if (e.end - e.start > 0 && wr < CAP) {
            if (rd < size)
                ring[(ring_w++) % CAP] = data.get(rd++);
            data.set(wr++, e);
        }
        while (ring_w - ring_r && wr < CAP) {
            if (rd < size)
                ring[(ring_w++) % CAP] = data.get(rd++);
            data.set(wr++, ring[(ring_r++) % CAP]);
        }
        size = wr + size - rd;
        if (size == CAP)
            end_ = dmin(end_, data.end(size - 1));
    }
};

// sample<n> (marching_steps.h:202-220) as a generator: border i+1 from border i
template <class S, int CAP>
struct BorderWalk {
    int cur;
    S prev;
    template <class L>
    __device__ __forceinline__ void begin(const L& sec)
    {
        cur = 0;
        prev = sec.size ? sec.data.start(0) : S(0);
    }
    template <class L>
    __device__ __forceinline__ S next(const L& sec)
    {
        if (sec.size == 0)
            return S(0);
        const S end = sec.data.end(cur);
        const S v = dmin(end, prev + sec.data.dt(cur));
        if (cur < sec.size - 1 && v == end)
            ++cur;
        prev = v;
        return v;
    }
};

template <class S>
struct Ray1d {
    S weight, centre, var;
};

template <class S>
__device__ __forceinline__ S dexp(S v);
template <>
__device__ __forceinline__ float dexp<float>(float v) { return expf(v); }
template <>
__device__ __forceinline__ double dexp<double>(double v) { return exp(v); }

template <class S>
__device__ __forceinline__ S derf(S v);
template <>
__device__ __forceinline__ float derf<float>(float v) { return erff(v); }
template <>
__device__ __forceinline__ double derf<double>(double v) { return erf(v); }

// restriction of N(mu, A^-1) to o + t d (stroke::gaussian::intersect_with_ray_inv_C
// as used at math.h:436): variance 1/(d^T A d), centre d^T A (mu - o) / (d^T A d),
// weight = integral of the normalised density along the ray
template <class S>
__device__ __forceinline__ Ray1d<S> intersect_ray(const S* mu, const S* A, const S* o, const S* d)
{
    const S Ad0 = (A[0] * d[0] + A[1] * d[1]) + A[2] * d[2];
    const S Ad1 = (A[1] * d[0] + A[3] * d[1]) + A[4] * d[2];
    const S Ad2 = (A[2] * d[0] + A[4] * d[1]) + A[5] * d[2];
    const S m0 = mu[0] - o[0], m1 = mu[1] - o[1], m2 = mu[2] - o[2];
    const S q = (d[0] * Ad0 + d[1] * Ad1) + d[2] * Ad2;
    const S C = S(1) / q;
    const S centre = ((Ad0 * m0 + Ad1 * m1) + Ad2 * m2) * C;
    const S r0 = (o[0] + centre * d[0]) - mu[0];
    const S r1 = (o[1] + centre * d[1]) - mu[1];
    const S r2 = (o[2] + centre * d[2]) - mu[2];
    const S Ar0 = (A[0] * r0 + A[1] * r1) + A[2] * r2;
    const S Ar1 = (A[1] * r0 + A[3] * r1) + A[4] * r2;
    const S Ar2 = (A[2] * r0 + A[4] * r1) + A[5] * r2;
    const S E = (r0 * Ar0 + r1 * Ar1) + r2 * Ar2;
    const S detA = A[0] * (A[3] * A[5] - A[4] * A[4]) - A[1] * (A[1] * A[5] - A[4] * A[2]) + A[2] * (A[1] * A[4] - A[3] * A[2]);
    const S norm3 = sqrt_rn(detA) / S(15.749609945722419); // sqrt((2 pi)^3)
    const S norm1 = S(1) / sqrt_rn(S(2 * kPi) * C);
    return { norm3 / norm1 * dexp(S(-0.5) * E), centre, C };
}

// adjoint w.r.t. mu and the stored inverse covariance
template <class S>
__device__ __forceinline__ void intersect_ray_grad(const S* mu, const S* A, const S* o, const S* d, S g_w, S g_centre, S g_var, S* g_mu,
    S* g_A)
{
    const S Ad[3] = { (A[0] * d[0] + A[1] * d[1]) + A[2] * d[2], (A[1] * d[0] + A[3] * d[1]) + A[4] * d[2],
        (A[2] * d[0] + A[4] * d[1]) + A[5] * d[2] };
    const S m[3] = { mu[0] - o[0], mu[1] - o[1], mu[2] - o[2] };
    const S q = dot3(d, Ad);
    const S C = S(1) / q;
    const S k = dot3(Ad, m);
    const S centre = k * C;
    const S r[3] = { (o[0] + centre * d[0]) - mu[0], (o[1] + centre * d[1]) - mu[1], (o[2] + centre * d[2]) - mu[2] };
    const S Ar[3] = { (A[0] * r[0] + A[1] * r[1]) + A[2] * r[2], (A[1] * r[0] + A[3] * r[1]) + A[4] * r[2],
        (A[2] * r[0] + A[4] * r[1]) + A[5] * r[2] };
    const S E = dot3(r, Ar);
    const S detA = A[0] * (A[3] * A[5] - A[4] * A[4]) - A[1] * (A[1] * A[5] - A[4] * A[2]) + A[2] * (A[1] * A[4] - A[3] * A[2]);
    const S W = sqrt_rn(detA) / S(15.749609945722419) * sqrt_rn(S(2 * kPi) * C) * dexp(S(-0.5) * E);
    const S gwW = g_w * W;
#pragma unroll
    for (int i = 0; i < 3; ++i)
        g_mu[i] = gwW * Ar[i] + g_centre * Ad[i] * C;
    // A^-1 via the adjugate
    const S inv_det = S(1) / detA;
    const S Ai[6] = { (A[3] * A[5] - A[4] * A[4]) * inv_det, (A[2] * A[4] - A[1] * A[5]) * inv_det, (A[1] * A[4] - A[2] * A[3]) * inv_det,
        (A[0] * A[5] - A[2] * A[2]) * inv_det, (A[2] * A[1] - A[0] * A[4]) * inv_det, (A[0] * A[3] - A[1] * A[1]) * inv_det };
    const int idx[6][2] = { { 0, 0 }, { 0, 1 }, { 0, 2 }, { 1, 1 }, { 1, 2 }, { 2, 2 } };
#pragma unroll
    for (int e = 0; e < 6; ++e) {
        const int i = idx[e][0], j = idx[e][1];
        const S ddT = d[i] * d[j];
        S G = gwW * S(0.5) * (Ai[e] - C * ddT - r[i] * r[j]) + g_centre * (C * S(0.5) * (d[i] * m[j] + m[i] * d[j]) - k * C * C * ddT)
            - g_var * C * C * ddT;
        g_A[e] = (i == j) ? G : S(2) * G;
    }
}

// stroke::gaussian::cdf_inv_SD: standard normal cdf of (x - centre) / sd
template <class S>
__device__ __forceinline__ S cdf_inv_sd(S centre, S inv_sd, S x)
{
    return S(0.5) * (S(1) + derf((x - centre) * inv_sd * S(0.70710678118654752440)));
}

// The float path's cdf, the marcher's innermost operation (one per bin and
// Gaussian): erf by Abramowitz & Stegun, Handbook of Mathematical Functions,
// formula 7.1.26, erf(x) = 1 - t (a1 + t (a2 + t (a3 + t (a4 + t a5)))) e^{-x^2}
// with t = 1 / (1 + p x), x >= 0, |error| <= 1.5e-7: one reciprocal and one
// exponential instead of a library erf.  The double path keeps erf.
template <>
__device__ __forceinline__ float cdf_inv_sd<float>(float centre, float inv_sd, float x)
{
    const float v = (x - centre) * inv_sd * 0.70710678118654752440f;
    const float a = fabsf(v);
    const float t = __builtin_amdgcn_rcpf(fmaf(0.3275911f, a, 1.f));
    const float poly = t * fmaf(t, fmaf(t, fmaf(t, fmaf(t, 1.061405429f, -1.453152027f), 1.421413741f), -0.284496736f), 0.254829592f);
    const float e = __builtin_amdgcn_exp2f(-a * a * 1.4426950408889634f);
    const float erf_abs = fmaf(-poly, e, 1.f);
    return 0.5f + 0.5f * copysignf(erf_abs, v);
}

// The cdf and the density at x of the normal (centre, 1 / inv_sd), for the
// adjoint of a bin's mass (d cdf / d centre = -inv_sd pdf, d cdf / d inv_sd =
// (x - centre) pdf).  Double: erf and exp.  Float: the cdf above, whose
// e^{-v^2} with v = z / sqrt 2 is the density's e^{-z^2 / 2}, so one
// exponential serves both.
template <class S>
__device__ __forceinline__ void cdf_pdf_inv_sd(S centre, S inv_sd, S x, S* cdf, S* pdf)
{
    const S z = (x - centre) * inv_sd;
    *cdf = S(0.5) * (S(1) + derf(z * S(0.70710678118654752440)));
    *pdf = dexp(S(-0.5) * z * z) * S(0.39894228040143267794);
}

template <>
__device__ __forceinline__ void cdf_pdf_inv_sd<float>(float centre, float inv_sd, float x, float* cdf, float* pdf)
{
    const float v = (x - centre) * inv_sd * 0.70710678118654752440f;
    const float a = fabsf(v);
    const float t = __builtin_amdgcn_rcpf(fmaf(0.3275911f, a, 1.f));
    const float poly = t * fmaf(t, fmaf(t, fmaf(t, fmaf(t, 1.061405429f, -1.453152027f), 1.421413741f), -0.284496736f), 0.254829592f);
    const float e = __builtin_amdgcn_exp2f(-a * a * 1.4426950408889634f);
    const float erf_abs = fmaf(-poly, e, 1.f);
    *cdf = 0.5f + 0.5f * copysignf(erf_abs, v);
    *pdf = e * 0.39894228040143267794f;
}

// pixel ray, vol_marcher_forward.cu:308-315
template <class S>
__device__ __forceinline__ void pixel_ray(const DevCamera<S>& cam, int px, int py, S* dir)
{
    const S pix0 = S(px) + S(0.5), pix1 = S(py) + S(0.5);
    const S ndc[3] = { pix0 * S(2) / S(cam.width) - S(1), pix1 * S(2) / S(cam.height) - S(1), S(0) };
    S v[4];
    transform_point(cam.inv_proj, ndc, v, 4);
    S p[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        p[i] = v[i] / v[3] - cam.pos[i];
    const S inv_len = S(1) / sqrt_rn(dot3(p, p));
#pragma unroll
    for (int i = 0; i < 3; ++i)
        dir[i] = p[i] * inv_len;
}

template <int CTRL, int ROW_MASK = 0xf, int BANK_MASK = 0xf>
__device__ __forceinline__ double dpp_f64(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, int(b & 0xffffffffll), CTRL, ROW_MASK, BANK_MASK, true);
    const int hi = __builtin_amdgcn_update_dpp(0, int(b >> 32), CTRL, ROW_MASK, BANK_MASK, true);
    return __longlong_as_double((long long)(uint32_t(lo)) | ((long long)hi << 32));
}

__device__ __forceinline__ double wave_sum_to_last(double v)
{
    v += dpp_f64<0x111>(v);
    v += dpp_f64<0x112>(v);
    v += dpp_f64<0x114>(v);
    v += dpp_f64<0x118>(v);
    v += dpp_f64<0x142, 0xa>(v);
    v += dpp_f64<0x143, 0xc>(v);
    return v;
}

} // namespace dgmr
