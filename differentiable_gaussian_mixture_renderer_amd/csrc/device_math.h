// SPDX-License-Identifier: GPL-3.0-or-later
// Device-side Gaussian math for gfx950, templated on the scalar type.
//
// Semantics follow the reference header math (src/dgmr/math.h,
// src/dgmr/grad/math.h).  Expressions that decide binning (culling, pixel
// extents, tile rectangles, depth keys) are written as explicit left-to-right
// sequences of IEEE operations; the library is compiled with -ffp-contract=off
// so no multiply-add is fused behind our back, which makes these values
// bit-identical to a host evaluation of the same sequence.  Hot loops that do
// not decide binning use fmaf explicitly.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dgmr {

// torch.sign semantics for the L1 gradient: 0 for equal values, NaN propagates
__device__ __forceinline__ float l1_sign(float d) { return d > 0.f ? 1.f : (d < 0.f ? -1.f : d); }

constexpr double kPi = 3.14159265358979323846;

// real SH constants (constants.h:26-43)
constexpr double SH_C0 = 0.28209479177387814;
constexpr double SH_C1 = 0.4886025119029199;
constexpr double SH_C2_0 = 1.0925484305920792, SH_C2_1 = -1.0925484305920792, SH_C2_2 = 0.31539156525252005,
                 SH_C2_3 = -1.0925484305920792, SH_C2_4 = 0.5462742152960396;
constexpr double SH_C3_0 = -0.5900435899266435, SH_C3_1 = 2.890611442640554, SH_C3_2 = -0.4570457994644658,
                 SH_C3_3 = 0.3731763325901154, SH_C3_4 = -0.4570457994644658, SH_C3_5 = 1.445305721320277,
                 SH_C3_6 = -0.5900435899266435;

// Camera as the kernels see it.  view/proj in the reference's tensor layout
// (tensor row i = glm column i): out[k] = sum_i m[4*i + k] * (x, y, z, 1)[i].
template <class S>
struct DevCamera {
    S view[16];
    S proj[16];
    S inv_proj[16];
    S pos[3];
    S bg[3];
    S tan_fovx, tan_fovy, focal_x, focal_y;
    int width, height;
    int grid_x, grid_y, tile_w, tile_h;
};

template <class S>
__device__ __forceinline__ S dmin(S a, S b) { return b < a ? b : a; }
template <class S>
__device__ __forceinline__ S dmax(S a, S b) { return a < b ? b : a; }
// a b + c rounded once (the library is built without implicit contraction)
__device__ __forceinline__ float dfma(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ double dfma(double a, double b, double c) { return fma(a, b, c); }

template <class S>
__device__ __forceinline__ void transform_point(const S* m, const S* p, S* out, int n_out)
{
    for (int k = 0; k < n_out; ++k)
        out[k] = ((m[0 + k] * p[0] + m[4 + k] * p[1]) + m[8 + k] * p[2]) + m[12 + k];
}

template <class S>
__device__ __forceinline__ S dot3(const S* a, const S* b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

template <class S>
__device__ __forceinline__ S sqrt_rn(S v);
template <>
__device__ __forceinline__ float sqrt_rn<float>(float v) { return __builtin_sqrtf(v); } // correctly rounded (__fsqrt_rn is the native approximation here)
template <>
__device__ __forceinline__ double sqrt_rn<double>(double v) { return __builtin_sqrt(v); }

template <class S>
__device__ __forceinline__ S div_rn(S a, S b);
template <>
__device__ __forceinline__ float div_rn<float>(float a, float b) { return __fdiv_rn(a, b); }
template <>
__device__ __forceinline__ double div_rn<double>(double a, double b) { return __ddiv_rn(a, b); }

template <class S>
__device__ __forceinline__ S length3(const S* v) { return sqrt_rn(dot3(v, v)); }

template <class S>
__device__ __forceinline__ S dceil(S v);
template <>
__device__ __forceinline__ float dceil<float>(float v) { return ceilf(v); }
template <>
__device__ __forceinline__ double dceil<double>(double v) { return ceil(v); }

// float -> int with NaN and overflow mapped to defined values (the reference's
// plain (int) cast is undefined there, util.h:42-52)
template <class S>
__device__ __forceinline__ int guarded_int(S v)
{
    if (!(v == v))
        return 0;
    if (v > S(1073741824))
        return 1073741824;
    if (v < S(-1073741824))
        return -1073741824;
    return int(v);
}

// util::getRect (util.h:42-52) for a tile of tile_w x tile_h pixels
template <class S>
__device__ __forceinline__ void get_rect(S px, S py, int ext_x, int ext_y, int grid_x, int grid_y, int tile_w, int tile_h,
    int* x0, int* y0, int* x1, int* y1)
{
    // tile sides are powers of two, so x / tw is exactly x * (1 / tw)
    const S tw = S(tile_w), th = S(tile_h);
    const S itw = div_rn(S(1), tw), ith = div_rn(S(1), th);
    const int a = guarded_int((px - S(ext_x)) * itw);
    const int b = guarded_int((py - S(ext_y)) * ith);
    const int c = guarded_int((((px + S(ext_x)) + tw) - S(1)) * itw);
    const int d = guarded_int((((py + S(ext_y)) + th) - S(1)) * ith);
    *x0 = min(grid_x, max(0, a));
    *y0 = min(grid_y, max(0, b));
    *x1 = min(grid_x, max(0, c));
    *y1 = min(grid_y, max(0, d));
}

template <class S>
__device__ __forceinline__ S det3(const S a[3][3])
{
    return a[0][0] * (a[1][1] * a[2][2] - a[2][1] * a[1][2])
        - a[1][0] * (a[0][1] * a[2][2] - a[2][1] * a[0][2])
        + a[2][0] * (a[0][1] * a[1][2] - a[1][1] * a[0][2]);
}

// glm::toMat3 of an un-normalised quaternion (w, x, y, z), math notation R[r][c]
template <class S>
__device__ __forceinline__ void quat_to_mat(const S* q, S R[3][3])
{
    const S w = q[0], x = q[1], y = q[2], z = q[3];
    const S xx = x * x, yy = y * y, zz = z * z;
    const S xz = x * z, xy = x * y, yz = y * z;
    const S wx = w * x, wy = w * y, wz = w * z;
    R[0][0] = S(1) - S(2) * (yy + zz);
    R[1][0] = S(2) * (xy + wz);
    R[2][0] = S(2) * (xz - wy);
    R[0][1] = S(2) * (xy - wz);
    R[1][1] = S(1) - S(2) * (xx + zz);
    R[2][1] = S(2) * (yz + wx);
    R[0][2] = S(2) * (xz + wy);
    R[1][2] = S(2) * (yz - wx);
    R[2][2] = S(1) - S(2) * (xx + yy);
}

// math::compute_cov (math.h:170-175), stored [xx, xy, xz, yy, yz, zz]
template <class S>
__device__ __forceinline__ void compute_cov(const S* s, const S* q, S* cov)
{
    S R[3][3], RS[3][3];
    quat_to_mat(q, R);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            RS[r][c] = R[r][c] * s[c];
    cov[0] = (RS[0][0] * RS[0][0] + RS[0][1] * RS[0][1]) + RS[0][2] * RS[0][2];
    cov[1] = (RS[0][0] * RS[1][0] + RS[0][1] * RS[1][1]) + RS[0][2] * RS[1][2];
    cov[2] = (RS[0][0] * RS[2][0] + RS[0][1] * RS[2][1]) + RS[0][2] * RS[2][2];
    cov[3] = (RS[1][0] * RS[1][0] + RS[1][1] * RS[1][1]) + RS[1][2] * RS[1][2];
    cov[4] = (RS[1][0] * RS[2][0] + RS[1][1] * RS[2][1]) + RS[1][2] * RS[2][2];
    cov[5] = (RS[2][0] * RS[2][0] + RS[2][1] * RS[2][1]) + RS[2][2] * RS[2][2];
}

// affine_transform_and_cut (math.h:65-73): top-left 2x2 of M C M^T
template <class S>
__device__ __forceinline__ void affine_cut(const S* c, const S M[3][3], S* out)
{
    const S a0 = M[0][0], a1 = M[0][1], a2 = M[0][2];
    const S b0 = M[1][0], b1 = M[1][1], b2 = M[1][2];
    out[0] = a0 * (c[0] * a0 + c[1] * a1 + c[2] * a2) + a1 * (c[1] * a0 + c[3] * a1 + c[4] * a2) + a2 * (c[2] * a0 + c[4] * a1 + c[5] * a2);
    out[1] = a0 * (c[0] * b0 + c[1] * b1 + c[2] * b2) + a1 * (c[1] * b0 + c[3] * b1 + c[4] * b2) + a2 * (c[2] * b0 + c[4] * b1 + c[5] * b2);
    out[2] = b0 * (c[0] * b0 + c[1] * b1 + c[2] * b2) + b1 * (c[1] * b0 + c[3] * b1 + c[4] * b2) + b2 * (c[2] * b0 + c[4] * b1 + c[5] * b2);
}

template <class S>
__device__ __forceinline__ void matmul3(const S a[3][3], const S b[3][3], S out[3][3])
{
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            out[r][c] = (a[r][0] * b[0][c] + a[r][1] * b[1][c]) + a[r][2] * b[2][c];
}

// intermediate values of the EWA splat (math.h:322-363)
template <class S>
struct SplatIn {
    S t_un[3];
    S t[3];
    S l;
    S J[3][3];
    S W[3][3];
    S T[3][3];
    S cov3[6];
    S cov2[3]; // unfiltered
};

// a / b and sqrt from the hardware reciprocal and square root (about an ulp)
// for float: for the recomputations inside gradients, where last-bit
// differences sit far below the tolerance.  Forward quantities that the
// binning depends on keep the correctly rounded div_rn / sqrt_rn.
__device__ __forceinline__ float fast_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double fast_div(double a, double b) { return a / b; }
__device__ __forceinline__ float fast_sqrt(float v) { return __builtin_amdgcn_sqrtf(v); }
__device__ __forceinline__ double fast_sqrt(double v) { return __builtin_sqrt(v); }

// view-space centre, Jacobian J, view rotation W and T = J W of the EWA
// splat (everything but the covariances).  FAST: fast_div / fast_sqrt
// instead of the correctly rounded forms (gradient recomputation only).
template <class S, bool FAST = false>
__device__ __forceinline__ void splat_projection(const S* mu, const DevCamera<S>& cam, SplatIn<S>* in)
{
    transform_point(cam.view, mu, in->t_un, 3);
    const S lim_x = S(1.3) * cam.tan_fovx * in->t_un[2];
    const S lim_y = S(1.3) * cam.tan_fovy * in->t_un[2];
    in->t[0] = dmin(dmax(in->t_un[0], -lim_x), lim_x);
    in->t[1] = dmin(dmax(in->t_un[1], -lim_y), lim_y);
    in->t[2] = in->t_un[2];
    const S* t = in->t;
    if constexpr (FAST) {
        in->l = fast_sqrt(dot3(t, t));
        const S inv_z = fast_div(S(1), t[2]), inv_l = fast_div(S(1), in->l);
        in->J[0][0] = cam.focal_x * inv_z;
        in->J[0][2] = -(cam.focal_x * t[0]) * (inv_z * inv_z);
        in->J[1][1] = cam.focal_y * inv_z;
        in->J[1][2] = -(cam.focal_y * t[1]) * (inv_z * inv_z);
        in->J[2][0] = t[0] * inv_l;
        in->J[2][1] = t[1] * inv_l;
        in->J[2][2] = t[2] * inv_l;
    } else {
        in->l = length3(in->t);
        in->J[0][0] = div_rn(cam.focal_x, t[2]);
        in->J[0][2] = div_rn(-(cam.focal_x * t[0]), t[2] * t[2]);
        in->J[1][1] = div_rn(cam.focal_y, t[2]);
        in->J[1][2] = div_rn(-(cam.focal_y * t[1]), t[2] * t[2]);
        in->J[2][0] = div_rn(t[0], in->l);
        in->J[2][1] = div_rn(t[1], in->l);
        in->J[2][2] = div_rn(t[2], in->l);
    }
    in->J[0][1] = 0;
    in->J[1][0] = 0;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            in->W[r][c] = cam.view[4 * c + r];
    matmul3(in->J, in->W, in->T);
}

template <class S, bool FAST = false>
__device__ __forceinline__ void splat_internals(const S* mu, const S* s, const S* q, const DevCamera<S>& cam, SplatIn<S>* in)
{
    splat_projection<S, FAST>(mu, cam, in);
    compute_cov(s, q, in->cov3);
    affine_cut(in->cov3, in->T, in->cov2);
}

// project + ndc2screen (math.h:89-104)
template <class S>
__device__ __forceinline__ void project_to_screen(const S* mu, const DevCamera<S>& cam, S* xy)
{
    S pp[4];
    transform_point(cam.proj, mu, pp, 4);
    const S d = pp[3] + S(0.0000001);
    const S nx = div_rn(pp[0], d), ny = div_rn(pp[1], d);
    xy[0] = ((nx + S(1)) * S(cam.width) - S(1)) * S(0.5);
    xy[1] = ((ny + S(1)) * S(cam.height) - S(1)) * S(0.5);
}

// larger2 (math.h:112-133) as indices
template <class S>
__device__ __forceinline__ void larger2_idx(const S* v, int* i0, int* i1)
{
    *i0 = 0;
    *i1 = 1;
    if (v[0] < v[1]) {
        if (v[0] < v[2]) {
            *i0 = 1;
            *i1 = 2;
        }
    } else if (v[1] < v[2]) {
        *i0 = 0;
        *i1 = 2;
    }
}

// v[i0], v[i1] of larger2_idx without indexing by a runtime value (which would
// place v in scratch memory); (i0, i1) is one of (0,1), (1,2), (0,2)
template <class S>
__device__ __forceinline__ void larger2_values(const S* v, S* a, S* b, int* i0, int* i1)
{
    larger2_idx(v, i0, i1);
    *a = *i0 == 0 ? v[0] : v[1];
    *b = *i1 == 1 ? v[1] : v[2];
}

// SH colour (math.h:233-271) of one channel from basis values
template <class S>
__device__ __forceinline__ void sh_basis(int deg, const S* dir, S* Y)
{
    const S x = dir[0], y = dir[1], z = dir[2];
    Y[0] = S(SH_C0);
#pragma unroll
    for (int k = 1; k < 16; ++k)
        Y[k] = 0;
    if (deg > 0) {
        Y[1] = -S(SH_C1) * y;
        Y[2] = S(SH_C1) * z;
        Y[3] = -S(SH_C1) * x;
    }
    if (deg > 1) {
        const S xx = x * x, yy = y * y, zz = z * z;
        Y[4] = S(SH_C2_0) * (x * y);
        Y[5] = S(SH_C2_1) * (y * z);
        Y[6] = S(SH_C2_2) * (S(2) * zz - xx - yy);
        Y[7] = S(SH_C2_3) * (x * z);
        Y[8] = S(SH_C2_4) * (xx - yy);
        if (deg > 2) {
            Y[9] = S(SH_C3_0) * y * (S(3) * xx - yy);
            Y[10] = S(SH_C3_1) * (x * y) * z;
            Y[11] = S(SH_C3_2) * y * (S(4) * zz - xx - yy);
            Y[12] = S(SH_C3_3) * z * (S(2) * zz - S(3) * xx - S(3) * yy);
            Y[13] = S(SH_C3_4) * x * (S(4) * zz - xx - yy);
            Y[14] = S(SH_C3_5) * z * (xx - yy);
            Y[15] = S(SH_C3_6) * x * (xx - S(3) * yy);
        }
    }
}

// basis derivatives d Y_k / d dir
template <class S>
__device__ __forceinline__ void sh_basis_grad(int deg, const S* dir, S dY[16][3])
{
#pragma clang fp contract(fast)
    const S x = dir[0], y = dir[1], z = dir[2];
#pragma unroll
    for (int k = 0; k < 16; ++k)
        dY[k][0] = dY[k][1] = dY[k][2] = 0;
    if (deg > 0) {
        dY[1][1] = -S(SH_C1);
        dY[2][2] = S(SH_C1);
        dY[3][0] = -S(SH_C1);
    }
    if (deg > 1) {
        dY[4][0] = S(SH_C2_0) * y; dY[4][1] = S(SH_C2_0) * x;
        dY[5][1] = S(SH_C2_1) * z; dY[5][2] = S(SH_C2_1) * y;
        dY[6][0] = S(SH_C2_2) * (S(-2) * x); dY[6][1] = S(SH_C2_2) * (S(-2) * y); dY[6][2] = S(SH_C2_2) * (S(4) * z);
        dY[7][0] = S(SH_C2_3) * z; dY[7][2] = S(SH_C2_3) * x;
        dY[8][0] = S(SH_C2_4) * (S(2) * x); dY[8][1] = S(SH_C2_4) * (S(-2) * y);
        if (deg > 2) {
            const S xx = x * x, yy = y * y, zz = z * z;
            dY[9][0] = S(SH_C3_0) * (S(6) * x * y); dY[9][1] = S(SH_C3_0) * (S(3) * xx - S(3) * yy);
            dY[10][0] = S(SH_C3_1) * (y * z); dY[10][1] = S(SH_C3_1) * (x * z); dY[10][2] = S(SH_C3_1) * (x * y);
            dY[11][0] = S(SH_C3_2) * (S(-2) * x * y); dY[11][1] = S(SH_C3_2) * (S(4) * zz - xx - S(3) * yy); dY[11][2] = S(SH_C3_2) * (S(8) * y * z);
            dY[12][0] = S(SH_C3_3) * (S(-6) * x * z); dY[12][1] = S(SH_C3_3) * (S(-6) * y * z); dY[12][2] = S(SH_C3_3) * (S(6) * zz - S(3) * xx - S(3) * yy);
            dY[13][0] = S(SH_C3_4) * (S(4) * zz - S(3) * xx - yy); dY[13][1] = S(SH_C3_4) * (S(-2) * x * y); dY[13][2] = S(SH_C3_4) * (S(8) * x * z);
            dY[14][0] = S(SH_C3_5) * (S(2) * x * z); dY[14][1] = S(SH_C3_5) * (S(-2) * y * z); dY[14][2] = S(SH_C3_5) * (xx - yy);
            dY[15][0] = S(SH_C3_6) * (S(3) * xx - S(3) * yy); dY[15][1] = S(SH_C3_6) * (S(-6) * x * y);
        }
    }
}

// g += sum_k w[k] dY_k/d dir over the basis of degree deg, without forming
// the 16 x 3 derivatives (the same terms as sh_basis_grad, each scaled by
// its weight once): the derivative table and the weights are never live
// together
template <class S>
__device__ __forceinline__ void sh_dir_grad(int deg, const S* dir, const S w[16], S g[3])
{
#pragma clang fp contract(fast)
    const S x = dir[0], y = dir[1], z = dir[2];
    if (deg > 0) {
        const S c1 = S(SH_C1);
        g[1] -= c1 * w[1];
        g[2] += c1 * w[2];
        g[0] -= c1 * w[3];
    }
    if (deg > 1) {
        const S a4 = S(SH_C2_0) * w[4], a5 = S(SH_C2_1) * w[5], a6 = S(SH_C2_2) * w[6], a7 = S(SH_C2_3) * w[7], a8 = S(SH_C2_4) * w[8];
        g[0] += a4 * y + a6 * (S(-2) * x) + a7 * z + a8 * (S(2) * x);
        g[1] += a4 * x + a5 * z + a6 * (S(-2) * y) + a8 * (S(-2) * y);
        g[2] += a5 * y + a6 * (S(4) * z) + a7 * x;
        if (deg > 2) {
            const S xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
            const S b9 = S(SH_C3_0) * w[9], b10 = S(SH_C3_1) * w[10], b11 = S(SH_C3_2) * w[11], b12 = S(SH_C3_3) * w[12];
            const S b13 = S(SH_C3_4) * w[13], b14 = S(SH_C3_5) * w[14], b15 = S(SH_C3_6) * w[15];
            g[0] += b9 * (S(6) * xy) + b10 * yz + b11 * (S(-2) * xy) + b12 * (S(-6) * xz) + b13 * (S(4) * zz - S(3) * xx - yy)
                + b14 * (S(2) * xz) + b15 * (S(3) * xx - S(3) * yy);
            g[1] += b9 * (S(3) * xx - S(3) * yy) + b10 * xz + b11 * (S(4) * zz - xx - S(3) * yy) + b12 * (S(-6) * yz) + b13 * (S(-2) * xy)
                + b14 * (S(-2) * yz) + b15 * (S(-6) * xy);
            g[2] += b10 * xy + b11 * (S(8) * yz) + b12 * (S(6) * zz - S(3) * xx - S(3) * yy) + b13 * (S(8) * xz) + b14 * (xx - yy);
        }
    }
}

// portable natural logarithm from basic IEEE operations (bit decomposition plus
// the atanh series log m = 2 atanh((m-1)/(m+1)); J.-M. Muller, "Elementary
// Functions").  The host test suite evaluates the identical sequence.
__device__ __forceinline__ float portable_log(float x)
{
    if (!(x > 0.f))
        return x == 0.f ? -__builtin_inff() : __builtin_nanf("");
    if (x == __builtin_inff())
        return x;
    int e_adj = 0;
    if (x < 1.17549435e-38f) {
        x = x * 8388608.f;
        e_adj = -23;
    }
    uint32_t bits = __float_as_uint(x);
    int e = int((bits >> 23) & 0xffu) - 127;
    bits = (bits & 0x007fffffu) | 0x3f800000u;
    float m = __uint_as_float(bits);
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e = e + 1;
    }
    e = e + e_adj;
    const float f = __fdiv_rn(m - 1.f, m + 1.f);
    const float f2 = f * f;
    float p = 0.11111111f;
    p = p * f2 + 0.14285715f;
    p = p * f2 + 0.2f;
    p = p * f2 + 0.33333334f;
    p = p * f2 + 1.f;
    const float ef = float(e);
    return ef * 0.693145752f + (ef * 1.42860677e-06f + (f + f) * p);
}

__device__ __forceinline__ double portable_log(double x)
{
    if (!(x > 0.0))
        return x == 0.0 ? -__builtin_inf() : __builtin_nan("");
    if (x == __builtin_inf())
        return x;
    int e_adj = 0;
    if (x < 2.2250738585072014e-308) {
        x = x * 4503599627370496.0;
        e_adj = -52;
    }
    uint64_t bits = uint64_t(__double_as_longlong(x));
    int e = int((bits >> 52) & 0x7ffu) - 1023;
    bits = (bits & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
    double m = __longlong_as_double((long long)bits);
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e = e + 1;
    }
    e = e + e_adj;
    const double f = __ddiv_rn(m - 1.0, m + 1.0);
    const double f2 = f * f;
    double p = 1.0 / 25.0;
    p = p * f2 + 1.0 / 23.0;
    p = p * f2 + 1.0 / 21.0;
    p = p * f2 + 1.0 / 19.0;
    p = p * f2 + 1.0 / 17.0;
    p = p * f2 + 1.0 / 15.0;
    p = p * f2 + 1.0 / 13.0;
    p = p * f2 + 1.0 / 11.0;
    p = p * f2 + 1.0 / 9.0;
    p = p * f2 + 1.0 / 7.0;
    p = p * f2 + 1.0 / 5.0;
    p = p * f2 + 1.0 / 3.0;
    p = p * f2 + 1.0;
    const double ed = double(e);
    return ed * 6.93147180369123816490e-01 + (ed * 1.90821492927058770002e-10 + (f + f) * p);
}

// ------------------------------------------------------------- wave64 helpers
__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }

// a value the caller knows to be equal on all lanes, moved to a scalar register
__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

template <int CTRL, int ROW_MASK = 0xf, int BANK_MASK = 0xf>
__device__ __forceinline__ float dpp_f32(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, BANK_MASK, true));
}

// Sum over the 64 lanes of a wave through DPP row shifts and row broadcasts;
// the total lands in lane 63 (read it with __builtin_amdgcn_readlane(v, 63)).
__device__ __forceinline__ float wave_sum_to_last(float v)
{
    v += dpp_f32<0x111>(v); // row_shr:1
    v += dpp_f32<0x112>(v); // row_shr:2
    v += dpp_f32<0x114>(v); // row_shr:4
    v += dpp_f32<0x118>(v); // row_shr:8
    v += dpp_f32<0x142, 0xa>(v); // row_bcast:15 -> rows 1, 3
    v += dpp_f32<0x143, 0xc>(v); // row_bcast:31 -> rows 2, 3
    return v;
}

// Sums eight per-lane values over the wave with three transposing halvings
// (lanes l / l+32 through v_permlane32_swap, rows 0,2 / 1,3 through
// v_permlane16_swap, lanes l / l^8 through row_ror:8) and a 3-step DPP sum in
// groups of 8 lanes.  Lane 8m then holds the total of value kSum8Order[m].
constexpr int kSum8Order[8] = { 0, 4, 2, 6, 1, 5, 3, 7 };

__device__ __forceinline__ float swap_sum32(float a, float b)
{
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

__device__ __forceinline__ float swap_sum16(float a, float b)
{
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

__device__ __forceinline__ float wave_sum8(float v0, float v1, float v2, float v3, float v4, float v5, float v6, float v7)
{
    const float w01 = swap_sum32(v0, v1), w23 = swap_sum32(v2, v3), w45 = swap_sum32(v4, v5), w67 = swap_sum32(v6, v7);
    const float u0 = swap_sum16(w01, w23); // rows hold v0, v2, v1, v3
    const float u1 = swap_sum16(w45, w67); // rows hold v4, v6, v5, v7
    const bool hi = (lane_id() & 8u) != 0;
    const float keep = hi ? u1 : u0, send = hi ? u0 : u1;
    float x = keep + dpp_f32<0x128>(send); // row_ror:8 pairs lane l with l^8
    x += dpp_f32<0x141>(x);                // row_half_mirror: l <-> 7-l in groups of 8
    x += dpp_f32<0x4e>(x);                 // quad_perm [2,3,0,1]
    x += dpp_f32<0xb1>(x);                 // quad_perm [1,0,3,2]
    return x;
}

__device__ __forceinline__ float wave_sum(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, wave_sum_to_last(v)), 63));
}

} // namespace dgmr
