// SPDX-License-Identifier: GPL-3.0-or-later
// Device-side adjoints of the per-Gaussian math (reference: src/dgmr/grad/math.h).
// Symmetric matrices are stored [xx,xy,xz,yy,yz,zz] / [xx,xy,yy]; their
// gradients are taken w.r.t. the stored values (an off-diagonal value enters
// the matrix twice).
#pragma once

#include "device_math.h"

namespace dgmr {

template <class S>
__device__ __forceinline__ void det3_grad(const S a[3][3], S g, S out[3][3])
{
#pragma clang fp contract(fast)
    out[0][0] = g * (a[1][1] * a[2][2] - a[1][2] * a[2][1]);
    out[0][1] = g * (a[1][2] * a[2][0] - a[1][0] * a[2][2]);
    out[0][2] = g * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    out[1][0] = g * (a[0][2] * a[2][1] - a[0][1] * a[2][2]);
    out[1][1] = g * (a[0][0] * a[2][2] - a[0][2] * a[2][0]);
    out[1][2] = g * (a[0][1] * a[2][0] - a[0][0] * a[2][1]);
    out[2][0] = g * (a[0][1] * a[1][2] - a[0][2] * a[1][1]);
    out[2][1] = g * (a[0][2] * a[1][0] - a[0][0] * a[1][2]);
    out[2][2] = g * (a[0][0] * a[1][1] - a[0][1] * a[1][0]);
}

template <class S>
__device__ __forceinline__ void sym3_to_mat(const S* c, S m[3][3])
{
    m[0][0] = c[0]; m[0][1] = c[1]; m[0][2] = c[2];
    m[1][0] = c[1]; m[1][1] = c[3]; m[1][2] = c[4];
    m[2][0] = c[2]; m[2][1] = c[4]; m[2][2] = c[5];
}

template <class S>
__device__ __forceinline__ void sym3_grad_to_mat(const S* g, S m[3][3])
{
    m[0][0] = g[0]; m[0][1] = g[1] * S(0.5); m[0][2] = g[2] * S(0.5);
    m[1][0] = g[1] * S(0.5); m[1][1] = g[3]; m[1][2] = g[4] * S(0.5);
    m[2][0] = g[2] * S(0.5); m[2][1] = g[4] * S(0.5); m[2][2] = g[5];
}

template <class S>
__device__ __forceinline__ void mat_grad_to_sym3(const S m[3][3], S* g)
{
    g[0] = m[0][0];
    g[1] = m[0][1] + m[1][0];
    g[2] = m[0][2] + m[2][0];
    g[3] = m[1][1];
    g[4] = m[1][2] + m[2][1];
    g[5] = m[2][2];
}

template <class S>
__device__ __forceinline__ void sym3_inverse(const S* c, S* inv)
{
    S m[3][3];
    sym3_to_mat(c, m);
    const S id = S(1) / det3(m);
    inv[0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) * id;
    inv[1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * id;
    inv[2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * id;
    inv[3] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * id;
    inv[4] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * id;
    inv[5] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * id;
}

// adjoint of X^-1 (stroke::grad::inverse): dL/dX = -A G A, A = X^-1
template <class S>
__device__ __forceinline__ void sym3_inverse_grad(const S* inv, const S* g_inv, S* g_x)
{
#pragma clang fp contract(fast)
    S A[3][3], G[3][3], t[3][3], r[3][3];
    sym3_to_mat(inv, A);
    sym3_grad_to_mat(g_inv, G);
    matmul3(A, G, t);
    matmul3(t, A, r);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            r[i][j] = -r[i][j];
    mat_grad_to_sym3(r, g_x);
}

template <class S>
__device__ __forceinline__ void quat_to_mat_grad(const S* q, const S G[3][3], S* gq)
{
#pragma clang fp contract(fast)
    const S w = q[0], x = q[1], y = q[2], z = q[3];
    gq[0] = S(2) * (z * G[1][0] - y * G[2][0] - z * G[0][1] + x * G[2][1] + y * G[0][2] - x * G[1][2]);
    gq[1] = S(2) * (y * G[1][0] + z * G[2][0] + y * G[0][1] - S(2) * x * G[1][1] + w * G[2][1] + z * G[0][2] - w * G[1][2] - S(2) * x * G[2][2]);
    gq[2] = S(2) * (-S(2) * y * G[0][0] + x * G[1][0] - w * G[2][0] + x * G[0][1] + z * G[2][1] + w * G[0][2] + z * G[1][2] - S(2) * y * G[2][2]);
    gq[3] = S(2) * (-S(2) * z * G[0][0] + w * G[1][0] + x * G[2][0] - w * G[0][1] - S(2) * z * G[1][1] + y * G[2][1] + x * G[0][2] + y * G[1][2]);
}

// grad::compute_cov (grad/math.h:135-155)
template <class S>
__device__ __forceinline__ void compute_cov_grad(const S* s, const S* q, const S* g_cov, S* g_s, S* g_q)
{
#pragma clang fp contract(fast)
    S R[3][3], RS[3][3], G[3][3], gRS[3][3], gR[3][3];
    quat_to_mat(q, R);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            RS[r][c] = R[r][c] * s[c];
    sym3_grad_to_mat(g_cov, G);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            gRS[r][c] = S(2) * ((G[r][0] * RS[0][c] + G[r][1] * RS[1][c]) + G[r][2] * RS[2][c]);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        g_s[c] = (R[0][c] * gRS[0][c] + R[1][c] * gRS[1][c]) + R[2][c] * gRS[2][c];
#pragma unroll
        for (int r = 0; r < 3; ++r)
            gR[r][c] = gRS[r][c] * s[c];
    }
    quat_to_mat_grad(q, gR, g_q);
}

// grad::affine_transform_and_cut (grad/math.h:47-57)
template <class S>
__device__ __forceinline__ void affine_cut_grad(const S* c, const S M[3][3], const S* g2, S* g_cov, S g_M[3][3], bool want_M)
{
#pragma clang fp contract(fast)
    const S Gy[3][3] = { { g2[0], g2[1] * S(0.5), 0 }, { g2[1] * S(0.5), g2[2], 0 }, { 0, 0, 0 } };
    S MT[3][3], t[3][3], gc[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int k = 0; k < 3; ++k)
            MT[r][k] = M[k][r];
    matmul3(MT, Gy, t);
    matmul3(t, M, gc);
    mat_grad_to_sym3(gc, g_cov);
    if (want_M) {
        S C[3][3], MC[3][3];
        sym3_to_mat(c, C);
        matmul3(M, C, MC);
        matmul3(Gy, MC, t);
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int k = 0; k < 3; ++k)
                g_M[r][k] = S(2) * t[r][k];
    }
}

// The Ots splat's per-Gaussian inputs, the same for every view: the 3D
// covariance and the two largest scales (splat_ots_grad_view takes them, so
// a Gaussian seen from many views computes them once).
template <class S>
struct OtsPre {
    S cov3[6];
    S sa, sb;
    int i0, i1;
};

template <class S>
__device__ __forceinline__ OtsPre<S> ots_pre(const S* s, const S* q)
{
    OtsPre<S> o;
    compute_cov(s, q, o.cov3);
    larger2_values(s, &o.sa, &o.sb, &o.i0, &o.i1);
    return o;
}

// Adjoint of the Ots splat (math.h:389-397 and splat_internals) for one
// view: the weight, centroid and filtered covariance gradients (stored
// values) give g_w and g_mu for this view, and add dL/dcov3 into g_cov3 and
// dL/d(2 pi sa sb) into g_num.  Both are linear in what reaches the scales
// and rotation, so their sums over views go through splat_ots_grad_finish
// once.  Recomputation uses the hardware reciprocal (gradient tolerance).
template <class S>
__device__ __forceinline__ void splat_ots_grad_view(S w, const S* mu, const OtsPre<S>& pre, const DevCamera<S>& cam, S filter, S g_weight,
    const S* g_xy, const S* g_cov2, S* g_w, S* g_mu, S* g_cov3, S* g_num)
{
#pragma clang fp contract(fast)
    SplatIn<S> in;
    splat_projection<S, true>(mu, cam, &in);
#pragma unroll
    for (int k = 0; k < 6; ++k)
        in.cov3[k] = pre.cov3[k];
    affine_cut(in.cov3, in.T, in.cov2);
    const S cov_f[3] = { in.cov2[0] + filter, in.cov2[1], in.cov2[2] + filter };
    const S detJ = det3(in.J);
    const S num = S(2 * kPi) * pre.sa * pre.sb;
    const S det2 = cov_f[0] * cov_f[2] - cov_f[1] * cov_f[1];
    const S sq = fast_sqrt(det2);
    const S inv_i2 = fast_div(S(1), S(2 * kPi) * sq);
    *g_w = g_weight * detJ * num * inv_i2;
    const S g_det = g_weight * w * num * inv_i2;
    *g_num += g_weight * w * detJ * inv_i2;
    const S g_i2 = -g_weight * w * detJ * num * (inv_i2 * inv_i2);
    const S f = g_i2 * S(2 * kPi) * fast_div(S(1), S(2) * sq);
    S g_c2[3] = { g_cov2[0] + f * cov_f[2], g_cov2[1] + f * (S(-2) * cov_f[1]), g_cov2[2] + f * cov_f[0] };
    S g_J[3][3];
    det3_grad(in.J, g_det, g_J);
    S g_c3[6], g_T[3][3];
    affine_cut_grad(in.cov3, in.T, g_c2, g_c3, g_T, true);
#pragma unroll
    for (int k = 0; k < 6; ++k)
        g_cov3[k] += g_c3[k];
    // centroid through project + ndc2screen
    {
        S pp[4];
        transform_point(cam.proj, mu, pp, 4);
        const S inv_d = fast_div(S(1), pp[3] + S(0.0000001));
        const S g_nx = g_xy[0] * S(0.5) * S(cam.width);
        const S g_ny = g_xy[1] * S(0.5) * S(cam.height);
        const S g_pp[4] = { g_nx * inv_d, g_ny * inv_d, 0, -(g_nx * pp[0] + g_ny * pp[1]) * (inv_d * inv_d) };
#pragma unroll
        for (int i = 0; i < 3; ++i)
            g_mu[i] = ((cam.proj[4 * i + 0] * g_pp[0] + cam.proj[4 * i + 1] * g_pp[1]) + cam.proj[4 * i + 2] * g_pp[2]) + cam.proj[4 * i + 3] * g_pp[3];
    }
    // T = J W
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            g_J[r][c] += (g_T[r][0] * in.W[c][0] + g_T[r][1] * in.W[c][1]) + g_T[r][2] * in.W[c][2];
    // make_jakobian adjoint (grad/math.h:94-133)
    const S* t = in.t;
    const S inv_l = fast_div(S(1), in.l);
    const S inv_z = fast_div(S(1), t[2]);
    const S inv_z2 = inv_z * inv_z;
    S g_t[3];
    g_t[0] = g_J[0][2] * (-cam.focal_x * inv_z2) + g_J[2][0] * inv_l;
    g_t[1] = g_J[1][2] * (-cam.focal_y * inv_z2) + g_J[2][1] * inv_l;
    g_t[2] = g_J[0][0] * (-cam.focal_x * inv_z2) + g_J[1][1] * (-cam.focal_y * inv_z2)
        + g_J[0][2] * (S(2) * cam.focal_x * t[0] * (inv_z2 * inv_z)) + g_J[1][2] * (S(2) * cam.focal_y * t[1] * (inv_z2 * inv_z))
        + g_J[2][2] * inv_l;
    const S g_l = -(g_J[2][0] * t[0] + g_J[2][1] * t[1] + g_J[2][2] * t[2]) * (inv_l * inv_l);
#pragma unroll
    for (int i = 0; i < 3; ++i)
        g_t[i] += g_l * t[i] * inv_l;
    // clamp_to_fov passes the gradient inside the limits
    const S lim_x = S(1.3) * cam.tan_fovx * in.t_un[2];
    const S lim_y = S(1.3) * cam.tan_fovy * in.t_un[2];
    if (in.t_un[0] < -lim_x || in.t_un[0] > lim_x)
        g_t[0] = 0;
    if (in.t_un[1] < -lim_y || in.t_un[1] > lim_y)
        g_t[1] = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
        g_mu[i] += (in.W[0][i] * g_t[0] + in.W[1][i] * g_t[1]) + in.W[2][i] * g_t[2];
}

// scales and rotation gradients from the view sums of dL/dcov3 and
// dL/d(2 pi sa sb)
template <class S>
__device__ __forceinline__ void splat_ots_grad_finish(const S* s, const S* q, const OtsPre<S>& pre, const S* g_cov3, S g_num, S* g_s, S* g_q)
{
#pragma clang fp contract(fast)
    const S g_a = g_num * S(2 * kPi) * pre.sb, g_b = g_num * S(2 * kPi) * pre.sa;
    const S gs[3] = { pre.i0 == 0 ? g_a : S(0), pre.i0 == 1 ? g_a : (pre.i1 == 1 ? g_b : S(0)), pre.i1 == 2 ? g_b : S(0) };
    S gs2[3];
    compute_cov_grad(s, q, g_cov3, gs2, g_q);
#pragma unroll
    for (int i = 0; i < 3; ++i)
        g_s[i] = gs[i] + gs2[i];
}

} // namespace dgmr
