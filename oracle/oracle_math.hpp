// SPDX-License-Identifier: GPL-3.0-or-later
// dgmr oracle -- TEST INFRASTRUCTURE ONLY.
//
// A plain-C++ restatement of the reference's per-Gaussian math and its
// hand-written adjoints, templated on the scalar type so the same code runs in
// float (a mirror of the GPU arithmetic, used for bit-exact binning checks) and
// in double (the accuracy oracle).  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load this; the product never links it.
//
// Conventions (reference: src/dgmr/math.h, python_bindings/vol_marcher.cu:32-40):
//  * 4x4 camera matrices are stored exactly as the torch tensors the reference
//    receives: element (i, j) at m[4*i + j]; tensor row i is glm column i, so a
//    point transforms as out[k] = sum_i m[4*i + k] * (x, y, z, 1)[i].
//  * 3x3 matrices here are in math notation, a[r][c].
//  * symmetric 3x3 matrices are stored [xx, xy, xz, yy, yz, zz], 2x2 as
//    [xx, xy, yy] (stroke's Cov3/Cov2 packing, evidenced by math.h:68-72).
//  * gradients of symmetric matrices are taken w.r.t. the stored values (an
//    off-diagonal stored value enters the matrix twice).
//  * quaternions are (w, x, y, z) (GLM_FORCE_QUAT_DATA_WXYZ).
//
// Arithmetic order: every expression that decides binning (rects, culling,
// depth keys) is written as an explicit left-to-right sequence of IEEE
// operations with no fused multiply-add; the GPU kernels evaluate the same
// sequence under -ffp-contract=off, which makes float binning bit-exact.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>

namespace orc {

constexpr double kPi = 3.14159265358979323846;

// SH constants, constants.h:26-43 (real spherical harmonics, degree <= 3).
constexpr double SH_C0 = 0.28209479177387814;
constexpr double SH_C1 = 0.4886025119029199;
constexpr double SH_C2[5] = { 1.0925484305920792, -1.0925484305920792, 0.31539156525252005, -1.0925484305920792, 0.5462742152960396 };
constexpr double SH_C3[7] = { -0.5900435899266435, 2.890611442640554, -0.4570457994644658, 0.3731763325901154, -0.4570457994644658, 1.445305721320277, -0.5900435899266435 };

template <class S>
struct Camera {
    S view[16];
    S proj[16];
    S pos[3];
    S tan_fovx, tan_fovy;
    S focal_x, focal_y;
    int width, height;
};

// focal lengths as the reference derives them (splat_forward.cu:35-36, vol_marcher_forward.cu:56-57)
template <class S>
inline Camera<S> make_camera(const S* view, const S* proj, const S* pos, S tan_fovx, S tan_fovy, int width, int height)
{
    Camera<S> c;
    for (int i = 0; i < 16; ++i) {
        c.view[i] = view[i];
        c.proj[i] = proj[i];
    }
    for (int i = 0; i < 3; ++i)
        c.pos[i] = pos[i];
    c.tan_fovx = tan_fovx;
    c.tan_fovy = tan_fovy;
    c.focal_y = S(height) / (S(2) * tan_fovy);
    c.focal_x = S(width) / (S(2) * tan_fovx);
    c.width = width;
    c.height = height;
    return c;
}

// ---------------------------------------------------------------- elementary
template <class S>
inline void transform_point(const S* m, const S* p, S* out, int n_out)
{
    for (int k = 0; k < n_out; ++k)
        out[k] = ((m[0 + k] * p[0] + m[4 + k] * p[1]) + m[8 + k] * p[2]) + m[12 + k];
}

template <class S>
inline S dot3(const S* a, const S* b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

template <class S>
inline S length3(const S* v) { return std::sqrt(dot3(v, v)); }

// natural logarithm from basic IEEE operations only (bit decomposition and the
// atanh series log(m) = 2 atanh((m-1)/(m+1)), e.g. J.-M. Muller, "Elementary
// Functions", ch. on logarithms).  The GPU evaluates the identical sequence, so
// the marcher's iso-line rectangles agree bit for bit between host and device.
template <class S>
inline S portable_log(S x);

template <>
inline float portable_log<float>(float x)
{
    if (!(x > 0.f))
        return x == 0.f ? -std::numeric_limits<float>::infinity() : std::numeric_limits<float>::quiet_NaN();
    if (x == std::numeric_limits<float>::infinity())
        return x;
    int e_adj = 0;
    if (x < 1.17549435e-38f) {
        x = x * 8388608.f;
        e_adj = -23;
    }
    uint32_t bits;
    std::memcpy(&bits, &x, 4);
    int e = int((bits >> 23) & 0xffu) - 127;
    bits = (bits & 0x007fffffu) | 0x3f800000u;
    float m;
    std::memcpy(&m, &bits, 4);
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e = e + 1;
    }
    e = e + e_adj;
    const float f = (m - 1.f) / (m + 1.f);
    const float f2 = f * f;
    float p = 0.11111111f;
    p = p * f2 + 0.14285715f;
    p = p * f2 + 0.2f;
    p = p * f2 + 0.33333334f;
    p = p * f2 + 1.f;
    const float ef = float(e);
    return ef * 0.693145752f + (ef * 1.42860677e-06f + (f + f) * p);
}

template <>
inline double portable_log<double>(double x)
{
    if (!(x > 0.0))
        return x == 0.0 ? -std::numeric_limits<double>::infinity() : std::numeric_limits<double>::quiet_NaN();
    if (x == std::numeric_limits<double>::infinity())
        return x;
    int e_adj = 0;
    if (x < 2.2250738585072014e-308) {
        x = x * 4503599627370496.0;
        e_adj = -52;
    }
    uint64_t bits;
    std::memcpy(&bits, &x, 8);
    int e = int((bits >> 52) & 0x7ffu) - 1023;
    bits = (bits & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
    double m;
    std::memcpy(&m, &bits, 8);
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e = e + 1;
    }
    e = e + e_adj;
    const double f = (m - 1.0) / (m + 1.0);
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

// float -> int conversion guarded against NaN / overflow (the reference's plain
// (int) cast is undefined there; util.h:42-52)
template <class S>
inline int guarded_int(S v)
{
    if (!(v == v))
        return 0;
    if (v > S(1073741824))
        return 1073741824;
    if (v < S(-1073741824))
        return -1073741824;
    return int(v);
}

// util::getRect, util.h:42-52, generalised to a tile of tile_w x tile_h pixels
template <class S>
inline void get_rect(S px, S py, int ext_x, int ext_y, int grid_x, int grid_y, int tile_w, int tile_h,
    int* min_x, int* min_y, int* max_x, int* max_y)
{
    const S tw = S(tile_w), th = S(tile_h);
    const int a = guarded_int((px - S(ext_x)) / tw);
    const int b = guarded_int((py - S(ext_y)) / th);
    const int c = guarded_int(((px + S(ext_x)) + tw - S(1)) / tw);
    const int d = guarded_int(((py + S(ext_y)) + th - S(1)) / th);
    *min_x = std::min(grid_x, std::max(0, a));
    *min_y = std::min(grid_y, std::max(0, b));
    *max_x = std::min(grid_x, std::max(0, c));
    *max_y = std::min(grid_y, std::max(0, d));
}

// util::getHigherMsb, util.h:56-70: one past the highest set bit
inline uint32_t higher_msb(uint32_t n)
{
    uint32_t r = 0;
    while (r < 32 && (n >> r) != 0)
        ++r;
    return r;
}

// ---------------------------------------------------------------- 3x3 helpers
template <class S>
inline S det3(const S a[3][3])
{
    return a[0][0] * (a[1][1] * a[2][2] - a[2][1] * a[1][2])
        - a[1][0] * (a[0][1] * a[2][2] - a[2][1] * a[0][2])
        + a[2][0] * (a[0][1] * a[1][2] - a[1][1] * a[0][2]);
}

// d det(A) / dA = cofactor matrix
template <class S>
inline void det3_grad(const S a[3][3], S g, S out[3][3])
{
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
inline void sym3_to_mat(const S* c, S m[3][3])
{
    m[0][0] = c[0]; m[0][1] = c[1]; m[0][2] = c[2];
    m[1][0] = c[1]; m[1][1] = c[3]; m[1][2] = c[4];
    m[2][0] = c[2]; m[2][1] = c[4]; m[2][2] = c[5];
}

// gradient w.r.t. the stored values -> symmetric matrix gradient (off-diagonals halved)
template <class S>
inline void sym3_grad_to_mat(const S* g, S m[3][3])
{
    m[0][0] = g[0]; m[0][1] = g[1] / 2; m[0][2] = g[2] / 2;
    m[1][0] = g[1] / 2; m[1][1] = g[3]; m[1][2] = g[4] / 2;
    m[2][0] = g[2] / 2; m[2][1] = g[4] / 2; m[2][2] = g[5];
}

// matrix gradient -> gradient w.r.t. the stored values of a symmetric matrix
template <class S>
inline void mat_grad_to_sym3(const S m[3][3], S* g)
{
    g[0] = m[0][0];
    g[1] = m[0][1] + m[1][0];
    g[2] = m[0][2] + m[2][0];
    g[3] = m[1][1];
    g[4] = m[1][2] + m[2][1];
    g[5] = m[2][2];
}

template <class S>
inline void matmul3(const S a[3][3], const S b[3][3], S out[3][3])
{
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            out[r][c] = (a[r][0] * b[0][c] + a[r][1] * b[1][c]) + a[r][2] * b[2][c];
}

template <class S>
inline void sym3_inverse(const S* c, S* inv)
{
    S m[3][3];
    sym3_to_mat(c, m);
    const S d = det3(m);
    const S id = S(1) / d;
    inv[0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) * id;
    inv[1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * id;
    inv[2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * id;
    inv[3] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * id;
    inv[4] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * id;
    inv[5] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * id;
}

// adjoint of inv = X^-1 for symmetric X (stored-value gradients in and out):
// dL/dX = -X^-1 G X^-1 (stroke::grad::inverse, used at vol_marcher_backward.cu:365)
template <class S>
inline void sym3_inverse_grad(const S* inv, const S* g_inv, S* g_x)
{
    S A[3][3], G[3][3], t[3][3], r[3][3];
    sym3_to_mat(inv, A);
    sym3_grad_to_mat(g_inv, G);
    matmul3(A, G, t);
    matmul3(t, A, r);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r[i][j] = -r[i][j];
    mat_grad_to_sym3(r, g_x);
}

// ---------------------------------------------------------------- covariance
// glm::toMat3 for an (un-normalised) quaternion (w, x, y, z), in math notation R[r][c]
template <class S>
inline void quat_to_mat(const S* q, S R[3][3])
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

template <class S>
inline void quat_to_mat_grad(const S* q, const S G[3][3], S* gq)
{
    const S w = q[0], x = q[1], y = q[2], z = q[3];
    // R00 = 1-2(yy+zz), R10 = 2(xy+wz), R20 = 2(xz-wy), R01 = 2(xy-wz), R11 = 1-2(xx+zz),
    // R21 = 2(yz+wx), R02 = 2(xz+wy), R12 = 2(yz-wx), R22 = 1-2(xx+yy)
    gq[0] = S(2) * (z * G[1][0] - y * G[2][0] - z * G[0][1] + x * G[2][1] + y * G[0][2] - x * G[1][2]);
    gq[1] = S(2) * (y * G[1][0] + z * G[2][0] + y * G[0][1] - S(2) * x * G[1][1] + w * G[2][1] + z * G[0][2] - w * G[1][2] - S(2) * x * G[2][2]);
    gq[2] = S(2) * (-S(2) * y * G[0][0] + x * G[1][0] - w * G[2][0] + x * G[0][1] + z * G[2][1] + w * G[0][2] + z * G[1][2] - S(2) * y * G[2][2]);
    gq[3] = S(2) * (-S(2) * z * G[0][0] + w * G[1][0] + x * G[2][0] - w * G[0][1] - S(2) * z * G[1][1] + y * G[2][1] + x * G[0][2] + y * G[1][2]);
}

// math::compute_cov, math.h:170-175: Sigma = (R S)(R S)^T
template <class S>
inline void compute_cov(const S* s, const S* q, S* cov)
{
    S R[3][3], RS[3][3];
    quat_to_mat(q, R);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            RS[r][c] = R[r][c] * s[c];
    const auto e = [&](int i, int j) { return (RS[i][0] * RS[j][0] + RS[i][1] * RS[j][1]) + RS[i][2] * RS[j][2]; };
    cov[0] = e(0, 0);
    cov[1] = e(0, 1);
    cov[2] = e(0, 2);
    cov[3] = e(1, 1);
    cov[4] = e(1, 2);
    cov[5] = e(2, 2);
}

// grad::compute_cov, grad/math.h:135-155
template <class S>
inline void compute_cov_grad(const S* s, const S* q, const S* g_cov, S* g_s, S* g_q)
{
    S R[3][3], RS[3][3], G[3][3], gRS[3][3], gR[3][3];
    quat_to_mat(q, R);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            RS[r][c] = R[r][c] * s[c];
    sym3_grad_to_mat(g_cov, G);
    // dL/dRS = (G + G^T) RS
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            gRS[r][c] = S(2) * ((G[r][0] * RS[0][c] + G[r][1] * RS[1][c]) + G[r][2] * RS[2][c]);
    for (int c = 0; c < 3; ++c) {
        g_s[c] = (R[0][c] * gRS[0][c] + R[1][c] * gRS[1][c]) + R[2][c] * gRS[2][c];
        for (int r = 0; r < 3; ++r)
            gR[r][c] = gRS[r][c] * s[c];
    }
    quat_to_mat_grad(q, gR, g_q);
}

// affine_transform_and_cut, math.h:65-73: top-left 2x2 of M Sigma M^T, M in math notation
template <class S>
inline void affine_transform_and_cut(const S* c, const S M[3][3], S* out)
{
    const S a0 = M[0][0], a1 = M[0][1], a2 = M[0][2];
    const S b0 = M[1][0], b1 = M[1][1], b2 = M[1][2];
    out[0] = a0 * (c[0] * a0 + c[1] * a1 + c[2] * a2) + a1 * (c[1] * a0 + c[3] * a1 + c[4] * a2) + a2 * (c[2] * a0 + c[4] * a1 + c[5] * a2);
    out[1] = a0 * (c[0] * b0 + c[1] * b1 + c[2] * b2) + a1 * (c[1] * b0 + c[3] * b1 + c[4] * b2) + a2 * (c[2] * b0 + c[4] * b1 + c[5] * b2);
    out[2] = b0 * (c[0] * b0 + c[1] * b1 + c[2] * b2) + b1 * (c[1] * b0 + c[3] * b1 + c[4] * b2) + b2 * (c[2] * b0 + c[4] * b1 + c[5] * b2);
}

// grad::affine_transform_and_cut, grad/math.h:47-57: g2 are stored-value gradients of the 2x2 result
template <class S>
inline void affine_transform_and_cut_grad(const S* c, const S M[3][3], const S* g2, S* g_cov, S g_M[3][3])
{
    // Y = M C M^T (3x3), only Y00, Y01, Y11 are kept; Gy = matrix gradient of Y
    S Gy[3][3] = { { g2[0], g2[1] / 2, 0 }, { g2[1] / 2, g2[2], 0 }, { 0, 0, 0 } };
    S C[3][3], t[3][3], MT[3][3], gc[3][3];
    sym3_to_mat(c, C);
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k)
            MT[r][k] = M[k][r];
    // dL/dC = M^T Gy M
    matmul3(MT, Gy, t);
    matmul3(t, M, gc);
    mat_grad_to_sym3(gc, g_cov);
    // dL/dM = 2 Gy M C (Gy and C symmetric)
    S MC[3][3];
    matmul3(M, C, MC);
    matmul3(Gy, MC, t);
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k)
            g_M[r][k] = S(2) * t[r][k];
}

// ---------------------------------------------------------------- projection
// make_jakobian, math.h:157-168 (math notation J[r][c])
template <class S>
inline void make_jacobian(const S* t, S l, S fx, S fy, S J[3][3])
{
    J[0][0] = fx / t[2];
    J[0][1] = 0;
    J[0][2] = -(fx * t[0]) / (t[2] * t[2]);
    J[1][0] = 0;
    J[1][1] = fy / t[2];
    J[1][2] = -(fy * t[1]) / (t[2] * t[2]);
    J[2][0] = t[0] / l;
    J[2][1] = t[1] / l;
    J[2][2] = t[2] / l;
}

// grad::make_jakobian, grad/math.h:94-133
template <class S>
inline void make_jacobian_grad(const S* t, S l, S fx, S fy, const S gJ[3][3], S* g_t, S* g_l)
{
    const S tz2 = t[2] * t[2];
    g_t[0] = gJ[0][2] * (-fx / tz2) + gJ[2][0] / l;
    g_t[1] = gJ[1][2] * (-fy / tz2) + gJ[2][1] / l;
    g_t[2] = gJ[0][0] * (-fx / tz2) + gJ[1][1] * (-fy / tz2)
        + gJ[0][2] * (S(2) * fx * t[0] / (tz2 * t[2])) + gJ[1][2] * (S(2) * fy * t[1] / (tz2 * t[2]))
        + gJ[2][2] / l;
    *g_l = -(gJ[2][0] * t[0] + gJ[2][1] * t[1] + gJ[2][2] * t[2]) / (l * l);
}

// project + ndc2screen, math.h:89-104
template <class S>
inline void project_to_screen(const S* mu, const Camera<S>& cam, S* xy)
{
    const S p[3] = { mu[0], mu[1], mu[2] };
    S pp[4];
    transform_point(cam.proj, p, pp, 4);
    const S d = pp[3] + S(0.0000001);
    const S nx = pp[0] / d, ny = pp[1] / d;
    xy[0] = ((nx + S(1)) * S(cam.width) - S(1)) * S(0.5);
    xy[1] = ((ny + S(1)) * S(cam.height) - S(1)) * S(0.5);
}

// grad::ndc2screen + grad::project, grad/math.h:34-45, 80-92
template <class S>
inline void project_to_screen_grad(const S* mu, const Camera<S>& cam, const S* g_xy, S* g_mu)
{
    const S p[3] = { mu[0], mu[1], mu[2] };
    S pp[4];
    transform_point(cam.proj, p, pp, 4);
    const S d = pp[3] + S(0.0000001);
    const S g_nx = g_xy[0] * S(0.5) * S(cam.width);
    const S g_ny = g_xy[1] * S(0.5) * S(cam.height);
    S g_pp[4] = { g_nx / d, g_ny / d, 0, -(g_nx * pp[0] + g_ny * pp[1]) / (d * d) };
    for (int i = 0; i < 3; ++i)
        g_mu[i] = ((cam.proj[4 * i + 0] * g_pp[0] + cam.proj[4 * i + 1] * g_pp[1]) + cam.proj[4 * i + 2] * g_pp[2]) + cam.proj[4 * i + 3] * g_pp[3];
}

template <class S>
inline void view_rotation(const Camera<S>& cam, S W[3][3])
{
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            W[r][c] = cam.view[4 * c + r];
}

// larger2 / smaller2, math.h:112-155 (returned as indices into the 3-vector)
template <class S>
inline void larger2_idx(const S* v, int* i0, int* i1)
{
    if (v[0] < v[1]) {
        if (v[0] < v[2]) {
            *i0 = 1; *i1 = 2;
            return;
        }
    } else {
        if (v[1] < v[2]) {
            *i0 = 0; *i1 = 2;
            return;
        }
    }
    *i0 = 0; *i1 = 1;
}

template <class S>
inline void smaller2_idx(const S* v, int* i0, int* i1)
{
    if (v[0] > v[1]) {
        if (v[0] > v[2]) {
            *i0 = 1; *i1 = 2;
            return;
        }
    } else {
        if (v[1] > v[2]) {
            *i0 = 0; *i1 = 2;
            return;
        }
    }
    *i0 = 0; *i1 = 1;
}

// ---------------------------------------------------------------- screen-space splat
// The five weight formulations of math.h:322-410 (types.h:26-33).
enum class Formulation : int { Opacity = 0, Mass = 1, Density = 2, Ots = 3, Ols = 4 };

template <class S>
struct Splat2d {
    S weight;
    S xy[2];
    S cov[3];
};

// integral of exp(-1/2 x^T C^-1 x) over the plane: 2 pi sqrt(det C)
template <class S>
inline S integrate_exp_2d(const S* c) { return S(2 * kPi) * std::sqrt(c[0] * c[2] - c[1] * c[1]); }

template <class S>
inline void integrate_exp_2d_grad(const S* c, S g, S* gc)
{
    const S det = c[0] * c[2] - c[1] * c[1];
    const S f = g * S(2 * kPi) / (S(2) * std::sqrt(det));
    gc[0] = f * c[2];
    gc[1] = f * (S(-2) * c[1]);
    gc[2] = f * c[0];
}

// intermediate values of math::splat shared by forward and adjoint
template <class S>
struct SplatInternals {
    S t_unclamped[3];
    S t[3];
    S l;
    S J[3][3];
    S W[3][3];
    S T[3][3];
    S cov3[6];
    S cov2_unfiltered[3];
};

template <class S>
inline void splat_internals(const S* mu, const S* s, const S* q, const Camera<S>& cam, SplatInternals<S>* in)
{
    transform_point(cam.view, mu, in->t_unclamped, 3);
    const S lim_x = S(1.3) * cam.tan_fovx * in->t_unclamped[2];
    const S lim_y = S(1.3) * cam.tan_fovy * in->t_unclamped[2];
    in->t[0] = std::min(std::max(in->t_unclamped[0], -lim_x), lim_x);
    in->t[1] = std::min(std::max(in->t_unclamped[1], -lim_y), lim_y);
    in->t[2] = in->t_unclamped[2];
    in->l = length3(in->t);
    make_jacobian(in->t, in->l, cam.focal_x, cam.focal_y, in->J);
    view_rotation(cam, in->W);
    matmul3(in->J, in->W, in->T);
    compute_cov(s, q, in->cov3);
    affine_transform_and_cut(in->cov3, in->T, in->cov2_unfiltered);
}

// math::splat<F>, math.h:322-410
template <class S>
inline Splat2d<S> splat(Formulation F, S w, const S* mu, const S* s, const S* q, const Camera<S>& cam, S filter)
{
    SplatInternals<S> in;
    splat_internals(mu, s, q, cam, &in);
    Splat2d<S> g;
    project_to_screen(mu, cam, g.xy);
    g.cov[0] = in.cov2_unfiltered[0] + filter;
    g.cov[1] = in.cov2_unfiltered[1];
    g.cov[2] = in.cov2_unfiltered[2] + filter;
    switch (F) {
    case Formulation::Opacity: {
        const S* c = in.cov2_unfiltered;
        const S det_c = c[0] * c[2] - c[1] * c[1];
        const S det_n = g.cov[0] * g.cov[2] - g.cov[1] * g.cov[1];
        g.weight = w * std::sqrt(std::max(S(0.000025), det_c / det_n));
        break;
    }
    case Formulation::Mass: {
        g.weight = w * det3(in.J) * (S(1) / integrate_exp_2d(g.cov));
        break;
    }
    case Formulation::Density: {
        const S i3 = S(std::sqrt(std::pow(2 * kPi, 3.0))) * s[0] * s[1] * s[2];
        g.weight = w * det3(in.J) * i3 / integrate_exp_2d(g.cov);
        break;
    }
    case Formulation::Ots:
    case Formulation::Ols: {
        int i0, i1;
        if (F == Formulation::Ots)
            larger2_idx(s, &i0, &i1);
        else
            smaller2_idx(s, &i0, &i1);
        const S i2prime = S(2 * kPi) * s[i0] * s[i1];
        g.weight = w * det3(in.J) * i2prime / integrate_exp_2d(g.cov);
        break;
    }
    }
    return g;
}

// grad::splat<F>, grad/math.h:333-527 (Mass/Density share the Ots structure).
// g_xy, g_cov (stored), g_weight: incoming gradients of the 2D Gaussian.
template <class S>
inline void splat_grad(Formulation F, S w, const S* mu, const S* s, const S* q, const Camera<S>& cam, S filter,
    S g_weight, const S* g_xy, const S* g_cov2, S* out_g_w, S* out_g_mu, S* out_g_s, S* out_g_q)
{
    SplatInternals<S> in;
    splat_internals(mu, s, q, cam, &in);
    S cov_f[3] = { in.cov2_unfiltered[0] + filter, in.cov2_unfiltered[1], in.cov2_unfiltered[2] + filter };
    S g_c2[3] = { 0, 0, 0 };
    S g_J[3][3] = {};
    S g_s[3] = { 0, 0, 0 };
    S g_w = 0;
    switch (F) {
    case Formulation::Opacity: {
        const S* c = in.cov2_unfiltered;
        const S det_c = c[0] * c[2] - c[1] * c[1];
        const S det_n = cov_f[0] * cov_f[2] - cov_f[1] * cov_f[1];
        const S factor = det_c / det_n;
        g_w = g_weight * std::sqrt(std::max(S(0.000025), factor));
        for (int i = 0; i < 3; ++i)
            g_c2[i] = g_cov2[i];
        if (!(factor < S(0.000025))) {
            const S g_factor = g_weight * w / (S(2) * std::sqrt(factor));
            const S g_det_c = g_factor / det_n;
            const S g_det_n = -g_factor * det_c / (det_n * det_n);
            // d det / d [a, b, c] = [c, -2b, a]
            g_c2[0] += g_det_c * c[2] + g_det_n * cov_f[2];
            g_c2[1] += g_det_c * (S(-2) * c[1]) + g_det_n * (S(-2) * cov_f[1]);
            g_c2[2] += g_det_c * c[0] + g_det_n * cov_f[0];
        }
        break;
    }
    case Formulation::Mass:
    case Formulation::Density:
    case Formulation::Ots:
    case Formulation::Ols: {
        const S detJ = det3(in.J);
        S num = 1;
        int i0 = 0, i1 = 1;
        if (F == Formulation::Density)
            num = S(std::sqrt(std::pow(2 * kPi, 3.0))) * s[0] * s[1] * s[2];
        if (F == Formulation::Ots)
            larger2_idx(s, &i0, &i1);
        if (F == Formulation::Ols)
            smaller2_idx(s, &i0, &i1);
        if (F == Formulation::Ots || F == Formulation::Ols)
            num = S(2 * kPi) * s[i0] * s[i1];
        const S i2 = integrate_exp_2d(cov_f);
        g_w = g_weight * detJ * num / i2;
        const S g_det = g_weight * w * num / i2;
        const S g_num = g_weight * w * detJ / i2;
        const S g_i2 = -g_weight * w * detJ * num / (i2 * i2);
        S gc[3];
        integrate_exp_2d_grad(cov_f, g_i2, gc);
        for (int i = 0; i < 3; ++i)
            g_c2[i] = g_cov2[i] + gc[i];
        if (F == Formulation::Density) {
            const S f = S(std::sqrt(std::pow(2 * kPi, 3.0)));
            g_s[0] = g_num * f * s[1] * s[2];
            g_s[1] = g_num * f * s[0] * s[2];
            g_s[2] = g_num * f * s[0] * s[1];
        }
        if (F == Formulation::Ots || F == Formulation::Ols) {
            g_s[i0] = g_num * S(2 * kPi) * s[i1];
            g_s[i1] = g_num * S(2 * kPi) * s[i0];
        }
        det3_grad(in.J, g_det, g_J);
        break;
    }
    }
    // cov2 = cut(T Sigma T^T)
    S g_cov3[6], g_T[3][3];
    affine_transform_and_cut_grad(in.cov3, in.T, g_c2, g_cov3, g_T);
    // screen centroid
    S g_mu[3];
    project_to_screen_grad(mu, cam, g_xy, g_mu);
    // T = J W  ->  dJ += dT W^T
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            g_J[r][c] += (g_T[r][0] * in.W[c][0] + g_T[r][1] * in.W[c][1]) + g_T[r][2] * in.W[c][2];
    S g_t[3], g_l;
    make_jacobian_grad(in.t, in.l, cam.focal_x, cam.focal_y, g_J, g_t, &g_l);
    for (int i = 0; i < 3; ++i)
        g_t[i] += g_l * in.t[i] / in.l;
    // clamp_to_fov: the gradient passes where the value is inside the limits
    const S lim_x = S(1.3) * cam.tan_fovx * in.t_unclamped[2];
    const S lim_y = S(1.3) * cam.tan_fovy * in.t_unclamped[2];
    S g_tu[3] = { g_t[0], g_t[1], g_t[2] };
    if (in.t_unclamped[0] < -lim_x || in.t_unclamped[0] > lim_x)
        g_tu[0] = 0;
    if (in.t_unclamped[1] < -lim_y || in.t_unclamped[1] > lim_y)
        g_tu[1] = 0;
    // t = W mu + translation  ->  d mu += W^T g
    for (int i = 0; i < 3; ++i)
        g_mu[i] += (in.W[0][i] * g_tu[0] + in.W[1][i] * g_tu[1]) + in.W[2][i] * g_tu[2];
    S gs2[3], gq[4];
    compute_cov_grad(s, q, g_cov3, gs2, gq);
    for (int i = 0; i < 3; ++i) {
        out_g_mu[i] = g_mu[i];
        out_g_s[i] = g_s[i] + gs2[i];
    }
    for (int i = 0; i < 4; ++i)
        out_g_q[i] = gq[i];
    *out_g_w = g_w;
}

// ---------------------------------------------------------------- SH
// sh_to_colour, math.h:233-271.  sh: 16 x 3 coefficients.
template <class S>
inline void sh_to_colour(const S* sh, int deg, const S* dir, S* rgb, bool* clamped_ok)
{
    const auto c = [&](int k, int ch) { return sh[3 * k + ch]; };
    const S x = dir[0], y = dir[1], z = dir[2];
    for (int ch = 0; ch < 3; ++ch) {
        S r = S(SH_C0) * c(0, ch);
        if (deg > 0) {
            r = r - S(SH_C1) * y * c(1, ch) + S(SH_C1) * z * c(2, ch) - S(SH_C1) * x * c(3, ch);
            if (deg > 1) {
                const S xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
                r = r + (S(SH_C2[0]) * xy * c(4, ch) + S(SH_C2[1]) * yz * c(5, ch) + S(SH_C2[2]) * (S(2) * zz - xx - yy) * c(6, ch)
                    + S(SH_C2[3]) * xz * c(7, ch) + S(SH_C2[4]) * (xx - yy) * c(8, ch));
                if (deg > 2) {
                    r = r + (S(SH_C3[0]) * y * (S(3) * xx - yy) * c(9, ch) + S(SH_C3[1]) * xy * z * c(10, ch)
                        + S(SH_C3[2]) * y * (S(4) * zz - xx - yy) * c(11, ch) + S(SH_C3[3]) * z * (S(2) * zz - S(3) * xx - S(3) * yy) * c(12, ch)
                        + S(SH_C3[4]) * x * (S(4) * zz - xx - yy) * c(13, ch) + S(SH_C3[5]) * z * (xx - yy) * c(14, ch)
                        + S(SH_C3[6]) * x * (xx - S(3) * yy) * c(15, ch));
                }
            }
        }
        r = r + S(0.5);
        clamped_ok[ch] = r >= S(0);
        rgb[ch] = std::max(r, S(0));
    }
}

// basis values Y_k(dir) (with the reference's signs) and their direction derivatives
template <class S>
inline void sh_basis(int deg, const S* dir, S* Y, S dY[16][3])
{
    const S x = dir[0], y = dir[1], z = dir[2];
    for (int k = 0; k < 16; ++k) {
        Y[k] = 0;
        dY[k][0] = dY[k][1] = dY[k][2] = 0;
    }
    Y[0] = S(SH_C0);
    if (deg > 0) {
        Y[1] = -S(SH_C1) * y; dY[1][1] = -S(SH_C1);
        Y[2] = S(SH_C1) * z; dY[2][2] = S(SH_C1);
        Y[3] = -S(SH_C1) * x; dY[3][0] = -S(SH_C1);
    }
    if (deg > 1) {
        const S xx = x * x, yy = y * y, zz = z * z;
        Y[4] = S(SH_C2[0]) * x * y; dY[4][0] = S(SH_C2[0]) * y; dY[4][1] = S(SH_C2[0]) * x;
        Y[5] = S(SH_C2[1]) * y * z; dY[5][1] = S(SH_C2[1]) * z; dY[5][2] = S(SH_C2[1]) * y;
        Y[6] = S(SH_C2[2]) * (S(2) * zz - xx - yy); dY[6][0] = S(SH_C2[2]) * (-S(2) * x); dY[6][1] = S(SH_C2[2]) * (-S(2) * y); dY[6][2] = S(SH_C2[2]) * (S(4) * z);
        Y[7] = S(SH_C2[3]) * x * z; dY[7][0] = S(SH_C2[3]) * z; dY[7][2] = S(SH_C2[3]) * x;
        Y[8] = S(SH_C2[4]) * (xx - yy); dY[8][0] = S(SH_C2[4]) * S(2) * x; dY[8][1] = S(SH_C2[4]) * (-S(2) * y);
    }
    if (deg > 2) {
        const S xx = x * x, yy = y * y, zz = z * z;
        Y[9] = S(SH_C3[0]) * y * (S(3) * xx - yy);
        dY[9][0] = S(SH_C3[0]) * S(6) * x * y; dY[9][1] = S(SH_C3[0]) * (S(3) * xx - S(3) * yy);
        Y[10] = S(SH_C3[1]) * x * y * z;
        dY[10][0] = S(SH_C3[1]) * y * z; dY[10][1] = S(SH_C3[1]) * x * z; dY[10][2] = S(SH_C3[1]) * x * y;
        Y[11] = S(SH_C3[2]) * y * (S(4) * zz - xx - yy);
        dY[11][0] = S(SH_C3[2]) * (-S(2) * x * y); dY[11][1] = S(SH_C3[2]) * (S(4) * zz - xx - S(3) * yy); dY[11][2] = S(SH_C3[2]) * S(8) * y * z;
        Y[12] = S(SH_C3[3]) * z * (S(2) * zz - S(3) * xx - S(3) * yy);
        dY[12][0] = S(SH_C3[3]) * (-S(6) * x * z); dY[12][1] = S(SH_C3[3]) * (-S(6) * y * z); dY[12][2] = S(SH_C3[3]) * (S(6) * zz - S(3) * xx - S(3) * yy);
        Y[13] = S(SH_C3[4]) * x * (S(4) * zz - xx - yy);
        dY[13][0] = S(SH_C3[4]) * (S(4) * zz - S(3) * xx - yy); dY[13][1] = S(SH_C3[4]) * (-S(2) * x * y); dY[13][2] = S(SH_C3[4]) * S(8) * x * z;
        Y[14] = S(SH_C3[5]) * z * (xx - yy);
        dY[14][0] = S(SH_C3[5]) * S(2) * x * z; dY[14][1] = S(SH_C3[5]) * (-S(2) * y * z); dY[14][2] = S(SH_C3[5]) * (xx - yy);
        Y[15] = S(SH_C3[6]) * x * (xx - S(3) * yy);
        dY[15][0] = S(SH_C3[6]) * (S(3) * xx - S(3) * yy); dY[15][1] = S(SH_C3[6]) * (-S(6) * x * y);
    }
}

// grad::sh_to_colour, grad/math.h:203-282: g_sh (16 x 3) and g_dir
template <class S>
inline void sh_to_colour_grad(const S* sh, int deg, const S* dir, const S* g_rgb, const bool* clamped_ok, S* g_sh, S* g_dir)
{
    S Y[16], dY[16][3];
    sh_basis(deg, dir, Y, dY);
    S gr[3];
    for (int ch = 0; ch < 3; ++ch)
        gr[ch] = clamped_ok[ch] ? g_rgb[ch] : S(0);
    g_dir[0] = g_dir[1] = g_dir[2] = 0;
    for (int k = 0; k < 16; ++k) {
        for (int ch = 0; ch < 3; ++ch) {
            g_sh[3 * k + ch] = Y[k] * gr[ch];
            for (int a = 0; a < 3; ++a)
                g_dir[a] += dY[k][a] * sh[3 * k + ch] * gr[ch];
        }
    }
}

} // namespace orc
