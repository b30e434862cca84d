// SPDX-License-Identifier: GPL-3.0-or-later
// Whether a screen-space Gaussian can reach alpha >= 1/255 inside a tile: the
// blends stage only the Gaussians of a batch that can, since the others blend
// nowhere in the tile.  Plain float arithmetic, usable on the host as well
// (the CPU tests check it against a brute-force evaluation over the pixels).
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define DGMR_HOST_DEVICE __host__ __device__
#else
#define DGMR_HOST_DEVICE
#endif

namespace dgmr {

// Screen centre (x, y), conic (a, b, c) and opacity o; box [bx0, bx1] x
// [by0, by1] in integer pixel coordinates, as the blend evaluates them.
// alpha = o e^-Q with Q(d) = (a dx^2 + 2 b dx dy + c dy^2) / 2, so the
// Gaussian reaches 1/255 in the box iff the minimum of Q over the box is at
// most ln(255 o).  Q is convex: the minimum is 0 when the centre lies in the
// box, else it lies on an edge, where Q is a 1D quadratic minimised at its
// stationary point clamped to the edge.  The test is conservative: a
// relative 1e-4 and absolute 1e-3 margin covers the float evaluation of Q in
// the blend, and a conic that is not positive definite (or NaN) reaches.
DGMR_HOST_DEVICE inline bool reaches_box(float x, float y, float a, float b, float c, float o, float bx0, float bx1, float by0, float by1)
{
    if (!(a > 0.f && c > 0.f && a * c - b * b > 0.f))
        return true;
    const float lim = logf(o) + 5.5412635451584258f; // ln(255 o)
    const float ex0 = bx0 - x, ex1 = bx1 - x, ey0 = by0 - y, ey1 = by1 - y;
    if (ex0 <= 0.f && ex1 >= 0.f && ey0 <= 0.f && ey1 >= 0.f)
        return lim >= -1e-3f;
    const float ia = 1.f / a, ic = 1.f / c;
    const auto q = [&](float dx, float dy) { return 0.5f * (a * dx * dx + 2.f * b * dx * dy + c * dy * dy); };
    const auto clampf = [](float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); };
    const float q_x0 = q(ex0, clampf(-b * ex0 * ic, ey0, ey1));
    const float q_x1 = q(ex1, clampf(-b * ex1 * ic, ey0, ey1));
    const float q_y0 = q(clampf(-b * ey0 * ia, ex0, ex1), ey0);
    const float q_y1 = q(clampf(-b * ey1 * ia, ex0, ex1), ey1);
    const float qmin = fminf(fminf(q_x0, q_x1), fminf(q_y0, q_y1));
    return qmin * (1.f - 1e-4f) - 1e-3f <= lim;
}

} // namespace dgmr
