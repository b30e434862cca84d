// SPDX-License-Identifier: GPL-3.0-or-later
// Error reporting, camera set-up and workspace sizing shared by the C ABI.
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "host_common.h"
#include "primitives.h"

namespace dgmr {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string& msg) { g_last_error = msg; }

bool valid_tiling(const dgmr_tiling* t)
{
    if (!t)
        return false;
    const bool shape = (t->tile_w == 16 && t->tile_h == 16) || (t->tile_w == 8 && t->tile_h == 8) || (t->tile_w == 32 && t->tile_h == 8);
    const bool batch = t->batch == 128 || t->batch == 256 || t->batch == 512;
    return shape && batch;
}

uint32_t higher_msb(uint32_t n)
{
    uint32_t r = 0;
    while (r < 32 && (n >> r) != 0)
        ++r;
    return r;
}

// 4x4 inverse in double (Gauss-Jordan with partial pivoting)
static bool invert4(const double* m, double* out)
{
    double a[4][8];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 8; ++c)
            a[r][c] = c < 4 ? m[4 * r + c] : (c - 4 == r ? 1.0 : 0.0);
    for (int c = 0; c < 4; ++c) {
        int piv = c;
        for (int r = c + 1; r < 4; ++r)
            if (std::fabs(a[r][c]) > std::fabs(a[piv][c]))
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

template <class S>
DevCamera<S> make_dev_camera(const dgmr_camera& c, const dgmr_tiling& t)
{
    DevCamera<S> d;
    // matrices arrive as the scene dtype would hold them
    double proj_s[16];
    for (int i = 0; i < 16; ++i) {
        d.view[i] = S(c.view[i]);
        d.proj[i] = S(c.proj[i]);
        proj_s[i] = double(d.proj[i]);
    }
    double inv[16];
    if (!invert4(proj_s, inv))
        std::memset(inv, 0, sizeof(inv));
    for (int i = 0; i < 16; ++i)
        d.inv_proj[i] = S(inv[i]);
    for (int i = 0; i < 3; ++i) {
        d.pos[i] = S(c.position[i]);
        d.bg[i] = S(c.background[i]);
    }
    // the reference passes tan_fov as float even for double scenes (vol_marcher.cu:229-230)
    d.tan_fovx = S(float(c.tan_fovx));
    d.tan_fovy = S(float(c.tan_fovy));
    d.focal_y = S(c.height) / (S(2) * d.tan_fovy);
    d.focal_x = S(c.width) / (S(2) * d.tan_fovx);
    d.width = c.width;
    d.height = c.height;
    d.tile_w = t.tile_w;
    d.tile_h = t.tile_h;
    d.grid_x = (c.width + t.tile_w - 1) / t.tile_w;
    d.grid_y = (c.height + t.tile_h - 1) / t.tile_h;
    return d;
}

template DevCamera<float> make_dev_camera<float>(const dgmr_camera&, const dgmr_tiling&);
template DevCamera<double> make_dev_camera<double>(const dgmr_camera&, const dgmr_tiling&);

} // namespace dgmr

extern "C" {

const char* dgmr_last_error(void) { return dgmr::g_last_error.c_str(); }

const char* dgmr_version(void) { return "dgmr-mi355x 0.1 (gfx950)"; }

} // extern "C"
