// Host build of the blend's tile-reach test (tile_reach.h) for the CPU tests.
#include <cstdint>

#include "../../differentiable_gaussian_mixture_renderer_amd/csrc/tile_reach.h"

extern "C" void reach_box_batch(int n, const float* g, const float* box, uint8_t* out)
{
    // g: [n, 6] x, y, a, b, c, o; box: [n, 4] x0, x1, y0, y1
    for (int i = 0; i < n; ++i) {
        const float* p = g + 6 * i;
        const float* bx = box + 4 * i;
        out[i] = dgmr::reaches_box(p[0], p[1], p[2], p[3], p[4], p[5], bx[0], bx[1], bx[2], bx[3]) ? 1 : 0;
    }
}
