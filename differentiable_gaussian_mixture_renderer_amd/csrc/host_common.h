// SPDX-License-Identifier: GPL-3.0-or-later
// Host-side helpers shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/dgmr.h"
#include "device_math.h"

namespace dgmr {

void set_error(const std::string& msg);

// largest image side accepted: tile grids stay below 2^16 tiles per axis (the
// pair emission packs a grid width and a rectangle width into 16 bits each)
constexpr int32_t kMaxImageSide = 1 << 15;

inline bool valid_image(int32_t w, int32_t h) { return w > 0 && h > 0 && w <= kMaxImageSide && h <= kMaxImageSide; }

#define DGMR_HIP_CHECK(expr)                                                                                   \
    do {                                                                                                       \
        hipError_t e_ = (expr);                                                                                \
        if (e_ != hipSuccess) {                                                                                \
            ::dgmr::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));                             \
            return DGMR_ERR_DEVICE;                                                                            \
        }                                                                                                      \
    } while (0)

#define DGMR_ARG_CHECK(cond, msg)                                                                              \
    do {                                                                                                       \
        if (!(cond)) {                                                                                         \
            ::dgmr::set_error(msg);                                                                            \
            return DGMR_ERR_ARGUMENT;                                                                          \
        }                                                                                                      \
    } while (0)

// bump allocator over a caller-provided workspace (256-byte aligned slices)
struct Carver {
    char* base;
    size_t cap;
    size_t used = 0;
    Carver(void* b, size_t c) : base(static_cast<char*>(b)), cap(c) {}
    template <class T>
    T* take(size_t count)
    {
        const size_t bytes = ((count * sizeof(T) + 255) / 256) * 256;
        T* p = reinterpret_cast<T*>(base ? base + used : nullptr);
        used += bytes;
        return p;
    }
    bool fits() const { return used <= cap; }
};

inline int64_t ceil_div64(int64_t a, int64_t b) { return (a + b - 1) / b; }

bool valid_tiling(const dgmr_tiling* t);

// fills the device camera: focal lengths as splat_forward.cu:35-36, inverse
// projection for the marcher's rays (vol_marcher_forward.cu:302)
template <class S>
DevCamera<S> make_dev_camera(const dgmr_camera& c, const dgmr_tiling& t);

uint32_t higher_msb(uint32_t n);

// stage timing (dgmr_profile_enable): a scope records a start event on
// construction and a stop event on destruction when profiling is on
struct StageScope {
    StageScope(const char* name, hipStream_t stream, double bytes = 0);
    ~StageScope();
    int slot = -1;
    hipStream_t stream;
};
// sets the byte model of a recorded stage once it is known (no-op for slot < 0)
void stage_set_bytes(int slot, double bytes);

// The pair count R read back without stalling the stream: it is copied to
// pinned host memory right after the scan that produces it, the kernels that
// depend on it read it on the device (grids sized by the pair capacity), and
// the host waits for the copy only when the value is needed, after every
// launch of the entry point has been queued.  Per thread: one pending count
// at a time.
struct PendingCount {
    const uint32_t* device = nullptr; // the count on the device, readable by later kernels
    const volatile uint32_t* host = nullptr;
    hipEvent_t done = nullptr;
    // stages whose byte model depends on R: bytes = per_pair * R + fixed
    int slot[4] = { -1, -1, -1, -1 };
    double per_pair[4] = { 0, 0, 0, 0 }, fixed[4] = { 0, 0, 0, 0 };
    int n_slots = 0;
    void add_stage(int s, double pp, double fx)
    {
        if (n_slots < 4) {
            slot[n_slots] = s;
            per_pair[n_slots] = pp;
            fixed[n_slots] = fx;
            ++n_slots;
        }
    }
    // R, or -1 when waiting failed; fills the stage byte models
    int64_t wait();
};
// queues the copy of *dev_count; returns DGMR_OK or DGMR_ERR_DEVICE
int enqueue_count(const uint32_t* dev_count, hipStream_t st, PendingCount* pc);

// multi-view binning: Gaussian ids span n_views blocks of n_per_view; view v's
// tiles occupy [views[v].z, views[v].z + views[v].x * views[v].y) of the
// global tile range (views: device array of (grid_x, grid_y, first tile, -))
// Pairs are sorted on the view-local tile id only: emission is in (view,
// depth) order and the sort is stable, so each (view, tile) run stays
// contiguous and in depth order; the tile ranges add the view's first tile.
struct BatchBinning {
    const int4* views;
    int64_t n_per_view;
    int n_views;
    int64_t total_tiles;
    int64_t max_view_tiles; // largest grid_x * grid_y over the views
};

// binning shared by both renderers (implemented with the splatter): depth sort,
// scan, emission, pair sort, tile ranges.  The preprocess must have filled the
// depth-key and id buffers returned by binning_key_buffers, and its records
// too when bin_rects_written (otherwise they are derived from rects and xy):
// the bin rectangles, or with cull (tight culling) the reach records.
// R comes back through *pending (wait() after the dependent launches); work
// sized by R is bounded by bins->pair_capacity, and R > capacity leaves a
// truncated, in-bounds binning that the caller reports as DGMR_ERR_CAPACITY.
int bin_gaussians(int64_t n, const int32_t* rects, const uint32_t* tiles, const void* xy, bool xy_double, bool bin_rects_written, bool cull,
    int grid_x, int grid_y, int tile_w, int tile_h, dgmr_binning* bins, Carver& cv, PendingCount* pending, hipStream_t st,
    const BatchBinning* batch);
// the preprocess's outputs in the workspace: depth keys, ids and the
// per-Gaussian records (cull: ReachRec; otherwise uint2 bin rectangles
// (first tile y0 * grid_x + x0, (y1 - y0) << 16 | (x1 - x0)), zero when culled).
// ids comes back null when the binning numbers the entries itself (large n):
// the preprocess then writes none
void binning_key_buffers(void* workspace, int64_t n, bool cull, uint32_t** keys, uint32_t** ids, void** records = nullptr);
size_t splat_workspace_bytes(int64_t n, int64_t cap, bool cull = false);

} // namespace dgmr
