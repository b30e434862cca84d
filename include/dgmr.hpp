/* SPDX-License-Identifier: GPL-3.0-or-later
 *
 * C++ façade over the C ABI (dgmr.h), for C++ callers of the reference's
 * C++ interface: dgmr::splat_forward (splat_forward.h:61),
 * dgmr::vol_marcher::forward<T> / backward<T> (vol_marcher_forward.h:44,
 * vol_marcher_backward.h:37), dgmr::vol_raster_forward
 * (vol_raster_forward.h:70) and dgmr::sibr_interfaces::{splat, vol_march}
 * (sibr_interfaces.h:24-27).  The data structs keep the reference's field
 * names; glm and whack types become plain arrays and device-pointer views:
 *
 *   * a TensorView<const glm::vec3, 1> of N Gaussians is a View<const T>
 *     over 3N scalars (quaternions 4N, w first; SHs<3> 48N; weights N);
 *   * a glm::mat4 is a Mat4<T>, 16 scalars column-major (glm's layout, the
 *     bytes of the INRIA world_view_transform / full_proj_transform tensors);
 *   * a TensorView<T, 3> framebuffer is a Framebuffer<T>, [3, H, W] device.
 *
 * Where the reference returns torch tensors (vol_marcher::ForwardCache,
 * Gradients) the façade returns DeviceBuffers it owns.  Scratch (workspace,
 * pair buffers) lives in a Renderer, grown on demand and reused; the free
 * functions use one Renderer per thread.  Errors throw dgmr::Error with the
 * library's message.  Header only; link with -ldgmr_hip.
 */
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>

#include "dgmr.h"

namespace dgmr {

class Error : public std::runtime_error {
public:
    Error(int status, const std::string& what) : std::runtime_error(what), status_(status) { }
    int status() const { return status_; }

private:
    int status_;
};

inline void check(int rc, const char* where)
{
    if (rc != DGMR_OK)
        throw Error(rc, std::string(where) + ": " + dgmr_last_error());
}

inline void check_hip(hipError_t e, const char* where)
{
    if (e != hipSuccess)
        throw Error(DGMR_ERR_DEVICE, std::string(where) + ": " + hipGetErrorString(e));
}

/* Owned device memory (hipMalloc / hipFree), grown on demand, never shrunk. */
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t bytes) { reserve(bytes); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept : p_(std::exchange(o.p_, nullptr)), bytes_(std::exchange(o.bytes_, 0)) { }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept
    {
        if (this != &o) {
            release();
            p_ = std::exchange(o.p_, nullptr);
            bytes_ = std::exchange(o.bytes_, 0);
        }
        return *this;
    }
    ~DeviceBuffer() { release(); }
    void reserve(size_t bytes)
    {
        if (bytes <= bytes_)
            return;
        release();
        check_hip(hipMalloc(&p_, bytes), "hipMalloc");
        bytes_ = bytes;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p_); }
    void* data() const { return p_; }
    size_t bytes() const { return bytes_; }

private:
    void release()
    {
        if (p_)
            (void)hipFree(p_);
        p_ = nullptr;
        bytes_ = 0;
    }
    void* p_ = nullptr;
    size_t bytes_ = 0;
};

template <class T>
struct View {
    T* data = nullptr;
    int64_t size = 0; /* scalars */
};

template <class T>
using Mat4 = std::array<T, 16>;
template <class T>
using Vec3 = std::array<T, 3>;

template <class T>
struct Framebuffer {
    T* data = nullptr; /* [3, height, width] */
    int height = 0, width = 0;
};

/* ------------------------------------------------------------------ splatter */
struct Statistics {
    unsigned n_rendered = 0; /* (tile, Gaussian) pairs; the reference returns 0 */
};

/* splat_forward.h:36-52 */
struct SplatForwardData {
    View<const float> gm_centroids;     /* [N,3] */
    View<const float> gm_sh_params;     /* [N,16,3] */
    View<const float> gm_weights;       /* [N] */
    View<const float> gm_cov_scales;    /* [N,3] */
    View<const float> gm_cov_rotations; /* [N,4] w,x,y,z */

    Framebuffer<float> framebuffer;
    Mat4<float> view_matrix {};
    Mat4<float> proj_matrix {};
    Vec3<float> cam_poition {};
    Vec3<float> background {};
    int sh_degree = 3;
    float cov_scale_multiplier = 1.f;
    float tan_fovy = 0.f;
    float tan_fovx = 0.f;
    bool debug = true;
};

/* ----------------------------------------------------------------- rasteriser */
struct VolRasterStatistics {
    unsigned n_rendered = 0; /* (tile, Gaussian) pairs; the reference returns 0 */
};

/* vol_raster_forward.h:31-57 */
struct VolRasterForwardData {
    View<const float> gm_centroids;
    View<const float> gm_sh_params;
    View<const float> gm_weights;
    View<const float> gm_cov_scales;
    View<const float> gm_cov_rotations;

    Framebuffer<float> framebuffer;
    Mat4<float> view_matrix {};
    Mat4<float> proj_matrix {};
    Vec3<float> cam_poition {};
    Vec3<float> background {};
    int sh_degree = 3;
    float cov_scale_multiplier = 1.f;
    float tan_fovy = 0.f;
    float tan_fovx = 0.f;
    float max_depth = 20.f;
    bool debug = true;
    enum class RenderMode { Full, Bins, Depth };
    RenderMode debug_render_mode = RenderMode::Full;
    int debug_render_bin = -1;
};

/* -------------------------------------------------------------------- marcher */
namespace vol_marcher {

/* vol_marcher.h:29-57 */
template <class scalar_t>
struct ForwardData {
    View<const scalar_t> gm_centroids;
    View<const scalar_t> gm_sh_params;
    View<const scalar_t> gm_weights;
    View<const scalar_t> gm_cov_scales;
    View<const scalar_t> gm_cov_rotations;

    Mat4<scalar_t> view_matrix {};
    Mat4<scalar_t> proj_matrix {};
    Vec3<scalar_t> cam_poition {};
    Vec3<scalar_t> background {};
    int sh_degree = 3;
    scalar_t cov_scale_multiplier = 1;
    scalar_t tan_fovy = 0;
    scalar_t tan_fovx = 0;
    scalar_t max_depth = 20;
    bool debug = true;
    enum class RenderMode { Full, Bins, Depth };
    RenderMode debug_render_mode = RenderMode::Full;
    int debug_render_bin = -1;
};

/* vol_marcher_forward.h:28-41, arrays of the scene scalar unless noted */
struct ForwardCache {
    DeviceBuffer rects;                  /* int32 [N,2] */
    DeviceBuffer rgb;                    /* [N,3] */
    DeviceBuffer rgb_sh_clamped;         /* uint8 [N,3] */
    DeviceBuffer depths;                 /* [N] */
    DeviceBuffer points_xy_image;        /* [N,2] */
    DeviceBuffer inverse_filtered_cov3d; /* [N,6] */
    DeviceBuffer filtered_masses;        /* [N] */
    DeviceBuffer tiles_touched;          /* uint32 [N] */
    DeviceBuffer point_offsets;          /* uint32 [N] */
    DeviceBuffer i_ranges;               /* uint32 [tiles,2] */
    DeviceBuffer b_point_list;           /* uint32 [n_pairs] */
    DeviceBuffer remaining_transparency; /* [H,W] */
    int64_t n_gaussians = 0;
    int64_t n_pairs = 0;
    int width = 0, height = 0;
};

/* vol_marcher_backward.h:28-34 */
struct Gradients {
    DeviceBuffer gm_centroids;     /* [N,3] */
    DeviceBuffer gm_sh_params;     /* [N,16,3] */
    DeviceBuffer gm_weights;       /* [N] */
    DeviceBuffer gm_cov_scales;    /* [N,3] */
    DeviceBuffer gm_cov_rotations; /* [N,4] */
};

} // namespace vol_marcher

namespace detail {

template <class T>
inline dgmr_gaussians gaussians(const View<const T>& centroids, const View<const T>& sh, const View<const T>& weights, const View<const T>& scales,
    const View<const T>& rotations, int sh_degree, double multiplier)
{
    dgmr_gaussians g {};
    g.n = weights.size;
    g.dtype = std::is_same<T, double>::value ? DGMR_F64 : DGMR_F32;
    g.means = centroids.data;
    g.scales = scales.data;
    g.rotations = rotations.data;
    g.opacity = weights.data;
    g.sh = sh.data;
    g.sh_coeffs = g.n > 0 ? int32_t(sh.size / (3 * g.n)) : 16;
    g.sh_degree = sh_degree;
    g.scale_multiplier = multiplier;
    g.activation = DGMR_ACT_NONE;
    if (centroids.size != 3 * g.n || scales.size != 3 * g.n || rotations.size != 4 * g.n || (g.n > 0 && sh.size != 3 * g.n * g.sh_coeffs))
        throw Error(DGMR_ERR_ARGUMENT, "dgmr: Gaussian arrays disagree on N");
    return g;
}

template <class T>
inline dgmr_camera camera(const Mat4<T>& view, const Mat4<T>& proj, const Vec3<T>& pos, const Vec3<T>& bg, T tan_fovx, T tan_fovy, int width,
    int height)
{
    dgmr_camera c {};
    for (int i = 0; i < 16; ++i) {
        c.view[i] = double(view[i]);
        c.proj[i] = double(proj[i]);
    }
    for (int i = 0; i < 3; ++i) {
        c.position[i] = double(pos[i]);
        c.background[i] = double(bg[i]);
    }
    c.tan_fovx = double(tan_fovx);
    c.tan_fovy = double(tan_fovy);
    c.width = width;
    c.height = height;
    return c;
}

} // namespace detail

/* Scratch owner: workspace and per-frame buffers reused across calls.  Calls
 * on one Renderer are queued on its stream in order; a Renderer is not
 * thread-safe. */
class Renderer {
public:
    explicit Renderer(hipStream_t stream = nullptr) : stream_(stream) { }

    Statistics splat_forward(SplatForwardData& d)
    {
        const dgmr_gaussians g = detail::gaussians(d.gm_centroids, d.gm_sh_params, d.gm_weights, d.gm_cov_scales, d.gm_cov_rotations,
            d.sh_degree, d.cov_scale_multiplier);
        const dgmr_camera cam = detail::camera(d.view_matrix, d.proj_matrix, d.cam_poition, d.background, d.tan_fovx, d.tan_fovy,
            d.framebuffer.width, d.framebuffer.height);
        const dgmr_tiling tiling { 16, 16, 64 };
        const int64_t n = g.n;
        splat_geometry(n);
        final_T_.reserve(sizeof(float) * size_t(cam.width) * size_t(cam.height));
        dgmr_splat_geometry geom { sg_rects_.as<int32_t>(), sg_tiles_.as<uint32_t>(), sg_xy_.as<float>(), sg_conic_.as<float>(),
            sg_rgb_.as<float>(), sg_clamped_.as<uint8_t>(), sg_depths_.as<float>() };
        dgmr_stats stats {};
        run_with_capacity(n, cam, tiling, [&](dgmr_binning* bins) {
            return dgmr_splat_forward(&g, &cam, &tiling, &geom, bins, d.framebuffer.data, final_T_.as<float>(), ws_.data(), ws_.bytes(), &stats,
                stream_);
        }, &stats, "dgmr_splat_forward");
        return Statistics { unsigned(stats.n_pairs) };
    }

    VolRasterStatistics vol_raster_forward(VolRasterForwardData& d)
    {
        const dgmr_gaussians g = detail::gaussians(d.gm_centroids, d.gm_sh_params, d.gm_weights, d.gm_cov_scales, d.gm_cov_rotations,
            d.sh_degree, d.cov_scale_multiplier);
        const dgmr_camera cam = detail::camera(d.view_matrix, d.proj_matrix, d.cam_poition, d.background, d.tan_fovx, d.tan_fovy,
            d.framebuffer.width, d.framebuffer.height);
        const dgmr_tiling tiling { 16, 16, 64 };
        const int64_t n = g.n;
        splat_geometry(n);
        raster_inv_cov_.reserve(sizeof(float) * 6 * size_t(n > 0 ? n : 1));
        raster_mass_.reserve(sizeof(float) * size_t(n > 0 ? n : 1));
        dgmr_splat_geometry geom { sg_rects_.as<int32_t>(), sg_tiles_.as<uint32_t>(), sg_xy_.as<float>(), sg_conic_.as<float>(),
            sg_rgb_.as<float>(), sg_clamped_.as<uint8_t>(), sg_depths_.as<float>() };
        const dgmr_raster_options opt { int32_t(d.debug_render_mode), d.debug_render_bin, d.max_depth };
        dgmr_stats stats {};
        run_with_capacity(n, cam, tiling, [&](dgmr_binning* bins) {
            return dgmr_raster_forward(&g, &cam, &opt, &geom, bins, raster_inv_cov_.as<float>(), raster_mass_.as<float>(), d.framebuffer.data,
                ws_.data(), ws_.bytes(), &stats, stream_);
        }, &stats, "dgmr_raster_forward");
        return VolRasterStatistics { unsigned(stats.n_pairs) };
    }

    template <class scalar_t>
    vol_marcher::ForwardCache march_forward(Framebuffer<scalar_t> fb, const vol_marcher::ForwardData<scalar_t>& d, int max_large_steps = 0)
    {
        const dgmr_gaussians g = detail::gaussians(d.gm_centroids, d.gm_sh_params, d.gm_weights, d.gm_cov_scales, d.gm_cov_rotations,
            d.sh_degree, double(d.cov_scale_multiplier));
        const dgmr_camera cam = detail::camera(d.view_matrix, d.proj_matrix, d.cam_poition, d.background, d.tan_fovx, d.tan_fovy, fb.width,
            fb.height);
        const dgmr_tiling tiling { 16, 16, 64 };
        const int64_t n = g.n, nn = n > 0 ? n : 1;
        const size_t S = sizeof(scalar_t);
        vol_marcher::ForwardCache c;
        c.n_gaussians = n;
        c.width = fb.width;
        c.height = fb.height;
        c.rects.reserve(8 * size_t(nn));
        c.rgb.reserve(3 * S * size_t(nn));
        c.rgb_sh_clamped.reserve(3 * size_t(nn));
        c.depths.reserve(S * size_t(nn));
        c.points_xy_image.reserve(2 * S * size_t(nn));
        c.inverse_filtered_cov3d.reserve(6 * S * size_t(nn));
        c.filtered_masses.reserve(S * size_t(nn));
        c.tiles_touched.reserve(4 * size_t(nn));
        c.point_offsets.reserve(4 * size_t(nn));
        c.i_ranges.reserve(8 * size_t(tiles(cam)));
        c.remaining_transparency.reserve(S * size_t(fb.width) * size_t(fb.height));
        dgmr_march_geometry geom { c.rects.as<int32_t>(), c.rgb.data(), c.rgb_sh_clamped.as<uint8_t>(), c.depths.data(), c.points_xy_image.data(),
            c.inverse_filtered_cov3d.data(), c.filtered_masses.data(), c.tiles_touched.as<uint32_t>() };
        dgmr_march_options opt {};
        opt.max_large_steps = max_large_steps;
        opt.render_mode = int32_t(d.debug_render_mode);
        opt.debug_render_bin = d.debug_render_bin;
        opt.max_depth = double(d.max_depth);
        dgmr_stats stats {};
        int64_t cap = capacity_ > 0 ? capacity_ : 4 * nn + 4096;
        for (int attempt = 0;; ++attempt) {
            c.b_point_list.reserve(4 * size_t(cap));
            size_t ws_bytes = 0;
            check(dgmr_workspace_size(n, cam.width, cam.height, &tiling, cap, &ws_bytes), "dgmr_workspace_size");
            ws_.reserve(ws_bytes);
            dgmr_binning bins { c.point_offsets.as<uint32_t>(), c.b_point_list.as<uint32_t>(), c.i_ranges.as<uint32_t>(), cap };
            const int rc = dgmr_march_forward(&g, &cam, &tiling, &opt, &geom, &bins, fb.data, c.remaining_transparency.data(), ws_.data(),
                ws_.bytes(), &stats, stream_);
            if (rc == DGMR_ERR_CAPACITY && attempt < 2) {
                cap = stats.n_pairs + stats.n_pairs / 4 + 1024;
                continue;
            }
            check(rc, "dgmr_march_forward");
            break;
        }
        capacity_ = cap > capacity_ ? cap : capacity_;
        c.n_pairs = stats.n_pairs;
        return c;
    }

    template <class scalar_t>
    vol_marcher::Gradients march_backward(Framebuffer<const scalar_t> fb, const vol_marcher::ForwardData<scalar_t>& d,
        const vol_marcher::ForwardCache& c, const scalar_t* grad, int max_large_steps = 0)
    {
        const dgmr_gaussians g = detail::gaussians(d.gm_centroids, d.gm_sh_params, d.gm_weights, d.gm_cov_scales, d.gm_cov_rotations,
            d.sh_degree, double(d.cov_scale_multiplier));
        const dgmr_camera cam = detail::camera(d.view_matrix, d.proj_matrix, d.cam_poition, d.background, d.tan_fovx, d.tan_fovy, fb.width,
            fb.height);
        const dgmr_tiling tiling { 16, 16, 64 };
        const int64_t nn = g.n > 0 ? g.n : 1;
        const size_t S = sizeof(scalar_t);
        vol_marcher::Gradients out;
        out.gm_centroids.reserve(3 * S * size_t(nn));
        out.gm_cov_scales.reserve(3 * S * size_t(nn));
        out.gm_cov_rotations.reserve(4 * S * size_t(nn));
        out.gm_weights.reserve(S * size_t(nn));
        out.gm_sh_params.reserve(3 * S * size_t(nn) * size_t(g.sh_coeffs));
        render_grads_.reserve(13 * S * size_t(nn));
        dgmr_march_geometry geom { c.rects.as<int32_t>(), c.rgb.data(), c.rgb_sh_clamped.as<uint8_t>(), c.depths.data(), c.points_xy_image.data(),
            c.inverse_filtered_cov3d.data(), c.filtered_masses.data(), c.tiles_touched.as<uint32_t>() };
        dgmr_binning bins { c.point_offsets.as<uint32_t>(), c.b_point_list.as<uint32_t>(), c.i_ranges.as<uint32_t>(), c.n_pairs };
        dgmr_march_options opt {};
        opt.max_large_steps = max_large_steps;
        check(dgmr_march_backward(&g, &cam, &tiling, &opt, &geom, &bins, c.n_pairs, fb.data, c.remaining_transparency.data(), grad,
                  render_grads_.data(), out.gm_centroids.data(), out.gm_cov_scales.data(), out.gm_cov_rotations.data(), out.gm_weights.data(),
                  out.gm_sh_params.data(), stream_),
            "dgmr_march_backward");
        return out;
    }

    hipStream_t stream() const { return stream_; }

private:
    static int64_t tiles(const dgmr_camera& cam) { return int64_t((cam.width + 15) / 16) * ((cam.height + 15) / 16); }

    void splat_geometry(int64_t n)
    {
        const size_t nn = size_t(n > 0 ? n : 1);
        sg_rects_.reserve(8 * nn);
        sg_tiles_.reserve(4 * nn);
        sg_xy_.reserve(8 * nn);
        sg_conic_.reserve(16 * nn);
        sg_rgb_.reserve(12 * nn);
        sg_clamped_.reserve(3 * nn);
        sg_depths_.reserve(4 * nn);
        offsets_.reserve(4 * nn);
    }

    template <class Call>
    void run_with_capacity(int64_t n, const dgmr_camera& cam, const dgmr_tiling& tiling, Call call, dgmr_stats* stats, const char* where)
    {
        ranges_.reserve(8 * size_t(tiles(cam)));
        int64_t cap = capacity_ > 0 ? capacity_ : 4 * (n > 0 ? n : 1) + 4096;
        for (int attempt = 0;; ++attempt) {
            point_list_.reserve(4 * size_t(cap));
            size_t ws_bytes = 0;
            check(dgmr_workspace_size(n, cam.width, cam.height, &tiling, cap, &ws_bytes), "dgmr_workspace_size");
            ws_.reserve(ws_bytes);
            dgmr_binning bins { offsets_.as<uint32_t>(), point_list_.as<uint32_t>(), ranges_.as<uint32_t>(), cap };
            const int rc = call(&bins);
            if (rc == DGMR_ERR_CAPACITY && attempt < 2) {
                cap = stats->n_pairs + stats->n_pairs / 4 + 1024;
                continue;
            }
            check(rc, where);
            break;
        }
        capacity_ = cap > capacity_ ? cap : capacity_;
    }

    hipStream_t stream_;
    int64_t capacity_ = 0;
    DeviceBuffer ws_, offsets_, point_list_, ranges_, final_T_;
    DeviceBuffer sg_rects_, sg_tiles_, sg_xy_, sg_conic_, sg_rgb_, sg_clamped_, sg_depths_;
    DeviceBuffer raster_inv_cov_, raster_mass_, render_grads_;
};

namespace detail {
inline Renderer& thread_renderer()
{
    static thread_local Renderer r;
    return r;
}
} // namespace detail

/* splat_forward.h:61 */
inline Statistics splat_forward(SplatForwardData& forward_data) { return detail::thread_renderer().splat_forward(forward_data); }

/* vol_raster_forward.h:70 */
inline VolRasterStatistics vol_raster_forward(VolRasterForwardData& forward_data)
{
    return detail::thread_renderer().vol_raster_forward(forward_data);
}

namespace vol_marcher {
/* vol_marcher_forward.h:44 */
template <class scalar_t>
ForwardCache forward(Framebuffer<scalar_t> framebuffer, ForwardData<scalar_t>& forward_data)
{
    return detail::thread_renderer().march_forward(framebuffer, forward_data);
}

/* vol_marcher_backward.h:37: grad is the [3, H, W] device gradient of the framebuffer */
template <class scalar_t>
Gradients backward(const Framebuffer<const scalar_t>& framebuffer, const ForwardData<scalar_t>& forward_data, const ForwardCache& forward_cache,
    const scalar_t* grad)
{
    return detail::thread_renderer().march_backward(framebuffer, forward_data, forward_cache, grad);
}
} // namespace vol_marcher

/* sibr_interfaces.h:24-27 */
namespace sibr_interfaces {
inline void splat(SplatForwardData& forward_data) { dgmr::splat_forward(forward_data); }
inline void vol_march(Framebuffer<float> framebuffer, vol_marcher::ForwardData<float>& forward_data)
{
    (void)vol_marcher::forward(framebuffer, forward_data);
}
} // namespace sibr_interfaces

} // namespace dgmr
