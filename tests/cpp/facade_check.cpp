// SPDX-License-Identifier: GPL-3.0-or-later
// Drives the C++ façade (include/dgmr.hpp) the way a C++ caller of the
// reference's interface would, for tests/test_gpu_facade.py, which loads
// this file built as a shared library and calls
//   dgmr_facade_check("splat" | "raster" | "march", input path, output path)
// Input (little endian): int64 n, int32 width, height, sh_degree, then float32
// means[3n], scales[3n], rotations[4n], opacity[n], sh[48n], view[16],
// proj[16], position[3], background[3], tan_fovx, tan_fovy.
// Output: float32 framebuffer [3,H,W]; for march also the five gradients
// (centroids, sh, weights, scales, rotations) of L = mean(framebuffer).
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <vector>

#include "dgmr.hpp"

namespace {

template <class T>
void read(std::ifstream& f, T* p, size_t count)
{
    f.read(reinterpret_cast<char*>(p), std::streamsize(sizeof(T) * count));
    if (!f)
        throw std::runtime_error("short input file");
}

struct Upload {
    dgmr::DeviceBuffer buf;
    Upload(const std::vector<float>& v) : buf(sizeof(float) * (v.empty() ? 1 : v.size()))
    {
        dgmr::check_hip(hipMemcpy(buf.data(), v.data(), sizeof(float) * v.size(), hipMemcpyHostToDevice), "hipMemcpy");
    }
    dgmr::View<const float> view(size_t n) const { return { buf.as<const float>(), int64_t(n) }; }
};

std::vector<float> download(const dgmr::DeviceBuffer& b, size_t count)
{
    std::vector<float> h(count);
    dgmr::check_hip(hipMemcpy(h.data(), b.data(), sizeof(float) * count, hipMemcpyDeviceToHost), "hipMemcpy");
    return h;
}

} // namespace

extern "C" int dgmr_facade_check(const char* mode_c, const char* in_path, const char* out_path)
{
    try {
        const std::string mode = mode_c;
        std::ifstream in(in_path, std::ios::binary);
        int64_t n = 0;
        int32_t W = 0, H = 0, deg = 0;
        read(in, &n, 1);
        read(in, &W, 1);
        read(in, &H, 1);
        read(in, &deg, 1);
        std::vector<float> means(3 * n), scales(3 * n), rots(4 * n), opac(n), sh(48 * n);
        read(in, means.data(), means.size());
        read(in, scales.data(), scales.size());
        read(in, rots.data(), rots.size());
        read(in, opac.data(), opac.size());
        read(in, sh.data(), sh.size());
        dgmr::Mat4<float> view, proj;
        dgmr::Vec3<float> pos, bg;
        float tanx = 0, tany = 0;
        read(in, view.data(), 16);
        read(in, proj.data(), 16);
        read(in, pos.data(), 3);
        read(in, bg.data(), 3);
        read(in, &tanx, 1);
        read(in, &tany, 1);
        const Upload d_means(means), d_scales(scales), d_rots(rots), d_opac(opac), d_sh(sh);
        const size_t plane = size_t(W) * size_t(H);
        dgmr::DeviceBuffer fb(sizeof(float) * 3 * plane);
        dgmr::check_hip(hipMemset(fb.data(), 0, sizeof(float) * 3 * plane), "hipMemset");
        std::vector<std::vector<float>> outputs;
        if (mode == "splat") {
            dgmr::SplatForwardData d;
            d.gm_centroids = d_means.view(means.size());
            d.gm_sh_params = d_sh.view(sh.size());
            d.gm_weights = d_opac.view(opac.size());
            d.gm_cov_scales = d_scales.view(scales.size());
            d.gm_cov_rotations = d_rots.view(rots.size());
            d.framebuffer = { fb.as<float>(), H, W };
            d.view_matrix = view;
            d.proj_matrix = proj;
            d.cam_poition = pos;
            d.background = bg;
            d.sh_degree = deg;
            d.tan_fovx = tanx;
            d.tan_fovy = tany;
            dgmr::sibr_interfaces::splat(d);
        } else if (mode == "raster") {
            dgmr::VolRasterForwardData d;
            d.gm_centroids = d_means.view(means.size());
            d.gm_sh_params = d_sh.view(sh.size());
            d.gm_weights = d_opac.view(opac.size());
            d.gm_cov_scales = d_scales.view(scales.size());
            d.gm_cov_rotations = d_rots.view(rots.size());
            d.framebuffer = { fb.as<float>(), H, W };
            d.view_matrix = view;
            d.proj_matrix = proj;
            d.cam_poition = pos;
            d.background = bg;
            d.sh_degree = deg;
            d.tan_fovx = tanx;
            d.tan_fovy = tany;
            dgmr::vol_raster_forward(d);
        } else if (mode == "march") {
            dgmr::vol_marcher::ForwardData<float> d;
            d.gm_centroids = d_means.view(means.size());
            d.gm_sh_params = d_sh.view(sh.size());
            d.gm_weights = d_opac.view(opac.size());
            d.gm_cov_scales = d_scales.view(scales.size());
            d.gm_cov_rotations = d_rots.view(rots.size());
            d.view_matrix = view;
            d.proj_matrix = proj;
            d.cam_poition = pos;
            d.background = bg;
            d.sh_degree = deg;
            d.tan_fovx = tanx;
            d.tan_fovy = tany;
            const dgmr::vol_marcher::ForwardCache cache = dgmr::vol_marcher::forward(dgmr::Framebuffer<float> { fb.as<float>(), H, W }, d);
            const Upload grad(std::vector<float>(3 * plane, 1.f / float(3 * plane)));
            const dgmr::vol_marcher::Gradients g = dgmr::vol_marcher::backward(
                dgmr::Framebuffer<const float> { fb.as<const float>(), H, W }, d, cache, grad.buf.as<const float>());
            outputs.push_back(download(g.gm_centroids, 3 * n));
            outputs.push_back(download(g.gm_sh_params, 48 * n));
            outputs.push_back(download(g.gm_weights, n));
            outputs.push_back(download(g.gm_cov_scales, 3 * n));
            outputs.push_back(download(g.gm_cov_rotations, 4 * n));
        } else {
            std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
            return 2;
        }
        dgmr::check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
        std::ofstream out(out_path, std::ios::binary);
        const std::vector<float> img = download(fb, 3 * plane);
        out.write(reinterpret_cast<const char*>(img.data()), std::streamsize(sizeof(float) * img.size()));
        for (const auto& o : outputs)
            out.write(reinterpret_cast<const char*>(o.data()), std::streamsize(sizeof(float) * o.size()));
        std::printf("facade_check %s: n=%lld %dx%d ok\n", mode.c_str(), (long long)n, W, H);
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "facade_check: %s\n", e.what());
        return 1;
    }
}
