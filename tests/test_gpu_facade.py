"""The C++ façade (include/dgmr.hpp) against the Python path on the same
inputs: tests/cpp/facade_check.cpp fills the reference's C++ structs
(SplatForwardData, VolRasterForwardData, vol_marcher::ForwardData<float>)
and calls sibr_interfaces::splat, vol_raster_forward and
vol_marcher::forward/backward.  Images must be identical (same kernels, same
inputs); the marcher's gradients within 1e-5 relative (float atomics)."""
import ctypes
import math
import os

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "tests", "cpp", "_build", "libfacade_check.so")


def _inputs(tmp_path):
    from differentiable_gaussian_mixture_renderer_amd import scenes
    import oracle as O
    sc = scenes.make_scene(3000, seed=21, extent=1.0, sh_degree=3)
    tan = math.tan(math.radians(30))
    cam = scenes.look_at_camera([0.5, -0.3, -4.0], [0.0, 0.0, 0.0], 96, 80, tan * 96 / 80, tan)
    means, scales, rots, opac, sh16, deg = O.scene_arrays(sc)
    bg = np.array([0.1, 0.2, 0.3], np.float32)
    path = tmp_path / "in.bin"
    with open(path, "wb") as f:
        f.write(np.array([means.shape[0]], np.int64).tobytes())
        f.write(np.array([cam.width, cam.height, deg], np.int32).tobytes())
        for a in (means, scales, rots, np.reshape(opac, -1), sh16, cam.world_view_transform, cam.full_proj_transform, cam.camera_center, bg,
                  [cam.tan_fovx, cam.tan_fovy]):
            f.write(np.ascontiguousarray(a, np.float32).tobytes())
    return path, (means, scales, rots, opac, sh16, deg), cam, bg


def _run(mode, path, tmp_path, n, plane):
    lib = ctypes.CDLL(LIB)
    lib.dgmr_facade_check.restype = ctypes.c_int
    lib.dgmr_facade_check.argtypes = [ctypes.c_char_p] * 3
    out = tmp_path / f"{mode}.bin"
    print(f"[gpu facade] {mode}", flush=True)
    assert lib.dgmr_facade_check(mode.encode(), str(path).encode(), str(out).encode()) == 0
    data = np.fromfile(out, np.float32)
    img = data[:3 * plane]
    return img, data[3 * plane:]


def _torch_args(arrs, cam, bg):
    import torch
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float32, device=dev)
    means, scales, rots, opac, sh16, deg = arrs
    return [t(sh16), t(np.reshape(opac, (-1, 1))), t(means), t(scales), t(rots)], (
        t(cam.world_view_transform), t(cam.full_proj_transform), t(cam.camera_center), t(bg), cam.width, cam.height, cam.tan_fovx,
        cam.tan_fovy, deg)


def test_facade_matches_python_path(tmp_path):
    if not gpu_available():
        pytest.skip("no GPU")
    if not os.path.exists(LIB):
        pytest.fail("tests/cpp/_build/libfacade_check.so is missing: run __graft_entry__.build()")
    import torch
    from differentiable_gaussian_mixture_renderer_amd.python_bindings import splatter, vol_marcher, vol_raster
    path, arrs, cam, bg = _inputs(tmp_path)
    n = arrs[0].shape[0]
    H, W = cam.height, cam.width
    plane = H * W
    params, cam_args = _torch_args(arrs, cam, bg)

    img, _ = _run("splat", path, tmp_path, n, plane)
    want = splatter.Splatter.apply(*params, *cam_args).detach().cpu().numpy().reshape(-1)
    np.testing.assert_array_equal(img, want)

    img, _ = _run("raster", path, tmp_path, n, plane)
    want = vol_raster.vol_raster_forward(*params, *cam_args).cpu().numpy().reshape(-1)
    np.testing.assert_array_equal(img, want)

    img, grads = _run("march", path, tmp_path, n, plane)
    saved = dict(vol_marcher.DEFAULT_OPTIONS)
    vol_marcher.DEFAULT_OPTIONS["max_large_steps"] = 0
    try:
        ps = [p.clone().requires_grad_(True) for p in params]
        out = vol_marcher.VolMarcher.apply(*ps, *cam_args)
        got_g = torch.autograd.grad(out.mean(), ps)
    finally:
        vol_marcher.DEFAULT_OPTIONS.update(saved)
    np.testing.assert_array_equal(img, out.detach().cpu().numpy().reshape(-1))
    sizes = [3 * n, 48 * n, n, 3 * n, 4 * n]  # centroids, sh, weights, scales, rotations
    order = [2, 0, 1, 3, 4]                   # index into (sh, opacity, means, scales, rotations)
    off = 0
    for size, k in zip(sizes, order):
        g = grads[off:off + size]
        off += size
        ref = got_g[k].detach().cpu().numpy().reshape(-1)
        scale = np.abs(ref).max() + 1e-30
        assert np.abs(g - ref).max() <= 1e-5 * scale, (k, np.abs(g - ref).max(), scale)
