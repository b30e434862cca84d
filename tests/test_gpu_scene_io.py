"""A scene loaded from the INRIA-layout fixture (tests/golden/inria_scene.ply,
written byte by byte by make_inria_ply.py) and its cameras.json, rendered on
the GPU and compared with the oracle on the same loaded values:
  * splatter: binning bit-exact against the float restatement, image within
    1e-4 of the FP64 oracle on pixels without a float-ambiguous blend
    decision, gradients of a random dL/dimg (zeroed on flagged pixels) within
    1e-3 relative (max and L2) per parameter tensor;
  * marcher (float32): image within 1e-4 of the FP64 oracle with the float32
    ray and border placement, on pixels whose thresholds are not flagged.
"""
import os

import numpy as np
import pytest

import oracle as O
from conftest import gpu_available

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _loaded():
    from differentiable_gaussian_mixture_renderer_amd import scene_io, scenes
    ply = scene_io.load_ply(os.path.join(GOLDEN, "inria_scene.ply"))
    sh, op, mu, s, q = ply.activated()
    scene = scenes.Scene(means=mu, scales=s, rotations=q, opacity=op, sh=sh, sh_degree=ply.sh_degree,
                         log_scales=ply.log_scales, opacity_logits=ply.opacity_logits)
    cams = scene_io.load_cameras_json(os.path.join(GOLDEN, "inria_cameras.json"))
    return scene, cams


@pytest.mark.parametrize("view", [0, 1])
def test_loaded_scene_splat_forward_backward(view):
    if not gpu_available():
        pytest.skip("no GPU")
    import torch
    from differentiable_gaussian_mixture_renderer_amd.python_bindings import splatter
    scene, cams = _loaded()
    cam = cams[view]
    bg = np.array([0.05, 0.1, 0.15], np.float32)
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.ascontiguousarray(a), device=dev)
    args = [t(scene.sh), t(scene.opacity), t(scene.means), t(scene.scales), t(scene.rotations)]
    cam_args = [t(cam.world_view_transform), t(cam.full_proj_transform), t(cam.camera_center), t(bg), cam.width, cam.height,
                cam.tan_fovx, cam.tan_fovy, scene.sh_degree]
    print(f"[gpu scene io] splat view {view} {cam.width}x{cam.height}", flush=True)
    cache = splatter.cpp_binding.forward(*args, *cam_args)
    out = {k: v.cpu().numpy() for k, v in zip(splatter.CACHE_FIELDS, cache)}
    ref64 = O.SplatOracle(O.scene_arrays(scene), cam, dbl=True, float_binning=True)
    b = ref64.binning()
    assert b["point_list"].size > 1000
    np.testing.assert_array_equal(out["b_point_list"].view(np.uint32), b["point_list"])
    np.testing.assert_array_equal(out["i_ranges"].view(np.uint32), b["ranges"])
    fb64, T64, amb, _ = ref64.render(bg)
    ok = amb == 0
    err = np.abs(out["framebuffer"] - fb64).max(axis=0)
    print(f"flagged {(~ok).mean():.4f}, max err {err[ok].max():.3g}", flush=True)
    assert (~ok).mean() <= 0.02 and err[ok].max() <= 1e-4
    grad = np.random.default_rng(view).normal(size=fb64.shape) / fb64.size
    grad[:, ~ok] = 0.0
    g = splatter.cpp_binding.backward(*args, *cam_args, t(grad.astype(np.float32)), cache)
    got = dict(zip(("sh", "opacity", "means", "scales", "rotations"), [x.cpu().numpy() for x in g]))
    want = ref64.backward(bg, fb64, T64, grad)
    for name in ("means", "scales", "rotations", "opacity", "sh"):
        a = got[name].reshape(scene.n, -1)
        w = want[name].reshape(scene.n, -1)[:, :a.shape[1]]
        max_rel = np.abs(a - w).max() / np.abs(w).max()
        l2_rel = np.linalg.norm(a - w) / np.linalg.norm(w)
        print(f"{name}: max rel {max_rel:.3g}, l2 rel {l2_rel:.3g}", flush=True)
        assert max_rel <= 1e-3 and l2_rel <= 1e-3, name


def test_loaded_scene_marcher_forward():
    if not gpu_available():
        pytest.skip("no GPU")
    import torch
    from differentiable_gaussian_mixture_renderer_amd.python_bindings import vol_marcher as vm
    scene, cams = _loaded()
    cam = cams[0]
    bg = np.array([0.05, 0.1, 0.15], np.float32)
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.ascontiguousarray(a), device=dev)
    args = [t(scene.sh), t(scene.opacity), t(scene.means), t(scene.scales), t(scene.rotations), t(cam.world_view_transform),
            t(cam.full_proj_transform), t(cam.camera_center), t(bg), cam.width, cam.height, cam.tan_fovx, cam.tan_fovy, scene.sh_degree]
    print("[gpu scene io] marcher view 0", flush=True)
    fb = vm.cpp_binding.forward(*args)[0].cpu().numpy()
    ref64 = O.MarchOracle(O.scene_arrays(scene), cam, dbl=True, float_binning=True)
    fb64, _, _, amb = ref64.render(bg, with_ambiguity=True, float_sampling=True)
    ok = amb == 0
    err = np.abs(fb - fb64).max(axis=0)
    print(f"flagged {(~ok).mean():.4f}, max err {err[ok].max():.3g}", flush=True)
    assert (~ok).mean() <= 0.02 and err[ok].max() <= 1e-4
