"""GPU parity of the tile splatter against the host oracle.

Bars (stated per check):
  * binning (rects, tiles touched, point offsets, sorted point list, tile
    ranges): bit-exact against the float restatement;
  * image: max |gpu - oracle_fp64| <= 1e-4 over pixels whose discrete blend
    decisions are not within float rounding of a threshold (the oracle flags
    those; they are counted and bounded);
  * gradients: with dL/dimg zeroed on flagged pixels, per parameter tensor
    max|gpu - fp64| <= 1e-3 * max|fp64| and ||gpu - fp64|| <= 1e-3 ||fp64||.
"""
import numpy as np
import pytest

import oracle as O
from conftest import gpu_available

pytestmark = pytest.mark.gpu

IMG_TOL = 1e-4
GRAD_TOL = 1e-3


def _torch():
    import torch
    if not gpu_available():
        pytest.skip("no GPU")
    return torch


def _run_gpu(scene, cam, bg, tile=(16, 16), batch=64, grad_fb=None, cull=False):
    torch = _torch()
    from differentiable_gaussian_mixture_renderer_amd.python_bindings import splatter
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.ascontiguousarray(a), device=dev)
    args = [t(scene.sh), t(scene.opacity), t(scene.means), t(scene.scales), t(scene.rotations)]
    cam_args = [t(cam.world_view_transform), t(cam.full_proj_transform), t(cam.camera_center), t(bg), cam.width, cam.height,
                cam.tan_fovx, cam.tan_fovy, scene.sh_degree]
    print(f"[gpu splat] n={scene.n} {cam.width}x{cam.height} tile={tile} batch={batch} cull={cull}", flush=True)
    cache = splatter.cpp_binding.forward(*args, *cam_args, tile=tile, batch=batch, cull=cull)
    out = {k: v.cpu().numpy() for k, v in zip(splatter.CACHE_FIELDS, cache)}
    if grad_fb is not None:
        grads = splatter.cpp_binding.backward(*args, *cam_args, t(grad_fb.astype(np.float32)), cache, tile=tile, batch=batch)
        out["grads"] = dict(zip(("sh", "opacity", "means", "scales", "rotations"), [g.cpu().numpy() for g in grads]))
    torch.cuda.synchronize()
    return out


def _check_binning(out, ref_f32):
    g = ref_f32.geometry()
    b = ref_f32.binning()
    vis = g["tiles"] > 0
    np.testing.assert_array_equal(out["tiles_touched"].view(np.uint32), g["tiles"])
    np.testing.assert_array_equal(out["rects"][vis], g["rects"][vis])
    np.testing.assert_array_equal(out["point_offsets"].view(np.uint32), b["point_offsets"])
    np.testing.assert_array_equal(out["b_point_list"].view(np.uint32), b["point_list"])
    np.testing.assert_array_equal(out["i_ranges"].view(np.uint32), b["ranges"])
    return vis


def _scene_small():
    from differentiable_gaussian_mixture_renderer_amd import scenes
    return scenes.small_rgb_frame()


def _scene_sh3(n=60_000, w=480, h=272):
    from differentiable_gaussian_mixture_renderer_amd import scenes
    return scenes.training_frame(n=n, width=w, height=h)


@pytest.mark.parametrize("which", ["small", "sh3"])
def test_forward_parity(which):
    scene, cam = _scene_small() if which == "small" else _scene_sh3()
    bg = np.array([0.25, 0.5, 0.75], np.float32)
    out = _run_gpu(scene, cam, bg)
    ref32 = O.SplatOracle(O.scene_arrays(scene), cam, dbl=False)
    vis = _check_binning(out, ref32)
    assert vis.sum() > 0
    g32 = ref32.geometry()
    np.testing.assert_allclose(out["points_xy_image"][vis], g32["xy"][vis], rtol=0, atol=0)
    np.testing.assert_allclose(out["depths"][vis], g32["depth"][vis], rtol=0, atol=0)
    ref64 = O.SplatOracle(O.scene_arrays(scene), cam, dbl=True, float_binning=True)
    fb64, T64, amb, _ = ref64.render(bg)
    ok = amb == 0
    err = np.abs(out["framebuffer"] - fb64).max(axis=0)
    print(f"flagged pixels {int((~ok).sum())} of {ok.size}; max err (unflagged) {err[ok].max():.3g}; max err (all) {err.max():.3g}")
    assert (~ok).mean() < 0.02
    assert err[ok].max() <= IMG_TOL
    assert np.abs(out["final_transparency"] - T64)[ok].max() <= IMG_TOL


@pytest.mark.parametrize("tile,batch", [((16, 16), 64), ((8, 8), 128), ((32, 8), 256)])
def test_tilings_match_oracle_of_same_tiling(tile, batch):
    scene, cam = _scene_sh3(n=30_000, w=352, h=200)
    bg = np.zeros(3, np.float32)
    out = _run_gpu(scene, cam, bg, tile=tile, batch=batch)
    ref32 = O.SplatOracle(O.scene_arrays(scene), cam, dbl=False, tile=tile)
    _check_binning(out, ref32)
    ref64 = O.SplatOracle(O.scene_arrays(scene), cam, dbl=True, tile=tile)
    fb64, _, amb, _ = ref64.render(bg)
    ok = amb == 0
    assert np.abs(out["framebuffer"] - fb64).max(axis=0)[ok].max() <= IMG_TOL


def test_binning_with_32bit_tile_keys():
    """2304x2304 with 8x8 tiles is 82 944 tiles: pair keys need 17 bits, so
    the binning takes its 32-bit key path (16-bit otherwise); bit-exact
    against the oracle."""
    from differentiable_gaussian_mixture_renderer_amd import scenes
    import math
    scene = scenes.make_scene(8_000, seed=21, extent=1.0, sh_mode="rgb")
    t = math.tan(math.radians(30.0))
    cam = scenes.look_at_camera([0.0, 0.0, -4.0], [0.0, 0.0, 0.0], 2304, 2304, t, t)
    out = _run_gpu(scene, cam, np.zeros(3, np.float32), tile=(8, 8), batch=64)
    ref32 = O.SplatOracle(O.scene_arrays(scene), cam, dbl=False, tile=(8, 8))
    assert ref32.n_pairs > 0
    _check_binning(out, ref32)


@pytest.mark.parametrize("tile,batch", [((16, 16), 64), ((8, 8), 128), ((32, 8), 256)])
def test_tight_culling_same_image_fewer_pairs(tile, batch):
    """cull=1 against the reference binning on the same inputs: image and
    transmittance bit-identical, gradients equal up to the order of the
    atomic sums (max 1e-4 of the largest, L2 1e-5: sums that cancel
    amplify the order's rounding); every tile's list an ordered sublist of the
    reference's, no shorter than the oracle's count of Gaussians whose alpha
    reaches 1/255 at some pixel of that tile, and at least 5 % fewer pairs."""
    scene, cam = _scene_sh3(n=30_000, w=352, h=200)
    bg = np.array([0.3, 0.1, 0.2], np.float32)
    grad = np.random.default_rng(3).uniform(-1, 1, (3, cam.height, cam.width)) / (3 * cam.height * cam.width)
    ref = _run_gpu(scene, cam, bg, tile=tile, batch=batch, grad_fb=grad)
    got = _run_gpu(scene, cam, bg, tile=tile, batch=batch, grad_fb=grad, cull=True)
    assert np.array_equal(got["framebuffer"], ref["framebuffer"])
    assert np.array_equal(got["final_transparency"], ref["final_transparency"])
    for name, a in ref["grads"].items():
        d = got["grads"][name] - a
        print(f"{name}: max {np.abs(d).max() / np.abs(a).max():.3g}, l2 {np.linalg.norm(d) / np.linalg.norm(a):.3g}")
        assert np.abs(d).max() <= 1e-4 * np.abs(a).max() and np.linalg.norm(d) <= 1e-5 * np.linalg.norm(a), name
    _, _, reach = O.SplatOracle(O.scene_arrays(scene), cam, dbl=False, tile=tile).tile_work(reachable=True)
    ra, rb = ref["i_ranges"].view(np.uint32), got["i_ranges"].view(np.uint32)
    pa, pb = ref["b_point_list"].view(np.uint32), got["b_point_list"].view(np.uint32)
    for t in range(ra.shape[0]):
        la, lb = pa[ra[t, 0]:ra[t, 1]], pb[rb[t, 0]:rb[t, 1]]
        assert len(lb) >= reach[t], t
        pos = {int(g): i for i, g in enumerate(la)}
        at = [pos.get(int(g), -1) for g in lb]
        assert min(at, default=0) >= 0 and all(x < y for x, y in zip(at, at[1:])), t
    print(f"pairs: reference {len(pa)}, tight {len(pb)}, reachable {int(reach.sum())}")
    assert len(pb) <= 0.95 * len(pa)


@pytest.mark.parametrize("which", ["small", "sh3"])
def test_backward_parity(which):
    scene, cam = _scene_small() if which == "small" else _scene_sh3(n=20_000, w=256, h=160)
    bg = np.array([0.1, 0.2, 0.3], np.float32)
    rng = np.random.default_rng(7)
    ref64 = O.SplatOracle(O.scene_arrays(scene), cam, dbl=True, float_binning=True)
    fb64, T64, amb, _ = ref64.render(bg)
    target = rng.uniform(0, 1, fb64.shape)
    grad = np.sign(fb64 - target) / fb64.size
    grad[:, amb != 0] = 0
    out = _run_gpu(scene, cam, bg, grad_fb=grad)
    r = ref64.backward(bg, fb64, T64, grad)
    n = scene.n
    k = scene.sh.shape[1]
    pairs = [("means", out["grads"]["means"], r["means"]), ("scales", out["grads"]["scales"], r["scales"]),
             ("rotations", out["grads"]["rotations"], r["rotations"]),
             ("opacity", out["grads"]["opacity"].reshape(n), r["opacity"]),
             ("sh", out["grads"]["sh"], r["sh"][:, :k])]
    for name, got, want in pairs:
        scale = np.abs(want).max()
        assert scale > 0, name
        max_rel = np.abs(got - want).max() / scale
        l2_rel = np.linalg.norm(got - want) / np.linalg.norm(want)
        print(f"{name}: max rel {max_rel:.3g}, l2 rel {l2_rel:.3g}")
        assert max_rel <= GRAD_TOL and l2_rel <= GRAD_TOL, name


def test_autograd_function_and_empty_view():
    torch = _torch()
    from differentiable_gaussian_mixture_renderer_amd import scenes
    from differentiable_gaussian_mixture_renderer_amd.python_bindings import splatter
    scene, cam = scenes.small_rgb_frame()
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.ascontiguousarray(a), device=dev)
    params = [t(scene.sh).requires_grad_(), t(scene.opacity).requires_grad_(), t(scene.means).requires_grad_(),
              t(scene.scales).requires_grad_(), t(scene.rotations).requires_grad_()]
    bg = t(np.array([0.0, 0.0, 0.0], np.float32))
    img = splatter.apply(*params, t(cam.world_view_transform), t(cam.full_proj_transform), t(cam.camera_center), bg, cam.width,
                         cam.height, cam.tan_fovx, cam.tan_fovy, 0)
    (img * img).sum().backward()
    for p in params:
        assert p.grad is not None and p.grad.shape == p.shape and torch.isfinite(p.grad).all()
    # every Gaussian behind the camera: background image, zero gradients
    params2 = [p.detach().clone().requires_grad_() for p in params]
    with torch.no_grad():
        params2[2][:, 2] = -10.0
    img2 = splatter.apply(*params2, t(cam.world_view_transform), t(cam.full_proj_transform), t(cam.camera_center), bg, cam.width,
                          cam.height, cam.tan_fovx, cam.tan_fovy, 0)
    img2.sum().backward()
    assert float(img2.abs().max()) == 0.0
    assert all(float(p.grad.abs().max()) == 0.0 for p in params2)


def test_fused_activations_match_torch_activations():
    """SplatterActivated (exp / normalise / sigmoid inside the kernels) against
    torch activations followed by Splatter.apply.  The two differ by an ulp in
    the activated values, which can flip a blend decision at a pixel sitting on
    a threshold, so: at most 0.05 % of pixels differ by more than 2e-6, and per
    parameter tensor ||g_fused - g_torch|| <= 1e-3 ||g_torch|| (a wrong chain
    rule is off by O(1))."""
    torch = _torch()
    from differentiable_gaussian_mixture_renderer_amd import scenes
    from differentiable_gaussian_mixture_renderer_amd.python_bindings import splatter
    scene, cam = scenes.training_frame(n=30_000, width=320, height=200)
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.ascontiguousarray(a), device=dev)
    raw = [t(scene.sh), t(scene.opacity_logits), t(scene.means), t(scene.log_scales), t(scene.rotations * 1.7)]
    cam_args = [t(cam.world_view_transform), t(cam.full_proj_transform), t(cam.camera_center), t(np.zeros(3, np.float32)), cam.width,
                cam.height, cam.tan_fovx, cam.tan_fovy, 3]
    target = torch.rand((3, cam.height, cam.width), generator=torch.Generator(device=dev).manual_seed(5), device=dev)

    def run(fused):
        ps = [r.clone().requires_grad_() for r in raw]
        if fused:
            img = splatter.SplatterActivated.apply(*ps, *cam_args)
        else:
            sh, logit, mu, ls, q = ps
            img = splatter.apply(sh, torch.sigmoid(logit), mu, torch.exp(ls), torch.nn.functional.normalize(q, dim=-1), *cam_args)
        loss = (img - target).abs().mean()
        grads = torch.autograd.grad(loss, ps)
        return img.detach(), grads

    print("[gpu splat] fused activations 30k 320x200", flush=True)
    img_f, g_f = run(True)
    img_t, g_t = run(False)
    off = int(((img_f - img_t).abs() > 2e-6).any(dim=0).sum())
    assert off <= 0.0005 * cam.width * cam.height, off
    for name, a, b in zip(("sh", "opacity_logits", "means", "log_scales", "rotations"), g_f, g_t):
        nb = float(b.norm())
        assert nb > 0, name
        assert float((a - b).norm()) <= 1e-3 * nb, (name, float((a - b).norm()), nb)


def test_compacted_depth_sort_single_view_equals_uncompacted_subset():
    """4.3M Gaussians (at least 4M: the binning compacts the visible entries
    and numbers them itself, the preprocess writes no ids) around a camera at
    the origin looking along +z, so every Gaussian with z <= 0 is culled.  The
    same render of only the Gaussians with z > 0 (2.15M, no compaction, ids
    written by the preprocess) gives the same image and transmittance bit for
    bit, the same tile ranges, and the same point list up to the subset's
    numbering: dropping culled Gaussians changes nothing else."""
    torch = _torch()
    from differentiable_gaussian_mixture_renderer_amd import scenes
    from differentiable_gaussian_mixture_renderer_amd.python_bindings import splatter
    scene = scenes.make_scene(4_300_000, seed=31, extent=5.0, sh_degree=0, sh_mode="rgb")
    cam = scenes.look_at_camera((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 320, 240, 0.6, 0.45)
    keep = np.nonzero(scene.means[:, 2] > 0.0)[0]
    assert keep.size < (4 << 20) <= scene.n
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.ascontiguousarray(a), device=dev)
    cam_args = [t(cam.world_view_transform), t(cam.full_proj_transform), t(cam.camera_center), t(np.zeros(3, np.float32)), cam.width,
                cam.height, cam.tan_fovx, cam.tan_fovy, 0]
    outs = []
    for sel in (slice(None), keep):
        args = [t(scene.sh[sel]), t(scene.opacity[sel]), t(scene.means[sel]), t(scene.scales[sel]), t(scene.rotations[sel])]
        print(f"[gpu splat] compaction check n={args[2].shape[0]}", flush=True)
        cache = splatter.cpp_binding.forward(*args, *cam_args, cull=False)
        outs.append([c.cpu() for c in cache])
    torch.cuda.synchronize()
    full, sub = outs
    assert torch.equal(full[0], sub[0]) and torch.equal(full[11], sub[11])  # framebuffer, transmittance
    assert int(full[12][0]) == int(sub[12][0]) > 0                          # pair count
    assert torch.equal(full[9], sub[9])                                      # tile ranges
    np.testing.assert_array_equal(full[10].numpy(), keep[sub[10].numpy()].astype(np.int32))


def test_backward_against_finite_differences_of_the_fp64_forward():
    """The HIP splat backward (float32) against central differences of the
    FP64 render (float binning, as the GPU bins), 32 random parameters over
    the five tensors (the reference's pattern, grad_inria_splatter_render.cu:160-228).
    dL/dimg is zeroed on pixels the oracle flags as float-ambiguous; h = 1e-6
    max(1, |theta|) (1e-3 for SH coefficients, which enter linearly) moves no
    discrete decision (the two step sizes agree).  Bar per parameter:
    |gpu - fd| <= 1e-3 |fd| + 1e-4 s + e (s the largest |fd| drawn for that
    tensor, covering float accumulation on small gradients, as the max-relative
    bar of test_backward_parity; e the quotient's round-off)."""
    torch = _torch()
    from differentiable_gaussian_mixture_renderer_amd import scenes
    scene = scenes.make_scene(2_000, seed=47, extent=1.0, sh_degree=3)
    scene.scales = (scene.scales * 2.0).astype(np.float32)
    cam = scenes.look_at_camera([0.3, -0.2, -4.0], [0.0, 0.1, 0.0], 96, 64, 0.6, 0.4)
    bg = np.array([0.2, 0.1, 0.3], np.float32)
    base = [scene.means.astype(np.float64), scene.scales.astype(np.float64), scene.rotations.astype(np.float64),
            scene.opacity.reshape(-1).astype(np.float64), O.scene_arrays(scene)[4].astype(np.float64)]
    ref = O.SplatOracle((*base, 3), cam, dbl=True, float_binning=True)
    fb64, _, amb, _ = ref.render(bg)
    rng = np.random.default_rng(23)
    grad = rng.normal(size=fb64.shape)
    grad[:, amb != 0] = 0.0
    out = _run_gpu(scene, cam, bg, grad_fb=grad)

    def loss(params):
        fb, _, _, _ = O.SplatOracle((*params, 3), cam, dbl=True, float_binning=True).render(bg)
        return float((fb * grad).sum())

    names = ("means", "scales", "rotations", "opacity", "sh")
    gpu = {k: out["grads"][k].reshape(scene.n, -1) for k in names}
    rows = []
    for r in range(32):
        t = r % 5
        name = names[t]
        live = np.nonzero(np.abs(gpu[name]).max(axis=1) > 0)[0]
        i = int(live[rng.integers(live.size)])
        j = int(rng.integers(gpu[name].shape[1]))
        v = float(base[t].reshape(scene.n, -1)[i, j])
        fds = []
        for step in ((1e-3, 5e-4) if name == "sh" else (1e-6, 5e-7)):
            hh = step * max(1.0, abs(v))
            vals = []
            for sgn in (1, -1):
                p2 = [p.copy() for p in base]
                p2[t].reshape(scene.n, -1)[i, j] = v + sgn * hh
                vals.append(loss(p2))
            fds.append((vals[0] - vals[1]) / (2 * hh))
        e = 8 * np.sqrt(fb64.size) * np.finfo(np.float64).eps * np.abs(grad * fb64).max() / hh
        rows.append((name, i, j, fds[0], fds[1], float(gpu[name][i, j]), e))
    for name in names:
        sel = [r for r in rows if r[0] == name]
        s = max(abs(r[3]) for r in sel)
        assert s > 0, name
        for _, i, j, fd, fd2, g, e in sel:
            print(f"{name}[{i},{j}]: fd {fd:.7g} (half step {fd2:.7g}), gpu {g:.7g}, rel {abs(g - fd) / max(abs(fd), 1e-30):.2g}")
            assert abs(fd - fd2) <= 1e-5 * abs(fd) + e, (name, i, j, "decision moved")
            assert abs(g - fd) <= 1e-3 * abs(fd) + 1e-4 * s + e, (name, i, j, fd, g)
