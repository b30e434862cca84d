"""GPU parity of the volumetric ray marcher against the host oracle.

Bars:
  * binning (rects from the iso-line of 1/255, tiles, point offsets, sorted
    point list, tile ranges): bit-exact against the restatement of the same
    precision;
  * double GPU vs double oracle: image and gradients within 1e-9 (relative to
    the largest magnitude), i.e. the same algorithm up to libm rounding;
  * float GPU vs the FP64 oracle with the float32 placement (the float
    pipeline's binning, ray direction and bin borders, everything evaluated on
    them in FP64): image and transmittance max |gpu - fp64| <= 1e-4 over the
    pixels whose mass or transmittance thresholds are not within float
    rounding (the oracle flags them, <= 2 %); gradients with dL/dimg zeroed on
    flagged pixels, per parameter tensor max |gpu - fp64| <= 1e-3 max |fp64|
    and ||gpu - fp64|| <= 1e-3 ||fp64||.  Each precision placing its own ray
    and borders is not comparable at that level: the reference unprojects NDC
    z = 0, about 2 znear from the camera, so a float32 ray direction carries
    ~1e-5 relative error, and its border walk lands on section ends by design;
  * double gradients against central finite differences of the GPU forward on
    fixed sample sections (the reference's gradient-check mode), 32 random
    parameters, relative error <= 1e-4.
"""
import numpy as np
import pytest

import oracle as O
from conftest import gpu_available

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not gpu_available():
        pytest.skip("no GPU")
    return torch


def small_scene(n=3000, extent=1.5, w=64, h=48, seed=3):
    from differentiable_gaussian_mixture_renderer_amd import scenes
    scene = scenes.make_scene(n, seed=seed, extent=extent, sh_degree=3)
    scene.scales = (scene.scales * 2.0).astype(np.float32)
    t = 0.5
    cam = scenes.look_at_camera([0.2, -0.1, -6.0], [0.0, 0.0, 0.0], w, h, t * w / h, t)
    return scene, cam


def run_gpu(scene, cam, bg, dtype, grad_fb=None, max_large_steps=0, frozen=None, tile=(16, 16), replay_steps=0, render_mode=None):
    """Forward (and backward with grad_fb) through the reference binding;
    replay_steps > 0 gives both calls a replay store keeping that many large
    steps, whose per-tile step counts come back as out["tile_steps"]."""
    torch = _torch()
    from differentiable_gaussian_mixture_renderer_amd.python_bindings import vol_marcher as vm
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.ascontiguousarray(a), device=dev, dtype=dtype)
    args = [t(scene.sh), t(scene.opacity), t(scene.means), t(scene.scales), t(scene.rotations)]
    cam_args = [t(cam.world_view_transform), t(cam.full_proj_transform), t(cam.camera_center), t(bg), cam.width, cam.height,
                cam.tan_fovx, cam.tan_fovy, scene.sh_degree]
    print(f"[gpu march] n={scene.n} {cam.width}x{cam.height} {dtype} frozen={frozen is not None} replay={replay_steps}", flush=True)
    replay = None
    if replay_steps:
        replay = vm.make_replay(cam.width, cam.height, dtype, dev, tile=tile, steps=replay_steps, max_bytes=1 << 40)
        assert replay is not None and replay.steps == replay_steps
    cache = vm.cpp_binding.forward(*args, *cam_args, tile=tile, max_large_steps=max_large_steps, frozen=frozen, replay=replay,
                                   render_mode=render_mode)
    names = ("framebuffer", "rects", "rgb", "clamped", "depths", "xy", "inv_cov", "masses", "tiles", "offsets", "ranges",
             "point_list", "remaining")
    out = {k: v.cpu().numpy() for k, v in zip(names, cache)}
    if grad_fb is not None:
        grads = vm.cpp_binding.backward(*args, *cam_args, t(grad_fb), cache, tile=tile, max_large_steps=max_large_steps, frozen=frozen,
                                        replay=replay)
        out["grads"] = dict(zip(("sh", "opacity", "means", "scales", "rotations"), [g.cpu().numpy() for g in grads]))
    if replay is not None:
        n_tiles = out["ranges"].shape[0]
        out["tile_steps"] = replay.buffer[:4 * n_tiles].cpu().numpy().view(np.uint32).copy()
    torch.cuda.synchronize()
    return out


def check_binning(out, ref):
    g = ref.geometry()
    b = ref.binning()
    vis = g["tiles"] > 0
    np.testing.assert_array_equal(out["tiles"].view(np.uint32), g["tiles"])
    np.testing.assert_array_equal(out["rects"][vis], g["rects"][vis])
    np.testing.assert_array_equal(out["offsets"].view(np.uint32), b["point_offsets"])
    np.testing.assert_array_equal(out["point_list"].view(np.uint32), b["point_list"])
    np.testing.assert_array_equal(out["ranges"].view(np.uint32), b["ranges"])
    return vis


def test_forward_double_matches_double_oracle():
    torch = _torch()
    scene, cam = small_scene()
    bg = np.array([0.2, 0.3, 0.4])
    out = run_gpu(scene, cam, bg, torch.float64)
    ref = O.MarchOracle(O.scene_arrays(scene), cam, dbl=True, float_binning=False)
    vis = check_binning(out, ref)
    assert vis.sum() > 100
    g = ref.geometry()
    np.testing.assert_allclose(out["inv_cov"][vis], g["inv_cov"][vis], rtol=1e-12)
    np.testing.assert_allclose(out["masses"][vis], g["mass"][vis], rtol=1e-12)
    fb, T, nl = ref.render(bg)
    err = np.abs(out["framebuffer"] - fb).max()
    print(f"double: max image err {err:.3g}, large steps max {nl.max()}")
    assert err <= 1e-9
    assert np.abs(out["remaining"] - T).max() <= 1e-9


IMG_TOL = 1e-4
GRAD_TOL = 1e-3


def test_forward_float_binning_and_image():
    torch = _torch()
    scene, cam = small_scene()
    bg = np.array([0.2, 0.3, 0.4], np.float32)
    out = run_gpu(scene, cam, bg, torch.float32)
    ref32 = O.MarchOracle(O.scene_arrays(scene), cam, dbl=False)
    check_binning(out, ref32)
    ref64 = O.MarchOracle(O.scene_arrays(scene), cam, dbl=True, float_binning=True)
    fb, T, _, amb = ref64.render(bg, with_ambiguity=True, float_sampling=True)
    ok = amb == 0
    err = np.abs(out["framebuffer"] - fb).max(axis=0)
    own, _, _ = ref64.render(bg)
    err_own = np.abs(out["framebuffer"] - own).max(axis=0)
    print(f"float vs FP64 (float placement): flagged {(~ok).mean():.4f}, max (unflagged) {err[ok].max():.3g}, max (all) {err.max():.3g}; "
          f"vs FP64 own placement: median {np.median(err_own):.3g}, max {err_own.max():.3g}")
    assert (~ok).mean() <= 0.02
    assert err[ok].max() <= IMG_TOL
    assert np.abs(out["remaining"] - T)[ok].max() <= IMG_TOL


def _check_grads(got_all, want_all, n, k, label):
    for name in ("means", "scales", "rotations", "opacity", "sh"):
        got = got_all[name].reshape(n, -1)
        want = want_all[name].reshape(n, -1)[:, :got.shape[1]]
        scale = np.abs(want).max()
        assert scale > 0, name
        max_rel = np.abs(got - want).max() / scale
        l2_rel = np.linalg.norm(got - want) / np.linalg.norm(want)
        print(f"{label} {name}: max rel {max_rel:.3g}, l2 rel {l2_rel:.3g}")
        assert max_rel <= GRAD_TOL and l2_rel <= GRAD_TOL, (label, name)


def test_backward_double_matches_double_oracle():
    torch = _torch()
    scene, cam = small_scene(n=1500, w=48, h=32)
    bg = np.array([0.2, 0.3, 0.4])
    ref = O.MarchOracle(O.scene_arrays(scene), cam, dbl=True, float_binning=False)
    fb, T, _ = ref.render(bg)
    rng = np.random.default_rng(3)
    grad = rng.normal(size=fb.shape) / fb.size
    out = run_gpu(scene, cam, bg, torch.float64, grad_fb=grad)
    r = ref.backward(bg, fb, T, grad)
    n, k = scene.n, scene.sh.shape[1]
    for name, got, want in [("means", out["grads"]["means"], r["means"]), ("scales", out["grads"]["scales"], r["scales"]),
                            ("rotations", out["grads"]["rotations"], r["rotations"]),
                            ("opacity", out["grads"]["opacity"].reshape(n), r["opacity"]), ("sh", out["grads"]["sh"], r["sh"][:, :k])]:
        scale = np.abs(want).max()
        assert scale > 0, name
        rel = np.abs(got - want).max() / scale
        print(f"{name}: max rel {rel:.3g}")
        assert rel <= 1e-9, name


def test_backward_against_finite_differences():
    """Frozen sample sections make the forward smooth (vol_marcher.h:25); 32
    random parameters are perturbed (in float64) by central differences of the
    GPU forward and compared with the GPU backward."""
    torch = _torch()
    scene, cam = small_scene(n=400, extent=1.0, w=32, h=32, seed=5)
    for name in ("means", "scales", "rotations", "opacity", "sh"):
        setattr(scene, name, getattr(scene, name).astype(np.float64))
    frozen = [(3.5, 6.0, 2.5 / 31), (6.0, 8.5, 2.5 / 31)]
    bg = np.array([0.2, 0.3, 0.4])
    rng = np.random.default_rng(11)
    out0 = run_gpu(scene, cam, bg, torch.float64, frozen=frozen)
    grad = rng.normal(size=out0["framebuffer"].shape)
    out = run_gpu(scene, cam, bg, torch.float64, grad_fb=grad, frozen=frozen)
    names = ["means", "scales", "rotations", "opacity", "sh"]
    vis = np.nonzero(out0["tiles"] > 0)[0]
    errs = []
    for _ in range(32):
        name = names[rng.integers(len(names))]
        i = int(vis[rng.integers(len(vis))])
        base = getattr(scene, name)
        j = int(rng.integers(3 if name == "sh" else base[i].size))
        v = float(base[i].reshape(-1)[j])
        h = 1e-5 * max(1.0, abs(v))
        vals = []
        for sgn in (1, -1):
            pert = base.copy()
            pert[i].reshape(-1)[j] = v + sgn * h
            setattr(scene, name, pert)
            vals.append(float((run_gpu(scene, cam, bg, torch.float64, frozen=frozen)["framebuffer"] * grad).sum()))
        setattr(scene, name, base)
        fd = (vals[0] - vals[1]) / (2 * h)
        errs.append((name, fd, float(out["grads"][name][i].reshape(-1)[j])))
    scale = max(abs(fd) for _, fd, _ in errs)
    worst = max(abs(fd - an) for _, fd, an in errs) / scale
    print("worst relative FD error", worst)
    assert worst <= 1e-4, errs


def test_backward_float_matches_oracle():
    """float GPU gradients against the FP64 oracle with the float32 placement,
    dL/dimg zeroed on the pixels it flags."""
    torch = _torch()
    scene, cam = small_scene(n=1500, w=48, h=32)
    bg = np.array([0.2, 0.3, 0.4], np.float32)
    rng = np.random.default_rng(5)
    ref64 = O.MarchOracle(O.scene_arrays(scene), cam, dbl=True, float_binning=True)
    fb64, T64, _, amb = ref64.render(bg, with_ambiguity=True, float_sampling=True)
    grad = (rng.normal(size=fb64.shape) / fb64.size).astype(np.float32)
    grad[:, amb != 0] = 0
    out = run_gpu(scene, cam, bg, torch.float32, grad_fb=grad)
    r64 = ref64.backward(bg, fb64, T64, grad.astype(np.float64), float_sampling=True)
    _check_grads(out["grads"], r64, scene.n, scene.sh.shape[1], "f32 vs f64")


@pytest.mark.parametrize("dtype_name", ["float64", "float32"])
def test_backward_with_fewer_sh_coefficients(dtype_name):
    """A caller passing 4 SH coefficients at active degree 3: the forward pads
    the rest with zeros and the backward's preprocess reads them as ones
    (python_bindings/vol_marcher.cu:92, :177); the gradient holds 4 rows."""
    torch = _torch()
    dtype = getattr(torch, dtype_name)
    scene, cam = small_scene(n=1200, w=48, h=32, seed=9)
    scene.sh = np.ascontiguousarray(scene.sh[:, :4])
    bg = np.array([0.2, 0.3, 0.4])
    dbl = dtype_name == "float64"
    ref = O.MarchOracle(O.scene_arrays(scene), cam, dbl=True, float_binning=not dbl)
    if dbl:
        fb, T, _ = ref.render(bg)
        amb = np.zeros(T.shape, np.uint8)
    else:
        fb, T, _, amb = ref.render(bg, with_ambiguity=True, float_sampling=True)
    grad = np.random.default_rng(13).normal(size=fb.shape) / fb.size
    grad[:, amb != 0] = 0
    out = run_gpu(scene, cam, bg.astype(np.float32) if not dbl else bg, dtype, grad_fb=grad.astype(np.float32) if not dbl else grad)
    assert out["grads"]["sh"].shape == (scene.n, 4, 3)
    want = ref.backward(bg, fb, T, grad, float_sampling=not dbl, sh_stored=4)
    if dbl:
        for name in ("means", "scales", "rotations", "opacity", "sh"):
            got = out["grads"][name].reshape(scene.n, -1)
            w = want[name].reshape(scene.n, -1)[:, :got.shape[1]]
            rel = np.abs(got - w).max() / np.abs(w).max()
            print(f"{name}: max rel {rel:.3g}")
            assert rel <= 1e-9, name
        plain = ref.backward(bg, fb, T, grad)
        assert np.abs(plain["means"] - want["means"]).max() > 1e-6 * np.abs(want["means"]).max()   # the padding is visible
    else:
        _check_grads(out["grads"], want, scene.n, 4, "f32 sh4")


def test_volmarcher_autograd_and_large_step_cap():
    torch = _torch()
    from differentiable_gaussian_mixture_renderer_amd.python_bindings import vol_marcher as vm
    scene, cam = small_scene(n=800, w=32, h=32)
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.ascontiguousarray(a), device=dev)
    params = [t(scene.sh).requires_grad_(), t(scene.opacity).requires_grad_(), t(scene.means).requires_grad_(),
              t(scene.scales).requires_grad_(), t(scene.rotations).requires_grad_()]
    img = vm.apply(*params, t(cam.world_view_transform), t(cam.full_proj_transform), t(cam.camera_center),
                   t(np.zeros(3, np.float32)), cam.width, cam.height, cam.tan_fovx, cam.tan_fovy, 3)
    img.sum().backward()
    assert all(p.grad is not None and torch.isfinite(p.grad).all() for p in params)
    # with the replay store turned on, autograd keeps one between its forward
    # and backward; the gradients are the recomputing backward's
    grads = [p.grad.cpu().numpy() for p in params]
    saved = dict(vm.DEFAULT_OPTIONS)
    try:
        vm.DEFAULT_OPTIONS["replay"] = True
        for p in params:
            p.grad = None
        img = vm.apply(*params, t(cam.world_view_transform), t(cam.full_proj_transform), t(cam.camera_center),
                       t(np.zeros(3, np.float32)), cam.width, cam.height, cam.tan_fovx, cam.tan_fovy, 3)
        img.sum().backward()
    finally:
        vm.DEFAULT_OPTIONS.update(saved)
    for p, want, name in zip(params, grads, ("sh", "opacity", "means", "scales", "rotations")):
        assert np.abs(p.grad.cpu().numpy() - want).max() <= 2e-5 * np.abs(want).max(), name
    # a cap of 2 large steps (<= 254 bins per ray) matches the oracle under the same cap
    bg = np.zeros(3, np.float32)
    out = run_gpu(scene, cam, bg, torch.float64, max_large_steps=2)
    ref = O.MarchOracle(O.scene_arrays(scene), cam, dbl=True, float_binning=False)
    fb, _, nl = ref.render(bg, max_large_steps=2)
    assert nl.max() <= 2
    assert np.abs(out["framebuffer"] - fb).max() <= 1e-9


def test_chunk_size_and_tiles_do_not_change_results():
    """The float marcher keeps 16 or 32 bins in registers per pass; a chunk is
    skipped for a Gaussian only where every bin would fall under the per-bin
    threshold, so the image and remaining transparency are bit-identical."""
    import torch
    from conftest import gpu_available
    if not gpu_available():
        pytest.skip("no GPU")
    from differentiable_gaussian_mixture_renderer_amd import scenes
    from differentiable_gaussian_mixture_renderer_amd.python_bindings import vol_marcher
    scene = scenes.make_scene(3_000, seed=21, extent=1.5, sh_degree=1)
    cam = scenes.look_at_camera([0.0, 0.0, -4.0], [0.0, 0.0, 0.0], 96, 64, 0.6, 0.4)
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.ascontiguousarray(a), device=dev)
    args = [t(scene.sh), t(scene.opacity), t(scene.means), t(scene.scales), t(scene.rotations), t(cam.world_view_transform),
            t(cam.full_proj_transform), t(cam.camera_center), t(np.zeros(3, np.float32)), cam.width, cam.height, cam.tan_fovx,
            cam.tan_fovy, scene.sh_degree]
    print("[gpu march] chunk sizes 16/32", flush=True)
    saved = dict(vol_marcher.DEFAULT_OPTIONS)
    try:
        out = {}
        for chunk in (16, 32):
            vol_marcher.DEFAULT_OPTIONS["bins_per_chunk"] = chunk
            cache = vol_marcher.cpp_binding.forward(*args, max_large_steps=2)
            out[chunk] = (cache[0].cpu().numpy(), cache[-1].cpu().numpy())
    finally:
        vol_marcher.DEFAULT_OPTIONS.update(saved)
    np.testing.assert_array_equal(out[16][0], out[32][0])
    np.testing.assert_array_equal(out[16][1], out[32][1])


def test_compacted_binning_equals_uncompacted_subset():
    """The marcher above the compaction threshold (4.3M Gaussians; the
    preprocess leaves the ids to the compaction) against the same render of
    only the Gaussians with z > 0 (no compaction), camera at the origin
    looking along +z so that every dropped Gaussian is culled: the same image
    and remaining transmittance bit for bit, the same tile ranges, and the
    same point list up to the subset's numbering."""
    torch = _torch()
    from differentiable_gaussian_mixture_renderer_amd import scenes
    from differentiable_gaussian_mixture_renderer_amd.python_bindings import vol_marcher as vm
    scene = scenes.make_scene(4_300_000, seed=37, extent=5.0, sh_degree=0, sh_mode="rgb")
    cam = scenes.look_at_camera((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 96, 64, 0.2, 0.15)  # ~40k Gaussians in view
    keep = np.nonzero(scene.means[:, 2] > 0.0)[0]
    assert keep.size < (4 << 20) <= scene.n
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.ascontiguousarray(a), device=dev)
    cam_args = [t(cam.world_view_transform), t(cam.full_proj_transform), t(cam.camera_center), t(np.zeros(3, np.float32)), cam.width,
                cam.height, cam.tan_fovx, cam.tan_fovy, 0]
    outs = []
    for sel in (slice(None), keep):
        args = [t(scene.sh[sel]), t(scene.opacity[sel]), t(scene.means[sel]), t(scene.scales[sel]), t(scene.rotations[sel])]
        print(f"[gpu march] compaction check n={args[2].shape[0]}", flush=True)
        cache = vm.cpp_binding.forward(*args, *cam_args)
        outs.append([c.cpu() for c in cache])
    torch.cuda.synchronize()
    full, sub = outs
    assert torch.equal(full[0], sub[0]) and torch.equal(full[12], sub[12])  # framebuffer, remaining transmittance
    assert torch.equal(full[10], sub[10])                                    # tile ranges
    assert full[11].numel() == sub[11].numel() > 0
    np.testing.assert_array_equal(full[11].numpy(), keep[sub[11].numpy()].astype(np.int32))


@pytest.mark.parametrize("mode,bin_", [(1, 0), (1, 40), (1, -1), (2, -1)])
def test_debug_render_modes(mode, bin_):
    """The reference's debug modes (vol_marcher_forward.cu:443-484): Bins shows
    one bin's (r m + g m + b m + m) * max_depth (green when empty) with T = 0;
    Depth multiplies exp(-m) over every bin and shows end() / max_depth (+inf
    until the section list is full, and for pixels whose tile marches on after
    they stopped).  Double GPU against the double oracle: equal infinities,
    finite values within 1e-9 relative."""
    torch = _torch()
    from differentiable_gaussian_mixture_renderer_amd.python_bindings import vol_marcher as vm
    scene, cam = small_scene()
    bg = np.array([0.2, 0.3, 0.4])
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.ascontiguousarray(a), device=dev, dtype=torch.float64)
    args = [t(scene.sh), t(scene.opacity), t(scene.means), t(scene.scales), t(scene.rotations), t(cam.world_view_transform),
            t(cam.full_proj_transform), t(cam.camera_center), t(bg), cam.width, cam.height, cam.tan_fovx, cam.tan_fovy, scene.sh_degree]
    print(f"[gpu march] debug mode {mode} bin {bin_}", flush=True)
    cache = vm.cpp_binding.forward(*args, render_mode=mode, debug_render_bin=bin_, max_depth=15.0)
    fb, T = cache[0].cpu().numpy(), cache[12].cpu().numpy()
    ref = O.MarchOracle(O.scene_arrays(scene), cam, dbl=True, float_binning=False)
    fb64, T64, _ = ref.render(bg, render_mode=mode, debug_render_bin=bin_, max_depth=15.0)
    assert np.array_equal(np.isinf(fb), np.isinf(fb64))
    fin = np.isfinite(fb64)
    scale = max(1.0, float(np.abs(fb64[fin]).max())) if fin.any() else 1.0
    assert np.abs(fb[fin] - fb64[fin]).max(initial=0.0) <= 1e-9 * scale
    assert np.abs(T - T64).max() <= 1e-9
    if mode == 1:
        assert (T == 0).all()
        assert fin.all() and (fb64 != 0).any()


@pytest.mark.parametrize("dtype_name,tile", [("float64", (16, 16)), ("float32", (16, 16)), ("float32", (8, 8))])
def test_replay_store_matches_recomputation(dtype_name, tile):
    """With a replay store the forward keeps each pixel's bins and borders and
    the backward reads them back for the tiles whose block took at most
    replay_steps large steps, recomputing the others.  Keeping the fewest
    steps any tile took (a mix of both kinds of tile) and the most (every
    tile read back) give the recomputing backward's gradients: within float64 rounding in double, and
    within the float atomics' summation order in float; the forward's image
    is bit-identical.  Double is also held to the oracle at 1e-9."""
    torch = _torch()
    dtype = getattr(torch, dtype_name)
    scene, cam = small_scene(n=3000, w=64, h=48, seed=4)
    dbl = dtype_name == "float64"
    bg = np.array([0.2, 0.3, 0.4]) if dbl else np.array([0.2, 0.3, 0.4], np.float32)
    rng = np.random.default_rng(8)
    grad = rng.normal(size=(3, cam.height, cam.width)) / (3 * cam.height * cam.width)
    grad = grad if dbl else grad.astype(np.float32)
    base = run_gpu(scene, cam, bg, dtype, grad_fb=grad, tile=tile)
    # the tiles' large-step counts (written whatever the store keeps)
    counts = run_gpu(scene, cam, bg, dtype, tile=tile, replay_steps=1)["tile_steps"]
    assert counts.max() > counts.min()
    for steps in (int(counts.min()), int(counts.max())):
        out = run_gpu(scene, cam, bg, dtype, grad_fb=grad, tile=tile, replay_steps=steps)
        np.testing.assert_array_equal(out["framebuffer"], base["framebuffer"])
        np.testing.assert_array_equal(out["remaining"], base["remaining"])
        ts = out["tile_steps"]
        np.testing.assert_array_equal(ts, counts)
        print(f"replay {steps}: tiles read back {(ts <= steps).sum()} of {ts.size}, most steps {ts.max()}")
        tol = 1e-12 if dbl else 2e-5
        for name in ("means", "scales", "rotations", "opacity", "sh"):
            got, want = out["grads"][name], base["grads"][name]
            rel = np.abs(got - want).max() / np.abs(want).max()
            print(f"  {name}: max rel {rel:.3g}")
            assert rel <= tol, (steps, name)
    if dbl:
        ref = O.MarchOracle(O.scene_arrays(scene), cam, dbl=True, float_binning=False)
        fb, T, _ = ref.render(bg)
        r = ref.backward(bg, fb, T, grad)
        n, k = scene.n, scene.sh.shape[1]
        for name in ("means", "scales", "rotations", "opacity", "sh"):
            got = out["grads"][name].reshape(n, -1)
            want = r[name].reshape(n, -1)[:, :got.shape[1]]
            assert np.abs(got - want).max() <= 1e-9 * np.abs(want).max(), name


def test_replay_store_after_a_debug_render_keeps_nothing():
    """A debug-mode forward marks every tile as not kept, so a backward given
    the same store recomputes everything (and matches the plain backward)."""
    torch = _torch()
    scene, cam = small_scene(n=1500, w=48, h=32, seed=6)
    bg = np.array([0.2, 0.3, 0.4], np.float32)
    out = run_gpu(scene, cam, bg, torch.float32, replay_steps=2, render_mode=1)
    assert (out["tile_steps"] == 0xFFFFFFFF).all()


def test_lists_ending_in_negative_depths():
    """Large Gaussians just in front of the camera have a negative depth
    (distance minus 1.5 of their largest scale), and their float32 keys sort
    after every positive one: a tile whose list ends with one takes the
    section pass's full walk (a Gaussian that close covers every tile here),
    where the other tests' tiles visit only the wave's candidates.  It holds
    the float-placed FP64 bar."""
    torch = _torch()
    scene, cam = small_scene(n=2000, w=48, h=32, seed=12)
    rng = np.random.default_rng(12)
    near = 12
    c = np.asarray(cam.camera_center, np.float64)
    fwd = -c / np.linalg.norm(c)
    pts = c + fwd * rng.uniform(0.45, 0.7, near)[:, None] + rng.normal(scale=0.08, size=(near, 3))
    scene.means[:near] = pts.astype(np.float32)
    scene.scales[:near] = 0.45
    scene.opacity[:near] = 0.3
    depth = np.linalg.norm(scene.means[:near] - c, axis=1) - 1.5 * 0.45
    assert (depth < 0).any()
    bg = np.array([0.2, 0.3, 0.4], np.float32)
    out = run_gpu(scene, cam, bg, torch.float32)
    ref32 = O.MarchOracle(O.scene_arrays(scene), cam, dbl=False)
    check_binning(out, ref32)
    depths = ref32.geometry()["depth"]
    pl = out["point_list"].view(np.uint32)
    ranges = out["ranges"].view(np.uint32)
    ends = [depths[pl[e - 1]] for s_, e in ranges if e > s_]
    assert any(d < 0 for d in ends)
    ref64 = O.MarchOracle(O.scene_arrays(scene), cam, dbl=True, float_binning=True)
    fb, T, _, amb = ref64.render(bg, with_ambiguity=True, float_sampling=True)
    ok = amb == 0
    err = np.abs(out["framebuffer"] - fb).max(axis=0)
    print(f"flagged {(~ok).mean():.4f}, max (unflagged) {err[ok].max():.3g}")
    assert (~ok).mean() <= 0.05
    assert err[ok].max() <= IMG_TOL
    assert np.abs(out["remaining"] - T)[ok].max() <= IMG_TOL
