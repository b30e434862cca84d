"""Batched multi-view splatting (one persistent launch per stage) against the
one-camera-per-launch path and the oracle: every view's image and the summed
gradients agree with rendering each camera alone."""
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


def scene_and_cams(n=20_000, n_cams=6, seed=4):
    from differentiable_gaussian_mixture_renderer_amd import scenes
    scene = scenes.make_scene(n, seed=seed, extent=5.0, sh_degree=3)
    cams = scenes.sphere_cameras(n_cams, seed=seed, resolutions=[(64, 64), (128, 96), (160, 90), (200, 120)])
    return scene, cams


@pytest.mark.parametrize("tile,batch,cull", [((16, 16), 64, False), ((8, 8), 128, False), ((32, 8), 256, False),
                                             ((16, 16), 64, True), ((8, 8), 256, True)])
def test_batch_matches_single_views(tile, batch, cull):
    """cull=True (tight culling in the batch) against the plain path with the
    reference binning: the same images and gradients."""
    torch = _torch()
    from differentiable_gaussian_mixture_renderer_amd.batching import BatchSplatter, render_plain
    scene, cams = scene_and_cams()
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.ascontiguousarray(a), device=dev)
    params = [t(scene.sh), t(scene.opacity), t(scene.means), t(scene.scales), t(scene.rotations)]
    bg = (0.1, 0.2, 0.3)
    print(f"[gpu batch] {len(cams)} views tile={tile} batch={batch} cull={cull}", flush=True)
    bs = BatchSplatter(tile=tile, batch=batch, cull=cull)
    fbs, Ts = bs.forward(*params, cams, bg=bg)
    rng = np.random.default_rng(1)
    grads = [torch.tensor(rng.normal(size=(3, c.height, c.width)).astype(np.float32) / (3 * c.height * c.width), device=dev) for c in cams]
    gb = bs.backward(grads)
    gi = iter(grads)
    imgs, gp = render_plain(*params, cams, bg=bg, grad_fn=lambda img: next(gi), tile=tile, batch=batch)
    torch.cuda.synchronize()
    for v, cam in enumerate(cams):
        np.testing.assert_allclose(fbs[v].cpu().numpy(), imgs[v].cpu().numpy(), rtol=0, atol=1e-6)
        ref = O.SplatOracle(O.scene_arrays(scene), cam, dbl=False, tile=tile)
        fb_ref, _, amb, _ = ref.render(np.array(bg, np.float32))
        assert np.abs(fbs[v].cpu().numpy() - fb_ref).max(axis=0)[amb == 0].max() <= 1e-4
    for a, b in zip(gb, gp):
        a, b = a.cpu().numpy(), b.cpu().numpy()
        assert np.linalg.norm(a - b) <= 1e-4 * np.linalg.norm(b) + 1e-12


def test_gradient_groups_match_per_group_plain_sums():
    """view groups [0, 0, -1, 1, 1, 1]: group k's gradients equal the plain
    path's gradients summed over its views (||diff|| <= 1e-4 ||ref||); the
    view without a group contributes to none.  Views without a gradient image
    pass NULL through dgmr_splat_batch_backward (view 2 in group -1, and view
    4 in group 1, which then adds nothing, as a zero dL/dimage)."""
    torch = _torch()
    from differentiable_gaussian_mixture_renderer_amd.batching import BatchSplatter, render_plain
    scene, cams = scene_and_cams()
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.ascontiguousarray(a), device=dev)
    params = [t(scene.sh), t(scene.opacity), t(scene.means), t(scene.scales), t(scene.rotations)]
    print(f"[gpu batch] gradient groups over {len(cams)} views", flush=True)
    bs = BatchSplatter()
    fbs, _ = bs.forward(*params, cams)
    rng = np.random.default_rng(2)
    grads = [torch.tensor(rng.normal(size=(3, c.height, c.width)).astype(np.float32) / (3 * c.height * c.width), device=dev) for c in cams]
    groups = [0, 0, -1, 1, 1, 1]
    out = bs.backward([g if k >= 0 and v != 4 else None for v, (g, k) in enumerate(zip(grads, groups))], groups=groups)
    assert len(out) == 2
    for k, members in ((0, [0, 1]), (1, [3, 5])):
        gi = iter([grads[v] for v in members])
        _, ref = render_plain(*params, [cams[v] for v in members], grad_fn=lambda img: next(gi))
        for a, b in zip(out[k], ref):
            a, b = a.cpu().numpy(), b.cpu().numpy()
            assert np.linalg.norm(a - b) <= 1e-4 * np.linalg.norm(b) + 1e-12, k


@pytest.mark.parametrize("cull", [False, True])
def test_large_batch_with_compacted_depth_sort(cull):
    """40 views of 120k Gaussians (4.8M view-Gaussian entries): the batched
    binning compacts the visible entries before its depth sort; images and
    gradients must still equal the one-camera-per-launch path's (reference
    binning), with either binning in the batch."""
    torch = _torch()
    from differentiable_gaussian_mixture_renderer_amd.batching import BatchSplatter, render_plain
    scene, cams = scene_and_cams(n=120_000, n_cams=40, seed=6)
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.ascontiguousarray(a), device=dev)
    params = [t(scene.sh), t(scene.opacity), t(scene.means), t(scene.scales), t(scene.rotations)]
    print(f"[gpu batch] {len(cams)} views x {scene.n} Gaussians (compacted depth sort) cull={cull}", flush=True)
    bs = BatchSplatter(cull=cull)
    fbs, _ = bs.forward(*params, cams)
    rng = np.random.default_rng(3)
    grads = [torch.tensor(rng.normal(size=(3, c.height, c.width)).astype(np.float32) / (3 * c.height * c.width), device=dev) for c in cams]
    gb = bs.backward(grads)
    gi = iter(grads)
    imgs, gp = render_plain(*params, cams, grad_fn=lambda img: next(gi))
    torch.cuda.synchronize()
    for a, b in zip(fbs, imgs):
        np.testing.assert_allclose(a.cpu().numpy(), b.cpu().numpy(), rtol=0, atol=1e-6)
    for a, b in zip(gb, gp):
        a, b = a.cpu().numpy(), b.cpu().numpy()
        assert np.linalg.norm(a - b) <= 1e-4 * np.linalg.norm(b) + 1e-12


def test_step_l1_equals_forward_loss_backward():
    """BatchSplatter.step_l1 (buffers prepared before the forward's wait) is
    forward + losses.l1_grad_many + backward: the same images, the same
    gradients up to the order of the atomic sums."""
    torch = _torch()
    from differentiable_gaussian_mixture_renderer_amd import losses
    from differentiable_gaussian_mixture_renderer_amd.batching import BatchSplatter
    scene, cams = scene_and_cams(seed=9)
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.ascontiguousarray(a), device=dev)
    params = [t(scene.sh), t(scene.opacity), t(scene.means), t(scene.scales), t(scene.rotations)]
    gen = torch.Generator(device=dev).manual_seed(5)
    targets = [torch.rand((3, c.height, c.width), generator=gen, device=dev) for c in cams]
    print(f"[gpu batch] step_l1 over {len(cams)} views", flush=True)
    bs = BatchSplatter()
    fbs, _ = bs.forward(*params, cams)
    ref = bs.backward(losses.l1_grad_many(fbs, targets))
    imgs, got = BatchSplatter().step_l1(*params, cams, targets)
    torch.cuda.synchronize()
    for a, b in zip(imgs, fbs):
        assert torch.equal(a, b)
    for a, b in zip(got, ref):
        a, b = a.cpu().numpy(), b.cpu().numpy()
        assert np.linalg.norm(a - b) <= 1e-5 * np.linalg.norm(b) + 1e-12


def test_repeated_backward_after_one_forward():
    """The forward leaves the screen gradients zeroed and the first backward
    uses them without a clearing pass; a second and third backward after the
    same forward clear them again: all three give the same gradients (up to
    the order of the atomic sums), and equal a fresh forward + backward."""
    torch = _torch()
    from differentiable_gaussian_mixture_renderer_amd.batching import BatchSplatter
    scene, cams = scene_and_cams(seed=13)
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.ascontiguousarray(a), device=dev)
    params = [t(scene.sh), t(scene.opacity), t(scene.means), t(scene.scales), t(scene.rotations)]
    rng = np.random.default_rng(3)
    grads = [torch.tensor(rng.normal(size=(3, c.height, c.width)).astype(np.float32), device=dev) for c in cams]
    print(f"[gpu batch] three backwards after one forward over {len(cams)} views", flush=True)
    bs = BatchSplatter()
    bs.forward(*params, cams)
    runs = [bs.backward(grads) for _ in range(3)]
    fresh = BatchSplatter()
    fresh.forward(*params, cams)
    ref = fresh.backward(grads)
    torch.cuda.synchronize()
    for run in runs:
        for a, b in zip(run, ref):
            a, b = a.cpu().numpy(), b.cpu().numpy()
            assert np.linalg.norm(a - b) <= 1e-5 * np.linalg.norm(b) + 1e-12


def test_backward_l1_matches_separate_loss_gradient():
    """backward_l1 (L1 gradient formed inside the blend backward) against
    backward(l1_grad_many(images, targets)) with groups [0, 0, -1, 1, 1, 1]
    and no target for the ungrouped view: the per-pixel gradient is the same
    float expression, so the sums agree up to the order of the atomic adds.
    A plain backward after it (grad images again) gives its own result."""
    torch = _torch()
    from differentiable_gaussian_mixture_renderer_amd import losses
    from differentiable_gaussian_mixture_renderer_amd.batching import BatchSplatter
    scene, cams = scene_and_cams(seed=21)
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.ascontiguousarray(a), device=dev)
    params = [t(scene.sh), t(scene.opacity), t(scene.means), t(scene.scales), t(scene.rotations)]
    gen = torch.Generator(device=dev).manual_seed(8)
    targets = [torch.rand((3, c.height, c.width), generator=gen, device=dev) for c in cams]
    groups = [0, 0, -1, 1, 1, 1]
    print(f"[gpu batch] fused L1 backward over {len(cams)} views", flush=True)
    bs = BatchSplatter()
    fbs, _ = bs.forward(*params, cams)
    fused = bs.backward_l1([tg if k >= 0 else None for tg, k in zip(targets, groups)], groups=groups)
    gimg = losses.l1_grad_many(fbs, targets)
    ref = bs.backward([g if k >= 0 else None for g, k in zip(gimg, groups)], groups=groups)
    rng = np.random.default_rng(6)
    other = [torch.tensor(rng.normal(size=(3, c.height, c.width)).astype(np.float32), device=dev) for c in cams]
    plain = bs.backward(other)
    fresh = BatchSplatter()
    fresh.forward(*params, cams)
    plain_ref = fresh.backward(other)
    torch.cuda.synchronize()
    assert len(fused) == 2
    for k in range(2):
        for a, b in zip(fused[k], ref[k]):
            a, b = a.cpu().numpy(), b.cpu().numpy()
            assert np.linalg.norm(a - b) <= 1e-5 * np.linalg.norm(b) + 1e-12, k
    for a, b in zip(plain, plain_ref):
        a, b = a.cpu().numpy(), b.cpu().numpy()
        assert np.linalg.norm(a - b) <= 1e-5 * np.linalg.norm(b) + 1e-12


def test_id_slot_order_matches_spatial_order():
    """dgmr_tiling.slot_order 1 (Gaussians in id order) renders the same
    images and gradients as the default Morton slot order; the two differ
    only in which thread handles which Gaussian (and in the order of equal
    depths, which this scene does not have on a shared pixel)."""
    torch = _torch()
    from differentiable_gaussian_mixture_renderer_amd.batching import BatchSplatter
    scene, cams = scene_and_cams(n=30_000, n_cams=8, seed=7)
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.ascontiguousarray(a), device=dev)
    params = [t(scene.sh), t(scene.opacity), t(scene.means), t(scene.scales), t(scene.rotations)]
    rng = np.random.default_rng(5)
    grads = [torch.tensor(rng.normal(size=(3, c.height, c.width)).astype(np.float32) / (3 * c.height * c.width), device=dev) for c in cams]
    out = {}
    for spatial in (True, False):
        print(f"[gpu batch] slot order spatial={spatial}", flush=True)
        bs = BatchSplatter(spatial_order=spatial)
        fbs, _ = bs.forward(*params, cams)
        gb = bs.backward(grads, groups=[0, 0, 0, 1, 1, -1, 2, 2])
        torch.cuda.synchronize()
        out[spatial] = ([f.cpu().numpy() for f in fbs], [[g.cpu().numpy() for g in grp] for grp in gb])
    for a, b in zip(out[True][0], out[False][0]):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-6)
    for ga, gb_ in zip(out[True][1], out[False][1]):
        for a, b in zip(ga, gb_):
            assert np.linalg.norm(a - b) <= 1e-5 * np.linalg.norm(b) + 1e-12
