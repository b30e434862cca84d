"""Batched multi-view splatting through one persistent launch per stage.

`BatchSplatter` renders one scene from many cameras of uneven resolution with
the C ABI's dgmr_splat_batch_forward / dgmr_splat_batch_backward: one
preprocess launch over all (view, Gaussian) pairs, one binning pass over the
global tile range, one persistent blend launch pulling tiles of every view from
a device queue.  The workspace carries geometry and binning from the forward to
the backward.

`render_plain` is the one-camera-per-launch path (a Splatter call per view)
that the batched path is measured against.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence

import torch

from . import _common as C
from . import _native as N


def _cam_struct(cam, bg):
    return N.camera_struct(cam.world_view_transform, cam.full_proj_transform, cam.camera_center, bg, cam.tan_fovx, cam.tan_fovy,
                           cam.width, cam.height)


class BatchSplatter:
    """Forward and backward of one scene over a list of cameras (scenes.Camera)."""

    def __init__(self, tile=(16, 16), batch: int = 64, device="cuda:0", cull: bool = False, spatial_order: bool = True):
        # cull: bin into the tiles where alpha can reach 1/255 only (same images and gradients, about a
        # fifth fewer pairs; over many small views the per-row reach arithmetic costs more than it saves)
        # spatial_order: the Gaussians take their slots in the Morton order of their centres, so a wave's
        # Gaussians are culled together in each view (dgmr_tiling.slot_order 0); in id order instead, a
        # backward writes its gradient rows contiguously, which pays when it returns many gradient groups
        self.tiling = C.tiling_struct(tile, batch, cull)
        self.spatial_order = spatial_order
        self.device = torch.device(device)
        self.capacity = 0
        self._ws = None
        self._views = None
        self._n_pairs = 0
        self._keep = []

    def _view_array(self, cams, bg, fbs, Ts):
        arr = (N.BatchView * len(cams))()
        for v, cam in enumerate(cams):
            arr[v].cam = _cam_struct(cam, bg)
            arr[v].framebuffer = fbs[v].data_ptr()
            arr[v].final_transmittance = Ts[v].data_ptr()
            arr[v].grad_framebuffer = None
            arr[v].l1_target = None
        return arr

    def _prepare(self, shs, opacity, means, scales, rotations, cameras, bg, sh_degree):
        shs, opacity, means, scales, rotations = C.prepare_inputs(shs, opacity, means, scales, rotations, True)
        dev = means.device
        self._g = C.gaussians_struct(shs, opacity, means, scales, rotations, sh_degree)
        self._keep = [shs, opacity, means, scales, rotations]
        self.cameras = list(cameras)
        self.bg = tuple(float(b) for b in bg)
        fbs = [torch.empty((3, c.height, c.width), dtype=torch.float32, device=dev) for c in self.cameras]
        Ts = [torch.empty((c.height, c.width), dtype=torch.float32, device=dev) for c in self.cameras]
        self._views = self._view_array(self.cameras, self.bg, fbs, Ts)
        self._fbs, self._Ts = fbs, Ts
        return dev

    @property
    def spatial_order(self) -> bool:
        return self.tiling.slot_order == 0

    @spatial_order.setter
    def spatial_order(self, on: bool):
        # read by the forward and by the backward that follows it: change it between steps only
        self.tiling.slot_order = 0 if on else 1

    def _launch_forward(self, dev):
        n = self._keep[2].shape[0]
        lib = N.load()
        # pair capacity from the largest pairs-per-pixel density seen so far
        px = sum(c.width * c.height for c in self.cameras)
        cap = max(int(C.pairs_per_pixel(dev) * px * 1.25) + 65536, 1 << 16)
        stats = N.Stats()
        stream = C.stream_handle(dev)
        with torch.cuda.device(dev):
            for _attempt in range(3):
                size = ctypes.c_size_t(0)
                N.check(lib.dgmr_batch_workspace_size(n, self._views, len(self.cameras), ctypes.byref(self.tiling), cap, ctypes.byref(size)),
                        "dgmr_batch_workspace_size")
                if self._ws is None or self._ws.numel() < size.value:
                    self._ws = None
                    self._ws = torch.empty(int(size.value * 1.05) + 256, dtype=torch.uint8, device=dev)
                rc = lib.dgmr_splat_batch_forward(ctypes.byref(self._g), self._views, len(self.cameras), ctypes.byref(self.tiling),
                                                  self._ws.data_ptr(), self._ws.numel(), cap, ctypes.byref(stats), stream)
                if rc == N.DGMR_ERR_CAPACITY:
                    cap = int(stats.n_pairs * 1.25) + 4096
                    continue
                N.check(rc, "dgmr_splat_batch_forward")
                break
            else:
                raise N.DgmrError(N.DGMR_ERR_CAPACITY, "dgmr_splat_batch_forward", "pair capacity kept growing")
        self.capacity = cap
        self._n_pairs = int(stats.n_pairs)
        C.remember_pairs_per_pixel(dev, self._n_pairs / max(px, 1))

    def forward(self, shs, opacity, means, scales, rotations, cameras: Sequence, bg=(0.0, 0.0, 0.0), sh_degree: int = 3):
        dev = self._prepare(shs, opacity, means, scales, rotations, cameras, bg, sh_degree)
        self._launch_forward(dev)
        return self._fbs, self._Ts

    def step_l1(self, shs, opacity, means, scales, rotations, cameras: Sequence, targets: Sequence[torch.Tensor], bg=(0.0, 0.0, 0.0),
                sh_degree: int = 3):
        """One multi-view training step with the L1 loss: the forward of every
        view and the gradients summed over the views, with dL/dimage =
        sign(image - target) / numel formed inside the blend backward from
        the image and the target (dgmr_batch_view.l1_target), so no gradient
        image is written or read.  Equals forward + losses.l1_grad_many +
        backward; every descriptor is prepared before the forward waits for
        its pair count, so the backward launches follow that wait at once.
        Returns (images, (g_sh, g_opacity, g_means, g_scales, g_rotations))."""
        dev = self._prepare(shs, opacity, means, scales, rotations, cameras, bg, sh_degree)
        if len(targets) != len(self.cameras):
            raise ValueError("step_l1: one target per camera")
        targets = [t.to(device=dev, dtype=torch.float32).contiguous() for t in targets]
        for f, t in zip(self._fbs, targets):
            if f.shape != t.shape:
                raise ValueError("step_l1: each target has its view's image shape")
        for v, (f, t) in enumerate(zip(self._fbs, targets)):
            self._views[v].grad_framebuffer = None
            self._views[v].l1_target = t.data_ptr()
            self._views[v].l1_scale = 1.0 / max(f.numel(), 1)  # the scale losses.l1_items passes
        out = self._grad_buffers(1)
        self._launch_forward(dev)
        self._launch_backward(None, 1, out)
        self._step_keep = targets  # alive until the next step's launches are queued behind these
        return self._fbs, tuple(t[0] for t in out)

    @property
    def n_pairs(self) -> int:
        return self._n_pairs

    def backward(self, grad_images: List[Optional[torch.Tensor]], groups: Optional[Sequence[int]] = None):
        """Gradients of the scene parameters.  groups=None: summed over all
        views, returned as (g_sh, g_opacity, g_means, g_scales, g_rotations).
        Otherwise groups[v] in {-1, 0..G-1} (a group's views consecutive, -1 =
        no gradient, its grad image may be None) and a list of G such tuples is
        returned, one per group."""
        return self._backward(grad_images, None, groups)

    def backward_l1(self, targets: List[Optional[torch.Tensor]], groups: Optional[Sequence[int]] = None):
        """backward() with dL/dimage = sign(image - target) / numel of each
        view formed inside the blend backward (dgmr_batch_view.l1_target):
        the same gradients as backward(losses.l1_grad_many(images, targets)),
        without writing or reading gradient images.  targets[v] None: view v
        has a zero gradient (its tiles are skipped)."""
        if len(targets) != len(self.cameras):
            raise ValueError("backward_l1: one target (or None) per camera")
        return self._backward([None] * len(targets), targets, groups)

    def _backward(self, grad_images, targets, groups):
        G = 1 if groups is None else max(1, max(groups) + 1)
        dev = self._fbs[0].device
        keep = []
        for v, g in enumerate(grad_images):  # the forward's descriptors; only the gradient inputs are new
            t = None if targets is None else targets[v]
            if t is not None:
                t = t.to(device=dev, dtype=torch.float32).contiguous()
                if t.shape != self._fbs[v].shape:
                    raise ValueError("backward_l1: each target has its view's image shape")
                self._views[v].grad_framebuffer = None
                self._views[v].l1_target = t.data_ptr()
                self._views[v].l1_scale = 1.0 / max(t.numel(), 1)  # the scale losses.l1_items passes
                keep.append(t)
                continue
            self._views[v].l1_target = None
            if g is None:  # no gradient image: the view adds nothing (dgmr_splat_batch_backward skips its tiles)
                self._views[v].grad_framebuffer = None
                continue
            g = g.to(torch.float32).contiguous()
            keep.append(g)
            self._views[v].grad_framebuffer = g.data_ptr()
        out = self._grad_buffers(G)
        self._launch_backward(groups, G, out)
        out = [(g_sh, g_op, g_means, g_scales, g_rot) for g_sh, g_op, g_means, g_scales, g_rot in zip(*[list(t) for t in out])]
        return out[0] if groups is None else out

    def _grad_buffers(self, G):
        shs, opacity, means, _, _ = self._keep
        n = means.shape[0]
        f32 = dict(dtype=torch.float32, device=means.device)
        return (torch.empty((G, *shs.shape), **f32), torch.empty((G, *opacity.shape), **f32), torch.empty((G, n, 3), **f32),
                torch.empty((G, n, 3), **f32), torch.empty((G, n, 4), **f32))

    def _launch_backward(self, groups, G, out):
        g_sh, g_op, g_means, g_scales, g_rot = out
        dev = g_sh.device
        grp = None if groups is None else (ctypes.c_int32 * len(groups))(*[int(x) for x in groups])
        with torch.cuda.device(dev):
            rc = N.load().dgmr_splat_batch_backward(ctypes.byref(self._g), self._views, len(self.cameras), ctypes.byref(self.tiling),
                                                    self._ws.data_ptr(), self._ws.numel(), self.capacity, self._n_pairs, grp, G,
                                                    g_means.data_ptr(), g_scales.data_ptr(), g_rot.data_ptr(), g_op.data_ptr(),
                                                    g_sh.data_ptr(), C.stream_handle(dev))
            N.check(rc, "dgmr_splat_batch_backward")


def render_plain(shs, opacity, means, scales, rotations, cameras: Sequence, bg=(0.0, 0.0, 0.0), sh_degree: int = 3,
                 grad_fn=None, tile=(16, 16), batch: int = 64, cull: bool = False):
    """One Splatter forward (and backward when grad_fn maps image -> dL/dimage)
    per camera; gradients summed over views."""
    from .python_bindings import splatter
    bg_t = torch.tensor(bg, dtype=torch.float32)
    images = []
    acc = None
    for cam in cameras:
        args = (shs, opacity, means, scales, rotations, torch.tensor(cam.world_view_transform), torch.tensor(cam.full_proj_transform),
                torch.tensor(cam.camera_center), bg_t, cam.width, cam.height, cam.tan_fovx, cam.tan_fovy, sh_degree)
        cache = splatter.cpp_binding.forward(*args, tile=tile, batch=batch, cull=cull)
        images.append(cache[0])
        if grad_fn is not None:
            g = splatter.cpp_binding.backward(*args, grad_fn(cache[0]), cache, tile=tile, batch=batch)
            acc = g if acc is None else [a + b for a, b in zip(acc, g)]
    return images, acc
