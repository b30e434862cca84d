"""Drop-in for the reference's python_bindings/vol_marcher.py.

`VolMarcher.apply(shs, opacity, means3D, scales, rotations, world_view_transform,
full_proj_transform, camera_center, bg_color, image_width, image_height,
tanfovx, tanfovy, active_sh_degree) -> image [3,H,W]` with the same argument
order, dtypes (float32 or float64, all inputs alike) and gradients
(python_bindings/vol_marcher.py:19-65).  `cpp_binding.forward` returns the
reference's 13-tensor list (python_bindings/vol_marcher.cu:109-122):
[framebuffer, rects, rgb, rgb_sh_clamped, depths, points_xy_image,
inverse_filtered_cov3d, filtered_masses, tiles_touched, point_offsets,
i_ranges, b_point_list, remaining_transparency]; `cpp_binding.backward`
returns [grad_sh, grad_weights, grad_means, grad_scales, grad_rotations].

The kernels run on MI355X through libdgmr_hip.so (include/dgmr.h); there is no
CPU fallback.
"""
from __future__ import annotations

import ctypes
import types

import torch

from .. import _common as C
from .. import _native as N

DEFAULT_TILING = ((16, 16), 256)
# march options: large-step cap (0 = unbounded, as the reference) and frozen
# sample sections for gradient checks (DGMR_TORCH_GRAD_CHECK_CONST_SAMPLES)
# and the reference's debug render modes (ForwardData::debug_render_mode,
# debug_render_bin, max_depth; vol_marcher.h:44-55), forward only
# and the replay store (off by default: the backward recomputes every bin, as
# the reference's does; "replay": True makes VolMarcher.apply keep the
# forward's bins for its backward, "replay_steps" large steps of them, 0 =
# the large-step cap or 4, within "replay_max_bytes")
DEFAULT_OPTIONS = {"max_large_steps": 0, "frozen": None, "bins_per_chunk": 0, "render_mode": 0, "debug_render_bin": -1,
                   "max_depth": 20.0, "replay": False, "replay_steps": 0, "replay_max_bytes": 16 << 30}
RENDER_MODES = {"full": 0, "bins": 1, "depth": 2}


class MarchReplay:
    """Device memory in which the forward keeps each pixel's bins and bin
    borders for the first `steps` large steps (dgmr_march_options.replay),
    so that the backward reads them back instead of recomputing them.  Tiles
    whose rays took more large steps are recomputed; results are the same
    either way.  No counterpart in the reference, whose backward recomputes
    every bin (vol_marcher_backward.cu:126-318)."""

    def __init__(self, buffer: torch.Tensor, steps: int):
        self.buffer = buffer
        self.steps = int(steps)


def replay_bytes(width: int, height: int, tile, batch: int, steps: int, dtype) -> int:
    lib = N.load()
    out = ctypes.c_size_t(0)
    N.check(lib.dgmr_march_replay_size(int(width), int(height), ctypes.byref(C.tiling_struct(tile, batch)), int(steps),
                                       N.DGMR_F64 if dtype == torch.float64 else N.DGMR_F32, ctypes.byref(out)), "dgmr_march_replay_size")
    return int(out.value)


def make_replay(width: int, height: int, dtype, device, tile=None, batch: int = 0, steps=None, max_bytes=None):
    """A replay store for one frame, or None when none fits: `steps` large
    steps (default: the large-step cap when set, else 4, at most 4), fewer
    while the store would exceed max_bytes or a quarter of the free device
    memory."""
    tile = tile or DEFAULT_TILING[0]
    batch = batch or DEFAULT_TILING[1]
    if steps is None:
        steps = int(DEFAULT_OPTIONS.get("replay_steps", 0))
    if steps <= 0:
        cap = int(DEFAULT_OPTIONS["max_large_steps"])
        steps = min(cap, 4) if cap > 0 else 4
    limit = int(DEFAULT_OPTIONS.get("replay_max_bytes", 0) if max_bytes is None else max_bytes)
    free, _total = torch.cuda.mem_get_info(device)
    limit = min(limit, free // 4)
    while steps >= 1:
        nbytes = replay_bytes(width, height, tile, batch, steps, dtype)
        if nbytes <= limit:
            return MarchReplay(torch.empty((nbytes,), dtype=torch.uint8, device=device), steps)
        steps -= 1
    return None


def _replay_options(o: N.MarchOptions, replay) -> None:
    if replay is None:
        return
    if replay.buffer.data_ptr() % 256:
        raise ValueError("replay store must be 256-byte aligned")
    o.replay = replay.buffer.data_ptr()
    o.replay_bytes = replay.buffer.numel()
    o.replay_steps = replay.steps


def _options(max_large_steps=None, frozen=None, render_mode=None, debug_render_bin=None, max_depth=None) -> N.MarchOptions:
    o = N.MarchOptions()
    mode = DEFAULT_OPTIONS["render_mode"] if render_mode is None else render_mode
    o.render_mode = int(RENDER_MODES.get(mode, mode) if isinstance(mode, str) else mode)
    o.debug_render_bin = int(DEFAULT_OPTIONS["debug_render_bin"] if debug_render_bin is None else debug_render_bin)
    o.max_depth = float(DEFAULT_OPTIONS["max_depth"] if max_depth is None else max_depth)
    o.max_large_steps = int(DEFAULT_OPTIONS["max_large_steps"] if max_large_steps is None else max_large_steps)
    fr = DEFAULT_OPTIONS["frozen"] if frozen is None else frozen
    fr = list(fr or [])
    if len(fr) > 8:
        raise ValueError("at most 8 frozen sections")
    o.n_frozen = len(fr)
    o.bins_per_chunk = int(DEFAULT_OPTIONS.get("bins_per_chunk", 0))
    for i, (a, b, c) in enumerate(fr):
        o.frozen[i][0], o.frozen[i][1], o.frozen[i][2] = float(a), float(b), float(c)
    return o


def _forward(sh_params, weights, centroids, cov_scales, cov_rotations, view_matrix, proj_matrix, cam_position, background,
             render_width: int, render_height: int, tan_fovx: float, tan_fovy: float, sh_degree: int,
             tile=None, batch: int = 0, max_large_steps=None, frozen=None, scale_multiplier: float = 1.0, render_mode=None,
             debug_render_bin=None, max_depth=None, replay=None):
    tile = tile or DEFAULT_TILING[0]
    batch = batch or DEFAULT_TILING[1]
    sh_params, weights, centroids, cov_scales, cov_rotations = C.prepare_inputs(sh_params, weights, centroids, cov_scales, cov_rotations,
                                                                                 False)
    dev = centroids.device
    dt = centroids.dtype
    n = centroids.shape[0]
    W, H = int(render_width), int(render_height)
    lib = N.load()
    tiling = C.tiling_struct(tile, batch)
    cam = C.camera_struct(view_matrix, proj_matrix, cam_position, background, W, H, tan_fovx, tan_fovy)
    g = C.gaussians_struct(sh_params, weights, centroids, cov_scales, cov_rotations, sh_degree, scale_multiplier)
    opt = _options(max_large_steps, frozen, render_mode, debug_render_bin, max_depth)
    _replay_options(opt, replay)
    gx, gy = (W + tiling.tile_w - 1) // tiling.tile_w, (H + tiling.tile_h - 1) // tiling.tile_h
    fo = dict(dtype=dt, device=dev)
    io = dict(dtype=torch.int32, device=dev)
    with torch.cuda.device(dev):
        fb = torch.zeros((3, H, W), **fo)
        remaining = torch.ones((H, W), **fo)
        rects = torch.empty((n, 2), **io)
        rgb = torch.empty((n, 3), **fo)
        clamped = torch.empty((n, 3), dtype=torch.bool, device=dev)
        depths = torch.empty((n,), **fo)
        xy = torch.empty((n, 2), **fo)
        inv_cov = torch.empty((n, 6), **fo)
        masses = torch.empty((n,), **fo)
        tiles = torch.empty((n,), **io)
        offsets = torch.empty((n,), **io)
        ranges = torch.empty((gx * gy, 2), **io)
        geom = N.MarchGeometry(rects.data_ptr(), rgb.data_ptr(), clamped.data_ptr(), depths.data_ptr(), xy.data_ptr(),
                               inv_cov.data_ptr(), masses.data_ptr(), tiles.data_ptr())
        cap = C.initial_capacity(dev, "march", n)
        stats = N.Stats()
        stream = C.stream_handle(dev)
        for _attempt in range(3):
            point_list = torch.empty((max(cap, 1),), **io)
            bins = N.Binning(offsets.data_ptr(), point_list.data_ptr(), ranges.data_ptr(), cap)
            ws = C.workspace(dev, n, W, H, tiling, cap)
            rc = lib.dgmr_march_forward(ctypes.byref(g), ctypes.byref(cam), ctypes.byref(tiling), ctypes.byref(opt), ctypes.byref(geom),
                                        ctypes.byref(bins), fb.data_ptr(), remaining.data_ptr(), ws.data_ptr(), ws.numel(),
                                        ctypes.byref(stats), stream)
            if rc == N.DGMR_ERR_CAPACITY:
                cap = int(stats.n_pairs * 1.25) + 1024
                C.remember_capacity(dev, "march", cap)
                continue
            N.check(rc, "dgmr_march_forward")
            break
        else:
            raise N.DgmrError(N.DGMR_ERR_CAPACITY, "dgmr_march_forward", "pair capacity kept growing")
        C.remember_capacity(dev, "march", cap)
        R = int(stats.n_pairs)
    return [fb, rects, rgb, clamped, depths, xy, inv_cov, masses, tiles, offsets, ranges, point_list[:R], remaining]


def _backward(sh_params, weights, centroids, cov_scales, cov_rotations, view_matrix, proj_matrix, cam_position, background,
              render_width: int, render_height: int, tan_fovx: float, tan_fovy: float, sh_degree: int, grad_framebuffer, cache_vector,
              tile=None, batch: int = 0, max_large_steps=None, frozen=None, scale_multiplier: float = 1.0, replay=None):
    tile = tile or DEFAULT_TILING[0]
    batch = batch or DEFAULT_TILING[1]
    sh_params, weights, centroids, cov_scales, cov_rotations = C.prepare_inputs(sh_params, weights, centroids, cov_scales, cov_rotations,
                                                                                 False)
    fb, rects, rgb, clamped, depths, xy, inv_cov, masses, tiles, offsets, ranges, point_list, remaining = cache_vector
    dev = centroids.device
    dt = centroids.dtype
    n = centroids.shape[0]
    W, H = int(render_width), int(render_height)
    lib = N.load()
    tiling = C.tiling_struct(tile, batch)
    cam = C.camera_struct(view_matrix, proj_matrix, cam_position, background, W, H, tan_fovx, tan_fovy)
    g = C.gaussians_struct(sh_params, weights, centroids, cov_scales, cov_rotations, sh_degree, scale_multiplier)
    opt = _options(max_large_steps, frozen)
    _replay_options(opt, replay)
    grad_framebuffer = grad_framebuffer.to(dt).contiguous()
    fo = dict(dtype=dt, device=dev)
    with torch.cuda.device(dev):
        rg = torch.empty((n, 13), **fo)
        g_means = torch.empty((n, 3), **fo)
        g_scales = torch.empty((n, 3), **fo)
        g_rot = torch.empty((n, 4), **fo)
        g_w = torch.empty((n, 1), **fo)
        g_sh = torch.empty(sh_params.shape, **fo)
        geom = N.MarchGeometry(rects.data_ptr(), rgb.data_ptr(), clamped.data_ptr(), depths.data_ptr(), xy.data_ptr(),
                               inv_cov.data_ptr(), masses.data_ptr(), tiles.data_ptr())
        bins = N.Binning(offsets.data_ptr(), point_list.data_ptr(), ranges.data_ptr(), point_list.numel())
        rc = lib.dgmr_march_backward(ctypes.byref(g), ctypes.byref(cam), ctypes.byref(tiling), ctypes.byref(opt), ctypes.byref(geom),
                                     ctypes.byref(bins), point_list.numel(), fb.data_ptr(), remaining.data_ptr(),
                                     grad_framebuffer.data_ptr(), rg.data_ptr(), g_means.data_ptr(), g_scales.data_ptr(),
                                     g_rot.data_ptr(), g_w.data_ptr(), g_sh.data_ptr(), C.stream_handle(dev))
        N.check(rc, "dgmr_march_backward")
    return [g_sh, g_w, g_means, g_scales, g_rot]


cpp_binding = types.SimpleNamespace(forward=_forward, backward=_backward)


class VolMarcher(torch.autograd.Function):
    """python_bindings/vol_marcher.py:19-62"""

    @staticmethod
    def forward(ctx, shs: torch.Tensor, opacity: torch.Tensor, means3D: torch.Tensor, scales: torch.Tensor, rotations: torch.Tensor,
                world_view_transform: torch.Tensor, full_proj_transform: torch.Tensor, camera_center: torch.Tensor,
                bg_color: torch.Tensor, image_width: int, image_height: int, tanfovx: float, tanfovy: float, active_sh_degree: int):
        # the replay store lives beside the reference's cache, for this call's backward only
        ctx.replay = None
        if DEFAULT_OPTIONS.get("replay") and any(ctx.needs_input_grad[:5]) and int(DEFAULT_OPTIONS["render_mode"]) == 0:
            ctx.replay = make_replay(int(image_width), int(image_height), means3D.dtype, means3D.device)
        cache = cpp_binding.forward(shs, opacity, means3D, scales, rotations, world_view_transform, full_proj_transform, camera_center,
                                    bg_color, image_width, image_height, tanfovx, tanfovy, active_sh_degree, replay=ctx.replay)
        ctx.save_for_backward(shs, opacity, means3D, scales, rotations, world_view_transform, full_proj_transform, camera_center,
                              bg_color, torch.tensor(image_width), torch.tensor(image_height), torch.tensor(tanfovx), torch.tensor(tanfovy),
                              torch.tensor(active_sh_degree), *cache)
        return cache[0]

    @staticmethod
    def backward(ctx, grad_output):
        if not grad_output.is_contiguous():
            grad_output = grad_output.contiguous()
        shs, opacity, means3D, scales, rotations, world_view_transform, full_proj_transform, camera_center, bg_color, image_width, \
            image_height, tanfovx, tanfovy, active_sh_degree, *cache = ctx.saved_tensors
        grad_sh, grad_weights, grad_positions, grad_scales, grad_rotations = cpp_binding.backward(
            shs, opacity, means3D, scales, rotations, world_view_transform, full_proj_transform, camera_center, bg_color,
            image_width.item(), image_height.item(), tanfovx.item(), tanfovy.item(), active_sh_degree.item(), grad_output, cache,
            replay=ctx.replay)
        ctx.replay = None
        return grad_sh, grad_weights, grad_positions, grad_scales, grad_rotations, None, None, None, None, None, None, None, None, None


apply = VolMarcher.apply
