"""Tile splatter behind the reference's Python operator interface.

The reference exposes its splatter only to C++ (sibr_interfaces.cpp:26-29,
splat_forward.cu:30); this module gives it the argument list of the
reference's `VolMarcher` (python_bindings/vol_marcher.py:19-65) so a caller can
switch renderers by name:

    image = Splatter.apply(shs, opacity, means3D, scales, rotations,
                           world_view_transform, full_proj_transform, camera_center,
                           bg_color, image_width, image_height, tanfovx, tanfovy,
                           active_sh_degree)

`cpp_binding.forward` / `cpp_binding.backward` mirror the extension functions of
python_bindings/vol_marcher.cu:217-304 (same argument order; the cache is a
list of tensors that round-trips through save_for_backward).
"""
from __future__ import annotations

import ctypes
import types

import torch

from .. import _common as C
from .. import _native as N

# tile shape and LDS batch used when the caller passes none
DEFAULT_TILING = ((16, 16), 64)
# binning used when the caller passes none: False keeps the reference's tile
# rectangles (its b_point_list / i_ranges); True bins only the tiles where
# alpha can reach 1/255 (same image and gradients, fewer pairs)
DEFAULT_CULL = False

# cache layout returned by forward (index: content)
CACHE_FIELDS = ("framebuffer", "rects", "rgb", "rgb_sh_clamped", "depths", "points_xy_image", "conic_opacity",
                "tiles_touched", "point_offsets", "i_ranges", "b_point_list", "final_transparency", "n_pairs")


def _forward(shs, weights, centroids, cov_scales, cov_rotations, view_matrix, proj_matrix, cam_position, background,
             render_width: int, render_height: int, tan_fovx: float, tan_fovy: float, sh_degree: int,
             tile=None, batch: int = 0, scale_multiplier: float = 1.0, activation: int = N.DGMR_ACT_NONE, cull=None):
    """cull=True bins each Gaussian only into the tiles where its alpha can
    reach 1/255 (dgmr_tiling.cull): same image, transmittance and gradients,
    fewer pairs; b_point_list and i_ranges then differ from the reference's.
    None: DEFAULT_CULL."""
    cull = DEFAULT_CULL if cull is None else bool(cull)
    tile = tile or DEFAULT_TILING[0]
    batch = batch or DEFAULT_TILING[1]
    shs, weights, centroids, cov_scales, cov_rotations = C.prepare_inputs(shs, weights, centroids, cov_scales, cov_rotations, True)
    dev = centroids.device
    n = centroids.shape[0]
    W, H = int(render_width), int(render_height)
    lib = N.load()
    tiling = C.tiling_struct(tile, batch, cull)
    cam = C.camera_struct(view_matrix, proj_matrix, cam_position, background, W, H, tan_fovx, tan_fovy)
    g = C.gaussians_struct(shs, weights, centroids, cov_scales, cov_rotations, sh_degree, scale_multiplier, activation)
    gx, gy = (W + tiling.tile_w - 1) // tiling.tile_w, (H + tiling.tile_h - 1) // tiling.tile_h
    f32 = dict(dtype=torch.float32, device=dev)
    with torch.cuda.device(dev):
        fb = torch.zeros((3, H, W), **f32)
        T = torch.empty((H, W), **f32)
        rects = torch.empty((n, 2), dtype=torch.int32, device=dev)
        tiles = torch.empty((n,), dtype=torch.int32, device=dev)
        xy = torch.empty((n, 2), **f32)
        conic = torch.empty((n, 4), **f32)
        rgb = torch.empty((n, 3), **f32)
        clamped = torch.empty((n, 3), dtype=torch.bool, device=dev)
        depths = torch.empty((n,), **f32)
        offsets = torch.empty((n,), dtype=torch.int32, device=dev)
        ranges = torch.empty((gx * gy, 2), dtype=torch.int32, device=dev)
        geom = N.SplatGeometry(rects.data_ptr(), tiles.data_ptr(), xy.data_ptr(), conic.data_ptr(), rgb.data_ptr(),
                               clamped.data_ptr(), depths.data_ptr())
        cap = C.initial_capacity(dev, "splat", n)
        stats = N.Stats()
        stream = C.stream_handle(dev)
        for _attempt in range(3):
            point_list = torch.empty((max(cap, 1),), dtype=torch.int32, device=dev)
            bins = N.Binning(offsets.data_ptr(), point_list.data_ptr(), ranges.data_ptr(), cap)
            ws = C.workspace(dev, n, W, H, tiling, cap)
            rc = lib.dgmr_splat_forward(ctypes.byref(g), ctypes.byref(cam), ctypes.byref(tiling), ctypes.byref(geom),
                                        ctypes.byref(bins), fb.data_ptr(), T.data_ptr(), ws.data_ptr(), ws.numel(),
                                        ctypes.byref(stats), stream)
            if rc == N.DGMR_ERR_CAPACITY:
                cap = int(stats.n_pairs * 1.25) + 1024
                C.remember_capacity(dev, "splat", cap)
                continue
            N.check(rc, "dgmr_splat_forward")
            break
        else:
            raise N.DgmrError(N.DGMR_ERR_CAPACITY, "dgmr_splat_forward", "pair capacity kept growing")
        R = int(stats.n_pairs)
        C.remember_capacity(dev, "splat", cap)
    return [fb, rects, rgb, clamped, depths, xy, conic, tiles, offsets, ranges, point_list[:R], T,
            torch.tensor([R], dtype=torch.int64)]


def _backward(shs, weights, centroids, cov_scales, cov_rotations, view_matrix, proj_matrix, cam_position, background,
              render_width: int, render_height: int, tan_fovx: float, tan_fovy: float, sh_degree: int, grad_framebuffer, cache,
              tile=None, batch: int = 0, scale_multiplier: float = 1.0, activation: int = N.DGMR_ACT_NONE, cull=None):
    del cull  # the backward walks whatever lists the forward binned
    tile = tile or DEFAULT_TILING[0]
    batch = batch or DEFAULT_TILING[1]
    shs, weights, centroids, cov_scales, cov_rotations = C.prepare_inputs(shs, weights, centroids, cov_scales, cov_rotations, True)
    fb, rects, rgb, clamped, depths, xy, conic, tiles, offsets, ranges, point_list, T, n_pairs = cache
    dev = centroids.device
    n = centroids.shape[0]
    W, H = int(render_width), int(render_height)
    lib = N.load()
    tiling = C.tiling_struct(tile, batch)
    cam = C.camera_struct(view_matrix, proj_matrix, cam_position, background, W, H, tan_fovx, tan_fovy)
    g = C.gaussians_struct(shs, weights, centroids, cov_scales, cov_rotations, sh_degree, scale_multiplier, activation)
    grad_framebuffer = grad_framebuffer.to(torch.float32).contiguous()
    f32 = dict(dtype=torch.float32, device=dev)
    with torch.cuda.device(dev):
        screen = torch.empty((n, 9), **f32)
        g_means = torch.empty((n, 3), **f32)
        g_scales = torch.empty((n, 3), **f32)
        g_rot = torch.empty((n, 4), **f32)
        g_op = torch.empty(weights.shape, **f32)
        g_sh = torch.empty(shs.shape, **f32)
        geom = N.SplatGeometry(rects.data_ptr(), tiles.data_ptr(), xy.data_ptr(), conic.data_ptr(), rgb.data_ptr(),
                               clamped.data_ptr(), depths.data_ptr())
        bins = N.Binning(offsets.data_ptr(), point_list.data_ptr(), ranges.data_ptr(), point_list.numel())
        rc = lib.dgmr_splat_backward(ctypes.byref(g), ctypes.byref(cam), ctypes.byref(tiling), ctypes.byref(geom), ctypes.byref(bins),
                                     int(n_pairs.reshape(-1)[0]), fb.data_ptr(), T.data_ptr(), grad_framebuffer.data_ptr(),
                                     screen.data_ptr(), g_means.data_ptr(), g_scales.data_ptr(), g_rot.data_ptr(), g_op.data_ptr(),
                                     g_sh.data_ptr(), C.stream_handle(dev))
        N.check(rc, "dgmr_splat_backward")
    return [g_sh, g_op, g_means, g_scales, g_rot]


cpp_binding = types.SimpleNamespace(forward=_forward, backward=_backward)


class Splatter(torch.autograd.Function):
    """Same signature as the reference's VolMarcher (python_bindings/vol_marcher.py:19-62)."""

    @staticmethod
    def forward(ctx, shs, opacity, means3D, scales, rotations, world_view_transform, full_proj_transform, camera_center, bg_color,
                image_width: int, image_height: int, tanfovx: float, tanfovy: float, active_sh_degree: int):
        cache = cpp_binding.forward(shs, opacity, means3D, scales, rotations, world_view_transform, full_proj_transform, camera_center,
                                    bg_color, image_width, image_height, tanfovx, tanfovy, active_sh_degree)
        ctx.scalars = (int(image_width), int(image_height), float(tanfovx), float(tanfovy), int(active_sh_degree))
        ctx.save_for_backward(shs, opacity, means3D, scales, rotations, world_view_transform, full_proj_transform, camera_center,
                              bg_color, *cache)
        return cache[0]

    @staticmethod
    def backward(ctx, grad_output):
        if not grad_output.is_contiguous():
            grad_output = grad_output.contiguous()
        shs, opacity, means3D, scales, rotations, wv, fp, cc, bg, *cache = ctx.saved_tensors
        W, H, tx, ty, deg = ctx.scalars
        g_sh, g_op, g_means, g_scales, g_rot = cpp_binding.backward(shs, opacity, means3D, scales, rotations, wv, fp, cc, bg, W, H, tx, ty,
                                                                    deg, grad_output, cache)
        return g_sh, g_op, g_means, g_scales, g_rot, None, None, None, None, None, None, None, None, None


apply = Splatter.apply


class SplatterActivated(torch.autograd.Function):
    """Splatter.apply's argument list, with the Gaussians given the way a 3DGS
    trainer stores them: opacity logits, log-scales and unnormalised
    quaternions.  exp, q/|q| and sigmoid run inside the preprocess and their
    derivatives inside the preprocess backward (the same values as
    torch.exp / F.normalize / torch.sigmoid before Splatter.apply, without the
    extra kernels); gradients are w.r.t. the stored values."""

    @staticmethod
    def forward(ctx, shs, opacity_logits, means3D, log_scales, rotations, world_view_transform, full_proj_transform, camera_center,
                bg_color, image_width: int, image_height: int, tanfovx: float, tanfovy: float, active_sh_degree: int):
        cache = cpp_binding.forward(shs, opacity_logits, means3D, log_scales, rotations, world_view_transform, full_proj_transform,
                                    camera_center, bg_color, image_width, image_height, tanfovx, tanfovy, active_sh_degree,
                                    activation=N.DGMR_ACT_LOG_NORM_SIGMOID)
        ctx.scalars = (int(image_width), int(image_height), float(tanfovx), float(tanfovy), int(active_sh_degree))
        ctx.save_for_backward(shs, opacity_logits, means3D, log_scales, rotations, world_view_transform, full_proj_transform,
                              camera_center, bg_color, *cache)
        return cache[0]

    @staticmethod
    def backward(ctx, grad_output):
        if not grad_output.is_contiguous():
            grad_output = grad_output.contiguous()
        shs, logits, means3D, log_scales, rotations, wv, fp, cc, bg, *cache = ctx.saved_tensors
        W, H, tx, ty, deg = ctx.scalars
        g = cpp_binding.backward(shs, logits, means3D, log_scales, rotations, wv, fp, cc, bg, W, H, tx, ty, deg, grad_output, cache,
                                 activation=N.DGMR_ACT_LOG_NORM_SIGMOID)
        return (*g, None, None, None, None, None, None, None, None, None)
