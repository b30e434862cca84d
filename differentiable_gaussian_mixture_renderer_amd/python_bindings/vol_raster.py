"""Volumetric rasteriser (forward only), the reference's third renderer.

The reference exposes `dgmr::vol_raster_forward(VolRasterForwardData&)`
(vol_raster_forward.h:68, vol_raster_forward.cu:36-520) to C++ only, with
the fields of VolRasterForwardData (vol_raster_forward.h:31-57); this module
gives it the argument list of the reference's `VolMarcher`
(python_bindings/vol_marcher.py:19-65) plus that struct's debug fields:

    image = vol_raster_forward(shs, opacity, means3D, scales, rotations,
                               world_view_transform, full_proj_transform, camera_center,
                               bg_color, image_width, image_height, tanfovx, tanfovy,
                               active_sh_degree)

The image is [3, H, W] float32; with no Gaussian binned it stays zero, as the
reference returns before rendering.  There is no backward (the reference has
none).
"""
from __future__ import annotations

import ctypes

import torch

from .. import _common as C
from .. import _native as N

RENDER_FULL, RENDER_BINS, RENDER_DEPTH = 0, 1, 2


def vol_raster_forward(shs, weights, centroids, cov_scales, cov_rotations, view_matrix, proj_matrix, cam_position, background,
                       render_width: int, render_height: int, tan_fovx: float, tan_fovy: float, sh_degree: int,
                       render_mode: int = RENDER_FULL, debug_bin: int = -1, max_depth: float = 20.0, scale_multiplier: float = 1.0,
                       return_cache: bool = False):
    shs, weights, centroids, cov_scales, cov_rotations = C.prepare_inputs(shs, weights, centroids, cov_scales, cov_rotations, True)
    dev = centroids.device
    n = centroids.shape[0]
    W, H = int(render_width), int(render_height)
    lib = N.load()
    tiling = C.tiling_struct((16, 16), 64)
    cam = C.camera_struct(view_matrix, proj_matrix, cam_position, background, W, H, tan_fovx, tan_fovy)
    g = C.gaussians_struct(shs, weights, centroids, cov_scales, cov_rotations, sh_degree, scale_multiplier, N.DGMR_ACT_NONE)
    opt = N.RasterOptions(int(render_mode), int(debug_bin), float(max_depth))
    gx, gy = (W + 15) // 16, (H + 15) // 16
    f32 = dict(dtype=torch.float32, device=dev)
    with torch.cuda.device(dev):
        fb = torch.zeros((3, H, W), **f32)
        rects = torch.empty((n, 2), dtype=torch.int32, device=dev)
        tiles = torch.empty((n,), dtype=torch.int32, device=dev)
        xy = torch.empty((n, 2), **f32)
        conic = torch.empty((n, 4), **f32)
        rgb = torch.empty((n, 3), **f32)
        clamped = torch.empty((n, 3), dtype=torch.bool, device=dev)
        depths = torch.empty((n,), **f32)
        inv_cov = torch.zeros((n, 6), **f32)
        mass = torch.zeros((n,), **f32)
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
            fb.zero_()
            rc = lib.dgmr_raster_forward(ctypes.byref(g), ctypes.byref(cam), ctypes.byref(opt), ctypes.byref(geom), ctypes.byref(bins),
                                         inv_cov.data_ptr(), mass.data_ptr(), fb.data_ptr(), ws.data_ptr(), ws.numel(),
                                         ctypes.byref(stats), stream)
            if rc == N.DGMR_ERR_CAPACITY:
                cap = int(stats.n_pairs * 1.25) + 1024
                C.remember_capacity(dev, "splat", cap)
                continue
            N.check(rc, "dgmr_raster_forward")
            break
        else:
            raise N.DgmrError(N.DGMR_ERR_CAPACITY, "dgmr_raster_forward", "pair capacity kept growing")
        R = int(stats.n_pairs)
        C.remember_capacity(dev, "splat", cap)
    if return_cache:
        return fb, dict(rects=rects, tiles_touched=tiles, points_xy_image=xy, rgb=rgb, depths=depths, inverse_filtered_cov3d=inv_cov,
                        filtered_masses=mass, point_offsets=offsets, i_ranges=ranges, b_point_list=point_list[:R], n_pairs=R)
    return fb
