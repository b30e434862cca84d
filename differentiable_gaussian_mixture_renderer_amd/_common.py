"""Shared host plumbing of the Python bindings: argument marshalling into the
C-ABI descriptors, torch-owned workspaces and pair-buffer capacity."""
from __future__ import annotations

import ctypes
import threading

import torch

from . import _native as N

_state_lock = threading.Lock()
_capacity = {}   # (device index, kind) -> pair capacity that fitted last time
_ppp = {}        # device index -> largest (tile, Gaussian) pairs per pixel seen


def stream_handle(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def dtype_code(t: torch.Tensor) -> int:
    if t.dtype == torch.float32:
        return N.DGMR_F32
    if t.dtype == torch.float64:
        return N.DGMR_F64
    raise TypeError("All input vectors must be of the same type (float32 or float64)")


def prepare_inputs(shs, opacity, means3D, scales, rotations, float32_only: bool):
    """Validates devices/dtypes/shapes and returns contiguous views."""
    tensors = [shs, opacity, means3D, scales, rotations]
    dev = means3D.device
    if dev.type != "cuda":
        raise ValueError("Gaussian parameters must live on the GPU")
    dt = means3D.dtype
    for t in tensors:
        if t.device != dev:
            raise ValueError("all Gaussian parameters must be on the same device")
        if t.dtype != dt:
            raise TypeError("All input vectors must be of the same type (float32 or float64)")
    if float32_only and dt != torch.float32:
        raise TypeError("the splatter computes in float32")
    if dt not in (torch.float32, torch.float64):
        raise TypeError("All input vectors must be of the same type (float32 or float64)")
    n = means3D.shape[0]
    if means3D.shape != (n, 3) or scales.shape != (n, 3) or rotations.shape != (n, 4):
        raise ValueError("expected means [N,3], scales [N,3], rotations [N,4]")
    if opacity.numel() != n:
        raise ValueError("expected opacity [N,1]")
    if shs.dim() != 3 or shs.shape[0] != n or shs.shape[2] != 3 or not (1 <= shs.shape[1] <= 16):
        raise ValueError("expected shs [N,K,3] with 1 <= K <= 16")
    return [t.contiguous() for t in tensors]


def gaussians_struct(shs, opacity, means3D, scales, rotations, sh_degree: int, scale_multiplier: float = 1.0,
                     activation: int = N.DGMR_ACT_NONE) -> N.Gaussians:
    g = N.Gaussians()
    g.n = means3D.shape[0]
    g.dtype = dtype_code(means3D)
    g.means = means3D.data_ptr()
    g.scales = scales.data_ptr()
    g.rotations = rotations.data_ptr()
    g.opacity = opacity.data_ptr()
    g.sh = shs.data_ptr()
    g.sh_coeffs = shs.shape[1]
    g.sh_degree = int(sh_degree)
    g.scale_multiplier = float(scale_multiplier)
    g.activation = int(activation)
    return g


def camera_struct(world_view_transform, full_proj_transform, camera_center, bg_color, image_width, image_height, tanfovx, tanfovy):
    # the reference copies these to the host (python_bindings/vol_marcher.cu:16-40)
    return N.camera_struct(world_view_transform, full_proj_transform, camera_center, bg_color, tanfovx, tanfovy,
                           image_width, image_height)


def tiling_struct(tile=(16, 16), batch=64, cull: bool = False) -> N.Tiling:
    t = N.Tiling()
    t.tile_w, t.tile_h = int(tile[0]), int(tile[1])
    t.batch = int(batch)
    t.cull = int(bool(cull))
    return t


def initial_capacity(device: torch.device, kind: str, n: int) -> int:
    with _state_lock:
        return _capacity.get((device.index, kind), max(1 << 16, 4 * n))


def remember_capacity(device: torch.device, kind: str, cap: int) -> None:
    with _state_lock:
        key = (device.index, kind)
        _capacity[key] = max(_capacity.get(key, 0), int(cap))


def pairs_per_pixel(device: torch.device) -> float:
    with _state_lock:
        return _ppp.get(device.index, 16.0)


def remember_pairs_per_pixel(device: torch.device, ppp: float) -> None:
    with _state_lock:
        _ppp[device.index] = max(_ppp.get(device.index, 0.0), float(ppp))


def workspace(device: torch.device, n: int, width: int, height: int, tiling: N.Tiling, cap: int) -> torch.Tensor:
    size = ctypes.c_size_t(0)
    N.check(N.load().dgmr_workspace_size(n, width, height, ctypes.byref(tiling), cap, ctypes.byref(size)), "dgmr_workspace_size")
    return torch.empty(max(int(size.value), 256), dtype=torch.uint8, device=device)
