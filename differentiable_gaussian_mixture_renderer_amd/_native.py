"""ctypes view of the C ABI in include/dgmr.h (libdgmr_hip.so).

The library is the product: every render call goes through it.  Loading fails
loudly when the library is missing; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libdgmr_hip.so")

DGMR_OK = 0
DGMR_ERR_ARGUMENT = 1
DGMR_ERR_CAPACITY = 2
DGMR_ERR_WORKSPACE = 3
DGMR_ERR_DEVICE = 4
DGMR_F32 = 0
DGMR_F64 = 1
DGMR_ACT_NONE = 0
DGMR_ACT_LOG_NORM_SIGMOID = 1

c_void_p = ctypes.c_void_p


class Gaussians(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int64),
        ("dtype", ctypes.c_int32),
        ("means", c_void_p),
        ("scales", c_void_p),
        ("rotations", c_void_p),
        ("opacity", c_void_p),
        ("sh", c_void_p),
        ("sh_coeffs", ctypes.c_int32),
        ("sh_degree", ctypes.c_int32),
        ("scale_multiplier", ctypes.c_double),
        ("activation", ctypes.c_int32),
    ]


class Camera(ctypes.Structure):
    _fields_ = [
        ("view", ctypes.c_double * 16),
        ("proj", ctypes.c_double * 16),
        ("position", ctypes.c_double * 3),
        ("background", ctypes.c_double * 3),
        ("tan_fovx", ctypes.c_double),
        ("tan_fovy", ctypes.c_double),
        ("width", ctypes.c_int32),
        ("height", ctypes.c_int32),
    ]


class Tiling(ctypes.Structure):
    _fields_ = [("tile_w", ctypes.c_int32), ("tile_h", ctypes.c_int32), ("batch", ctypes.c_int32), ("cull", ctypes.c_int32),
                ("slot_order", ctypes.c_int32)]


class Stats(ctypes.Structure):
    _fields_ = [("n_pairs", ctypes.c_int64), ("n_visible", ctypes.c_int64)]


class SplatGeometry(ctypes.Structure):
    _fields_ = [
        ("rects", c_void_p),
        ("tiles_touched", c_void_p),
        ("xy", c_void_p),
        ("conic_opacity", c_void_p),
        ("rgb", c_void_p),
        ("rgb_clamped", c_void_p),
        ("depths", c_void_p),
    ]


class Binning(ctypes.Structure):
    _fields_ = [
        ("point_offsets", c_void_p),
        ("point_list", c_void_p),
        ("ranges", c_void_p),
        ("pair_capacity", ctypes.c_int64),
    ]


class MarchGeometry(ctypes.Structure):
    _fields_ = [
        ("rects", c_void_p),
        ("rgb", c_void_p),
        ("rgb_clamped", c_void_p),
        ("depths", c_void_p),
        ("xy", c_void_p),
        ("inverse_filtered_cov3d", c_void_p),
        ("filtered_masses", c_void_p),
        ("tiles_touched", c_void_p),
    ]


class AdamTensor(ctypes.Structure):
    _fields_ = [
        ("param", c_void_p),
        ("grad", c_void_p),
        ("exp_avg", c_void_p),
        ("exp_avg_sq", c_void_p),
        ("numel", ctypes.c_int64),
        ("lr", ctypes.c_float),
    ]


class MarchOptions(ctypes.Structure):
    _fields_ = [
        ("max_large_steps", ctypes.c_int32),
        ("n_frozen", ctypes.c_int32),
        ("frozen", (ctypes.c_double * 3) * 8),
        ("bins_per_chunk", ctypes.c_int32),
        ("render_mode", ctypes.c_int32),
        ("debug_render_bin", ctypes.c_int32),
        ("max_depth", ctypes.c_double),
        ("replay", c_void_p),
        ("replay_bytes", ctypes.c_size_t),
        ("replay_steps", ctypes.c_int32),
    ]


class BatchView(ctypes.Structure):
    _fields_ = [
        ("cam", Camera),
        ("framebuffer", c_void_p),
        ("final_transmittance", c_void_p),
        ("grad_framebuffer", c_void_p),
        ("l1_target", c_void_p),
        ("l1_scale", ctypes.c_float),
        ("reserved", ctypes.c_int32),
    ]


class StageTime(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("total_ms", ctypes.c_double), ("launches", ctypes.c_int64), ("bytes", ctypes.c_double)]


class RasterOptions(ctypes.Structure):
    _fields_ = [("render_mode", ctypes.c_int32), ("debug_bin", ctypes.c_int32), ("max_depth", ctypes.c_float)]


class L1Item(ctypes.Structure):
    _fields_ = [("image", c_void_p), ("target", c_void_p), ("grad", c_void_p), ("grad_loss", c_void_p), ("n", ctypes.c_int64),
                ("scale", ctypes.c_float), ("reserved", ctypes.c_int32)]


_P = ctypes.POINTER
# exported symbol -> (restype, argtypes); this list mirrors include/dgmr.h
SIGNATURES = {
    "dgmr_workspace_size": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, _P(Tiling), ctypes.c_int64, _P(ctypes.c_size_t)]),
    "dgmr_splat_forward": (ctypes.c_int, [_P(Gaussians), _P(Camera), _P(Tiling), _P(SplatGeometry), _P(Binning), c_void_p, c_void_p,
                                          c_void_p, ctypes.c_size_t, _P(Stats), c_void_p]),
    "dgmr_splat_backward": (ctypes.c_int, [_P(Gaussians), _P(Camera), _P(Tiling), _P(SplatGeometry), _P(Binning), ctypes.c_int64,
                                           c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                           c_void_p]),
    "dgmr_march_forward": (ctypes.c_int, [_P(Gaussians), _P(Camera), _P(Tiling), _P(MarchOptions), _P(MarchGeometry), _P(Binning),
                                          c_void_p, c_void_p, c_void_p, ctypes.c_size_t, _P(Stats), c_void_p]),
    "dgmr_march_replay_size": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, _P(Tiling), ctypes.c_int32, ctypes.c_int32,
                                              _P(ctypes.c_size_t)]),
    "dgmr_march_backward": (ctypes.c_int, [_P(Gaussians), _P(Camera), _P(Tiling), _P(MarchOptions), _P(MarchGeometry), _P(Binning),
                                           ctypes.c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                           c_void_p, c_void_p, c_void_p]),
    "dgmr_batch_workspace_size": (ctypes.c_int, [ctypes.c_int64, _P(BatchView), ctypes.c_int32, _P(Tiling), ctypes.c_int64,
                                                 _P(ctypes.c_size_t)]),
    "dgmr_splat_batch_forward": (ctypes.c_int, [_P(Gaussians), _P(BatchView), ctypes.c_int32, _P(Tiling), c_void_p, ctypes.c_size_t,
                                                ctypes.c_int64, _P(Stats), c_void_p]),
    "dgmr_splat_batch_backward": (ctypes.c_int, [_P(Gaussians), _P(BatchView), ctypes.c_int32, _P(Tiling), c_void_p, ctypes.c_size_t,
                                                 ctypes.c_int64, ctypes.c_int64, _P(ctypes.c_int32), ctypes.c_int32, c_void_p, c_void_p,
                                                 c_void_p, c_void_p, c_void_p, c_void_p]),
    "dgmr_profile_enable": (ctypes.c_int, [ctypes.c_int]),
    "dgmr_profile_read": (ctypes.c_int, [_P(StageTime), ctypes.c_int32, _P(ctypes.c_int32)]),
    "dgmr_last_error": (ctypes.c_char_p, []),
    "dgmr_version": (ctypes.c_char_p, []),
    "dgmr_adam_step": (ctypes.c_int, [_P(AdamTensor), ctypes.c_int32, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_int64, c_void_p]),
    "dgmr_raster_forward": (ctypes.c_int, [_P(Gaussians), _P(Camera), _P(RasterOptions), _P(SplatGeometry), _P(Binning), c_void_p,
                                           c_void_p, c_void_p, c_void_p, ctypes.c_size_t, _P(Stats), c_void_p]),
    "dgmr_l1_partials_size": (ctypes.c_int, [ctypes.c_int64, _P(ctypes.c_int32)]),
    "dgmr_l1_loss_forward": (ctypes.c_int, [c_void_p, c_void_p, ctypes.c_int64, ctypes.c_float, c_void_p, c_void_p]),
    "dgmr_l1_loss_backward": (ctypes.c_int, [c_void_p, c_void_p, ctypes.c_int64, ctypes.c_float, c_void_p, c_void_p, c_void_p]),
    "dgmr_l1_loss_backward_many": (ctypes.c_int, [_P(L1Item), ctypes.c_int32, c_void_p]),
    "dgmr_scan_workspace_size": (ctypes.c_int, [ctypes.c_int64, _P(ctypes.c_size_t)]),
    "dgmr_scan_u32": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, ctypes.c_int64, ctypes.c_int32, c_void_p, ctypes.c_size_t, c_void_p]),
    "dgmr_sort_workspace_size": (ctypes.c_int, [ctypes.c_int64, _P(ctypes.c_size_t)]),
    "dgmr_sort_pairs": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, c_void_p,
                                       ctypes.c_size_t, c_void_p]),
    "dgmr_sort_pairs_noalloc": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                               c_void_p]),
    "dgmr_selftest_wave_sum": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
}

_lib = None
_lock = threading.Lock()


class DgmrError(RuntimeError):
    def __init__(self, status: int, where: str, message: str):
        super().__init__(f"{where} failed with status {status}: {message}")
        self.status = status


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Loads the library once; raises if it has not been built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise ImportError(f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                              "or `python differentiable_gaussian_mixture_renderer_amd/csrc/build.py`")
        lib = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def exported_symbols(path: str = LIB_PATH) -> list:
    lib = ctypes.CDLL(path)
    return [name for name in SIGNATURES if hasattr(lib, name)]


def check(status: int, where: str) -> None:
    if status != DGMR_OK:
        msg = load().dgmr_last_error()
        raise DgmrError(status, where, msg.decode() if msg else "")


def ptr(t) -> int:
    """Device address of a torch tensor (0 for None)."""
    return 0 if t is None else t.data_ptr()


def camera_struct(view, proj, position, background, tan_fovx, tan_fovy, width, height) -> Camera:
    """view/proj: 16 values in the reference's tensor layout; the rest plain numbers."""
    c = Camera()
    for i, v in enumerate(_flat(view, 16)):
        c.view[i] = v
    for i, v in enumerate(_flat(proj, 16)):
        c.proj[i] = v
    for i, v in enumerate(_flat(position, 3)):
        c.position[i] = v
    for i, v in enumerate(_flat(background, 3)):
        c.background[i] = v
    c.tan_fovx = float(tan_fovx)
    c.tan_fovy = float(tan_fovy)
    c.width = int(width)
    c.height = int(height)
    return c


def _flat(x, n):
    try:
        import torch
        if isinstance(x, torch.Tensor):
            x = x.detach().to("cpu", torch.float64).reshape(-1).tolist()
    except ImportError:
        pass
    vals = [float(v) for v in (x.reshape(-1).tolist() if hasattr(x, "reshape") else x)]
    if len(vals) < n:
        raise ValueError(f"expected {n} values, got {len(vals)}")
    return vals[:n]
