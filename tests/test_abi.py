"""The C-ABI library loads and exports every entry point include/dgmr.h
declares; argument validation reports errors through status codes without a
GPU (no kernel is launched here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dgmr.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dgmr_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from differentiable_gaussian_mixture_renderer_amd import _native as N
    lib = ctypes.CDLL(N.LIB_PATH)
    missing = [f for f in declared_functions() if not hasattr(lib, f)]
    assert not missing, missing
    # and the Python view knows every one of them
    assert set(declared_functions()) <= set(N.SIGNATURES)


def test_version_and_error_strings():
    from differentiable_gaussian_mixture_renderer_amd import _native as N
    lib = N.load()
    assert b"gfx950" in lib.dgmr_version()
    assert lib.dgmr_last_error() is not None


def test_argument_validation_without_gpu():
    from differentiable_gaussian_mixture_renderer_amd import _native as N
    lib = N.load()
    t = N.Tiling(16, 16, 256)
    size = ctypes.c_size_t(0)
    assert lib.dgmr_workspace_size(1000, 64, 64, ctypes.byref(t), 10000, ctypes.byref(size)) == N.DGMR_OK
    assert size.value > 0
    bad = N.Tiling(12, 12, 256)
    assert lib.dgmr_workspace_size(1000, 64, 64, ctypes.byref(bad), 10000, ctypes.byref(size)) == N.DGMR_ERR_ARGUMENT
    assert b"bad argument" in lib.dgmr_last_error()
    for batch, ok in ((64, True), (128, True), (256, True), (32, False), (512, False)):
        tb = N.Tiling(16, 16, batch)
        rc = lib.dgmr_workspace_size(1000, 64, 64, ctypes.byref(tb), 10000, ctypes.byref(size))
        assert (rc == N.DGMR_OK) == ok, batch
    # image sides beyond 32768 are rejected (16-bit tile grid)
    assert lib.dgmr_workspace_size(1000, 40000, 64, ctypes.byref(t), 10000, ctypes.byref(size)) == N.DGMR_ERR_ARGUMENT
    # null pointers are rejected before any device work
    rc = lib.dgmr_splat_forward(None, None, ctypes.byref(t), None, None, None, None, None, 0, None, None)
    assert rc == N.DGMR_ERR_ARGUMENT
    g = N.Gaussians()
    g.n = 10
    g.dtype = N.DGMR_F64
    cam = N.Camera()
    geom = N.SplatGeometry()
    bins = N.Binning()
    rc = lib.dgmr_splat_forward(ctypes.byref(g), ctypes.byref(cam), ctypes.byref(t), ctypes.byref(geom), ctypes.byref(bins), 1, 1, 1, 0,
                                None, None)
    assert rc == N.DGMR_ERR_ARGUMENT and b"float32" in lib.dgmr_last_error()
    opt = N.MarchOptions()
    g.dtype = 7
    rc = lib.dgmr_march_forward(ctypes.byref(g), ctypes.byref(cam), ctypes.byref(t), ctypes.byref(opt), ctypes.byref(N.MarchGeometry()),
                                ctypes.byref(bins), 1, 1, 1, 0, None, None)
    assert rc == N.DGMR_ERR_ARGUMENT


def test_missing_library_fails_loudly(tmp_path):
    from differentiable_gaussian_mixture_renderer_amd import _native as N
    with pytest.raises(ImportError):
        _load_fresh(N, str(tmp_path / "nope.so"))


def _load_fresh(N, path):
    saved = N._lib
    N._lib = None
    try:
        N.load(path)
    finally:
        N._lib = saved


def test_tight_culling_flag_validation_and_workspace():
    """dgmr_tiling.cull takes 0 or 1; with 1 the workspace holds 32-byte
    reach records (preprocess and depth-ordered copies) instead of 8-byte
    rectangles; the marcher refuses it."""
    from differentiable_gaussian_mixture_renderer_amd import _native as N
    lib = N.load()
    n, cap = 100_000, 1_000_000
    sizes = {}
    for cull in (0, 1):
        t = N.Tiling(16, 16, 64, cull)
        size = ctypes.c_size_t(0)
        assert lib.dgmr_workspace_size(n, 256, 256, ctypes.byref(t), cap, ctypes.byref(size)) == N.DGMR_OK
        sizes[cull] = size.value
    assert sizes[1] - sizes[0] >= 2 * 24 * n
    bad = N.Tiling(16, 16, 64, 2)
    size = ctypes.c_size_t(0)
    assert lib.dgmr_workspace_size(n, 256, 256, ctypes.byref(bad), cap, ctypes.byref(size)) == N.DGMR_ERR_ARGUMENT
    g = N.Gaussians()
    g.n = 10
    g.dtype = N.DGMR_F32
    g.sh_coeffs, g.sh_degree = 1, 0
    cam = N.Camera()
    cam.width = cam.height = 64
    opt = N.MarchOptions()
    geom = N.MarchGeometry()
    bins = N.Binning()
    tight = N.Tiling(16, 16, 64, 1)
    buf = ctypes.create_string_buffer(64)
    rc = lib.dgmr_march_forward(ctypes.byref(g), ctypes.byref(cam), ctypes.byref(tight), ctypes.byref(opt), ctypes.byref(geom),
                                ctypes.byref(bins), buf, buf, buf, 64, None, None)
    assert rc == N.DGMR_ERR_ARGUMENT
    assert b"bad tiling" in lib.dgmr_last_error()


def test_slot_order_validation_and_batch_workspace():
    """dgmr_tiling.slot_order takes 0 (Morton slots) or 1 (id order); the
    batched workspace holds the order, its sort buffers and the parameters
    copied into slot order in either case, so one size serves both."""
    from differentiable_gaussian_mixture_renderer_amd import _native as N
    lib = N.load()
    views = (N.BatchView * 2)()
    for v in views:
        v.cam.width, v.cam.height = 320, 240
    n, cap = 50_000, 1_000_000
    sizes = {}
    for order in (0, 1):
        t = N.Tiling(16, 16, 64, 0, order)
        size = ctypes.c_size_t(0)
        assert lib.dgmr_batch_workspace_size(n, views, 2, ctypes.byref(t), cap, ctypes.byref(size)) == N.DGMR_OK
        sizes[order] = size.value
    assert sizes[0] == sizes[1]
    # order, three sort buffers and centre, scales, rotation, opacity in slot order
    t_small = N.Tiling(16, 16, 64, 0, 0)
    small = ctypes.c_size_t(0)
    assert lib.dgmr_batch_workspace_size(0, views, 2, ctypes.byref(t_small), cap, ctypes.byref(small)) == N.DGMR_OK
    assert sizes[0] - small.value >= n * (4 * 4 + 11 * 4)
    for bad_order in (2, -1):
        bad = N.Tiling(16, 16, 64, 0, bad_order)
        size = ctypes.c_size_t(0)
        assert lib.dgmr_batch_workspace_size(n, views, 2, ctypes.byref(bad), cap, ctypes.byref(size)) == N.DGMR_ERR_ARGUMENT
        assert lib.dgmr_workspace_size(n, 256, 256, ctypes.byref(bad), cap, ctypes.byref(size)) == N.DGMR_ERR_ARGUMENT


def test_l1_many_argument_checks_without_gpu():
    """losses.l1_grad_many rejects mismatched shapes and dtypes before any
    launch; the C entry point rejects null and misaligned arrays."""
    import torch
    from differentiable_gaussian_mixture_renderer_amd import _native as N
    from differentiable_gaussian_mixture_renderer_amd import losses
    with pytest.raises(ValueError):
        losses.l1_grad_many([torch.zeros(3, 4, 4)], [torch.zeros(3, 4, 5)])
    with pytest.raises(ValueError):
        losses.l1_grad_many([torch.zeros(3, 4, 4, dtype=torch.float64)], [torch.zeros(3, 4, 4, dtype=torch.float64)])
    assert losses.l1_grad_many([], []) == []
    lib = N.load()
    items = (N.L1Item * 1)()
    assert lib.dgmr_l1_loss_backward_many(items, 1, None) == N.DGMR_ERR_ARGUMENT  # null arrays
    items[0].image = items[0].target = items[0].grad = items[0].grad_loss = 4100  # not 16-byte aligned
    items[0].n = 8
    assert lib.dgmr_l1_loss_backward_many(items, 1, None) == N.DGMR_ERR_ARGUMENT
    assert lib.dgmr_l1_loss_backward_many(items, 0, None) == N.DGMR_OK  # nothing to do


def test_ctypes_descriptors_match_the_header(tmp_path):
    """Sizes and field offsets of the ctypes mirrors equal what a C compiler
    lays out from include/dgmr.h (gcc on a probe that prints offsetof)."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("no gcc")
    from differentiable_gaussian_mixture_renderer_amd import _native as N
    mirrors = {"dgmr_gaussians": N.Gaussians, "dgmr_camera": N.Camera, "dgmr_tiling": N.Tiling, "dgmr_stats": N.Stats,
               "dgmr_splat_geometry": N.SplatGeometry, "dgmr_binning": N.Binning, "dgmr_march_geometry": N.MarchGeometry,
               "dgmr_march_options": N.MarchOptions, "dgmr_batch_view": N.BatchView, "dgmr_adam_tensor": N.AdamTensor,
               "dgmr_stage_time": N.StageTime, "dgmr_raster_options": N.RasterOptions, "dgmr_l1_item": N.L1Item}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "dgmr.h"', 'int main(void) {']
    for cname, cls in mirrors.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ['return 0;', '}']
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict((" ".join(l.split()[:2]), int(l.split()[2])) for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                                                      text=True).stdout.splitlines())
    for cname, cls in mirrors.items():
        assert got[f"{cname} size"] == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[f"{cname} {fname}"] == getattr(cls, fname).offset, (cname, fname)
