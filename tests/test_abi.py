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
