"""The blends' tile-reach test (csrc/tile_reach.h, built for the host by
tests/cpp/build.py): a Gaussian the test drops must reach alpha >= 1/255 at
no integer pixel of the tile, with alpha evaluated in float32 the way the
blend evaluates it (exp2 of the log2(e)-scaled quadratic, min 0.99).  Checked
by brute force over every pixel of random 16x16, 8x8 and 32x8 boxes around
random Gaussians, including near-singular conics and opacities close to the
1/255 floor; the test must also drop most far boxes (it is useful)."""
import ctypes
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "cpp", "_build", "libreach_check.so")
LOG2E = np.float32(1.4426950408889634)


def _lib():
    if not os.path.exists(LIB):
        import subprocess
        import sys
        subprocess.run([sys.executable, os.path.join(HERE, "cpp", "build.py")], check=True, stdout=subprocess.DEVNULL)
    L = ctypes.CDLL(LIB)
    L.reach_box_batch.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    return L


def _fma(a, b, c):
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(np.float32)


def _alpha_any(g, box):
    """float32 alpha of every integer pixel of each box, as the blend computes it"""
    out = np.zeros(len(g), bool)
    for i, (p, bx) in enumerate(zip(g, box)):
        x, y, a, b, c, o = [np.float32(v) for v in p]
        xs = np.arange(bx[0], bx[1] + 1, dtype=np.float32)
        ys = np.arange(bx[2], bx[3] + 1, dtype=np.float32)
        dx = (xs[None, :] - x).astype(np.float32)
        dy = (ys[:, None] - y).astype(np.float32)
        dx, dy = np.broadcast_arrays(dx, dy)
        A = np.float32(-0.5) * LOG2E * a
        B = -LOG2E * b
        C = np.float32(-0.5) * LOG2E * c
        adx2 = (A * dx).astype(np.float32) * dx
        bdx = (B * dx).astype(np.float32)
        e = _fma(_fma(np.full_like(dy, C), dy, bdx), dy, adx2.astype(np.float32))
        alpha = np.minimum(np.float32(0.99), (o * np.exp2(e.astype(np.float64))).astype(np.float32))
        out[i] = bool((alpha >= np.float32(1.0 / 255.0)).any())
    return out


@pytest.mark.parametrize("tile", [(16, 16), (8, 8), (32, 8)])
def test_dropped_gaussians_reach_no_pixel(tile):
    L = _lib()
    rng = np.random.default_rng(tile[0] * 100 + tile[1])
    n = 4000
    # 2D covariances from random scales and angles, some nearly singular
    s1 = np.exp(rng.uniform(np.log(0.3), np.log(40.0), n))
    s2 = s1 * np.exp(rng.uniform(np.log(0.02), 0.0, n))
    th = rng.uniform(0, np.pi, n)
    ct, st = np.cos(th), np.sin(th)
    cov = np.stack([ct * ct * s1 ** 2 + st * st * s2 ** 2, ct * st * (s1 ** 2 - s2 ** 2), st * st * s1 ** 2 + ct * ct * s2 ** 2], 1)
    det = cov[:, 0] * cov[:, 2] - cov[:, 1] ** 2
    conic = np.stack([cov[:, 2] / det, -cov[:, 1] / det, cov[:, 0] / det], 1)
    o = np.where(rng.uniform(size=n) < 0.2, rng.uniform(1.0 / 255, 0.02, n), rng.uniform(0.02, 0.99, n))
    # boxes around the centre, at up to ~4 sigma of the larger axis
    cx, cy = rng.uniform(0, 1000, n), rng.uniform(0, 1000, n)
    off = rng.normal(size=(n, 2)) * s1[:, None] * 2.5
    bx0 = np.floor(cx + off[:, 0]) - tile[0] // 2
    by0 = np.floor(cy + off[:, 1]) - tile[1] // 2
    g = np.stack([cx, cy, conic[:, 0], conic[:, 1], conic[:, 2], o], 1).astype(np.float32)
    box = np.stack([bx0, bx0 + tile[0] - 1, by0, by0 + tile[1] - 1], 1).astype(np.float32)
    got = np.zeros(n, np.uint8)
    L.reach_box_batch(n, g.ctypes.data, box.ctypes.data, got.ctypes.data)
    truth = _alpha_any(g, box)
    missed = np.nonzero(truth & (got == 0))[0]
    assert missed.size == 0, missed[:10]
    dropped = (got == 0).sum()
    print(f"reaching {truth.sum()}, kept {(got == 1).sum()}, dropped {dropped} of {n}")
    assert dropped >= 0.1 * n
    # kept but unreachable: only a thin band at the 1/255 level
    assert (got.astype(bool) & ~truth).sum() <= 0.4 * (~truth).sum()


def test_degenerate_conics_always_reach():
    L = _lib()
    g = np.array([[0, 0, 1, 2, 1, 0.5], [0, 0, -1, 0, 1, 0.5], [0, 0, np.nan, 0, 1, 0.5], [0, 0, 1, 0, 0, 0.5]], np.float32)
    box = np.tile(np.array([[100, 115, 100, 115]], np.float32), (4, 1))
    got = np.zeros(4, np.uint8)
    L.reach_box_batch(4, g.ctypes.data, box.ctypes.data, got.ctypes.data)
    assert got.all()
