"""ctypes wrapper of the host oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")
_lib = None

vp = ctypes.c_void_p
i32 = ctypes.c_int
i64 = ctypes.c_int64
f64 = ctypes.c_double


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        create = [i32, i32, i32, i64, vp, vp, vp, vp, vp, i32, f64, vp, vp, vp, f64, f64, i32, i32, i32, i32]
        for pre in ("orc_splat", "orc_march"):
            getattr(L, pre + "_create").restype = vp
            getattr(L, pre + "_create").argtypes = create
            getattr(L, pre + "_free").argtypes = [vp]
            getattr(L, pre + "_num_pairs").restype = i64
            getattr(L, pre + "_num_pairs").argtypes = [vp]
            getattr(L, pre + "_get_binning").argtypes = [vp, vp, vp, vp, vp]
        L.orc_splat_get_geometry.argtypes = [vp] * 9
        L.orc_splat_render.argtypes = [vp, vp, i32, i32, i32, i32, vp, vp, vp, vp]
        L.orc_splat_backward.argtypes = [vp, vp, i32, i32, i32, i32] + [vp] * 12
        L.orc_splat_tile_work.argtypes = [vp, vp, vp, vp]
        L.orc_march_get_geometry.argtypes = [vp] * 9
        L.orc_march_render.argtypes = [vp, vp, i32, i32, i32, i32, i32, i32, vp, i32, i32, i32, f64, vp, vp, vp, vp]
        L.orc_march_backward.argtypes = [vp, vp, i32, i32, i32, i32, i32, i32, vp, i32, i32, vp, vp, vp] + [vp] * 9
        L.orc_time_splat_fwd_bwd.restype = f64
        L.orc_time_splat_fwd_bwd.argtypes = [vp, vp, vp, vp, vp, i32, i32, i32, i32, i32]
        L.orc_portable_log_f.restype = ctypes.c_float
        L.orc_portable_log_f.argtypes = [ctypes.c_float]
        L.orc_portable_log_d.restype = f64
        L.orc_portable_log_d.argtypes = [f64]
        L.orc_get_rect_f.argtypes = [ctypes.c_float, ctypes.c_float, i32, i32, i32, i32, i32, i32, vp]
        L.orc_higher_msb.restype = ctypes.c_uint32
        L.orc_higher_msb.argtypes = [ctypes.c_uint32]
        L.orc_sections_run.restype = i32
        L.orc_sections_run.argtypes = [i32, i32, f64, i32, vp, vp, vp]
        L.orc_sections_combine.argtypes = [i32, vp, vp, vp]
        L.orc_sections_sample.restype = i32
        L.orc_sections_sample.argtypes = [i32, i32, f64, i32, vp, i32, vp]
        L.orc_sections_randomised_check.restype = i32
        L.orc_sections_randomised_check.argtypes = [i32]
        L.orc_compute_cov.argtypes = [vp] * 3
        L.orc_compute_cov_grad.argtypes = [vp] * 5
        L.orc_splat2d.argtypes = [i32, f64, vp, vp, vp, vp, vp, vp, f64, f64, i32, i32, f64, vp]
        L.orc_splat2d_grad.argtypes = [i32, f64, vp, vp, vp, vp, vp, vp, f64, f64, i32, i32, f64, vp, vp]
        L.orc_sh_to_colour.argtypes = [vp, i32, vp, vp, vp]
        L.orc_sh_to_colour_grad.argtypes = [vp, i32, vp, vp, vp, vp, vp]
        L.orc_intersect_ray.argtypes = [vp] * 5
        L.orc_intersect_ray_grad.argtypes = [vp] * 7
        L.orc_sample_gaussian.argtypes = [f64] + [vp] * 7
        L.orc_sample_gaussian_grad.argtypes = [f64] + [vp] * 11
        L.orc_integrate_bins.argtypes = [vp, vp, vp]
        L.orc_integrate_bins_grad.argtypes = [vp, vp, f64, vp, vp, f64, vp]
        f32 = ctypes.c_float
        L.orc_raster_render.argtypes = [i64, vp, vp, vp, vp, vp, i32, f64, vp, vp, vp, f64, f64, i32, i32, vp, i32, i32, f32, i32, i32, i32,
                                        i32, vp, vp]
        L.orc_raster_render64.argtypes = [i64, vp, vp, vp, vp, vp, i32, f64, vp, vp, vp, f64, f64, i32, i32, vp, i32, i32, f32, i32, i32,
                                          i32, i32, vp, vp, vp]
        L.orc_binsizer_run.restype = i32
        L.orc_binsizer_run.argtypes = [i32, f32, f32, i32, vp, vp, vp, vp, vp]
        L.orc_pl_create.argtypes = [vp, vp, vp, vp, f64, vp]
        L.orc_pl_sample.argtypes = [vp, f64, vp]
        L.orc_pl_sample_many.argtypes = [vp, i64, vp, vp]
        L.orc_pl_volume_integrate.argtypes = [i32, vp, i32, vp]
        _lib = L
    return _lib


def P(a: np.ndarray):
    return a.ctypes.data_as(vp)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


class _Handle:
    kind = ""

    def __init__(self, scene_arrays, cam, dbl=True, float_binning=True, tile=(16, 16), scale_multiplier=1.0):
        means, scales, rots, opac, sh16, deg = scene_arrays
        self.src = np.float64 if dbl else np.float32
        self.dt = np.float64 if dbl else np.float32
        self.n = means.shape[0]
        self.W, self.H = cam.width, cam.height
        self.tile = tile
        self._keep = [_c(means, self.src), _c(scales, self.src), _c(rots, self.src), _c(np.reshape(opac, -1), self.src),
                      _c(sh16, self.src), _c(cam.world_view_transform, self.src), _c(cam.full_proj_transform, self.src),
                      _c(cam.camera_center, self.src)]
        k = self._keep
        fn = getattr(lib(), f"orc_{self.kind}_create")
        self.h = fn(int(dbl), int(dbl), int(float_binning), self.n, P(k[0]), P(k[1]), P(k[2]), P(k[3]), P(k[4]), deg,
                    scale_multiplier, P(k[5]), P(k[6]), P(k[7]), float(np.float32(cam.tan_fovx)), float(np.float32(cam.tan_fovy)),
                    self.W, self.H, tile[0], tile[1])

    def __del__(self):
        if getattr(self, "h", None):
            getattr(lib(), f"orc_{self.kind}_free")(self.h)
            self.h = None

    @property
    def n_pairs(self) -> int:
        return int(getattr(lib(), f"orc_{self.kind}_num_pairs")(self.h))

    def binning(self):
        R = self.n_pairs
        gx = (self.W + self.tile[0] - 1) // self.tile[0]
        gy = (self.H + self.tile[1] - 1) // self.tile[1]
        offsets = np.zeros(self.n, np.uint32)
        keys = np.zeros(max(R, 1), np.uint64)
        pl = np.zeros(max(R, 1), np.uint32)
        ranges = np.zeros((gx * gy, 2), np.uint32)
        getattr(lib(), f"orc_{self.kind}_get_binning")(self.h, P(offsets), P(keys), P(pl), P(ranges))
        return dict(point_offsets=offsets, keys=keys[:R], point_list=pl[:R], ranges=ranges)


class SplatOracle(_Handle):
    """Tile splatter restatement.  dbl: evaluate in float64; float_binning:
    discrete decisions from the float run (the GPU's binning)."""
    kind = "splat"

    def geometry(self):
        n, dt = self.n, self.dt
        g = dict(rects=np.zeros((n, 2), np.int32), tiles=np.zeros(n, np.uint32), xy=np.zeros((n, 2), dt), conic=np.zeros((n, 3), dt),
                 weight2d=np.zeros(n, dt), rgb=np.zeros((n, 3), dt), clamped=np.zeros((n, 3), np.uint8), depth=np.zeros(n, dt))
        lib().orc_splat_get_geometry(self.h, P(g["rects"]), P(g["tiles"]), P(g["xy"]), P(g["conic"]), P(g["weight2d"]), P(g["rgb"]),
                                     P(g["clamped"]), P(g["depth"]))
        return g

    def render(self, bg, crop=None):
        x0, y0, x1, y1 = crop or (0, 0, self.W, self.H)
        fb = np.zeros((3, self.H, self.W), self.dt)
        T = np.ones((self.H, self.W), self.dt)
        amb = np.zeros((self.H, self.W), np.uint8)
        nev = np.zeros(1, np.uint64)
        bgc = _c(bg, self.dt)
        lib().orc_splat_render(self.h, P(bgc), x0, y0, x1, y1, P(fb), P(T), P(amb), P(nev))
        return fb, T, amb, int(nev[0])

    def backward(self, bg, fb, T, grad_fb, crop=None):
        x0, y0, x1, y1 = crop or (0, 0, self.W, self.H)
        n, dt = self.n, self.dt
        out = dict(means=np.zeros((n, 3), dt), scales=np.zeros((n, 3), dt), rotations=np.zeros((n, 4), dt), opacity=np.zeros(n, dt),
                   sh=np.zeros((n, 16, 3), dt), xy=np.zeros((n, 2), dt), conic=np.zeros((n, 3), dt), weight2d=np.zeros(n, dt),
                   rgb=np.zeros((n, 3), dt))
        bgc, fbc, Tc, gc = _c(bg, dt), _c(fb, dt), _c(T, dt), _c(grad_fb, dt)
        lib().orc_splat_backward(self.h, P(bgc), x0, y0, x1, y1, P(fbc), P(Tc), P(gc), P(out["means"]), P(out["scales"]),
                                 P(out["rotations"]), P(out["opacity"]), P(out["sh"]), P(out["xy"]), P(out["conic"]),
                                 P(out["weight2d"]), P(out["rgb"]))
        return out

    def tile_work(self, reachable=False):
        """Per tile: list entries traversed before the tile saturates, entries
        blended, and (optionally) entries reaching alpha >= 1/255 anywhere in the tile."""
        gx = (self.W + self.tile[0] - 1) // self.tile[0]
        gy = (self.H + self.tile[1] - 1) // self.tile[1]
        trav = np.zeros(gx * gy, np.uint32)
        blend = np.zeros(gx * gy, np.uint32)
        reach = np.zeros(gx * gy, np.uint32) if reachable else None
        lib().orc_splat_tile_work(self.h, P(trav), P(blend), P(reach) if reachable else None)
        return (trav, blend, reach) if reachable else (trav, blend)

    def time_fwd_bwd(self, bg, grad_fb, crop=None, threads: int = 1) -> float:
        """Wall seconds of preprocess + binning + blend + backward (float32 restatement) on `threads` host threads."""
        assert self.dt == np.float32
        x0, y0, x1, y1 = crop or (0, 0, self.W, self.H)
        fb = np.zeros((3, self.H, self.W), np.float32)
        T = np.ones((self.H, self.W), np.float32)
        bgc, gc = _c(bg, np.float32), _c(grad_fb, np.float32)
        return float(lib().orc_time_splat_fwd_bwd(self.h, P(bgc), P(gc), P(fb), P(T), x0, y0, x1, y1, int(threads)))


class MarchOracle(_Handle):
    """Volumetric marcher restatement."""
    kind = "march"

    def geometry(self):
        n, dt = self.n, self.dt
        g = dict(rects=np.zeros((n, 2), np.int32), tiles=np.zeros(n, np.uint32), xy=np.zeros((n, 2), dt), depth=np.zeros(n, dt),
                 rgb=np.zeros((n, 3), dt), clamped=np.zeros((n, 3), np.uint8), inv_cov=np.zeros((n, 6), dt), mass=np.zeros(n, dt))
        lib().orc_march_get_geometry(self.h, P(g["rects"]), P(g["tiles"]), P(g["xy"]), P(g["depth"]), P(g["rgb"]), P(g["clamped"]),
                                     P(g["inv_cov"]), P(g["mass"]))
        return g

    @staticmethod
    def _frozen(frozen):
        if not frozen:
            return 0, np.zeros(3)
        arr = np.ascontiguousarray(frozen, np.float64).reshape(-1, 3)
        return arr.shape[0], arr

    def render(self, bg, crop=None, max_large_steps=0, frozen=None, with_ambiguity=False, float_sampling=False, render_mode=0,
               debug_render_bin=-1, max_depth=20.0):
        """(fb [3,H,W], T [H,W], large steps per pixel [H,W]) and, with
        with_ambiguity, a [H,W] uint8 mask of pixels whose discrete decisions
        lie within float32 rounding (meaningful for dbl=True): bit 0 section
        comparisons, 1 step sizes, 2 border snaps, 3 mass thresholds, 4 mass
        in the bins, 5 transmittance.  float_sampling (dbl=True): the bin
        borders come from the float32 pipeline, everything on them is FP64.
        render_mode 1 (Bins) / 2 (Depth): the reference's debug modes
        (vol_marcher_forward.cu:443-484) with debug_render_bin and max_depth;
        Depth needs whole tiles in the crop."""
        x0, y0, x1, y1 = crop or (0, 0, self.W, self.H)
        fb = np.zeros((3, self.H, self.W), self.dt)
        T = np.ones((self.H, self.W), self.dt)
        nl = np.zeros((self.H, self.W), np.int32)
        amb = np.zeros((self.H, self.W), np.uint8)
        nf, fr = self._frozen(frozen)
        bgc = _c(bg, self.dt)
        lib().orc_march_render(self.h, P(bgc), x0, y0, x1, y1, max_large_steps, nf, P(fr), int(float_sampling), int(render_mode),
                               int(debug_render_bin), float(max_depth), P(fb), P(T), P(nl),
                               P(amb) if with_ambiguity else None)
        return (fb, T, nl, amb) if with_ambiguity else (fb, T, nl)

    def backward(self, bg, fb, T, grad_fb, crop=None, max_large_steps=0, frozen=None, float_sampling=False, sh_stored=16):
        """Gradients of sum(grad_fb * image) (pixels with zero grad_fb are
        skipped).  sh_stored < 16: the caller passed that many SH coefficients
        per Gaussian, and the preprocess backward reads the rest as ones, as
        the reference binding pads them (python_bindings/vol_marcher.cu:177)."""
        x0, y0, x1, y1 = crop or (0, 0, self.W, self.H)
        n, dt = self.n, self.dt
        out = dict(means=np.zeros((n, 3), dt), scales=np.zeros((n, 3), dt), rotations=np.zeros((n, 4), dt), opacity=np.zeros(n, dt),
                   sh=np.zeros((n, 16, 3), dt), r_mass=np.zeros(n, dt), r_rgb=np.zeros((n, 3), dt), r_means=np.zeros((n, 3), dt),
                   r_inv_cov=np.zeros((n, 6), dt))
        nf, fr = self._frozen(frozen)
        bgc, fbc, Tc, gc = _c(bg, dt), _c(fb, dt), _c(T, dt), _c(grad_fb, dt)
        lib().orc_march_backward(self.h, P(bgc), x0, y0, x1, y1, max_large_steps, nf, P(fr), int(float_sampling), int(sh_stored), P(fbc),
                                 P(Tc), P(gc), P(out["means"]), P(out["scales"]), P(out["rotations"]), P(out["opacity"]), P(out["sh"]),
                                 P(out["r_mass"]), P(out["r_rgb"]), P(out["r_means"]), P(out["r_inv_cov"]))
        return out


def scene_arrays(scene):
    """(means, scales, rotations, opacity, sh16, degree) of a scenes.Scene"""
    n, k, _ = scene.sh.shape
    sh16 = np.zeros((n, 16, 3), scene.sh.dtype)
    sh16[:, :k] = scene.sh
    return scene.means, scene.scales, scene.rotations, scene.opacity, sh16, scene.sh_degree


RASTER_FULL, RASTER_BINS, RASTER_DEPTH = 0, 1, 2


def raster_render(scene_arrays, cam, bg, mode=RASTER_FULL, debug_bin=-1, max_depth=20.0, crop=None, scale_multiplier=1.0, dbl=False):
    """Float restatement of vol_raster_forward (vol_raster_forward.cu:36-520):
    planar [3, H, W] image (pixels outside crop = (x0, y0, x1, y1) stay 0, and
    all stay 0 when nothing is binned, as the reference returns before
    rendering) and the pair count of the binning.

    dbl=True evaluates it in FP64 on the float32 pipeline's placement (the
    scene's float32 values, the float32 binning and per-pixel ray, everything
    else in double) and returns (image float64, pair count, ambiguity [H, W]):
    the RasterTieKind bits of the decisions that lay within float rounding of
    their thresholds (mass 1, centre order 2, cheapest merge 4, fullness 8,
    approximation branch 16)."""
    means, scales, rots, opac, sh16, deg = scene_arrays
    f = np.float32
    keep = [_c(means, f), _c(scales, f), _c(rots, f), _c(np.reshape(opac, -1), f), _c(sh16, f), _c(cam.world_view_transform, f),
            _c(cam.full_proj_transform, f), _c(cam.camera_center, f), _c(bg, np.float64 if dbl else f)]
    W, H = cam.width, cam.height
    x0, y0, x1, y1 = crop if crop is not None else (0, 0, W, H)
    n_pairs = ctypes.c_int64(0)
    if dbl:
        fb = np.zeros((3, H, W), np.float64)
        amb = np.zeros((H, W), np.uint8)
        lib().orc_raster_render64(means.shape[0], P(keep[0]), P(keep[1]), P(keep[2]), P(keep[3]), P(keep[4]), deg, scale_multiplier,
                                  P(keep[5]), P(keep[6]), P(keep[7]), float(np.float32(cam.tan_fovx)), float(np.float32(cam.tan_fovy)),
                                  W, H, P(keep[8]), mode, debug_bin, max_depth, x0, y0, x1, y1, P(fb), P(amb), ctypes.byref(n_pairs))
        return fb, int(n_pairs.value), amb
    fb = np.zeros((3, H, W), np.float32)
    lib().orc_raster_render(means.shape[0], P(keep[0]), P(keep[1]), P(keep[2]), P(keep[3]), P(keep[4]), deg, scale_multiplier,
                            P(keep[5]), P(keep[6]), P(keep[7]), float(np.float32(cam.tan_fovx)), float(np.float32(cam.tan_fovy)), W, H,
                            P(keep[8]), mode, debug_bin, max_depth, x0, y0, x1, y1, P(fb), ctypes.byref(n_pairs))
    return fb, int(n_pairs.value)


def binsizer_run(n_bins, transmission_threshold, relevance_sigma, adds):
    """math::RasterBinSizer in float: is_full() after each add, then finalise();
    returns (full_after, begin, end, full_final)."""
    a = np.ascontiguousarray(np.asarray(adds, np.float32).reshape(-1, 3))
    full_after = np.zeros(len(a), np.uint8)
    begin = np.zeros(n_bins, np.float32)
    end = np.zeros(n_bins, np.float32)
    full_final = np.zeros(1, np.uint8)
    rc = lib().orc_binsizer_run(n_bins, transmission_threshold, relevance_sigma, len(a), P(a), P(full_after), P(begin), P(end),
                                P(full_final))
    if rc != 0:
        raise ValueError(f"unsupported bin count {n_bins}")
    return full_after.astype(bool), begin, end, bool(full_final[0])


def pl_create(percent_left, total, eval_left, eval_right, t_right):
    """piecewise_linear::create_approximation (double): 21 values d0, k0, t1, t2, k2 (x4), t_right"""
    args = [np.ascontiguousarray(v, np.float64) for v in (percent_left, total, eval_left, eval_right)]
    out = np.zeros(21, np.float64)
    lib().orc_pl_create(P(args[0]), P(args[1]), P(args[2]), P(args[3]), float(t_right), P(out))
    return out


def pl_sample(params, t):
    p = np.ascontiguousarray(params, np.float64)
    out = np.zeros(4, np.float64)
    lib().orc_pl_sample(P(p), float(t), P(out))
    return out


def pl_sample_many(params, ts):
    p = np.ascontiguousarray(params, np.float64)
    t = np.ascontiguousarray(ts, np.float64)
    out = np.zeros((len(t), 4), np.float64)
    lib().orc_pl_sample_many(P(p), len(t), P(t), P(out))
    return out


def pl_volume_integrate(groups, n_steps):
    p = np.ascontiguousarray(np.stack(groups), np.float64)
    out = np.zeros(3, np.float64)
    lib().orc_pl_volume_integrate(len(groups), P(p), n_steps, P(out))
    return out
