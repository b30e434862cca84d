"""The rasteriser's pieces in the oracle (CPU), pinned by the reference's own
known-answer tests and by what the mathematics fixes:

* piecewise_linear: unittests/piecewise_linear.cpp:28-143 (sample values,
  fitted end points and masses, the volume integral against a 1M-step
  numerical integration);
* RasterBinSizer: unittests/raster_bin_sizer.cpp:27-245 (4 bins, transmission
  threshold 0.1, relevance 1 sigma: fullness after each add, bins contiguous
  from 0 and ordered);
* the whole rasteriser: an empty view leaves the framebuffer as it was; a single
  Gaussian on the optical axis gives rgb * (1 - exp(-M)) + background, M the
  mass the bins capture along the ray (Full mode leaves T = 1,
  vol_raster_forward.cu:509-513).
"""
import math

import numpy as np
import pytest

import oracle as O

# ------------------------------------------------------------ piecewise linear


def _ivec(v, add=0.0):
    return np.trunc(np.asarray(v, np.float64) * 1000.0 + add).astype(np.int64)


def _group(d0, k0, t1, t2, k2, t_right=0.0):
    return np.concatenate([d0, k0, t1, t2, k2, [t_right]]).astype(np.float64)


def _check_group(p):
    d0, k0, t1, t2 = p[0:4], p[4:8], p[8:12], p[12:16]
    assert (d0 >= 0).all() and (t1 >= 0).all() and (t2 >= t1).all() and (d0 + t1 * k0 >= 0).all()


def test_piecewise_linear_function_group():
    p = _group([0.8, 0.4, 0.5, 0.2], [-0.8, -0.2, 0.5, 0.0], [1.0, 2.0, 0.5, 2.0], [2.0, 2.0, 0.5, 2.0], [0.1, 0.2, 0.3, 0.4])
    _check_group(p)
    want = {0.0: [0.8, 0.4, 0.50, 0.2], 0.5: [0.4, 0.3, 0.75, 0.2], 1.0: [0.0, 0.2, 0.90, 0.2], 2.0: [0.0, 0.0, 1.20, 0.2],
            3.0: [0.1, 0.2, 1.50, 0.6]}
    for t, w in want.items():
        assert (_ivec(O.pl_sample(p, t)) == _ivec(w)).all(), t


def _numeric_mass(p, t_right, td):
    ts = []
    t = 0.0
    while t < t_right:  # the reference's loop, t accumulated in double
        ts.append(t)
        t += td
    s = O.pl_sample_many(p, np.array(ts))
    return np.add.accumulate(s * td, axis=0)[-1]


@pytest.mark.parametrize("case", [1, 2])
def test_piecewise_linear_fitting(case):
    if case == 1:
        pl, total = [0.8, 0.4, 0.5, 0.2], [0.3, 0.6, 0.3, 0.8]
        el, er, t_right, td, add = [0.4, 0.7, 0.2, 0.9], [0.2, 0.1, 0.0, 0.1], 0.8, 0.001, 0.0
    else:
        pl, total = [0.5, 0.5, 0.5, 0.5], [0.8, 0.8, 0.8, 0.0]
        el, er, t_right, td, add = [0.0, 0.0, 0.2, 0.0], [0.0, 0.1, 0.0, 0.0], 0.8, 0.000001, 0.5
    p = O.pl_create(pl, total, el, er, t_right)
    _check_group(p)
    assert (_ivec(O.pl_sample(p, 0.0), 0.5) == _ivec(el, 0.5)).all()
    assert (_ivec(O.pl_sample(p, t_right), 0.5) == _ivec(er, 0.5)).all()
    assert (_ivec(_numeric_mass(p, t_right, td), add) == _ivec(total, add)).all()


def test_piecewise_linear_volume_integral():
    args = [
        ([0.0, 0.0, 0.0, 0.0], [0.3, 0.6, 0.3, 0.6], [0.0, 0.0, 0.0, 0.0], [0.2, 0.1, 0.0, 0.1], 0.8),
        ([0.8, 0.4, 0.5, 0.3], [0.5, 0.4, 0.8, 0.2], [0.2, 0.1, 0.0, 0.1], [0.4, 0.7, 0.2, 0.9], 0.5),
        ([0.6, 0.7, 0.1, 0.8], [0.1, 0.0, 0.1, 0.1], [0.4, 0.7, 0.2, 0.9], [0.8, 0.1, 0.8, 0.7], 1.5),
        ([0.1, 0.8, 0.7, 0.5], [0.1, 0.2, 0.0, 0.2], [0.8, 0.1, 0.8, 0.7], [0.4, 0.7, 0.2, 0.1], 2.5),
        ([0.5, 0.2, 0.6, 0.3], [0.5, 0.6, 0.7, 0.7], [0.4, 0.7, 0.2, 0.1], [0.4, 0.7, 0.2, 0.2], 0.5),
        ([0.7, 0.8, 0.2, 0.5], [0.5, 0.2, 0.3, 0.5], [0.4, 0.7, 0.2, 0.2], [0.4, 0.7, 0.2, 1.9], 0.8),
        ([0.8, 0.4, 0.8, 0.2], [0.3, 1.6, 1.3, 1.9], [0.4, 0.7, 0.2, 1.9], [0.4, 0.7, 0.2, 0.5], 0.1),
        ([0.1, 0.1, 0.4, 0.9], [0.8, 0.7, 0.3, 2.4], [0.4, 0.7, 0.2, 0.5], [0.2, 0.1, 0.0, 0.1], 1.2),
    ]
    groups = [O.pl_create(*a) for a in args]
    n_ref = 1_000_000
    max_t = 0.0
    for g in groups:
        max_t += g[20]
    dt = max_t / n_ref
    ts = np.empty(n_ref)
    t = 0.0
    for i in range(n_ref):  # accumulated as the reference's loop
        ts[i] = t
        t += dt
    # eval(t): walk the groups, subtracting each one's width (the reference's order)
    ev = np.zeros((n_ref, 4))
    rem = ts.copy()
    todo = np.ones(n_ref, bool)
    for g in groups:
        skip = todo & (g[20] < rem)
        here = todo & ~skip
        idx = np.nonzero(here)[0]
        if idx.size:
            ev[idx] = O.pl_sample_many(g, rem[idx])
        rem[skip] -= g[20]
        todo = skip
    dens = ev[:, 3] * dt
    m_before = np.concatenate([[0.0], np.add.accumulate(dens)[:-1]])
    ref = np.add.accumulate(ev[:, :3] * np.exp(-m_before)[:, None] * dt, axis=0)[-1]
    m_total = np.add.accumulate(dens)[-1]
    assert 0.0 < m_total < 20.0
    assert (ref >= 0.0).all() and (ref < 2.0).all()
    got = O.pl_volume_integrate(groups, 256)
    assert (np.abs(got - ref) < 1.0 / 256).all(), (got, ref)


# ------------------------------------------------------------ bin sizer

SIZER = dict(n_bins=4, transmission_threshold=0.1, relevance_sigma=1.0)


def _sizer(adds):
    return O.binsizer_run(SIZER["n_bins"], SIZER["transmission_threshold"], SIZER["relevance_sigma"], adds)


def _contiguous(begin, end):
    assert begin[0] == pytest.approx(0.0)
    for i in range(3):
        assert end[i] == pytest.approx(begin[i + 1])
    assert (begin <= end).all()


def test_bin_sizer_empty():
    full, begin, end, full_final = _sizer(np.zeros((0, 3)))
    assert not full_final
    np.testing.assert_allclose(begin, 0.0)
    np.testing.assert_allclose(end, 0.0)


CASES = {
    "single opaque gaussian": ([(10.52, 8, 4)], [True], True),
    "bad gaussian order": ([(0.25, 10, 2), (0.25, 8, 1), (0.25, 6, 2), (0.25, 4, 1), (0.25, 2, 2), (0.25, 0, 1)], [False] * 6, False),
    "one gaussian at 0.6": ([(0.6, 10, 2)], [False], False),
    "one gaussian at 0.3": ([(0.3, 10, 2)], None, None),
    "several, filling up": ([(0.35, 1.3, 1.0), (0.40, 4.7, 0.81), (0.6, 6.8, 1.0), (10.8, 9.6, 0.25)], [False, False, False, True], True),
    "one gap": ([(0.7, 4.7, 0.81), (0.6, 6.8, 1.0), (0.8, 9.6, 0.25)], None, None),
    "one gap, random order": ([(0.7, 4.7, 0.81), (0.8, 9.6, 0.25), (0.6, 6.8, 1.0)], None, None),
    "two gaps": ([(0.6, 6.8, 1.0), (0.9, 9.6, 0.25)], None, None),
    "many small, random order": ([(0.08, 1.8, 1.33), (0.05, 4.6, 0.25), (0.18, 2.8, 1.22), (0.06, 5.0, 0.33), (0.08, 3.8, 1.45),
                                  (0.23, 9.6, 0.25), (0.18, 6.0, 1.20), (0.05, 4.6, 0.25), (0.21, 6.8, 1.00), (0.38, 10.2, 0.25),
                                  (0.26, 10.8, 1.00), (0.28, 10.3, 0.25), (0.21, 10.5, 1.00), (0.37, 10.9, 0.25)], None, None),
}


@pytest.mark.parametrize("name", list(CASES))
def test_bin_sizer_cases(name):
    adds, full_after, full_final = CASES[name]
    full, begin, end, ff = _sizer(adds)
    if full_after is not None:
        assert list(full) == full_after
    if full_final is not None:
        assert ff == full_final
    _contiguous(begin, end)


def test_bin_sizer_merges_keep_order_and_mass():
    """More adds than slots: borders stay sorted and finite, the last border is
    the furthest relevance end, and fullness is the captured mass against
    -log(threshold) (computed here from the closed-form normal cdf)."""
    rng = np.random.default_rng(5)
    adds = np.stack([rng.uniform(0.05, 0.4, 40), rng.uniform(0.5, 20, 40), rng.uniform(0.1, 2, 40)], 1)
    full, begin, end, ff = O.binsizer_run(32, 0.005, 3.0, adds)
    assert np.isfinite(end).all() and (np.diff(end) >= 0).all()
    _ = begin
    captured = 0.0
    for m, c, s in adds.astype(np.float32).astype(np.float64):
        captured += m * (0.5 * math.erf(3 / math.sqrt(2)) - 0.5 * math.erf(-c / s / math.sqrt(2)))
    assert end[-1] == pytest.approx(float(np.max(adds[:, 1] + 3 * adds[:, 2])), rel=1e-6)
    assert ff == (captured >= -math.log(0.005))


# ------------------------------------------------------------ whole rasteriser

SH_C0 = 0.28209479177387814


def _one_gaussian_scene(rgb, opacity, scale, pos=(0.0, 0.0, 0.0)):
    means = np.array([pos], np.float32)
    scales = np.full((1, 3), scale, np.float32)
    rots = np.array([[1, 0, 0, 0]], np.float32)
    sh16 = np.zeros((1, 16, 3), np.float32)
    sh16[0, 0] = (np.asarray(rgb) - 0.5) / SH_C0
    return means, scales, rots, np.array([opacity], np.float32), sh16, 0


def test_raster_empty_view_leaves_framebuffer():
    from differentiable_gaussian_mixture_renderer_amd import scenes
    cam = scenes.look_at_camera([0.0, 0.0, -4.0], [0.0, 0.0, 0.0], 64, 64, math.tan(math.radians(30)), math.tan(math.radians(30)))
    sc = _one_gaussian_scene([0.2, 0.4, 0.6], 0.8, 0.3, pos=(0.0, 0.0, -10.0))  # behind the camera: culled
    bg = np.array([0.1, 0.2, 0.3], np.float32)
    fb, n_pairs = O.raster_render(sc, cam, bg)
    assert n_pairs == 0
    assert not fb.any()  # nothing binned: the reference returns before rendering


@pytest.mark.parametrize("opacity,scale", [(0.5, 0.3), (0.9, 0.2), (0.3, 0.5)])
def test_raster_single_gaussian_on_axis(opacity, scale):
    """Pixel (W/2, H/2) has NDC (0, 0): its ray is the optical axis through the
    Gaussian's centre, where the normalised 3D density integrates to
    1 / (2 pi var) along the ray, var = scale^2 + f^2 (the anti-aliasing
    filter).  The Ots mass is opacity * 2 pi (scale + f^2)^2, so the optical
    depth the bins capture over [0, centre + 3 sd] is
    M = opacity (scale + f^2)^2 / var * (Phi(3) - Phi(-d / sd))."""
    from differentiable_gaussian_mixture_renderer_amd import scenes
    W = H = 64
    tan = math.tan(math.radians(30))
    cam = scenes.look_at_camera([0.0, 0.0, -4.0], [0.0, 0.0, 0.0], W, H, tan, tan)
    rgb = np.array([0.2, 0.5, 0.9])
    bg = np.array([0.1, 0.2, 0.3], np.float32)
    fb, _ = O.raster_render(_one_gaussian_scene(rgb, opacity, scale), cam, bg)
    d = 4.0
    f = d * 0.55 * 2 * float(np.float32(tan)) / W
    var = scale ** 2 + f * f
    sd = math.sqrt(var)
    phi = lambda x: 0.5 * (1 + math.erf(x / math.sqrt(2)))
    M = opacity * (scale + f * f) ** 2 / var * (phi(3.0) - phi(-d / sd))
    want = rgb * (1 - math.exp(-M)) + bg
    np.testing.assert_allclose(fb[:, H // 2, W // 2], want, rtol=0, atol=2e-3)


@pytest.mark.parametrize("opacity,scale", [(0.5, 0.3), (0.9, 0.2)])
def test_raster_fp64_single_gaussian_matches_float(opacity, scale):
    """The FP64 evaluation (float32 placement) of the single on-axis Gaussian
    meets the same closed form, and agrees with the float restatement to
    float rounding; no decision of this pixel is near a threshold."""
    from differentiable_gaussian_mixture_renderer_amd import scenes
    W = H = 64
    tan = math.tan(math.radians(30))
    cam = scenes.look_at_camera([0.0, 0.0, -4.0], [0.0, 0.0, 0.0], W, H, tan, tan)
    rgb = np.array([0.2, 0.5, 0.9])
    bg = np.array([0.1, 0.2, 0.3])
    sc = _one_gaussian_scene(rgb, opacity, scale)
    f64, n_pairs, amb = O.raster_render(sc, cam, bg, dbl=True)
    f32, n32 = O.raster_render(sc, cam, bg.astype(np.float32))
    assert n_pairs == n32 > 0
    d = 4.0
    f = d * 0.55 * 2 * float(np.float32(tan)) / W
    var = scale ** 2 + f * f
    phi = lambda x: 0.5 * (1 + math.erf(x / math.sqrt(2)))
    M = opacity * (scale + f * f) ** 2 / var * (phi(3.0) - phi(-d / math.sqrt(var)))
    np.testing.assert_allclose(f64[:, H // 2, W // 2], rgb * (1 - math.exp(-M)) + bg, rtol=0, atol=2e-3)
    assert amb[H // 2, W // 2] == 0
    assert np.abs(f64 - f32).max() <= 2e-5


def test_raster_fp64_flags_cover_the_float_decisions():
    """On a scene of 6 000 Gaussians the float restatement stays within 1e-4
    of the FP64 evaluation on every pixel none of whose decisions (mass
    thresholds, the bin sizer's centre order, cheapest merge and fullness,
    the approximation's branch) lies within float rounding of its threshold;
    those are a few percent.  A mass on the ray planted at the 1e-3 threshold
    is flagged."""
    from differentiable_gaussian_mixture_renderer_amd import scenes
    sc = scenes.make_scene(6000, seed=12, extent=1.0, sh_degree=3)
    tan = math.tan(math.radians(30))
    cam = scenes.look_at_camera([0.0, 0.0, -4.0], [0.0, 0.0, 0.0], 96, 96, tan, tan)
    arrays = O.scene_arrays(sc)
    bg = np.array([0.1, 0.2, 0.3])
    f64, n_pairs, amb = O.raster_render(arrays, cam, bg, dbl=True)
    f32, n32 = O.raster_render(arrays, cam, bg.astype(np.float32))
    assert n_pairs == n32
    ok = amb == 0
    assert 0 < (~ok).mean() <= 0.1
    err = np.abs(f32 - f64).max(axis=0)
    assert err[ok].max() <= 1e-4, err[ok].max()
    # one Gaussian whose mass on the central ray sits at the 1e-3 cut
    W = H = 32
    cam1 = scenes.look_at_camera([0.0, 0.0, -4.0], [0.0, 0.0, 0.0], W, H, tan, tan)
    d = 4.0
    f = d * 0.55 * 2 * float(np.float32(tan)) / W
    scale = 0.3
    var = scale ** 2 + f * f
    # mass on the axis ray = opacity (scale + f^2)^2 / var
    opacity = 0.001 * var / (scale + f * f) ** 2
    _, _, amb1 = O.raster_render(_one_gaussian_scene([0.5, 0.5, 0.5], opacity, scale), cam1, bg, dbl=True)
    assert amb1[H // 2, W // 2] & 1
