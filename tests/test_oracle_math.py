"""Pins the oracle's per-Gaussian math and adjoints.

Checks are chosen so that a dropped term, a wrong sign or index, or a
transposed operand fails one of them:
  * closed forms / invariants (rotation convention against scipy, covariance
    eigenvalues, SH orthonormality over the sphere, the 1D ray restriction
    against numerical integration, the Opacity formulation's on-axis optical
    depth, compositing closed forms, log accuracy);
  * central finite differences of every adjoint (double precision), mirroring
    the reference's gradient checks (unittests/grad_math.cu).
"""
import math

import numpy as np
import pytest

import oracle as O

L = O.lib
P = O.P
rng = np.random.default_rng(1234)


def fd_check(f, x, grad, h=1e-6, rtol=1e-5, atol=1e-8):
    x = np.array(x, dtype=np.float64)
    num = np.zeros_like(x)
    for i in range(x.size):
        xp = x.copy()
        xm = x.copy()
        xp.flat[i] += h
        xm.flat[i] -= h
        num.flat[i] = (f(xp) - f(xm)) / (2 * h)
    np.testing.assert_allclose(grad, num, rtol=rtol, atol=atol)


def rand_quat():
    q = rng.normal(size=4)
    return q / np.linalg.norm(q)


# --------------------------------------------------------------- covariance
def compute_cov(s, q):
    out = np.zeros(6)
    L().orc_compute_cov(P(np.ascontiguousarray(s, np.float64)), P(np.ascontiguousarray(q, np.float64)), P(out))
    return out


def sym(c):
    return np.array([[c[0], c[1], c[2]], [c[1], c[3], c[4]], [c[2], c[4], c[5]]])


def test_rotation_convention_matches_scipy():
    from scipy.spatial.transform import Rotation
    for _ in range(20):
        q = rand_quat()
        s = rng.uniform(0.1, 2.0, 3)
        R = Rotation.from_quat([q[1], q[2], q[3], q[0]]).as_matrix()   # scipy: (x, y, z, w)
        np.testing.assert_allclose(sym(compute_cov(s, q)), R @ np.diag(s * s) @ R.T, atol=1e-12)


def test_covariance_eigenvalues_are_squared_scales():
    for q in [np.array([1.0, 0, 0, 0]), np.array([0.0, 1, 0, 0]), rand_quat()]:
        s = rng.uniform(0.1, 2.0, 3)
        ev = np.sort(np.linalg.eigvalsh(sym(compute_cov(s, q))))
        np.testing.assert_allclose(ev, np.sort(s * s), rtol=1e-12)


@pytest.mark.parametrize("q", [None, [1.0, 0, 0, 0], [0.0, 0, 1, 0]], ids=["random", "identity", "180deg"])
def test_compute_cov_grad(q):
    """grad_math.cu:249-282 (random, identity and 180-degree quaternions)"""
    q = rand_quat() if q is None else np.array(q)
    s = rng.uniform(0.1, 2.0, 3)
    g = rng.normal(size=6)
    gs, gq = np.zeros(3), np.zeros(4)
    L().orc_compute_cov_grad(P(s), P(q), P(g), P(gs), P(gq))
    fd_check(lambda x: compute_cov(x, q) @ g, s, gs)
    fd_check(lambda x: compute_cov(s, x) @ g, q, gq)


# --------------------------------------------------------------- EWA splat
def camera(width=64, height=48):
    from differentiable_gaussian_mixture_renderer_amd import scenes
    cam = scenes.look_at_camera([0.3, -0.2, -4.0], [0.0, 0.1, 0.0], width, height, 0.6, 0.45)
    return (np.ascontiguousarray(cam.world_view_transform, np.float64), np.ascontiguousarray(cam.full_proj_transform, np.float64),
            np.ascontiguousarray(cam.camera_center, np.float64), 0.6, 0.45, width, height)


def splat2d(form, w, mu, s, q, cam, filt):
    v, p, c, tx, ty, W, H = cam
    out = np.zeros(6)
    L().orc_splat2d(form, w, P(np.ascontiguousarray(mu, np.float64)), P(np.ascontiguousarray(s, np.float64)),
                    P(np.ascontiguousarray(q, np.float64)), P(v), P(p), P(c), tx, ty, W, H, filt, P(out))
    return out


@pytest.mark.parametrize("form", [0, 1, 2, 3, 4], ids=["Opacity", "Mass", "Density", "Ots", "Ols"])
@pytest.mark.parametrize("filt", [0.0, 0.3])
def test_splat_grad(form, filt):
    """grad_math.cu:33-65 and 435-474: all formulations, filter 0 and 0.3"""
    cam = camera()
    for _ in range(3):
        w = rng.uniform(0.2, 0.9)
        mu = rng.uniform(-0.5, 0.5, 3)
        s = rng.uniform(0.05, 0.4, 3)
        q = rand_quat()
        gin = rng.normal(size=6)
        out = np.zeros(11)
        v, p, c, tx, ty, W, H = cam
        L().orc_splat2d_grad(form, w, P(mu), P(s), P(q), P(v), P(p), P(c), tx, ty, W, H, filt, P(gin), P(out))
        f = lambda w_, mu_, s_, q_: splat2d(form, w_, mu_, s_, q_, cam, filt) @ gin
        fd_check(lambda x: f(x[0], mu, s, q), [w], out[0:1], rtol=1e-5)
        fd_check(lambda x: f(w, x, s, q), mu, out[1:4], rtol=2e-5, atol=1e-6)
        fd_check(lambda x: f(w, mu, x, q), s, out[4:7], rtol=2e-5, atol=1e-6)
        fd_check(lambda x: f(w, mu, s, x), q, out[7:11], rtol=2e-5, atol=1e-6)


def test_splat_centroid_is_the_projected_mean_and_cov_the_jacobian_sandwich():
    """EWA closed form: the 2D covariance is J W Sigma W^T J^T (top-left 2x2) with
    J the perspective Jacobian, checked against a finite-difference Jacobian of
    the pixel projection; the centroid is the projected mean."""
    cam = camera(96, 64)
    v, p, c, tx, ty, W, H = cam
    mu = np.array([0.2, -0.1, 0.3])
    s = np.array([0.1, 0.2, 0.05])
    q = rand_quat()
    g = splat2d(3, 0.5, mu, s, q, cam, 0.0)
    Vm = v.T  # math matrix: p_view = Vm @ [p, 1]
    Pm = p.T

    def pix(x):
        h = Pm @ np.append(x, 1.0)
        ndc = h[:2] / (h[3] + 1e-7)
        return ((ndc + 1) * np.array([W, H]) - 1) * 0.5

    np.testing.assert_allclose(g[1:3], pix(mu), rtol=1e-12)
    Jn = np.zeros((2, 3))
    for i in range(3):
        e = np.zeros(3)
        e[i] = 1e-6
        Jn[:, i] = (pix(mu + e) - pix(mu - e)) / 2e-6
    cov2 = Jn @ sym(compute_cov(s, q)) @ Jn.T
    np.testing.assert_allclose([g[3], g[4], g[5]], [cov2[0, 0], cov2[0, 1], cov2[1, 1]], rtol=2e-4)


# --------------------------------------------------------------- SH
def sh_colour(sh, deg, d):
    rgb = np.zeros(3)
    ok = np.zeros(3, np.uint8)
    L().orc_sh_to_colour(P(np.ascontiguousarray(sh, np.float64)), deg, P(np.ascontiguousarray(d, np.float64)), P(rgb), P(ok))
    return rgb, ok


def test_sh_basis_is_orthonormal_on_the_sphere():
    """Real SH of degree <= 3 with the reference's constants (constants.h:26-43)
    integrate to the identity over the unit sphere."""
    n = 20000
    i = np.arange(n) + 0.5
    phi = np.arccos(1 - 2 * i / n)
    th = math.pi * (1 + 5 ** 0.5) * i
    dirs = np.stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi), np.cos(phi)], 1)
    Y = np.zeros((n, 16))
    for k in range(16):
        sh = np.zeros((16, 3))
        sh[k, 0] = 0.01
        for j in range(n):
            Y[j, k] = (sh_colour(sh, 3, dirs[j])[0][0] - 0.5) / 0.01
    gram = Y.T @ Y * (4 * math.pi / n)
    np.testing.assert_allclose(gram, np.eye(16), atol=2e-3)


def _real_sh_scipy(l, m, d):
    """Real spherical harmonic (l, m) at unit direction d, built from scipy's
    complex Y_l^m (which carries the Condon-Shortley phase): sqrt(2) Re Y_l^m
    for m > 0, sqrt(2) Im Y_l^|m| for m < 0, Y_l^0 for m = 0."""
    from scipy.special import sph_harm_y
    theta = math.acos(max(-1.0, min(1.0, d[2])))
    phi = math.atan2(d[1], d[0])
    y = complex(sph_harm_y(l, abs(m), theta, phi))
    if m > 0:
        return math.sqrt(2.0) * y.real
    if m < 0:
        return math.sqrt(2.0) * y.imag
    return y.real


def test_sh_basis_equals_scipy_real_harmonics():
    """Each of the 16 basis functions of sh_to_colour (constants.h:26-43,
    math.h:233-271) equals, value and sign, the real harmonic (l, m) with
    coefficient k = l^2 + l + m, from scipy's complex harmonics; a flipped
    sign, a swapped pair or a wrong constant fails at the first direction."""
    for _ in range(40):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        for l in range(4):
            for m in range(-l, l + 1):
                k = l * l + l + m
                sh = np.zeros((16, 3))
                sh[k] = [0.1, -0.1, 0.05]
                rgb, ok = sh_colour(sh, 3, d)
                want = _real_sh_scipy(l, m, d)
                np.testing.assert_allclose((rgb - 0.5) / np.array([0.1, -0.1, 0.05]), want, rtol=0, atol=1e-12, err_msg=f"l={l} m={m}")
                assert ok.all()
    # the active degree cuts the series: coefficients above it contribute nothing
    d = np.array([0.48, -0.6, 0.64])
    sh = rng.normal(size=(16, 3)) * 0.05
    for deg in range(4):
        rgb, _ = sh_colour(sh, deg, d)
        kk = (deg + 1) ** 2
        want = sum(sh[k] * _real_sh_scipy(int(math.isqrt(k)), k - int(math.isqrt(k)) ** 2 - int(math.isqrt(k)), d) for k in range(kk))
        np.testing.assert_allclose(rgb, want + 0.5, atol=1e-12)


@pytest.mark.parametrize("scales,expect", [
    ((0.30, 0.20, 0.05), 1.0),             # smallest scale along the view axis: the weight is the opacity
    ((0.30, 0.05, 0.20), 0.30 * 0.20 / (0.30 * 0.05)),  # smallest across it: the screen footprint shrinks, the weight grows
    ((0.05, 0.25, 0.10), 0.25 * 0.10 / (0.05 * 0.25)),
])
@pytest.mark.parametrize("spin", [0.0, 0.7])
def test_ots_weight_closed_form_on_the_optical_axis(scales, expect, spin):
    """Ots weight = w det(SJ) I2(larger2(s)) / I2(cov2d) (math.h:389-397).
    For an axis-aligned Gaussian at the centre of a camera looking down the
    world z axis, with filter 0 (optionally spun about the view axis, which
    leaves the footprint's area unchanged), det(SJ) = f_x f_y / z^2 and
    I2(cov2d) = 2 pi (f_x s_x / z)(f_y s_y / z), so the weight is
    w s_a s_b / (s_x s_y) with (s_a, s_b) the two largest scales and (s_x, s_y)
    the in-image ones: w itself when the smallest scale faces the camera."""
    from differentiable_gaussian_mixture_renderer_amd import scenes
    c = scenes.look_at_camera([0.0, 0.0, -4.0], [0.0, 0.0, 0.0], 128, 96, 0.5, 0.375)
    cam = (np.ascontiguousarray(c.world_view_transform, np.float64), np.ascontiguousarray(c.full_proj_transform, np.float64),
           np.ascontiguousarray(c.camera_center, np.float64), 0.5, 0.375, 128, 96)
    q = np.array([math.cos(spin / 2), 0.0, 0.0, math.sin(spin / 2)])
    w = 0.37
    g = splat2d(3, w, [0.0, 0.0, 0.0], scales, q, cam, 0.0)
    np.testing.assert_allclose(g[0], w * expect, rtol=1e-9)
    # and with the 0.3 px^2 filter the footprint grows, so the weight drops
    assert splat2d(3, w, [0.0, 0.0, 0.0], scales, q, cam, 0.3)[0] < g[0]


@pytest.mark.parametrize("deg", [0, 1, 2, 3])
def test_sh_grad(deg):
    """grad_math.cu:392-431, 795-813"""
    for _ in range(3):
        sh = rng.normal(size=(16, 3)) * 0.3
        sh[0] += 1.0
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        g = rng.normal(size=3)
        _, ok = sh_colour(sh, deg, d)
        gsh, gdir = np.zeros((16, 3)), np.zeros(3)
        L().orc_sh_to_colour_grad(P(sh), deg, P(d), P(g), P(ok), P(gsh), P(gdir))
        fd_check(lambda x: sh_colour(x.reshape(16, 3), deg, d)[0] @ g, sh.reshape(-1), gsh.reshape(-1))
        fd_check(lambda x: sh_colour(sh, deg, x)[0] @ g, d, gdir)


# --------------------------------------------------------------- ray restriction
def intersect(mu, A, o, d):
    out = np.zeros(3)
    L().orc_intersect_ray(P(np.ascontiguousarray(mu, np.float64)), P(np.ascontiguousarray(A, np.float64)),
                          P(np.ascontiguousarray(o, np.float64)), P(np.ascontiguousarray(d, np.float64)), P(out))
    return out


def inv_sym(c):
    M = np.linalg.inv(sym(c))
    return np.array([M[0, 0], M[0, 1], M[0, 2], M[1, 1], M[1, 2], M[2, 2]])


def test_ray_restriction_matches_numerical_integration():
    from scipy import integrate
    for _ in range(5):
        cov = compute_cov(rng.uniform(0.1, 0.5, 3), rand_quat())
        A = inv_sym(cov)
        mu = rng.uniform(-0.5, 0.5, 3)
        o = np.array([0.1, -0.2, -3.0])
        d = mu - o + rng.normal(size=3) * 0.2
        d /= np.linalg.norm(d)
        w, centre, var = intersect(mu, A, o, d)
        S = sym(cov)
        norm = 1.0 / math.sqrt((2 * math.pi) ** 3 * np.linalg.det(S))
        dens = lambda t: norm * math.exp(-0.5 * (o + t * d - mu) @ np.linalg.solve(S, o + t * d - mu))
        m0 = integrate.quad(dens, -10, 20, points=[centre], limit=200)[0]
        m1 = integrate.quad(lambda t: t * dens(t), -10, 20, points=[centre], limit=200)[0]
        m2 = integrate.quad(lambda t: (t - m1 / m0) ** 2 * dens(t), -10, 20, points=[centre], limit=200)[0]
        np.testing.assert_allclose(w, m0, rtol=1e-7)
        np.testing.assert_allclose(centre, m1 / m0, rtol=1e-7, atol=1e-9)
        np.testing.assert_allclose(var, m2 / m0, rtol=1e-6)


def test_opacity_formulation_on_axis_optical_depth_equals_opacity():
    """Opacity formulation: mass = w 2 pi sqrt(det(view-plane covariance)); a
    view-axis ray through the centre then has optical depth w."""
    s = np.array([0.2, 0.3, 0.1])
    q = rand_quat()
    cov = compute_cov(s, q)
    A = inv_sym(cov)
    mu = np.array([0.0, 0.0, 2.0])
    S = sym(cov)
    mass = 0.7 * 2 * math.pi * math.sqrt(np.linalg.det(S[:2, :2]))
    w1d, _, _ = intersect(mu, A, np.zeros(3), np.array([0.0, 0.0, 1.0]))
    np.testing.assert_allclose(w1d * mass, 0.7, rtol=1e-12)


def test_ray_restriction_grad():
    for _ in range(5):
        A = inv_sym(compute_cov(rng.uniform(0.1, 0.5, 3), rand_quat()))
        mu = rng.uniform(-0.5, 0.5, 3)
        o = np.array([0.1, -0.2, -3.0])
        d = mu - o + rng.normal(size=3) * 0.2
        d /= np.linalg.norm(d)
        g = rng.normal(size=3)
        gmu, gA = np.zeros(3), np.zeros(6)
        L().orc_intersect_ray_grad(P(mu), P(A), P(o), P(d), P(g), P(gmu), P(gA))
        fd_check(lambda x: intersect(x, A, o, d) @ g, mu, gmu, rtol=1e-5, atol=1e-7)
        fd_check(lambda x: intersect(mu, x, o, d) @ g, A, gA, rtol=1e-5, atol=1e-7)


# --------------------------------------------------------------- bins
NB = 128


def sample_gaussian(mass, rgb, mu, A, o, d, borders):
    bins = np.zeros((NB - 1, 4))
    L().orc_sample_gaussian(mass, P(np.ascontiguousarray(rgb, np.float64)), P(np.ascontiguousarray(mu, np.float64)),
                            P(np.ascontiguousarray(A, np.float64)), P(o), P(d), P(np.ascontiguousarray(borders, np.float64)), P(bins))
    return bins


def _bin_setup():
    A = inv_sym(compute_cov(np.array([0.3, 0.2, 0.25]), rand_quat()))
    mu = np.array([0.05, -0.02, 0.1])
    o = np.array([0.0, 0.0, -3.0])
    d = np.array([0.0, 0.0, 1.0])
    w, c, var = intersect(mu, A, o, d)
    borders = np.linspace(c - 2.5 * math.sqrt(var), c + 2.5 * math.sqrt(var), NB)
    return A, mu, o, d, borders


def test_sample_gaussian_closed_form():
    """per-bin mass = mass_on_ray * (cdf(b_k+1) - cdf(b_k)) (math.h:430-468)"""
    from scipy.stats import norm
    A, mu, o, d, borders = _bin_setup()
    w, c, var = intersect(mu, A, o, d)
    mass = 2.0
    rgb = np.array([0.2, 0.5, 0.9])
    bins = sample_gaussian(mass, rgb, mu, A, o, d, borders)
    expect = np.diff(norm.cdf(borders, loc=c, scale=math.sqrt(var))) * w * mass
    expect[expect < 1e-5] = 0
    np.testing.assert_allclose(bins[:, 3], expect, rtol=1e-9, atol=1e-15)
    np.testing.assert_allclose(bins[:, :3], expect[:, None] * rgb[None, :], rtol=1e-9, atol=1e-15)


def test_sample_gaussian_grad():
    """grad_math.cu:745-793"""
    A, mu, o, d, borders = _bin_setup()
    rgb = np.array([0.2, 0.5, 0.9])
    mass = 2.0
    gb = rng.normal(size=(NB - 1, 4))
    gm, grgb, gmu, gA = np.zeros(1), np.zeros(3), np.zeros(3), np.zeros(6)
    L().orc_sample_gaussian_grad(mass, P(rgb), P(mu), P(A), P(o), P(d), P(borders), P(gb), P(gm), P(grgb), P(gmu), P(gA))
    f = lambda m_, r_, mu_, A_: (sample_gaussian(m_, r_, mu_, A_, o, d, borders) * gb).sum()
    fd_check(lambda x: f(x[0], rgb, mu, A), [mass], gm, rtol=1e-6)
    fd_check(lambda x: f(mass, x, mu, A), rgb, grgb, rtol=1e-6)
    fd_check(lambda x: f(mass, rgb, x, A), mu, gmu, rtol=2e-5, atol=1e-7)
    fd_check(lambda x: f(mass, rgb, mu, x), A, gA, rtol=2e-5, atol=1e-7)


def integrate_bins(C, T, bins):
    Cc = np.array(C, np.float64)
    Tc = np.array([T], np.float64)
    L().orc_integrate_bins(P(Cc), P(Tc), P(np.ascontiguousarray(bins, np.float64)))
    return Cc, Tc[0]


def test_integrate_bins_closed_form():
    bins = np.zeros((NB - 1, 4))
    bins[:10, 3] = rng.uniform(0.01, 0.1, 10)
    bins[:10, :3] = rng.uniform(0, 0.1, (10, 3))
    C, T = integrate_bins([0.1, 0.2, 0.3], 0.9, bins)
    Tk = 0.9 * np.concatenate([[1.0], np.cumprod(np.exp(-bins[:, 3]))])
    np.testing.assert_allclose(T, Tk[-1], rtol=1e-14)
    np.testing.assert_allclose(C, np.array([0.1, 0.2, 0.3]) + (bins[:, :3] * Tk[:-1, None]).sum(0), rtol=1e-14)


def test_integrate_bins_grad():
    """grad_math.cu:544-743 (bins chained with the remaining colour walk)"""
    bins = np.zeros((NB - 1, 4))
    k = 40
    bins[:k, 3] = rng.uniform(0.001, 0.05, k)
    bins[:k, :3] = rng.uniform(0, 0.05, (k, 3))
    C0, T0 = np.array([0.05, 0.02, 0.01]), 0.95
    gC = rng.normal(size=3)
    gT = rng.normal()
    Cf, Tf = integrate_bins(C0, T0, bins)
    # the adjoint walks forward from the final colour, normalised by T0
    Crem = (Cf - C0) / T0
    Cw = np.array(Crem, np.float64)
    Tw = np.array([T0])
    gbins = np.zeros_like(bins)
    L().orc_integrate_bins_grad(P(Cw), P(Tw), Tf, P(bins), P(gC), gT, P(gbins))
    loss = lambda b: (lambda C, T: C @ gC + T * gT)(*integrate_bins(C0, T0, b.reshape(bins.shape)))
    num = np.zeros_like(bins)
    h = 1e-7
    for i in range(k):
        for j in range(4):
            bp, bm = bins.copy(), bins.copy()
            bp[i, j] += h
            bm[i, j] -= h
            num[i, j] = (loss(bp) - loss(bm)) / (2 * h)
    np.testing.assert_allclose(gbins[:k], num[:k], rtol=1e-5, atol=1e-9)


# --------------------------------------------------------------- helpers
def test_portable_log_accuracy():
    xs = np.concatenate([rng.uniform(1e-6, 1, 2000), rng.uniform(1, 1e6, 2000), [1.0, 2.0, 0.5, 255.0, 1.0 / 255.0, 1e-40]])
    for x in xs:
        got = L().orc_portable_log_f(float(np.float32(x)))
        want = math.log(float(np.float32(x)))
        assert abs(got - want) <= 2 * np.spacing(np.float32(abs(want)) or np.float32(1e-7)) + 1e-7, (x, got, want)
        gd = L().orc_portable_log_d(float(x))
        assert abs(gd - math.log(x)) <= 4 * np.spacing(abs(math.log(x)) or 1e-16), (x, gd)


def test_higher_msb():
    """util.h:56-70: one past the highest set bit"""
    for n in [1, 2, 3, 4, 255, 256, 8160, 32400, 8100, 65535, 65536, 1 << 20]:
        assert L().orc_higher_msb(n) == int(n).bit_length()


def test_get_rect_truncation_and_clamping():
    """util.h:42-52: (int) truncation towards zero, clamped to [0, grid]"""
    out = np.zeros(4, np.int32)
    L().orc_get_rect_f(5.5, 3.0, 10, 2, 4, 4, 16, 16, P(out))
    # (5.5-10)/16 = -0.28 -> 0; (5.5+10+15)/16 = 1.9 -> 1; (3-2)/16 -> 0; (3+2+15)/16 = 1.25 -> 1
    assert list(out) == [0, 0, 1, 1]
    L().orc_get_rect_f(1000.0, 1000.0, 5, 5, 4, 4, 16, 16, P(out))
    assert list(out) == [4, 4, 4, 4]
    L().orc_get_rect_f(20.0, 40.0, 0, 0, 4, 4, 16, 16, P(out))
    # zero extent still covers the tile holding the centre: [1,2) x [2,3)
    assert list(out) == [1, 2, 2, 3]


def test_abramowitz_stegun_erf_constants():
    """The float marcher's erf (Abramowitz & Stegun 7.1.26): with its five
    coefficients and p, |erf_AS(x) - erf(x)| <= 1.5e-7 over [-6, 6] (the
    published bound), evaluated here in float64."""
    import math
    p, a = 0.3275911, (0.254829592, -0.284496736, 1.421413741, -1.453152027, 1.061405429)
    worst = 0.0
    for x in np.linspace(-6, 6, 24001):
        ax = abs(x)
        t = 1.0 / (1.0 + p * ax)
        poly = t * (a[0] + t * (a[1] + t * (a[2] + t * (a[3] + t * a[4]))))
        e = math.copysign(1.0 - poly * math.exp(-ax * ax), x)
        worst = max(worst, abs(e - math.erf(x)))
    assert worst <= 1.5e-7, worst
