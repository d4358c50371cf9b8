"""The far-field forms of the mixed-precision line sum (csrc/faddeeva.h,
csrc/voigt_sum.hip) checked against independent references on the CPU.

K(x, y) = Re w(x + iy).  References: scipy's ``wofz`` where it is accurate
(moderate |z|), and, for large |z|, the asymptotic series
K = (1/sqrt(pi)) sum_n (2n-1)!!/2^n sin((2n+1) theta) / |z|^(2n+1), theta = arg z,
whose real part is formed from sin((2n+1) theta) directly, so it keeps full
relative accuracy when y << x (where Re w is a tiny part of |w|).
"""

import math

import numpy as np
import pytest
from scipy.special import wofz

SQRT_PI = math.sqrt(math.pi)
S1_COEF = np.float32(164062.5)      # 1.05 / 6.4e-6, faddeeva.h kF32S1Coef
EPS_FAR = 6.4e-6                    # target of the one-node threshold
EPS_GH2 = 6.42e-6                   # the two-node form's error at |z| = 25 (max 6.41e-6)
P = 8                               # points per thread (kPointsMixed)


def k_series(x, y, terms=40):
    r = np.hypot(x, y)
    th = np.arctan2(y, x)
    k = np.zeros_like(r)
    coef = 1.0
    rp = r.copy()
    with np.errstate(over="ignore"):
        for n in range(terms):
            k += coef * np.sin((2 * n + 1) * th) / rp
            coef *= (2 * n + 1) / 2.0
            rp = rp * r * r
    return k / SQRT_PI


def s1_r2_min(y2):
    """faddeeva.h s1_r2_min, in float32 as the kernel evaluates it."""
    y2 = np.float32(y2)
    return np.maximum(np.float32(625.0), np.sqrt(S1_COEF * (np.float32(2.0) * y2 + np.float32(1.5))))


def k_s1(x, y):
    return y / (SQRT_PI * (x * x + y * y - 1.5))


def k_gh2(x, y):
    w = x * x + y * y - 0.5
    return y * (w + 1.0) / (SQRT_PI * (w * w + 2.0 * y * y))


def test_series_reference_agrees_with_wofz():
    rng = np.random.default_rng(11)
    x = rng.uniform(20.0, 200.0, 20000)
    y = 10.0 ** rng.uniform(-1.0, 2.0, 20000)
    ref = wofz(x + 1j * y).real
    assert np.max(np.abs(k_series(x, y) / ref - 1.0)) < 1e-12


@pytest.mark.parametrize("ylo,yhi", [(1e-4, 1e-2), (1e-2, 1.0), (1.0, 10.0), (10.0, 100.0),
                                     (100.0, 1000.0)])
def test_shifted_one_node_within_bound_where_used(ylo, yhi):
    """Where the kernel uses it (|z|^2 >= s1_r2_min(y^2)) the shifted one-node
    form is within 6.4e-6 relative of K, the two-node form's bound."""
    rng = np.random.default_rng(int(ylo * 1e4) + 7)
    n = 200000
    y = np.exp(rng.uniform(math.log(ylo), math.log(yhi), n))
    r_min = np.sqrt(s1_r2_min(y * y).astype(np.float64))
    r = r_min * np.exp(rng.uniform(0.0, math.log(200.0), n))
    r[: n // 10] = r_min[: n // 10]          # the boundary itself
    ok = r > y
    y, r = y[ok], r[ok]
    x = np.sqrt(r * r - y * y) * rng.choice([-1.0, 1.0], y.size)
    err = np.abs(k_s1(x, y) / k_series(x, y) - 1.0)
    assert err.max() <= EPS_FAR
    # and it is the form the thresholds were derived for: error ~ (2y^2 - 3/2)/|z|^4
    near = r < 1.01 * r_min[ok]
    assert err[near].max() > 0.2 * EPS_FAR


def test_shifted_one_node_error_model():
    """|err| <= 1.05 (3/2 + 2 y^2) / |z|^4 for |z| >= 20 (the bound the
    threshold is built from); a dropped or wrong shift fails it."""
    rng = np.random.default_rng(5)
    n = 400000
    y = np.exp(rng.uniform(math.log(1e-3), math.log(300.0), n))
    r = np.exp(rng.uniform(math.log(20.0), math.log(2000.0), n))
    ok = r > y * 1.0001
    y, r = y[ok], r[ok]
    x = np.sqrt(r * r - y * y)
    k = k_series(x, y)
    bound = 1.05 * (1.5 + 2.0 * y * y) / r ** 4
    assert np.all(np.abs(k_s1(x, y) / k - 1.0) <= bound)
    unshifted = y / (SQRT_PI * (x * x + y * y))
    assert np.any(np.abs(unshifted / k - 1.0) > bound)


def test_shifted_one_node_fp64_tier():
    """FP64 mode: where |z|^2 >= sqrt(1.05e11 (2 y^2 + 3/2)) (faddeeva.h
    s1_r2_min64) the shifted one-node form is within 1e-11 of K."""
    rng = np.random.default_rng(21)
    n = 400000
    y = np.exp(rng.uniform(math.log(1e-4), math.log(300.0), n))
    r_min = np.sqrt(np.sqrt(1.05e11 * (2.0 * y * y + 1.5)))
    r = r_min * np.exp(rng.uniform(0.0, math.log(50.0), n))
    r[: n // 10] = r_min[: n // 10]
    ok = r > y
    y, r = y[ok], r[ok]
    x = np.sqrt(r * r - y * y)
    err = np.abs(k_s1(x, y) / k_series(x, y) - 1.0)
    assert err.max() <= 1e-11
    assert err[r < 1.01 * r_min[ok]].max() > 2e-12


def k_s2(x, y):
    r = x * x + y * y
    return y * r / (SQRT_PI * (r * r - 1.5 * r + 2.0 * y * y - 1.5))


def test_second_order_shifted_form_error_model():
    """|err| <= 1.1 (21/4 + 9 y^2)/|z|^6 (+ rounding) for |z| >= 20; the
    one-node form (no 1/r correction) fails it."""
    rng = np.random.default_rng(8)
    n = 400000
    y = np.exp(rng.uniform(math.log(1e-4), math.log(1000.0), n))
    r = np.exp(rng.uniform(math.log(20.0), math.log(5000.0), n))
    ok = r > y * 1.0001
    y, r = y[ok], r[ok]
    x = np.sqrt(r * r - y * y)
    k = k_series(x, y)
    bound = 1.1 * (5.25 + 9.0 * y * y) / r ** 6 + 2e-15
    assert np.all(np.abs(k_s2(x, y) / k - 1.0) <= bound)
    assert np.any(np.abs(k_s1(x, y) / k - 1.0) > bound)


def test_second_order_fp64_tier():
    """Where the FP64 line sum uses it (|z|^2 >= cbrt(1.1e11 (9 y^2 + 21/4)),
    faddeeva.h s2_r2_min64) the second-order form is within 1e-11, and that
    radius lies below the one-node form's for every y."""
    rng = np.random.default_rng(13)
    n = 400000
    y = np.exp(rng.uniform(math.log(1e-4), math.log(1000.0), n))
    r_min = np.sqrt(np.cbrt(1.1e11 * (9.0 * y * y + 5.25)))
    assert np.all(r_min ** 2 < np.sqrt(1.05e11 * (2.0 * y * y + 1.5)))
    r = r_min * np.exp(rng.uniform(0.0, math.log(20.0), n))
    r[: n // 10] = r_min[: n // 10]
    ok = r > y
    y, r = y[ok], r[ok]
    x = np.sqrt(r * r - y * y)
    err = np.abs(k_s2(x, y) / k_series(x, y) - 1.0)
    assert err.max() <= 1e-11
    assert err[r < 1.01 * r_min[ok]].max() > 1e-12


def test_second_order_mixed_tier():
    """Mixed line sum: for y^2 < 155 the middle tier is the second-order form
    from |z|^2 = cbrt(171875 (9 y^2 + 21/4)) (faddeeva.h s2_r2_min32), inside
    6.4e-6 and below the two-node form's |z| = 25."""
    rng = np.random.default_rng(31)
    n = 300000
    y = np.exp(rng.uniform(math.log(1e-4), math.log(math.sqrt(155.0)), n)).astype(np.float32)
    y2 = (y * y).astype(np.float32)
    r2 = np.cbrt(np.float32(171875.0) * (np.float32(9.0) * y2 + np.float32(5.25))).astype(np.float64)
    assert np.all(r2 < 625.0)
    r = np.sqrt(r2) * np.exp(rng.uniform(0.0, math.log(3.0), n))
    r[: n // 10] = np.sqrt(r2[: n // 10])
    y = y.astype(np.float64)
    ok = r > y
    y, r = y[ok], r[ok]
    x = np.sqrt(r * r - y * y)
    err = np.abs(k_s2(x, y) / k_series(x, y) - 1.0)
    assert err.max() <= EPS_FAR


def test_two_node_within_bound_at_25():
    rng = np.random.default_rng(9)
    n = 200000
    th = rng.uniform(0.0, math.pi / 2, n)
    r = 25.0 * np.exp(rng.uniform(0.0, math.log(40.0), n))
    x, y = r * np.cos(th), np.maximum(r * np.sin(th), 1e-6)
    assert np.max(np.abs(k_gh2(x, y) / k_series(x, y) - 1.0)) <= EPS_GH2


@pytest.mark.parametrize("which", ["s1", "gh2"])
def test_mid_point_threshold_covers_the_whole_run(which):
    """|x_mid| >= hw (line_params.hip store_mixed) puts all P points of the
    run x_mid + (j - 3.5) hs inside the form's region."""
    rng = np.random.default_rng(3 if which == "s1" else 4)
    n = 100000
    y = np.exp(rng.uniform(math.log(1e-3), math.log(300.0), n)).astype(np.float32)
    hs = np.exp(rng.uniform(math.log(1e-3), math.log(30.0), n)).astype(np.float32)
    y2 = y * y
    r2 = s1_r2_min(y2) if which == "s1" else np.float32(625.0)
    hw = (np.float32(0.5 * (P - 1)) * hs + np.sqrt(np.maximum(r2 - y2, np.float32(0.0)))).astype(np.float32)
    xm = (hw * np.float32(1.0) + np.float32(1e-3) * hs * rng.uniform(0, 1, n).astype(np.float32))
    xm *= rng.choice([-1.0, 1.0], n).astype(np.float32)
    j = np.arange(P, dtype=np.float32) - np.float32(0.5 * (P - 1))
    x = xm[:, None] + j[None, :] * hs[:, None]
    z2 = x.astype(np.float64) ** 2 + y2[:, None].astype(np.float64)
    assert np.all(z2 >= np.asarray(r2, np.float64).reshape(-1, 1) * (1.0 - 1e-6))


def test_two_node_constants_from_c1_are_exact():
    """cw = c1 + 1 and d2 = 2 c1 + 3 reproduce y^2 - 1/2 and 2 y^2 exactly in
    float32 (c1 = y^2 - 3/2 as stored) for 3 <= y^2 < 2^23."""
    rng = np.random.default_rng(1)
    y2 = np.exp(rng.uniform(math.log(3.0), math.log(2.0 ** 23), 1000000)).astype(np.float32)
    c1 = (y2 - np.float32(1.5)).astype(np.float32)
    assert np.array_equal((c1 + np.float32(1.0)).astype(np.float32), (y2 - np.float32(0.5)).astype(np.float32))
    d2 = (np.float32(2.0) * c1 + np.float32(3.0)).astype(np.float32)
    assert np.array_equal(d2, (np.float32(2.0) * y2).astype(np.float32))


# 10-point Gauss-Hermite nodes and weights as faddeeva.h holds them (t > 0)
GH10_T = np.array([0.34290132722370461, 1.0366108297895137, 1.7566836492998818,
                   2.5327316742327898, 3.4361591188377376])
GH10_W = np.array([0.6108626337353258, 0.24013861108231469, 0.033874394455481063,
                   0.0013436457467812327, 7.6404328552326206e-6])


def k_gh10(x, y):
    """faddeeva.h gh_sum64 with the 10 nodes: the pair terms summed as one
    running quotient (numerator and product of the denominators)."""
    x2, y2 = x * x, y * y
    r, xmy, fxy = x2 + y2, x2 - y2, 4.0 * x2 * y2
    num = np.zeros_like(x)
    den = np.ones_like(x)
    for k, (t, w) in enumerate(zip(GH10_T, GH10_W)):
        u = xmy - t * t
        D = u * u + fxy
        num = w * (r + t * t) * den if k == 0 else num * D + w * (r + t * t) * den
        den = D if k == 0 else den * D
    return 2.0 * y / math.pi * num / den


def test_ten_node_constants_are_gauss_hermite():
    t, w = np.polynomial.hermite.hermgauss(10)
    np.testing.assert_allclose(GH10_T, t[t > 0], rtol=1e-15)
    np.testing.assert_allclose(GH10_W, w[t > 0], rtol=1e-13)
    assert abs(2 * GH10_W.sum() - SQRT_PI) < 1e-15


def test_ten_node_form_within_1e11_from_7():
    """FP64 near-centre pass A: the 10-node form holds K to 1e-11 for |z| >= 7
    and y >= 1e-8 (faddeeva.h kFp64Gh10R2, kFp64Gh10Ymin), against scipy's
    wofz; and it does not hold much nearer (so the radius is not slack)."""
    rng = np.random.default_rng(23)
    n = 400000
    y = np.exp(rng.uniform(math.log(1e-8), math.log(60.0), n))
    r = 7.0 * np.exp(rng.uniform(0.0, math.log(8.0), n))
    r[: n // 5] = 7.0
    ok = r > y
    y, r = y[ok], r[ok]
    x = np.sqrt(r * r - y * y)
    err = np.abs(k_gh10(x, y) / wofz(x + 1j * y).real - 1.0)
    assert err.max() <= 1e-11, (err.max(), x[err.argmax()], y[err.argmax()])
    xs = np.linspace(5.8, 6.4, 61)
    ys = np.full_like(xs, 1e-2)
    assert np.max(np.abs(k_gh10(xs, ys) / wofz(xs + 1j * ys).real - 1.0)) > 1e-10


# 6-point Gauss-Hermite nodes (squared) and weights as faddeeva.h holds them (FP32)
GH6_T2 = np.float32([0.19016350919348813, 1.7844927485432516, 5.5253437422632603])
GH6_W = np.float32([0.72462959522439252, 0.15706732032285664, 0.0045300099055088456])


def k_gh6_f32(x, y):
    """faddeeva.h gh6_sum32 times 2y/pi, evaluated in float32 like the kernel."""
    f32 = np.float32
    x2, y2 = f32(x) * f32(x), f32(y) * f32(y)
    r, xmy, fxy = x2 + y2, x2 - y2, f32(4) * x2 * y2
    for k in range(3):
        u = xmy - GH6_T2[k]
        D = u * u + fxy
        n = GH6_W[k] * (r + GH6_T2[k])
        num, den = (n, D) if k == 0 else (num * D + n * den, den * D)
    return f32(2 / np.pi) * f32(y) * num / den


def gh6_core_r2(y2):
    """faddeeva.h gh6_core_r2 (y^2 as the kernel forms it, c1 + 3/2 in FP32)."""
    return np.where(y2 < 1e-8, 49.0, np.where(y2 < 1e-6, 30.25, np.where(y2 < 4.0, 25.0, 12.96)))


def test_six_node_constants_are_gauss_hermite():
    t, w = np.polynomial.hermite.hermgauss(6)
    np.testing.assert_allclose(GH6_T2, (t[t > 0] ** 2).astype(np.float32), rtol=1e-7)
    np.testing.assert_allclose(GH6_W, w[t > 0].astype(np.float32), rtol=1e-7)


def test_six_node_form_outside_its_core_radius():
    """Mixed near-centre pass A: in FP32 the six-node form is within 5e-6 of
    scipy's wofz wherever |z|^2 >= gh6_core_r2(y^2) (y^2 carried as
    (y^2 - 3/2) + 3/2 in FP32, as the record holds it), for y from 1e-10 to
    50; and the radius is needed (the form misses 5e-6 just inside it for
    small y)."""
    f32 = np.float32
    rng = np.random.default_rng(29)
    n = 2_000_000
    y = np.exp(rng.uniform(math.log(1e-10), math.log(50.0), n))
    y2k = (f32(y) * f32(y) - f32(1.5)) + f32(1.5)        # the kernel's y^2
    rc = np.sqrt(gh6_core_r2(y2k.astype(np.float64)))
    r = rc * np.exp(rng.uniform(0.0, math.log(10.0), n))
    r[: n // 4] = rc[: n // 4] * (1 + 1e-6)
    ok = r > y
    y, r = y[ok], r[ok]
    x = np.sqrt(r * r - y * y)
    err = np.abs(k_gh6_f32(x, y).astype(np.float64) / wofz(x + 1j * y).real - 1.0)
    assert err.max() <= 5e-6, (err.max(), x[err.argmax()], y[err.argmax()])
    xs = np.linspace(4.3, 4.7, 41)
    ys = np.full_like(xs, 1e-2)
    assert np.max(np.abs(k_gh6_f32(xs, ys) / wofz(xs + 1j * ys).real - 1.0)) > 5e-6


def kgrad_pairs_f32(x, y, t2s, ws):
    """faddeeva.h kgrad_pairs32 (K, dK/dx, dK/dy of a Gauss-Hermite node-pair
    form, weights 2W/pi), evaluated in float32 like the kernel."""
    f32 = np.float32
    x, y = f32(x), f32(y)
    x2, y2 = x * x, y * y
    r = x2 + y2
    sf = sp = sy = f32(0)
    for t2, w in zip(t2s, ws):
        t2, c = f32(t2), f32(w) * f32(2 / np.pi)
        wv = r - t2
        d = f32(4) * t2 * y2
        rD = f32(1) / (wv * wv + d)
        F = (wv + f32(2) * t2) * rD
        Fp = (f32(4) * t2 * (t2 - x2) - wv * wv) * rD * rD
        sf = sf + c * F
        sp = sp + c * Fp
        sy = sy + c * (F - f32(2) * d * F * rD)
    return y * sf, f32(2) * x * y * sp, f32(2) * y2 * sp + sy


def _kgrad_exact(x, y):
    z = x + 1j * y
    w = wofz(z)
    wp = -2.0 * z * w + 2j / SQRT_PI
    return w.real, wp.real, -wp.imag


def test_six_node_gradient_outside_the_adjoint_core():
    """Adjoint drain, first pass (adjoint.hip adj_core_r2): in FP32 the
    six-node form's K is within 1e-6 of K and its dK/dx, dK/dy within 1e-6 of
    |grad K| wherever |z| >= 5.5 (|z| >= 7 for y < 1e-4), for y from 1e-10 to
    50 -- better than the four-node form from |z| = 8, where kgrad_mixed
    switches to it (1.5e-6 here in FP32); and both radii are needed: below
    y = 1e-4 the form leaves out exp(-x^2), which at |z| = 5.5 is 7% of K for
    y = 1e-10, and at |z| = 5 it misses 1e-6 for y = 1e-3."""
    rng = np.random.default_rng(31)
    n = 1_000_000

    def worst(t2s, ws, rad, ylo, yhi):
        y = np.exp(rng.uniform(math.log(ylo), math.log(yhi), n))
        r = rad * np.exp(rng.uniform(0.0, math.log(8.0), n))
        r[: n // 4] = rad * (1 + 1e-6)
        ok = r > y
        yy, rr = y[ok], r[ok]
        x = np.sqrt(rr * rr - yy * yy)
        K, Kx, Ky = (v.astype(np.float64) for v in kgrad_pairs_f32(x, yy, t2s, ws))
        E, Ex, Ey = _kgrad_exact(x, yy)
        g = np.hypot(Ex, Ey)
        return (np.max(np.abs(K / E - 1.0)),
                np.max(np.maximum(np.abs(Kx - Ex), np.abs(Ky - Ey)) / g))

    t4, w4 = np.polynomial.hermite.hermgauss(4)
    assert max(worst(t4[t4 > 0] ** 2, w4[t4 > 0], 8.0, 1e-10, 50.0)) <= 1.5e-6
    assert max(worst(GH6_T2, GH6_W, 7.0, 1e-10, 1e-4)) <= 1e-6
    assert max(worst(GH6_T2, GH6_W, 5.5, 1e-4, 50.0)) <= 1e-6
    assert worst(GH6_T2, GH6_W, 5.5, 1e-10, 1e-9)[0] > 1e-2
    assert max(worst(GH6_T2, GH6_W, 5.0, 1e-3, 2.0)) > 1e-6
