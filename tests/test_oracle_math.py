"""Pins the oracle's deterministic kernels before they are used as the GPU
path's checker: each check is a published value, a closed form or an
invariant that a dropped term, a wrong sign or a swapped operand breaks.

- Threefry-4x64-20: Random123's published known-answer vectors.
- solve_second (src/ssol_shape.c:864-896): roots of a(t-r1)(t-r2), hint
  ordering, the cancellation-free small root, the linear case.
- Ray/quadric (src/ssol_shape.c:970-1135): hit point on the surface z(x, y);
  focusing properties of the paraboloid (point focus), parabolic cylinder
  (focal line) and hyperboloid (focus to focus); plane distance; hemisphere
  radius.
- Sun directions (src/ssol_ranst_sun_dir.c:246-295): sin^2(theta) =
  u1 sin^2(theta_max) for the pillbox, sin(theta) = sigma sqrt(-2 ln u1) for
  the Gaussian, azimuth 2 pi u2; Buie: density normalisation and a
  Kolmogorov-Smirnov test against its own pdf.
- Spectrum lookup: the values of src/test_ssol_data.c:49-76.
- Piecewise-linear wavelength sampling: inverse of the closed-form CDF.
- BSDFs: specular and Lambertian closed forms; Fresnel reflectance at normal
  incidence; the thin-slab series 2R/(1+R); the glossy mirror's limit at zero
  roughness.
"""
import math

import numpy as np
import pytest

import oracle_ctypes as O
import ssol_scene as S


# Threefry ----------------------------------------------------------------------
def test_threefry_known_answers():
    # Random123 kat_vectors, threefry4x64_20
    assert O.threefry([0, 0, 0, 0], [0, 0, 0, 0]) == [
        0x09218ebde6c85537, 0x55941f5266d86105, 0x4bd25e16282434dc, 0xee29ec846bd2e40b]
    m = 2**64 - 1
    assert O.threefry([m] * 4, [m] * 4) == [
        0x29c24097942bba1b, 0x0371bbfb0f6f4e11, 0x3c231ffa33f83a1c, 0xcd29113fde32d168]
    # distinct words: each counter and key word lands in its own lane
    base = O.threefry([1, 2, 3, 4], [5, 6, 7, 8])
    assert O.threefry([2, 1, 3, 4], [5, 6, 7, 8]) != base
    assert O.threefry([1, 2, 3, 4], [6, 5, 7, 8]) != base


# solve_second --------------------------------------------------------------------
@pytest.mark.parametrize("a,r1,r2", [(1.0, 2.0, 5.0), (-3.0, -1.5, 4.0), (0.25, 1e-8, 1e8),
                                     (2.0, -7.0, -7.5)])
def test_solve_second_roots(a, r1, r2):
    b, c = -a * (r1 + r2), a * r1 * r2
    n, d = O.solve_second(a, b, c, hint=r1)
    assert n == 2
    assert d[0] == pytest.approx(r1, rel=1e-14, abs=0)
    assert d[1] == pytest.approx(r2, rel=1e-14, abs=0)
    n, d = O.solve_second(a, b, c, hint=r2)
    assert d[0] == pytest.approx(r2, rel=1e-14, abs=0)


def test_solve_second_degenerate():
    n, d = O.solve_second(0.0, 2.0, -6.0, hint=0.0)
    assert n == 1 and d[0] == 3.0
    assert O.solve_second(0.0, 0.0, 1.0, hint=0.0)[0] == 0
    assert O.solve_second(1.0, 0.0, 1.0, hint=0.0)[0] == 0  # negative discriminant: miss
    n, d = O.solve_second(1.0, -4.0, 4.0, hint=0.0)
    assert n == 1 and d[0] == 2.0


# Quadrics ------------------------------------------------------------------------
def _reflect(d, n):
    n = n / np.linalg.norm(n)
    return d - 2 * np.dot(d, n) * n


def _hit(q, data, org, d, hint):
    ok, t, g = O.quadric_intersect(q, data, org, d, hint)
    assert ok
    p = np.asarray(org) + t * np.asarray(d)
    assert p[2] == pytest.approx(O.quadric_z(q, data, p[:2]), abs=1e-9 * max(1, abs(p[2])))
    return t, p, g


def test_plane_distance():
    rng = np.random.default_rng(0)
    for _ in range(50):
        o = np.array([*rng.uniform(-5, 5, 2), rng.uniform(0.5, 20)])
        d = np.array([*rng.uniform(-0.5, 0.5, 2), -1.0])
        d /= np.linalg.norm(d)
        t, p, g = _hit(0, (0.0, 0.0), o, d, 0.0)
        assert t == pytest.approx(-o[2] / d[2], rel=1e-14)
        assert np.allclose(g / np.linalg.norm(g), [0, 0, 1])


@pytest.mark.parametrize("f", [0.5, 10.0, 250.0])
def test_paraboloid_point_focus(f):
    rng = np.random.default_rng(1)
    focus = np.array([0.0, 0.0, f])
    for _ in range(50):
        xy = rng.uniform(-f, f, 2)
        o = np.array([xy[0], xy[1], 3 * f])
        d = np.array([0.0, 0.0, -1.0])
        t, p, g = _hit(1, (f, 0.0), o, d, 2 * f)
        assert t == pytest.approx(3 * f - (xy @ xy) / (4 * f), rel=1e-13)
        r = _reflect(d, g)
        v = focus - p
        assert np.linalg.norm(np.cross(r, v)) <= 1e-12 * np.linalg.norm(v)
        assert np.dot(r, v) > 0


def test_paraboloid_two_focals_reduces_to_paraboloid():
    rng = np.random.default_rng(2)
    for _ in range(30):
        o = np.array([*rng.uniform(-4, 4, 2), 30.0])
        d = np.array([*rng.uniform(-0.2, 0.2, 2), -1.0])
        d /= np.linalg.norm(d)
        _, t1, g1 = O.quadric_intersect(1, (7.0, 0.0), o, d, 25.0)
        _, t2, g2 = O.quadric_intersect(2, (7.0, 7.0), o, d, 25.0)
        assert t1 == pytest.approx(t2, rel=1e-13)
        assert np.allclose(g1 / np.linalg.norm(g1), g2 / np.linalg.norm(g2), atol=1e-13)
    # distinct focal lengths: x-z and y-z sections focus at their own lengths
    for axis, f in ((0, 3.0), (1, 12.0)):
        o = np.zeros(3)
        o[axis] = 1.7
        o[2] = 50.0
        t, p, g = _hit(2, (3.0, 12.0), o, np.array([0.0, 0.0, -1.0]), 49.0)
        r = _reflect(np.array([0.0, 0.0, -1.0]), g)
        s = -p[axis] / r[axis]
        assert p[2] + s * r[2] == pytest.approx(f, rel=1e-12)


def test_parabolic_cylinder_focal_line():
    f = 4.0
    rng = np.random.default_rng(3)
    for _ in range(50):
        o = np.array([rng.uniform(-10, 10), rng.uniform(-6, 6), 20.0])
        d = np.array([0.0, 0.0, -1.0])
        t, p, g = _hit(4, (f, 0.0), o, d, 19.0)
        r = _reflect(d, g)
        assert abs(r[0]) < 1e-15  # stays in its x = const plane
        s = -p[1] / r[1] if abs(r[1]) > 1e-12 else (f - p[2]) / r[2]
        assert p[2] + s * r[2] == pytest.approx(f, rel=1e-12)


@pytest.mark.parametrize("img,real", [(1.0, 9.0), (100.0, 500.0), (2.0, 3.0)])
def test_hyperboloid_focus_to_focus(img, real):
    """Foci at z = img and z = -real in the quadric frame (vertex at 0): a ray
    aimed at one focus leaves the surface through the other."""
    rng = np.random.default_rng(4)
    f_in = np.array([0.0, 0.0, img])
    f_out = np.array([0.0, 0.0, -real])
    for _ in range(40):
        xy = rng.uniform(-img, img, 2)
        h = np.array([xy[0], xy[1], O.quadric_z(3, (img, real), xy)])
        d = f_out - h
        d /= np.linalg.norm(d)
        o = h - 2 * img * d
        t, p, g = _hit(3, (img, real), o, d, 2 * img)
        assert np.allclose(p, h, atol=1e-9 * (img + real))
        r = _reflect(d, g)
        v = f_in - p
        assert np.linalg.norm(np.cross(r, v)) <= 1e-9 * np.linalg.norm(v)


def test_hemisphere_radius():
    R = 3.0
    rng = np.random.default_rng(5)
    for _ in range(30):
        o = np.array([*rng.uniform(-1.5, 1.5, 2), 10.0])
        t, p, g = _hit(5, (R, 0.0), o, np.array([0.0, 0.0, -1.0]), 10.0)
        assert np.linalg.norm(p - [0, 0, R]) == pytest.approx(R, rel=1e-13)
        assert np.allclose(g / np.linalg.norm(g), ([0, 0, R] - p) / R, atol=1e-12)


# Frame and sun directions ---------------------------------------------------------
def test_basis_is_orthonormal_with_dir_as_z():
    rng = np.random.default_rng(6)
    for _ in range(50):
        d = rng.normal(size=3)
        B = O.basis(d)
        assert np.allclose(B.T @ B, np.eye(3), atol=1e-14)
        assert np.allclose(B[:, 2], d / np.linalg.norm(d), atol=1e-15)
        assert np.linalg.det(B) == pytest.approx(1.0, abs=1e-14)


def _angles(v, d):
    B = O.basis(d)
    loc = B.T @ v
    sin2 = loc[0] ** 2 + loc[1] ** 2
    phi = math.atan2(loc[1], loc[0]) % (2 * math.pi)
    return sin2, phi, loc


@pytest.mark.parametrize("half", [4.65e-3, 0.3, 1.0])
def test_pillbox_inverse_cdf(half):
    d = np.array([-0.5, 0.2, -0.84])
    rng = np.random.default_rng(7)
    for _ in range(100):
        u = rng.uniform(1e-9, 1, 2)
        v, used = O.sun_sample(1, half, d, u)
        assert used == 2
        assert np.linalg.norm(v) == pytest.approx(1, abs=1e-15)
        sin2, phi, loc = _angles(v, d)
        assert sin2 == pytest.approx(u[0] * math.sin(half) ** 2, rel=1e-9)
        assert loc[2] > 0
        assert math.cos(phi - 2 * math.pi * u[1]) == pytest.approx(1, abs=1e-9)


@pytest.mark.parametrize("sigma", [2.51e-3, 0.01, 0.4])
def test_gaussian_inverse_cdf(sigma):
    d = np.array([0.1, -0.3, -0.95])
    rng = np.random.default_rng(8)
    for _ in range(100):
        u = rng.uniform(1e-9, 1, 2)
        v, used = O.sun_sample(2, sigma, d, u)
        assert used == 2
        sin2, phi, _ = _angles(v, d)
        st = min(1.0, sigma * math.sqrt(-2 * math.log(u[0])))
        assert math.sqrt(sin2) == pytest.approx(st, rel=1e-9)
        assert math.cos(phi - 2 * math.pi * u[1]) == pytest.approx(1, abs=1e-9)


@pytest.mark.parametrize("csr", [0.01, 0.05, 0.1, 0.2])
def test_buie_density_and_samples(csr):
    # the density over theta in [0, 0.0436] (solar disc + circumsolar) integrates to ~1
    th = np.linspace(0, 0.0436, 200001)
    pdf = np.array([O.buie_pdf_theta(csr, t) for t in th[::20]])
    mass = np.trapezoid(pdf, th[::20])
    assert mass == pytest.approx(1.0, rel=2e-3)
    # Kolmogorov-Smirnov of the sampled angles against the density's CDF
    d = np.array([0.0, 0.0, -1.0])
    rng = np.random.default_rng(int(csr * 1000))
    n = 4000
    thetas = np.empty(n)
    for i in range(n):
        v, _ = O.sun_sample(3, csr, d, rng.uniform(0, 1, 64))
        thetas[i] = math.acos(min(1.0, -v[2]))
    cdf = np.concatenate([[0], np.cumsum(0.5 * (pdf[1:] + pdf[:-1]) * np.diff(th[::20]))])
    cdf /= cdf[-1]
    emp = np.searchsorted(np.sort(thetas), th[::20], side="right") / n
    assert np.max(np.abs(emp - cdf)) < 1.63 / math.sqrt(n)  # alpha = 1%


# Spectra ----------------------------------------------------------------------------
def test_spectrum_lookup_reference_values():
    # src/test_ssol_data.c:49-76: wavelengths i+1, intensities i+2, i in [0, 10)
    wl = np.arange(1, 11, dtype=float)
    val = wl + 1
    for i in range(10):
        assert O.spectrum_interpolate(wl, val, float(i + 1)) == float(i + 2)
    assert abs(O.spectrum_interpolate(wl, val, 1.5) - 2.5) <= 1e-6
    assert abs(O.spectrum_interpolate(wl, val, 1.25) - 2.25) <= 1e-6
    assert O.spectrum_interpolate(wl, val, 0.5) == 2
    assert O.spectrum_interpolate(wl, val, 0.1) == 2
    assert O.spectrum_interpolate(wl, val, 10) == 11
    assert O.spectrum_interpolate(wl, val, 10.1) == 11


def test_piecewise_linear_inverse_cdf():
    wl = np.array([0.3, 0.5, 0.8, 1.2, 2.0])
    val = np.array([0.2, 1.0, 0.7, 0.7, 0.0])
    area = 0.5 * (val[1:] + val[:-1]) * np.diff(wl)
    total = area.sum()

    def cdf(x):
        i = min(np.searchsorted(wl, x, side="right") - 1, len(wl) - 2)
        s = x - wl[i]
        b = (val[i + 1] - val[i]) / (wl[i + 1] - wl[i])
        return (area[:i].sum() + val[i] * s + 0.5 * b * s * s) / total

    for u in np.linspace(0.0005, 0.9995, 97):
        x = O.piecewise_linear_sample(wl, val, u)
        assert wl[0] <= x <= wl[-1]
        assert cdf(x) == pytest.approx(u, abs=1e-12)
    assert O.piecewise_linear_sample(np.array([0.55]), np.array([1.0]), 0.7) == 0.55


# BSDFs --------------------------------------------------------------------------------
def _dirs(rng, n):
    N = np.array([0.2, -0.1, 0.97])
    N /= np.linalg.norm(N)
    out = []
    for _ in range(n):
        w = rng.normal(size=3)
        w /= np.linalg.norm(w)
        if w @ N < 0.05:
            w = w - 2 * (w @ N) * N + 0.1 * N
            w /= np.linalg.norm(w)
        out.append(w)
    return N, out


def test_specular_mirror_closed_form():
    rng = np.random.default_rng(9)
    N, ws = _dirs(rng, 50)
    m = S.mirror(0.87)
    for wi in ws:
        w, wo, t = O.bsdf_sample(m, wi, N, rng.uniform(0, 1, 4))
        assert w == 0.87 and t == 1
        assert np.allclose(wo, 2 * (wi @ N) * N - wi, atol=1e-15)


def test_lambertian_cosine_moments():
    rng = np.random.default_rng(10)
    N, _ = _dirs(rng, 1)
    m = S.matte(0.3)
    cos = []
    for _ in range(20000):
        w, wo, t = O.bsdf_sample(m, N, N, rng.uniform(0, 1, 2))
        assert w == 0.3
        assert np.linalg.norm(wo) == pytest.approx(1, abs=1e-12)
        cos.append(wo @ N)
    cos = np.array(cos)
    assert cos.min() >= 0
    # cosine-weighted hemisphere: E[cos] = 2/3, E[cos^2] = 1/2
    assert cos.mean() == pytest.approx(2 / 3, abs=4 * np.sqrt(1 / 18 / len(cos)))
    assert (cos ** 2).mean() == pytest.approx(0.5, abs=4 * np.sqrt(1 / 12 / len(cos)))


def test_dielectric_fresnel_normal_incidence():
    n1, n2 = 1.0, 1.5
    R = ((n1 - n2) / (n1 + n2)) ** 2
    rng = np.random.default_rng(11)
    N = np.array([0.0, 0.0, 1.0])
    m = S.dielectric(S.Medium(), S.Medium(refractive_index=S.Data(1.5)))
    n = 40000
    refl = 0
    for _ in range(n):
        w, wo, t = O.bsdf_sample(m, N, N, rng.uniform(0, 1, 1), eta_out=n1, eta_in=n2)
        assert w == 1
        if t == 1:
            refl += 1
            assert np.allclose(wo, N)
        else:
            assert np.allclose(wo, -N)
    assert refl / n == pytest.approx(R, abs=4 * math.sqrt(R * (1 - R) / n))


def test_dielectric_snell_and_total_internal_reflection():
    N = np.array([0.0, 0.0, 1.0])
    m = S.dielectric(S.Medium(), S.Medium())
    th = 0.6
    wi = np.array([math.sin(th), 0.0, math.cos(th)])
    # u = 0.999 forces refraction unless F ~ 1
    w, wo, t = O.bsdf_sample(m, wi, N, [0.999], eta_out=1.0, eta_in=1.5)
    assert t == 2
    assert math.hypot(wo[0], wo[1]) == pytest.approx(math.sin(th) / 1.5, rel=1e-12)
    assert wo[2] < 0 and wo[0] < 0
    # from glass to air beyond the critical angle: always reflected
    th = 0.8
    wi = np.array([math.sin(th), 0.0, math.cos(th)])
    w, wo, t = O.bsdf_sample(m, wi, N, [0.999], eta_out=1.5, eta_in=1.0)
    assert t == 1 and np.allclose(wo, 2 * wi[2] * N - wi)


def test_thin_slab_series():
    R = 0.04
    N = np.array([0.0, 0.0, 1.0])
    m = S.thin_dielectric(S.Medium(), S.Medium(refractive_index=S.Data(1.5)), 3e-3)
    # lossless slab: weight 1, reflection share 2R/(1+R)
    w, wo, t = O.bsdf_sample(m, N, N, [0.0], slab_eta=1.5)
    assert w == pytest.approx(1.0, abs=1e-15) and t == 1
    Rt = 2 * R / (1 + R)
    _, _, t1 = O.bsdf_sample(m, N, N, [Rt * 0.999], slab_eta=1.5)
    _, wo2, t2 = O.bsdf_sample(m, N, N, [Rt * 1.001], slab_eta=1.5)
    assert t1 == 1 and t2 == 2 and np.allclose(wo2, -N)
    # absorbing slab at normal incidence: T = (1-R)^2 A / (1 - R^2 A^2)
    k = 50.0
    A = math.exp(-k * 3e-3)
    w, _, _ = O.bsdf_sample(m, N, N, [0.999], slab_eta=1.5, slab_abs=k)
    Tt = (1 - R) ** 2 * A / (1 - R * R * A * A)
    Rr = R + (1 - R) ** 2 * R * A * A / (1 - R * R * A * A)
    assert w == pytest.approx(Tt + Rr, rel=1e-12)


@pytest.mark.parametrize("distrib", [0, 1])
def test_glossy_mirror_limits(distrib):
    rng = np.random.default_rng(12 + distrib)
    N, ws = _dirs(rng, 20)
    sharp = S.mirror(0.92, roughness=1e-9, distribution=distrib)
    for wi in ws:
        w, wo, t = O.bsdf_sample(sharp, wi, N, rng.uniform(0, 1, 64))
        assert w == 0.92
        assert np.allclose(wo, 2 * (wi @ N) * N - wi, atol=1e-8)
    # with roughness m the half vector's tan(theta) is m sqrt(-ln(1-u)) (Beckmann)
    # or m sqrt(u) (pillbox): the spread of wo about the mirror direction
    # grows with m, and every direction leaves the surface
    m = 0.05
    rough = S.mirror(1.0, roughness=m, distribution=distrib)
    wi = N
    dev = []
    for _ in range(4000):
        w, wo, _ = O.bsdf_sample(rough, wi, N, rng.uniform(0, 1, 64))
        assert wo @ N > 0 and w == 1.0
        dev.append(math.acos(min(1.0, wo @ N)) / 2)  # = half-vector angle at normal incidence
    tan2 = np.tan(np.array(dev)) ** 2
    want = m * m if distrib == 0 else m * m / 2   # E[tan^2] = m^2 (Beckmann), m^2/2 (pillbox)
    assert tan2.mean() == pytest.approx(want, rel=0.06)
