"""Pins the oracle's BSDF samplers away from normal incidence, where the
forms that normal incidence cannot tell apart differ.

The oracle restates the Star-SF models that src/ssol_material.c:82-268
instantiates (Star-SF 0.6 is not in the image), so its samplers are pinned
here against the physics they implement, written in forms other than the
oracle's:

- Fresnel reflectance of a dielectric interface (the dielectric model,
  src/ssol_material.c:92-117): the sine and tangent forms of the Fresnel
  equations, Rs = (sin(ti - tt) / sin(ti + tt))^2 and Rp = (tan(ti - tt) /
  tan(ti + tt))^2, with Snell's law for tt; Rp = 0 at Brewster's angle
  tan(ti) = n2 / n1; textbook values at 45 degrees for n = 1.5; reciprocity
  R(ti, n1 -> n2) = R(tt, n2 -> n1). The oracle computes the amplitude
  (cosine) form, so a swapped index in rp, a dropped term or a wrong
  transmitted cosine fails one of these.
- Thin absorbing slab (src/ssol_material.c:150-180) at oblique incidence:
  the bounce-by-bounce sum of the incoherent internal reflections with path
  length thickness / cos(tt) per crossing, against the sampled weight and the
  reflection share.
- Glossy microfacet mirror (src/ssol_material.c:182-268, microfacet2 with a
  Beckmann or pillbox distribution) at 30 and 70 degrees incidence: the
  accepted half vectors must follow D(h) (h . N) restricted to the region
  where the reflected direction leaves the surface (the resampling branch
  microfacet2 implies), checked with Kolmogorov-Smirnov tests of the radial
  slope and the azimuth about the plane of incidence against CDFs integrated
  numerically from the slope-space densities (Beckmann: Gaussian slopes of
  variance m^2/2 per axis; pillbox: slopes uniform on the disk of radius m).

The sampler returns the reflection or transmission through a uniform it
compares with the reflectance; the reflectance is read back by bisecting
that uniform.
"""
import math

import numpy as np
import pytest

import oracle_ctypes as O
import ssol_scene as S

N = np.array([0.0, 0.0, 1.0])


def _wi(theta):
    return np.array([math.sin(theta), 0.0, math.cos(theta)])


def _threshold(sample):
    """The u at which sample(u) switches from lobe 1 (reflection) to lobe 2."""
    lo, hi = 0.0, 1.0
    if sample(hi) == 1:
        return 1.0
    for _ in range(64):
        mid = 0.5 * (lo + hi)
        if sample(mid) == 1:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def _fresnel_sine_tangent(ti, n1, n2):
    st = n1 / n2 * math.sin(ti)
    if st >= 1:
        return 1.0, 1.0, None
    tt = math.asin(st)
    rs = (math.sin(ti - tt) / math.sin(ti + tt)) ** 2
    rp = (math.tan(ti - tt) / math.tan(ti + tt)) ** 2
    return rs, rp, tt


def _dielectric_reflectance(theta, n1, n2):
    m = S.dielectric(S.Medium(), S.Medium())
    wi = _wi(theta)
    return _threshold(lambda u: O.bsdf_sample(m, wi, N, [u], eta_out=n1, eta_in=n2)[2])


@pytest.mark.parametrize("deg", [10, 30, 45, 60, 75, 85])
@pytest.mark.parametrize("n1,n2", [(1.0, 1.5), (1.5, 1.0), (1.0, 2.4), (1.33, 1.5)])
def test_fresnel_oblique_sine_tangent_forms(deg, n1, n2):
    ti = math.radians(deg)
    rs, rp, tt = _fresnel_sine_tangent(ti, n1, n2)
    F = _dielectric_reflectance(ti, n1, n2)
    if tt is None:
        assert F == 1.0
    else:
        assert F == pytest.approx(0.5 * (rs + rp), rel=1e-10, abs=1e-15)


def test_fresnel_brewster_angle():
    for n1, n2 in ((1.0, 1.5), (1.5, 1.0), (1.0, 2.4)):
        tb = math.atan(n2 / n1)
        rs, rp, tt = _fresnel_sine_tangent(tb, n1, n2)
        assert tt == pytest.approx(math.pi / 2 - tb, abs=1e-12)   # reflected _|_ refracted
        assert rp == pytest.approx(0.0, abs=1e-28)
        # only the s polarisation reflects: a swapped index in rp doubles F here
        assert _dielectric_reflectance(tb, n1, n2) == pytest.approx(0.5 * rs, rel=1e-10)
    # and Rp has its zero there: F/Rs is smallest at Brewster's angle
    tb = math.atan(1.5)
    ratio = [(_dielectric_reflectance(t, 1.0, 1.5) / _fresnel_sine_tangent(t, 1.0, 1.5)[0], t)
             for t in np.linspace(tb - 0.2, tb + 0.2, 41)]
    assert min(ratio)[1] == pytest.approx(tb, abs=0.011)


def test_fresnel_textbook_values_45_degrees():
    # n = 1.5 at 45 degrees: Rs = 0.0920, Rp = 0.0085 (Hecht, Optics, fig. 4.x)
    F = _dielectric_reflectance(math.radians(45), 1.0, 1.5)
    assert F == pytest.approx(0.5 * (0.0920 + 0.0085), abs=1e-4)
    # and the critical angle from glass: asin(1/1.5) = 41.81 degrees
    tc = math.asin(1 / 1.5)
    assert _dielectric_reflectance(tc + 1e-6, 1.5, 1.0) == 1.0
    assert _dielectric_reflectance(tc - 1e-3, 1.5, 1.0) < 1.0


@pytest.mark.parametrize("deg", [20, 40, 60, 80])
def test_fresnel_reciprocity(deg):
    ti = math.radians(deg)
    tt = math.asin(math.sin(ti) / 1.5)
    assert _dielectric_reflectance(ti, 1.0, 1.5) == pytest.approx(
        _dielectric_reflectance(tt, 1.5, 1.0), rel=1e-10)


def test_dielectric_refraction_direction_oblique():
    m = S.dielectric(S.Medium(), S.Medium())
    for deg in (15, 45, 75):
        ti = math.radians(deg)
        w, wo, t = O.bsdf_sample(m, _wi(ti), N, [0.999999], eta_out=1.0, eta_in=1.5)
        assert t == 2 and w == 1
        tt = math.asin(math.sin(ti) / 1.5)
        # transmitted ray: in the plane of incidence, on the other side, at tt
        assert wo == pytest.approx([-math.sin(tt), 0.0, -math.cos(tt)], abs=1e-12)


def _slab_by_bounces(ti, n, k, d):
    """Reflected and transmitted power of an absorbing slab, summed bounce by
    bounce: each crossing has length d / cos(tt) and keeps exp(-k d / cos tt)."""
    rs, rp, tt = _fresnel_sine_tangent(ti, 1.0, n)
    R = 0.5 * (rs + rp)
    T = 1 - R
    A = math.exp(-k * d / math.cos(tt))
    refl = R
    trans = 0.0
    inside = T          # power inside just after entering, heading down
    for _ in range(400):
        inside *= A     # cross the slab
        trans += inside * T
        inside *= R     # internal reflection at the far face
        inside *= A     # cross back
        refl += inside * T
        inside *= R
    return refl, trans


@pytest.mark.parametrize("deg", [30, 60])
@pytest.mark.parametrize("k", [0.0, 80.0])
def test_thin_slab_oblique(deg, k):
    d = 3e-3
    ti = math.radians(deg)
    m = S.thin_dielectric(S.Medium(), S.Medium(refractive_index=S.Data(1.5)), d)
    refl, trans = _slab_by_bounces(ti, 1.5, k, d)
    wi = _wi(ti)
    w, wo, t = O.bsdf_sample(m, wi, N, [0.999999], slab_eta=1.5, slab_abs=k)
    assert w == pytest.approx(refl + trans, rel=1e-12)
    assert t == 2 and wo == pytest.approx(-wi, abs=1e-15)   # no lateral shift
    share = _threshold(lambda u: O.bsdf_sample(m, wi, N, [u], slab_eta=1.5, slab_abs=k)[2])
    assert share == pytest.approx(refl / (refl + trans), rel=1e-10)
    w, wo, t = O.bsdf_sample(m, wi, N, [0.0], slab_eta=1.5, slab_abs=k)
    assert t == 1 and wo == pytest.approx(2 * wi[2] * N - wi, abs=1e-15)


def test_thin_slab_path_length_matters():
    # the same slab at 60 degrees absorbs more than at normal incidence by the
    # ratio of path lengths 1/cos(tt): T60 / T0 compared with the series
    d, k = 3e-3, 80.0
    m = S.thin_dielectric(S.Medium(), S.Medium(refractive_index=S.Data(1.5)), d)
    for deg in (45, 60, 75):
        ti = math.radians(deg)
        tt = math.asin(math.sin(ti) / 1.5)
        refl, trans = _slab_by_bounces(ti, 1.5, k, d)
        w = O.bsdf_sample(m, _wi(ti), N, [0.999999], slab_eta=1.5, slab_abs=k)[0]
        assert 1 - w == pytest.approx(1 - (refl + trans), rel=1e-10)
        # a slab that used the thickness d for a crossing, not d / cos(tt),
        # would keep measurably more power
        r_short, t_short = _slab_by_bounces(ti, 1.5, k * math.cos(tt), d)
        assert (r_short + t_short) - w > 1e-3


# Glossy mirror: accepted half vectors -----------------------------------------
def _slope_density(distrib, m, r):
    if distrib == 0:          # Beckmann: D(h)(h.N) d(omega) = exp(-|s|^2/m^2)/(pi m^2) d^2 s
        return np.exp(-(r / m) ** 2) / (math.pi * m * m)
    return np.where(r <= m, 1.0 / (math.pi * m * m), 0.0)   # pillbox


def _accepted_cdfs(distrib, m, theta_i):
    """Radial-slope and azimuth CDFs of the accepted half vectors, integrated
    over a polar grid in slope space."""
    rmax = 7 * m if distrib == 0 else m
    nr, nphi = 3000, 2048
    r = (np.arange(nr) + 0.5) * (rmax / nr)
    phi = (np.arange(nphi) + 0.5) * (2 * math.pi / nphi) - math.pi
    R_, P_ = np.meshgrid(r, phi, indexing="ij")
    sx, sy = R_ * np.cos(P_), R_ * np.sin(P_)
    inv = 1 / np.sqrt(1 + sx * sx + sy * sy)
    h = np.stack([sx * inv, sy * inv, inv], axis=-1)
    wi = _wi(theta_i)
    c = h @ wi
    wo_z = 2 * c * h[..., 2] - wi[2]
    ok = (c > 0) & (wo_z > 0)
    dens = _slope_density(distrib, m, R_) * R_ * ok          # polar Jacobian r
    pr = dens.sum(axis=1)
    pphi = dens.sum(axis=0)
    cdf_r = np.concatenate([[0.0], np.cumsum(pr)]) / pr.sum()
    cdf_phi = np.concatenate([[0.0], np.cumsum(pphi)]) / pphi.sum()
    edges_r = np.linspace(0, rmax, nr + 1)
    edges_phi = np.linspace(-math.pi, math.pi, nphi + 1)
    accept = dens.sum() * (rmax / nr) * (2 * math.pi / nphi)
    return (edges_r, cdf_r), (edges_phi, cdf_phi), accept


def _ks(samples, edges, cdf):
    x = np.sort(samples)
    F = np.interp(x, edges, cdf)
    n = len(x)
    i = np.arange(1, n + 1)
    return max(np.max(i / n - F), np.max(F - (i - 1) / n))


@pytest.mark.parametrize("distrib,m", [(0, 0.15), (1, 0.3)])
@pytest.mark.parametrize("deg", [30, 70])
def test_glossy_half_vectors_oblique_ks(distrib, m, deg):
    theta = math.radians(deg)
    wi = _wi(theta)
    mat = S.mirror(0.8, roughness=m, distribution=distrib)
    rng = np.random.default_rng(100 + 10 * distrib + deg)
    n = 12000
    rad, azi = np.empty(n), np.empty(n)
    for k in range(n):
        w, wo, t = O.bsdf_sample(mat, wi, N, rng.uniform(0, 1, 1024))
        # resampled, never killed: the weight is the reflectivity, the
        # direction above the surface
        assert w == 0.8 and t == 1 and wo[2] > 0
        h = wi + wo
        h /= np.linalg.norm(h)
        rad[k] = math.hypot(h[0], h[1]) / h[2]
        azi[k] = math.atan2(h[1], h[0])
    (er, cr), (ep, cp), accept = _accepted_cdfs(distrib, m, theta)
    crit = 1.95 / math.sqrt(n)       # alpha = 0.001
    assert _ks(rad, er, cr) < crit
    assert _ks(azi, ep, cp) < crit
    if deg == 70:
        # the resampling branch is exercised: a sizeable share of the first
        # draws is rejected at this incidence
        assert accept < 0.95


def test_glossy_rejection_shapes_the_distribution():
    # at 70 degrees the accepted half vectors lean toward the incoming
    # direction (positive x slope) against the unrestricted distribution;
    # a sampler that skipped the restriction would be symmetric
    theta = math.radians(70)
    m = 0.3
    mat = S.mirror(1.0, roughness=m, distribution=1)
    rng = np.random.default_rng(7)
    xs = []
    for _ in range(6000):
        _, wo, _ = O.bsdf_sample(mat, _wi(theta), N, rng.uniform(0, 1, 1024))
        h = _wi(theta) + wo
        xs.append(h[0] / h[2])
    (_, _), (ep, cp), _ = _accepted_cdfs(1, m, theta)
    mid = np.interp(0.0, ep, cp)  # P(phi < 0) is 1/2 by symmetry in y
    assert mid == pytest.approx(0.5, abs=1e-6)
    assert np.mean(xs) > 0.01
