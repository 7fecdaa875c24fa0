"""Scenes beyond the reference's known-answer tests, each with a closed form,
and the checks that hold both solvers (the GPU library and the oracle) to it.

- failing_dielectric: half the sampled area reflects onto a dielectric whose
  outside medium is not the air the ray arrives through. The reference fails
  such a path (src/ssol_material.c:97-101, RES_BAD_OP), cancels its tallies
  and leaves it out of the realisation count (src/ssol_solver.c:1229-1247);
  the other half is absorbed by a receiver. Absorbed per counted realisation
  is therefore exactly 2 DNI cos 45 with SE 0, and the failures are a
  binomial half of N.
- carved planes: punched planes cut by SUB and by non-convex AND/SUB polygons
  (src/ssol_shape.c:579-652, Star-CPR clipping) under a vertical sun, on an
  absorbing material: the sampled area is the polygon algebra's area and
  the absorbed flux is DNI times it with SE 0 (the w0 formula of
  src/ssol_solver.c:284-299 with the proxy equal to the surface).
- mirror_with_hole: a mirror with a SUB hole over a sampled absorber; the
  absorber is shadowed exactly where the carved mirror covers it.
- hemisphere bowls (src/ssol_shape.c:386-435, 1089-1111): an absorbing bowl
  under a vertical sun absorbs DNI times its flat proxy's area with SE 0; a
  mirror bowl of small aperture sends every ray through two stacked virtual
  receivers near its paraxial focus R/2 (like src/test_ssol_solver4.c's
  paraboloid).
"""
from __future__ import annotations

import math

import numpy as np

import ref_scenes as R
import ssol_scene as S

PI = math.pi
DNI = 1000.0


def poly_area(xy):
    xy = np.asarray(xy, dtype=np.float64)
    x, y = xy[:, 0], xy[:, 1]
    return 0.5 * abs(np.dot(x, np.roll(y, -1)) - np.dot(np.roll(x, -1), y))


# Failed paths ---------------------------------------------------------------------
def failing_dielectric(one_instance: bool = False):
    """Two 1x1 m horizontal mirrors under the sun (1, 0, -1): A at the
    origin reflects onto a dielectric at (2, 0, 2) whose outside medium is
    glass (n = 1.5) while the ray arrives through vacuum, so A's paths fail;
    B at y = 10 reflects onto an absorbing receiver at (2, 10, 2). With
    one_instance both mirrors are shaded shapes of one sampled instance
    (one sampled instance: the narrow wavefront refills slots in place);
    otherwise they are two sampled instances (chunked wavefront)."""
    mirror = S.mirror(1.0)
    sq_a = S.square_mesh(0.5)
    sq_b = S.rect_mesh(np.float32(-0.5), np.float32(0.5), np.float32(9.5), np.float32(10.5))
    glass = S.Medium(refractive_index=S.Data(1.5))
    diel = S.dielectric(glass, S.Medium())
    d_inst = S.Instance(S.Object([(S.square_mesh(0.5), diel, diel)]), R.t_pitch(PI, 2, 0, 2),
                        sample=0)
    absorber = S.matte(0.0)
    rcv = S.Instance(S.Object([(S.square_mesh(0.5), absorber, absorber)]),
                     R.t_pitch(PI, 2, 10, 2), receiver_mask=S.FRONT, sample=0)
    if one_instance:
        mirrors = [S.Instance(S.Object([(sq_a, mirror, mirror), (sq_b, mirror, mirror)]),
                              S.IDENTITY)]
    else:
        mirrors = [S.Instance(S.Object([(sq_a, mirror, mirror)]), S.IDENTITY),
                   S.Instance(S.Object([(sq_b, mirror, mirror)]), S.IDENTITY)]
    sun = S.Sun("directional", (1, 0, -1), DNI, R.spectrum3())
    name = "failing_dielectric_" + ("one" if one_instance else "two")
    return S.Scene(mirrors + [d_inst, rcv], sun, name=name), {"receiver": rcv}


FAILING_ABSORBED = 2 * DNI * math.cos(PI / 4)


def check_failures_unlimited(run, n, handles, rel=1e-9):
    """max_failed above N: RES_OK, failures a binomial half of N, left out of
    N, absorbed per counted realisation exact (to `rel`: 1e-9 for FP64
    weights, 1e-6 for the FP32 form's)."""
    assert run.res == 0, run.res
    f = run.failed_count
    assert run.realisation_count + f == n
    assert abs(f - n / 2) <= 5 * math.sqrt(n / 4), f
    g = run.global_()
    a = g["absorbed_by_receivers"]
    assert abs(a[0] - FAILING_ABSORBED) <= rel * FAILING_ABSORBED, a
    assert a[2] <= 1e-6 * FAILING_ABSORBED, a   # zero up to rounding of sum w^2 / N - E^2
    assert abs(g["cos_factor"][0] - math.cos(PI / 4)) <= 1e3 * rel, g["cos_factor"]
    for k in ("shadowed", "missing", "other_absorbed", "extinguished_by_atmosphere"):
        assert g[k][0] == 0, (k, g[k])
    rc = run.receiver(handles["receiver"])
    for k in ("incoming_flux", "absorbed_flux"):
        assert abs(rc[k][0] - FAILING_ABSORBED) <= rel * FAILING_ABSORBED, (k, rc[k])
    return f


# Carving ----------------------------------------------------------------------------
def _absorbing_plane(carvings, name):
    quad = S.Punched(S.Q_PLANE, (0, 0), list(S.IDENTITY), carvings)
    m = S.matte(0.0)
    inst = S.Instance(S.Object([(quad, m, m)]), S.IDENTITY, receiver_mask=S.FRONT)
    sun = S.Sun("directional", (0, 0, -1), DNI, R.spectrum3())
    return S.Scene([inst], sun, name=name), {"plane": inst, "shape": quad}


SUB_SQUARE = np.array([[-3.0, -1.0], [-3.0, 1.0], [-1.0, 1.0], [-1.0, -1.0]])
SUB_TRIANGLE = np.array([[1.0, -2.0], [4.0, -2.0], [2.5, 2.0]])
# an L (non-convex, clockwise) and an arrowhead (non-convex, counter-clockwise)
L_SHAPE = np.array([[-4.0, -4.0], [-4.0, 4.0], [-1.0, 4.0], [-1.0, -1.0], [4.0, -1.0],
                    [4.0, -4.0]])
ARROW = np.array([[0.0, 0.0], [1.5, -1.0], [0.75, 0.0], [1.5, 1.0]])


def sub_carved_plane():
    """10x10 m AND square minus a 2x2 m square and a triangle (SUB)."""
    carv = [S.rect_carving(5, 5), S.Carving(S.SUB, SUB_SQUARE), S.Carving(S.SUB, SUB_TRIANGLE)]
    area = 100 - poly_area(SUB_SQUARE) - poly_area(SUB_TRIANGLE)
    return _absorbing_plane(carv, "sub_carved_plane") + (area,)


def nonconvex_and_plane():
    """An L-shaped (non-convex) AND carving, minus a non-convex arrow (SUB)
    that lies inside its lower arm."""
    arrow = ARROW + np.array([1.0, -2.5])
    carv = [S.Carving(S.AND, L_SHAPE), S.Carving(S.SUB, arrow)]
    area = poly_area(L_SHAPE) - poly_area(arrow)
    return _absorbing_plane(carv, "nonconvex_and_plane") + (area,)


def check_absorbing_plane(run, n, area, rel=1e-9):
    assert run.res == 0 and run.realisation_count == n
    assert abs(run.sampled_area - area) <= 1e-6 * area, (run.sampled_area, area)
    g = run.global_()
    a = g["absorbed_by_receivers"]
    assert abs(a[0] - DNI * run.sampled_area) <= rel * DNI * area, a
    assert a[2] <= 1e-6 * DNI * area, a
    assert abs(g["cos_factor"][0] - 1) <= 1e3 * rel


def mirror_with_hole():
    """A 6x6 m mirror at z = 1 with a 2x2 m SUB hole (off centre) and a
    non-convex arrow SUB, over a 10x10 m absorbing receiver at z = 0; both
    sampled, sun vertical. The receiver is shadowed where the carved mirror
    covers it: E[shadowed] = DNI * mirror area."""
    hole = np.array([[0.5, 0.5], [0.5, 2.5], [2.5, 2.5], [2.5, 0.5]])
    arrow = ARROW + np.array([-2.5, -1.5])   # x in [-2.5, -1], y in [-2.5, -0.5]
    quad = S.Punched(S.Q_PLANE, (0, 0), list(S.IDENTITY),
                     [S.rect_carving(3, 3), S.Carving(S.SUB, hole), S.Carving(S.SUB, arrow)])
    m = S.mirror(1.0)
    a = S.matte(0.0)
    mir = S.Instance(S.Object([(quad, m, m)]), R.t_move(0, 0, 1))
    rcv = S.Instance(S.Object([(S.square_mesh(5.0), a, a)]), S.IDENTITY, receiver_mask=S.FRONT)
    sun = S.Sun("directional", (0, 0, -1), DNI, R.spectrum3())
    area = 36 - poly_area(hole) - poly_area(arrow)
    return S.Scene([mir, rcv], sun, name="mirror_with_hole"), {"mirror": mir, "receiver": rcv}, area


def check_mirror_with_hole(run, n, area, handles):
    assert run.res == 0 and run.realisation_count == n
    total = 100 + area
    assert abs(run.sampled_area - total) <= 1e-6 * total
    g = run.global_()
    sh = g["shadowed"]
    assert abs(sh[0] - DNI * area) <= 4 * sh[2], (sh, DNI * area)
    # what the mirror sends back up escapes; what it does not cover is absorbed
    rc = run.receiver(handles["receiver"])
    assert abs(rc["absorbed_flux"][0] - DNI * (100 - area)) <= 4 * rc["absorbed_flux"][2], rc
    assert abs(g["missing"][0] - DNI * area) <= 4 * g["missing"][2], g["missing"]


# Hemisphere -----------------------------------------------------------------------
def _slices(r, R_):
    return int(5.5 + 11 * r / R_)   # src/ssol_shape.c:1395-1403


def disk_carving(r, nverts):
    """Regular nverts-gon of radius r; with nverts a multiple of the slice
    count it holds the projected hemisphere mesh's outline exactly."""
    a = 2 * PI * np.arange(nverts) / nverts
    return S.Carving(S.AND, np.column_stack([r * np.cos(a), r * np.sin(a)]))


def hemisphere_proxy_area(radius, nslices):
    """Area of the regular nslices-gon inscribed in the carving circle: the
    outline of the projected hemisphere mesh."""
    return 0.5 * nslices * radius * radius * math.sin(2 * PI / nslices)


def hemisphere_absorber(R_=10.0, r=None, slices=None):
    """Absorbing hemisphere bowl z = R - sqrt(R^2 - x^2 - y^2), carved to a
    disk of radius r (the whole hemisphere when r is None), vertical sun."""
    rr = R_ if r is None else r
    n = slices if slices is not None else _slices(rr, R_)
    carv = [] if r is None else [disk_carving(r, 16 * n)]
    hint = S.SIZE_MAX if slices is None else slices
    bowl = S.Punched(S.Q_HEMISPHERE, (R_, 0.0), list(S.IDENTITY), carv, slices_hint=hint)
    m = S.matte(0.0)
    inst = S.Instance(S.Object([(bowl, m, m)]), S.IDENTITY, receiver_mask=S.FRONT | S.BACK)
    sun = S.Sun("directional", (0, 0, -1), DNI, R.spectrum3())
    return S.Scene([inst], sun, name="hemisphere_absorber"), {"bowl": inst}, \
        hemisphere_proxy_area(rr, n), cap_area(R_, rr)


def cap_area(R_, r):
    """Area of the spherical cap of radius R over a disk of radius r."""
    return 2 * PI * R_ * (R_ - math.sqrt(max(R_ * R_ - r * r, 0.0)))


def check_hemisphere_absorber(run, n, area, cap):
    """area: the flat proxy's (absorbed = DNI * area exactly); cap: the
    spherical cap's, which the lifted mesh (the estimator's sampled area)
    approaches from below."""
    assert run.res == 0 and run.realisation_count == n
    assert 0.75 * cap < run.sampled_area <= cap * (1 + 1e-6), (run.sampled_area, cap)
    g = run.global_()
    a = g["absorbed_by_receivers"]
    assert abs(a[0] - DNI * area) <= 1e-6 * DNI * area, (a, DNI * area)
    assert a[2] <= 1e-6 * DNI * area
    # cos factor: proxy area over the lifted mesh's area
    cf = g["cos_factor"]
    assert abs(cf[0] - area / run.sampled_area) <= 1e-6, (cf, area / run.sampled_area)


def hemisphere_mirror(R_=10.0, r=1.0):
    """Specular bowl of radius R carved to a disk of radius r = R/10 under a
    vertical sun; two stacked virtual 0.1x0.1 m receivers near the paraxial
    focus z = R/2 (every reflected ray crosses both: the spherical
    aberration at this aperture is R (1/(2 cos t) - 1/2) with sin t = r/R,
    a few mm here)."""
    n = _slices(r, R_)
    bowl = S.Punched(S.Q_HEMISPHERE, (R_, 0.0), list(S.IDENTITY), [disk_carving(r, 16 * n)])
    m = S.mirror(1.0)
    v = S.virtual()
    inst = S.Instance(S.Object([(bowl, m, m)]), S.IDENTITY)
    t_obj = S.Object([(S.square_mesh(0.05), v, v)])
    t1 = S.Instance(t_obj, R.t_pitch(PI, 0, 0, R_ / 2), receiver_mask=S.FRONT, sample=0)
    t2 = S.Instance(t_obj, R.t_pitch(PI, 0, 0, R_ / 2 + 0.02), receiver_mask=S.FRONT, sample=0)
    sun = S.Sun("directional", (0, 0, -1), DNI, R.spectrum3())
    return S.Scene([inst, t1, t2], sun, name="hemisphere_mirror"), {"t1": t1, "t2": t2}, \
        hemisphere_proxy_area(r, n)


def hemisphere_deep_mirror(R_=10.0, rho=0.8, zenith_deg=35.0):
    """The whole hemisphere bowl as a mirror of reflectivity rho under a sun
    zenith_deg off vertical: its rim shadows part of the bowl and reflected
    rays meet the bowl again, most of them several times, before they leave
    through the opening. The bowl is its own receiver on both sides, so its
    incoming flux counts every hit."""
    bowl = S.Punched(S.Q_HEMISPHERE, (R_, 0.0), list(S.IDENTITY), [], slices_hint=S.SIZE_MAX)
    m = S.mirror(rho)
    inst = S.Instance(S.Object([(bowl, m, m)]), S.IDENTITY, receiver_mask=S.FRONT | S.BACK)
    z = math.radians(zenith_deg)
    sun = S.Sun("directional", (math.sin(z), 0.0, -math.cos(z)), DNI, R.spectrum3())
    return S.Scene([inst], sun, name="hemisphere_deep_mirror"), {"bowl": inst}


def check_hemisphere_mirror(run, n, area, handles):
    assert run.res == 0 and run.realisation_count == n
    for t in ("t1", "t2"):
        rc = run.receiver(handles[t])
        assert abs(rc["incoming_flux"][0] - DNI * area) <= 1e-6 * DNI * area, (t, rc)
        assert rc["incoming_flux"][2] <= 1e-6 * DNI * area
