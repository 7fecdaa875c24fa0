"""The GPU library on the scenes of extra_scenes.py, against their closed
forms and against the oracle:

- failed paths (src/ssol_material.c:97-101) and max_failed
  (src/ssol_solver.c:1229-1247): failures left out of N; with the same
  streams on equal-area sampled triangles the GPU fails exactly the
  realisations the oracle fails; above the limit the solve returns
  RES_BAD_OP with its estimator, and a wavefront narrower than N starts no
  realisation after the limit is hit (chunked: no further chunk; refilled in
  place: no further refill);
- SUB, non-convex AND/SUB carvings and a mirror with holes
  (src/ssol_shape.c:579-652);
- the hemisphere quadric (src/ssol_shape.c:386-435, 1089-1111): absorbing
  bowls, a focusing mirror bowl and a deep mirror bowl whose rays meet it
  again (the trace skips the instance of a punched shape a ray leaves only
  where the filter could accept no candidate on it).
"""
import pytest

import extra_scenes as X
import ref_scenes as R
import ssol_ctypes as L
from conftest import within

pytestmark = pytest.mark.gpu

N_BIG = 2**62
NARROW = 4096


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("one", [False, True])
@pytest.mark.parametrize("slots", [0, NARROW])
def test_failed_paths_left_out_of_n(gpu, precision, one, slots):
    n = 200_000
    scn, h = X.failing_dielectric(one)
    sc = L.Scene(scn, precision=precision, path_slots=slots)
    e = sc.solve(n, seed=3, max_failed=N_BIG)

    class Run:  # the check reads the reference getters
        res = e.res
        failed_count = e.failed_count
        realisation_count = e.realisation_count
        global_ = e.global_

        @staticmethod
        def receiver(inst):
            return e.receiver(inst, 1)

    f = X.check_failures_unlimited(Run, n, h, rel=1e-9 if precision == "fp64" else 1e-6)
    o = R.Run("oracle", scn, n, seed=3, max_failed=N_BIG)
    if precision == "fp64":
        # same Threefry streams, equal-area triangles: the same realisations fail
        assert f == o.failed_count
    # (the FP32 form draws its start points from its own 24-bit uniforms:
    # a binomial half of N, checked above)
    sc.close()


@pytest.mark.parametrize("one", [False, True])
def test_max_failed_default_width(gpu, one):
    """Every realisation is in flight at once: all of them run, the limit
    turns the result into RES_BAD_OP, the estimator is returned."""
    n = 100_000
    scn, h = X.failing_dielectric(one)
    sc = L.Scene(scn)
    for k in (0, 1, 50):
        e = sc.solve(n, seed=3, max_failed=k)
        assert e.res == 5
        assert e.failed_count >= max(k, 1)
        assert e.realisation_count + e.failed_count == n
        a = e.global_()["absorbed_by_receivers"]
        assert abs(a[0] - X.FAILING_ABSORBED) <= 1e-9 * X.FAILING_ABSORBED
    sc.close()


def test_max_failed_stops_further_chunks(gpu):
    """Two sampled instances: a narrow wavefront runs the solve in chunks of
    NARROW realisations; the limit is hit in the first, which completes, and
    no other chunk starts (the reference skips every realisation not yet
    started)."""
    n = 200_000
    scn, h = X.failing_dielectric(False)
    sc = L.Scene(scn, path_slots=NARROW)
    e = sc.solve(n, seed=3, max_failed=10)
    assert e.res == 5
    assert e.stats.wavefronts == 1 and not e.stats.refilled
    assert e.realisation_count + e.failed_count == NARROW
    assert e.failed_count >= 10
    # and without a limit every chunk runs
    e = sc.solve(n, seed=3, max_failed=N_BIG)
    assert e.res == 0 and e.realisation_count + e.failed_count == n
    assert e.stats.wavefronts == -(-n // NARROW)
    sc.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_max_failed_stops_refills(gpu, precision):
    """One sampled instance: a narrow wavefront refills slots in place; once
    the limit is hit no slot takes a new realisation."""
    n = 400_000
    scn, h = X.failing_dielectric(True)
    sc = L.Scene(scn, precision=precision, path_slots=NARROW)
    e = sc.solve(n, seed=3, max_failed=10)
    assert e.res == 5 and e.stats.refilled
    done = e.realisation_count + e.failed_count
    assert NARROW <= done < n // 4, done
    assert e.failed_count >= 10
    sc.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("which", ["sub", "nonconvex"])
def test_carved_planes(gpu, precision, which):
    scn, h, area = (X.sub_carved_plane if which == "sub" else X.nonconvex_and_plane)()
    r = R.Run("gpu32" if precision == "fp32" else "gpu", scn, 100_000, seed=1)
    X.check_absorbing_plane(r, 100_000, area, rel=1e-9 if precision == "fp64" else 1e-6)
    # the same polygons, cut in another order: equal up to float rounding
    o = R.Run("oracle", scn, 10_000, seed=1)
    assert r.sampled_area == pytest.approx(o.sampled_area, rel=1e-7)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_mirror_with_holes(gpu, precision):
    scn, h, area = X.mirror_with_hole()
    n = 2_000_000
    r = R.Run("gpu32" if precision == "fp32" else "gpu", scn, n, seed=2)
    X.check_mirror_with_hole(r, n, area, h)
    o = R.Run("oracle", scn, n, seed=2)
    g, og = r.global_(), o.global_()
    for k in ("shadowed", "missing", "absorbed_by_receivers", "cos_factor"):
        floor = (1e-6 if precision == "fp32" else 1e-12) * abs(og[k][0])   # SE 0 for cos
        ok, tol = within(g[k], og[k], floor=floor)
        assert ok, (k, g[k], og[k], tol)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("r,slices", [(None, None), (4.0, None), (10.0, 9)])
def test_hemisphere_absorber(gpu, precision, r, slices):
    scn, h, area, cap = X.hemisphere_absorber(10.0, r, slices)
    run = R.Run("gpu32" if precision == "fp32" else "gpu", scn, 100_000, seed=4)
    X.check_hemisphere_absorber(run, 100_000, area, cap)
    o = R.Run("oracle", scn, 10_000, seed=4)
    assert run.sampled_area == pytest.approx(o.sampled_area, rel=1e-7)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_hemisphere_mirror_focus(gpu, precision):
    scn, h, area = X.hemisphere_mirror()
    run = R.Run("gpu32" if precision == "fp32" else "gpu", scn, 100_000, seed=5)
    X.check_hemisphere_mirror(run, 100_000, area, h)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_hemisphere_deep_mirror_rehits(gpu, precision):
    """Rays leaving the bowl meet it again: shadow rays toward the sun are
    stopped by the far rim, reflections hit the bowl several times. The GPU
    agrees with the oracle, whose trace enters every instance it crosses,
    on the bowl's incoming and absorbed flux (both sides) and the global
    sinks."""
    scn, h = X.hemisphere_deep_mirror()
    n = 400_000
    run = R.Run("gpu32" if precision == "fp32" else "gpu", scn, n, seed=6)
    o = R.Run("oracle", scn, n, seed=6)
    # a path traces its shadow ray, then (unless shadowed, a tenth of them)
    # its first reflection, about 1.9 segments without re-hits
    assert run.e.stats.segments > 2.3 * n
    g, og = run.global_(), o.global_()
    assert og["shadowed"][0] > 0.05 * og["absorbed_by_receivers"][0]   # the rim shadows
    for k in ("shadowed", "missing", "absorbed_by_receivers"):
        ok, tol = within(g[k], og[k], k=4.0)
        assert ok, (k, g[k], og[k], tol)
    for side in (X.S.FRONT, X.S.BACK):
        rg, ro = run.receiver(h["bowl"], side), o.receiver(h["bowl"], side)
        for f in ("incoming_flux", "absorbed_flux"):
            ok, tol = within(rg[f], ro[f], k=4.0)
            assert ok, (side, f, rg[f], ro[f], tol)
