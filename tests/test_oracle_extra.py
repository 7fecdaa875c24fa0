"""The oracle on scenes with closed forms that the reference's own tests do
not cover: failed paths and the max_failed abort (src/ssol_solver.c:
1229-1247), SUB and non-convex carvings (src/ssol_shape.c:579-652) and the
hemisphere quadric (src/ssol_shape.c:386-435, 1089-1111). The scenes and
their checks are in extra_scenes.py; the GPU library is held to the same
checks in test_gpu_extra.py."""
import math

import pytest

import extra_scenes as X
import oracle_ctypes as O
import ref_scenes as R

N_BIG = 2**62


@pytest.mark.parametrize("one", [False, True])
def test_failures_left_out_of_n(one):
    scn, h = X.failing_dielectric(one)
    r = R.Run("oracle", scn, 40_000, seed=3, max_failed=N_BIG)
    X.check_failures_unlimited(r, 40_000, h)


def test_max_failed_aborts_at_the_limit():
    scn, h = X.failing_dielectric(False)
    o = O.Oracle(scn)
    # one thread: realisations run in order, so the solve stops right at the
    # k-th failure
    for k in (1, 7, 100):
        res = o.solve(40_000, seed=3, max_failed=k, nthreads=1)
        assert res == 5                       # RES_BAD_OP, results still returned
        assert o.failed_count == k
        assert 0 < o.realisation_count + k < 40_000
        a = o.global_()["absorbed_by_receivers"]
        if o.realisation_count:
            assert a[0] == pytest.approx(X.FAILING_ABSORBED, rel=1e-9)
    # max_failed = 0: the first failure aborts as well
    assert o.solve(40_000, seed=3, max_failed=0, nthreads=1) == 5 and o.failed_count == 1
    # several threads stop within a thread's share of the limit
    assert o.solve(40_000, seed=3, max_failed=10, nthreads=4) == 5
    assert 10 <= o.failed_count <= 13


def test_sub_carving_area_and_flux():
    scn, h, area = X.sub_carved_plane()
    r = R.Run("oracle", scn, 20_000, seed=1)
    X.check_absorbing_plane(r, 20_000, area)


def test_nonconvex_carvings_area_and_flux():
    scn, h, area = X.nonconvex_and_plane()
    r = R.Run("oracle", scn, 20_000, seed=1)
    X.check_absorbing_plane(r, 20_000, area)


def test_mirror_with_hole_shadows_its_area():
    scn, h, area = X.mirror_with_hole()
    r = R.Run("oracle", scn, 400_000, seed=2)
    X.check_mirror_with_hole(r, 400_000, area, h)


@pytest.mark.parametrize("r,slices", [(None, None), (4.0, None), (10.0, 9)])
def test_hemisphere_absorber(r, slices):
    scn, h, area, cap = X.hemisphere_absorber(10.0, r, slices)
    run = R.Run("oracle", scn, 20_000, seed=4)
    X.check_hemisphere_absorber(run, 20_000, area, cap)


def test_hemisphere_mirror_focus():
    scn, h, area = X.hemisphere_mirror()
    run = R.Run("oracle", scn, 20_000, seed=5)
    X.check_hemisphere_mirror(run, 20_000, area, h)
