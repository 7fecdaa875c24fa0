"""Path recording on the GPU (ssol_solve with a path tracker, reference
src/ssol_solver.c:898-1127 and the getters of src/ssol_estimator.c:270-322):
the reference's API contract (src/test_ssol_solver1.c:178-211), the exact
identities between recorded paths and the tallies of the same solve, path by
path equality with the oracle where both draw the same random numbers
(specular scenes), and the oracle's statistics where they do not (roulette,
glossy mirrors)."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_ctypes as O
import ref_scenes as R
import ssol_ctypes as L
import ssol_scene as S
import tracker_checks as TC

pytestmark = pytest.mark.gpu

RES_OK, RES_BAD_ARG = 0, 1


def test_tracker_api_contract(gpu):
    """src/test_ssol_solver1.c:178-211 restated"""
    scene = R.solver1("all")[0]
    lib = L.Scene(scene)
    est = lib.solve(1, seed=0, tracker=(-1.0, -1.0))
    f = L.lib()
    n = C.c_size_t()
    assert f.ssol_estimator_get_tracked_paths_count(None, None) == RES_BAD_ARG
    assert f.ssol_estimator_get_tracked_paths_count(est.h, None) == RES_BAD_ARG
    assert f.ssol_estimator_get_tracked_paths_count(None, C.byref(n)) == RES_BAD_ARG
    assert f.ssol_estimator_get_tracked_paths_count(est.h, C.byref(n)) == RES_OK
    assert n.value == 1
    path = L.Path()
    assert f.ssol_estimator_get_tracked_path(None, 1, None) == RES_BAD_ARG
    assert f.ssol_estimator_get_tracked_path(est.h, 1, None) == RES_BAD_ARG
    assert f.ssol_estimator_get_tracked_path(None, 0, None) == RES_BAD_ARG
    assert f.ssol_estimator_get_tracked_path(est.h, 0, None) == RES_BAD_ARG
    assert f.ssol_estimator_get_tracked_path(None, 1, C.byref(path)) == RES_BAD_ARG
    assert f.ssol_estimator_get_tracked_path(est.h, 1, C.byref(path)) == RES_BAD_ARG
    assert f.ssol_estimator_get_tracked_path(None, 0, C.byref(path)) == RES_BAD_ARG
    assert f.ssol_estimator_get_tracked_path(est.h, 0, C.byref(path)) == RES_OK
    nv = C.c_size_t()
    assert f.ssol_path_get_vertices_count(None, None) == RES_BAD_ARG
    assert f.ssol_path_get_vertices_count(C.byref(path), None) == RES_BAD_ARG
    assert f.ssol_path_get_vertices_count(None, C.byref(nv)) == RES_BAD_ARG
    assert f.ssol_path_get_vertices_count(C.byref(path), C.byref(nv)) == RES_OK
    assert nv.value != 0
    v = L.PathVertex()
    cnt = nv.value
    assert f.ssol_path_get_vertex(None, cnt, None) == RES_BAD_ARG
    assert f.ssol_path_get_vertex(C.byref(path), cnt, None) == RES_BAD_ARG
    assert f.ssol_path_get_vertex(None, 0, None) == RES_BAD_ARG
    assert f.ssol_path_get_vertex(C.byref(path), 0, None) == RES_BAD_ARG
    assert f.ssol_path_get_vertex(None, cnt, C.byref(v)) == RES_BAD_ARG
    assert f.ssol_path_get_vertex(C.byref(path), cnt, C.byref(v)) == RES_BAD_ARG
    assert f.ssol_path_get_vertex(None, 0, C.byref(v)) == RES_BAD_ARG
    for i in range(cnt):
        assert f.ssol_path_get_vertex(C.byref(path), i, C.byref(v)) == RES_OK
    assert est.realisation_count == 1
    # a solve without tracker records no path
    est2 = lib.solve(10, seed=0)
    assert f.ssol_estimator_get_tracked_paths_count(est2.h, C.byref(n)) == RES_OK
    assert n.value == 0
    lib.close()


def _same_paths(gpu_paths, orc_paths, atol=1e-6, rtol=1e-9):
    """Fraction of paths equal in type, vertex count, positions and weights"""
    assert len(gpu_paths) == len(orc_paths)
    same = 0
    for (tg, vg), (to, vo) in zip(gpu_paths, orc_paths):
        to = L.PATH_TYPES[to]
        if tg != to or vg.shape != vo.shape:
            continue
        if np.all(np.abs(vg - vo) <= atol + rtol * np.abs(vo)):
            same += 1
    return same / len(orc_paths)


@pytest.mark.parametrize("name", ["solver1_all", "config1"])
def test_tracked_paths_equal_oracle_path_by_path(gpu, name):
    """Specular scenes with equal-area sampled triangles: the GPU and the
    oracle draw the same start point, sun direction and wavelength from the
    same Threefry streams and nothing else, so each realisation follows the
    same path in both (up to rounding at grazing hits)."""
    scene = R.solver1("all")[0] if name == "solver1_all" else S.config1()
    n = 4000
    lib = L.Scene(scene)
    est = lib.solve(n, seed=11, tracker=(2.5, 60.0))
    gp = est.tracked_paths()
    o = O.Oracle(scene)
    assert o.solve_tracked(n, seed=11, sun_ray_length=2.5, infinite_ray_length=60.0) == 0
    op = o.tracked_paths()
    frac = _same_paths(gp, op)
    print(f"{name}: {frac:.4f} of {n} paths equal")
    assert frac >= 0.995
    lib.close()


def test_tracked_paths_and_tallies(gpu):
    scene = R.solver1("all")[0]
    n = 50000
    lib = L.Scene(scene)
    est = lib.solve(n, seed=3, tracker=(3.0, 777.0))
    paths = est.tracked_paths()
    assert len(paths) == n
    g = est.global_()
    escaped = TC.check_paths(paths, n, scene.sun.dni, 3.0, 777.0, g, segments=est.stats.segments)
    counts = TC.type_counts(paths)
    assert counts["shadow"] > 0 and counts["success"] > 0 and escaped > 0, counts
    # same realisations, same tallies as the solve without tracker
    g2 = lib.solve(n, seed=3).global_()
    for k, v in g2.items():
        assert abs(v[0] - g[k][0]) <= 1e-9 * max(abs(v[0]), 1.0), (k, v, g[k])
    lib.close()


def test_tracker_default_and_zero_lengths(gpu):
    scene = R.solver1("all")[0]
    lib = L.Scene(scene)
    paths = lib.solve(500, seed=1, tracker=(-1.0, -1.0)).tracked_paths()
    for _, v in paths:
        assert abs(np.linalg.norm(v[1, :3] - v[0, :3]) - TC.solver1_extent()) < 1e-5
    est = lib.solve(2000, seed=2, tracker=(0.0, 0.0))
    paths = est.tracked_paths()
    TC.check_paths(paths, 2000, scene.sun.dni, 0.0, 0.0, est.global_())
    assert all(len(v) == 1 for ty, v in paths if ty == "shadow")
    lib.close()


def test_deep_paths_roulette_and_vertex_buffer_regrowth(gpu):
    """About 70 vertices per path (more than the first vertex buffer holds),
    and the roulette drops cancelled paths: the fraction kept agrees with
    the oracle's within 4 binomial standard errors."""
    gap = 0.11
    scene = TC.two_mirrors(gap)
    n = 20000
    lib = L.Scene(scene)
    est = lib.solve(n, seed=5, tracker=(1.0, 50.0))
    paths = est.tracked_paths()
    TC.check_alternating(paths, gap, 1.0, 50.0)
    assert sum(len(v) for _, v in paths) > 8 * n
    o = O.Oracle(scene)
    assert o.solve_tracked(n, seed=6, sun_ray_length=1.0, infinite_ray_length=50.0) == 0
    kg, ko = len(paths) / n, len(o.tracked_paths()) / n
    se = math.sqrt(kg * (1 - kg) / n + ko * (1 - ko) / n)
    print(f"kept: gpu {kg:.4f} oracle {ko:.4f} (se {se:.4f})")
    assert 0.4 < kg < 0.6
    assert abs(kg - ko) <= 4 * se
    lib.close()


def test_glossy_field_path_statistics_vs_oracle(gpu):
    """Glossy paraboloid heliostats (Beckmann): type fractions and the mean
    vertex count against the oracle's (independent streams)."""
    scene = S.config2(0, n_helio=12, radius=120.0, rcv_size=6.0, rcv_cells=8)
    n = 40000
    lib = L.Scene(scene)
    gp = lib.solve(n, seed=7, tracker=(5.0, 500.0)).tracked_paths()
    o = O.Oracle(scene, track_sampled_x_receiver=False)
    assert o.solve_tracked(n, seed=8, sun_ray_length=5.0, infinite_ray_length=500.0) == 0
    op = o.tracked_paths()
    cg, co = TC.type_counts(gp), TC.type_counts(op)
    print("gpu", cg, "oracle", co)
    for k in cg:
        pg, po = cg[k] / n, co[k] / n
        se = math.sqrt(pg * (1 - pg) / n + po * (1 - po) / n) + 1e-12
        assert abs(pg - po) <= 4 * se + 1e-9, (k, pg, po, se)
    mg = np.array([len(v) for _, v in gp], dtype=float)
    mo = np.array([len(v) for _, v in op], dtype=float)
    se = math.sqrt(mg.var() / n + mo.var() / n)
    assert abs(mg.mean() - mo.mean()) <= 4 * se + 1e-9, (mg.mean(), mo.mean(), se)
    lib.close()


def test_fp32_device_records_the_fp64_paths(gpu):
    """Path recording runs the FP64 form whatever the device precision."""
    scene = R.solver1("all")[0]
    a = L.Scene(scene)
    b = L.Scene(scene, precision="fp32")
    pa = a.solve(3000, seed=9, tracker=(2.0, 20.0)).tracked_paths()
    pb = b.solve(3000, seed=9, tracker=(2.0, 20.0)).tracked_paths()
    assert len(pa) == len(pb)
    for (ta, va), (tb, vb) in zip(pa, pb):
        assert ta == tb and np.array_equal(va, vb)
    a.close()
    b.close()
