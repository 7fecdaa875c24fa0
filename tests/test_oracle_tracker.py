"""The oracle's path recording (the reference's path tracker,
src/ssol_solver.c:898-1127) against the identities that tie recorded paths
to the tallies of the same solve, the default segment length
(src/ssol_solver.c:562-570, 1211-1220) in closed form, and the roulette's
dropping of cancelled paths (src/ssol_solver.c:1111-1117). This pins the
oracle before it checks the GPU tracker."""
import numpy as np

import oracle_ctypes as O
import ref_scenes as R
import tracker_checks as TC

N = 20000


def _solver1():
    scene = R.solver1("all")[0]
    o = O.Oracle(scene)
    return scene, o


def test_oracle_tracked_paths_and_tallies():
    scene, o = _solver1()
    assert o.solve_tracked(N, seed=3, sun_ray_length=3.0, infinite_ray_length=777.0) == 0
    paths = o.tracked_paths()
    assert len(paths) == N
    g = o.global_()
    escaped = TC.check_paths(paths, N, scene.sun.dni, 3.0, 777.0, g, segments=o.segments)
    counts = TC.type_counts(paths)
    assert counts["shadow"] > 0 and counts["success"] > 0 and escaped > 0, counts
    # same realisations as a plain (multi-threaded) solve
    o2 = O.Oracle(scene)
    assert o2.solve(N, seed=3) == 0
    for k, v in o2.global_().items():
        assert abs(v[0] - g[k][0]) <= 1e-9 * max(abs(v[0]), 1.0), (k, v, g[k])


def test_oracle_tracker_default_length():
    scene, o = _solver1()
    assert abs(o.tracked_extend() - TC.solver1_extent()) < 1e-6
    assert o.solve_tracked(200, seed=1) == 0
    for ty, v in o.tracked_paths():
        d = np.linalg.norm(v[1, :3] - v[0, :3])
        assert abs(d - TC.solver1_extent()) < 1e-6


def test_oracle_tracker_zero_lengths():
    scene, o = _solver1()
    assert o.solve_tracked(2000, seed=2, sun_ray_length=0.0, infinite_ray_length=0.0) == 0
    paths = o.tracked_paths()
    g = o.global_()
    TC.check_paths(paths, 2000, scene.sun.dni, 0.0, 0.0, g)
    assert all(len(v) == 1 for ty, v in paths if ty == 1)


def test_oracle_tracker_roulette_drops_cancelled_paths():
    gap = 0.11
    scene = TC.two_mirrors(gap)
    o = O.Oracle(scene)
    n = 4000
    assert o.solve_tracked(n, seed=5, sun_ray_length=1.0, infinite_ray_length=50.0) == 0
    paths = o.tracked_paths()
    # every path needs about 73 bounces: the roulette at depth 65 keeps half
    assert 0.4 * n < len(paths) < 0.6 * n
    assert all(ty == 0 for ty, _ in paths)  # no receiver: all missing
    TC.check_alternating(paths, gap, 1.0, 50.0)
    assert min(len(v) for _, v in paths) > 66
