"""Batched multi-sun solve (ssol_solve_sun_batch), the graph-captured
wavefront and the autotuner.

The RNG streams of a batch continue across suns, so sun j of a batch must
reproduce a single-sun solve whose rng counter starts at the sum of the
counts before j: same realisations, same path samples; only the order of
the FP64 tally additions differs (tolerance 1e-10 relative). The graph
mode runs the same kernels and must agree with the host-driven loop to
the same tolerance."""
import numpy as np
import pytest

import ref_scenes as R
import ssol_ctypes as L
import ssol_scene as S

pytestmark = pytest.mark.gpu


def _same(a, b, rel=1e-10):
    return abs(a[0] - b[0]) <= rel * max(abs(a[0]), abs(b[0])) + 1e-9 and \
        abs(a[2] - b[2]) <= 1e-6 * max(abs(a[2]), abs(b[2])) + 1e-9


@pytest.mark.parametrize("graph", [0, 8])
def test_batch_equals_consecutive_single_solves(gpu, graph):
    spec = S.config2(n_helio=20)
    dirs = [S.sun_dir_from(a, e) for a, e in ((180, 60), (120, 30), (250, 75), (90, 10))]
    counts = [150_000, 0, 60_000, 230_000]
    sc = L.Scene(spec, graph_iterations=graph)
    res, got = sc.solve_sun_batch(dirs, counts, seed=5, counter=1000)
    assert got == counts
    assert all(v == (0.0, 0.0, 0.0) for v in res[1].values())
    first = 1000
    for j, (d, n) in enumerate(zip(dirs, counts)):
        if n:
            one = S.Scene(spec.instances, S.Sun(spec.sun.kind, d, spec.sun.dni, spec.sun.spectrum,
                                                spec.sun.param), spec.atmosphere, name="one")
            lib = L.Scene(one)
            e = lib.solve(n, seed=5, counter=first)
            g = e.global_()
            for f in L.GLOBAL_FIELDS:
                assert _same(res[j][f], g[f]), (j, f, res[j][f], g[f])
            lib.close()
        first += n
    sc.close()


def test_graph_mode_matches_host_loop(gpu):
    spec = S.config3(n_helio=300)
    a = L.Scene(spec).solve(1_500_000, seed=3)
    b = L.Scene(spec, graph_iterations=16).solve(1_500_000, seed=3)
    assert a.realisation_count == b.realisation_count == 1_500_000
    ga, gb = a.global_(), b.global_()
    for f in L.GLOBAL_FIELDS:
        assert _same(ga[f], gb[f]), (f, ga[f], gb[f])
    rcv = spec.instances[-1]
    ra, rb = a.receiver(rcv, S.FRONT), b.receiver(rcv, S.FRONT)
    for f in L.RECEIVER_FIELDS:
        assert _same(ra[f], rb[f]), (f, ra[f], rb[f])
    assert a.stats.segments == b.stats.segments


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_graph_replayed_many_times_matches_host_loop(gpu, precision):
    """A narrow wavefront (chunks of 4,096 start-ordered realisations) with two
    iterations per graph: each solve replays its captured graph dozens of
    times, with a device-to-host copy and a synchronisation between replays,
    and three solves in a row capture it anew; every solve equals the
    host-driven loop's."""
    spec = S.config2(0, n_helio=20, radius=150.0)
    host = L.Scene(spec, precision=precision, path_slots=4096)
    graph = L.Scene(spec, precision=precision, path_slots=4096, graph_iterations=2)
    for seed in (1, 2, 3):
        a, b = host.solve(200_000, seed=seed), graph.solve(200_000, seed=seed)
        assert b.stats.iterations >= 60, b.stats.iterations  # 30+ replays
        assert a.realisation_count == b.realisation_count == 200_000
        assert a.stats.segments == b.stats.segments
        ga, gb = a.global_(), b.global_()
        for f in L.GLOBAL_FIELDS:
            assert _same(ga[f], gb[f]), (seed, f, ga[f], gb[f])
        rcv = spec.instances[-1]
        ra, rb = a.receiver(rcv, S.FRONT), b.receiver(rcv, S.FRONT)
        for f in L.RECEIVER_FIELDS:
            assert _same(ra[f], rb[f]), (seed, f, ra[f], rb[f])
    host.close()
    graph.close()


def test_batch_sweep_shape_and_physics(gpu):
    """The 512-sun sweep of the 100-paraboloid field: counts honoured, every
    sun's sinks plus shadowing balance its initial flux (DNI x sampled area x
    cos factor, to the sun's angular blur), and the cos factor peaks near the
    sun the field is aimed for."""
    spec, dirs, counts = S.config4(n_suns=64, n_az=8, n_el=8, log_min=4.0, log_max=5.0)
    sc = L.Scene(spec, graph_iterations=8)
    res, got = sc.solve_sun_batch(dirs, counts, seed=11)
    assert list(got) == [int(c) for c in counts]
    area = sc.solve(10_000, seed=1).sampled_area
    for j, g in enumerate(res):
        sinks = sum(g[f][0] for f in ("absorbed_by_receivers", "missing", "shadowed",
                                      "extinguished_by_atmosphere", "other_absorbed"))
        initial = spec.sun.dni * area * g["cos_factor"][0]
        assert abs(sinks - initial) <= 1e-3 * initial + 1e-9, (j, sinks, initial)
    cos = np.array([g["cos_factor"][0] for g in res])
    aimed = np.array(spec.sun.direction) / np.linalg.norm(spec.sun.direction)
    closeness = np.array([np.dot(d / np.linalg.norm(d), aimed) for d in dirs])
    assert cos[np.argmax(closeness)] >= np.percentile(cos, 75)
    sc.close()


def test_autotune_keeps_a_candidate_and_stays_correct(gpu):
    spec = S.config1()
    sc = L.Scene(spec)
    t = sc.autotune(400_000)
    assert t.candidates >= 3 and t.path_slots in (0, 1 << 20, 1 << 22, 1 << 24)
    assert t.paths_per_second > 0
    e = sc.solve(1_000_000, seed=2)
    a = e.global_()["absorbed_by_receivers"]
    assert abs(a[0] - S.CONFIG1_CLOSED_FORM["absorbed_by_receivers"]) <= 4 * a[2] + 1e-6 * 22500
    sc.close()


def test_autotune_batch_keeps_a_candidate_and_the_batch_stays_exact(gpu):
    """ssol_device_autotune_batch times widths x graph lengths on a probe
    batch and keeps the fastest; the sweep run with the pick equals the sweep
    run with the defaults (same realisations, FP64 additions reordered)."""
    spec, dirs, counts = S.config4(n_suns=32, log_min=4.0, log_max=5.0)
    ref_sc = L.Scene(spec)
    ref, ref_n = ref_sc.solve_sun_batch(dirs, counts, seed=4)
    ref_sc.close()
    sc = L.Scene(spec)
    t = sc.autotune_batch(dirs[::4], [max(1000, int(c) // 4) for c in counts[::4]])
    assert t.candidates >= 3 and t.path_slots in (0, 1 << 22, 1 << 24, 1 << 26)
    assert t.graph_iterations in (0, 16, 32) and t.paths_per_second > 0
    got, got_n = sc.solve_sun_batch(dirs, counts, seed=4)
    assert got_n == ref_n
    for a, b in zip(got, ref):
        for f in a:
            assert abs(a[f][0] - b[f][0]) <= 1e-10 * max(1.0, abs(b[f][0])), (f, a[f], b[f])
    sc.close()
