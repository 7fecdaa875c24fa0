"""The wavefront width changes speed, not results. By default every
realisation of a solve is in flight at once; with a narrower width
(ssol_device_set_wavefront_size) the solve either runs in chunks of that
many realisations, each all in flight (when there is something to order the
slots by: more than one sampled instance or sun), or keeps one wavefront
whose slots take further realisations in place, which runs the kernels'
refill code (one sampled instance and one sun). Realisation ids are static,
so both runs follow the same paths; only the order of the FP64 tally
additions may differ (tolerance 1e-10 relative). Covers the plain, FP32,
adjoint, tracked and batched solves in both narrow forms."""
import numpy as np
import pytest

import ref_scenes as R
import ssol_ctypes as L
import ssol_scene as S

pytestmark = pytest.mark.gpu

NARROW = 4096  # 32 blocks of 128 slots


def _close(a, b, rel=1e-10):
    return abs(a[0] - b[0]) <= rel * max(abs(a[0]), abs(b[0])) + 1e-9


def _same_globals(ea, eb):
    ga, gb = ea.global_(), eb.global_()
    for k in ga:
        assert _close(ga[k], gb[k]), (k, ga[k], gb[k])


SCENES = {
    "solver1": lambda: R.solver1("all")[0],            # chunked
    "solver1_primary": lambda: R.solver1("primary")[0],  # refilled in place
    "solver7_beam_down": lambda: R.solver7()[0],
    "paraboloid_field": lambda: S.config2(0, n_helio=20, radius=150.0, rcv_size=5.0, rcv_cells=8),
    "flat_field": lambda: S.config3(0, n_helio=200, r_min=50.0, r_max=250.0),
}


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("name", sorted(SCENES))
def test_narrow_wavefront_gives_the_default_result(gpu, name, precision):
    spec = SCENES[name]()
    n = 300_000
    wide = L.Scene(spec, precision=precision)
    narrow = L.Scene(spec, precision=precision, path_slots=NARROW)
    ew, en = wide.solve(n, seed=4), narrow.solve(n, seed=4)
    # a narrow solve takes more iterations, chunked or refilled
    assert en.stats.iterations > ew.stats.iterations
    assert ew.stats.wavefronts == 1 and not ew.stats.refilled
    assert en.stats.refilled == (en.stats.wavefronts == 1)
    if name == "solver1_primary":
        assert en.stats.refilled
    if name == "solver1":
        assert en.stats.wavefronts == -(-n // NARROW)
    assert ew.realisation_count == en.realisation_count == n
    _same_globals(ew, en)
    for inst in spec.instances:
        for side in (S.FRONT, S.BACK):
            if inst.receiver_mask & side:
                rw, rn = ew.receiver(inst, side), en.receiver(inst, side)
                for k in rw:
                    assert _close(rw[k], rn[k]), (k, rw[k], rn[k])
    wide.close()
    narrow.close()


def test_narrow_wavefront_adjoint(gpu):
    spec, aim, slope, axes = S.config5(0, n_helio=20, radius=150.0)
    helios = spec.instances[:-1]
    wide = L.Scene(spec)
    narrow = L.Scene(spec, path_slots=NARROW)
    gw = wide.solve_adjoint(200_000, helios, axes, d_absorbed=1.0, seed=2)
    gn = narrow.solve_adjoint(200_000, helios, axes, d_absorbed=1.0, seed=2)
    assert np.allclose(gw, gn, rtol=1e-9, atol=1e-9 * np.abs(gw).max())
    wide.close()
    narrow.close()


def test_narrow_wavefront_tracked_paths(gpu):
    spec = R.solver1("all")[0]
    wide = L.Scene(spec)
    narrow = L.Scene(spec, path_slots=NARROW)
    pw = wide.solve(20_000, seed=6, tracker=(2.0, 30.0)).tracked_paths()
    pn = narrow.solve(20_000, seed=6, tracker=(2.0, 30.0)).tracked_paths()
    assert len(pw) == len(pn) == 20_000
    for (tw, vw), (tn, vn) in zip(pw, pn):
        assert tw == tn and np.array_equal(vw, vn)
    wide.close()
    narrow.close()


def test_narrow_wavefront_batch(gpu):
    spec = S.config2(0, n_helio=20, radius=150.0)
    dirs = [S.sun_dir_from(a, e) for a, e in ((180, 60), (120, 30), (250, 75))]
    counts = [70_000, 0, 40_000]
    wide = L.Scene(spec)
    narrow = L.Scene(spec, path_slots=NARROW)
    rw, cw = wide.solve_sun_batch(dirs, counts, seed=3)
    rn, cn = narrow.solve_sun_batch(dirs, counts, seed=3)
    assert list(cw) == list(cn)
    for a, b in zip(rw, rn):
        for k in a:
            assert _close(a[k], b[k]), (k, a[k], b[k])
    wide.close()
    narrow.close()


def test_narrow_wavefront_refill_adjoint(gpu):
    """One heliostat, one sun: the adjoint's refill instantiation."""
    spec, aim, slope, axes = S.config5(0, n_helio=1, radius=60.0)
    helios = spec.instances[:-1]
    wide = L.Scene(spec)
    narrow = L.Scene(spec, path_slots=NARROW)
    gw = wide.solve_adjoint(100_000, helios, axes, d_absorbed=1.0, seed=2)
    gn = narrow.solve_adjoint(100_000, helios, axes, d_absorbed=1.0, seed=2)
    assert np.abs(gw).max() > 0
    assert np.allclose(gw, gn, rtol=1e-9, atol=1e-9 * np.abs(gw).max())
    wide.close()
    narrow.close()


def test_narrow_wavefront_refill_tracked_paths(gpu):
    """One sampled instance: the tracked solve refills slots in place."""
    spec = R.solver1("primary")[0]
    wide = L.Scene(spec)
    narrow = L.Scene(spec, path_slots=NARROW)
    en = narrow.solve(20_000, seed=6, tracker=(2.0, 30.0))
    assert en.stats.refilled
    pw = wide.solve(20_000, seed=6, tracker=(2.0, 30.0)).tracked_paths()
    pn = en.tracked_paths()
    assert len(pw) == len(pn) == 20_000
    for (tw, vw), (tn, vn) in zip(pw, pn):
        assert tw == tn and np.array_equal(vw, vn)
    wide.close()
    narrow.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_narrow_wavefront_batch_many_keys(gpu, precision):
    """800 suns x 21 instances = 16,800 (sun, instance) keys: the narrow
    batch runs in chunks of start-ordered realisations that span suns."""
    spec = S.config2(0, n_helio=20, radius=150.0)
    rng = np.random.default_rng(8)
    dirs = [S.sun_dir_from(a, e) for a, e in zip(rng.uniform(0, 360, 800), rng.uniform(20, 80, 800))]
    counts = rng.integers(0, 1200, 800)
    wide = L.Scene(spec, precision=precision)
    narrow = L.Scene(spec, precision=precision, path_slots=NARROW)
    rw, cw = wide.solve_sun_batch(dirs, counts, seed=3)
    rn, cn = narrow.solve_sun_batch(dirs, counts, seed=3)
    assert list(cw) == list(cn) == [int(c) for c in counts]
    for a, b in zip(rw, rn):
        for k in a:
            assert _close(a[k], b[k]), (k, a[k], b[k])
    wide.close()
    narrow.close()
