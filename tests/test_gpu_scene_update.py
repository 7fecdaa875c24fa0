"""Moving instances of a scene that has already been solved
(ssol_instance_set_transform between solves): the next solve takes new
instance frames and refits the top-level hierarchy on the GPU (rebuilding
it when the refit gets loose), and must give what a scene built anew with
the new transforms gives (same realisations; the FP64 tally
additions may come in another order, tolerance 1e-10 relative)."""
import numpy as np
import pytest

import ssol_ctypes as L
import ssol_scene as S

pytestmark = pytest.mark.gpu


def _close(a, b, rel=1e-10):
    return abs(a[0] - b[0]) <= rel * max(abs(a[0]), abs(b[0])) + 1e-9


def _compare(ea, eb, spec_a, spec_b, rcv_index):
    for k, v in ea.global_().items():
        assert _close(v, eb.global_()[k]), (k, v, eb.global_()[k])
    ra = ea.receiver(spec_a.instances[rcv_index], S.FRONT)
    rb = eb.receiver(spec_b.instances[rcv_index], S.FRONT)
    for k in ra:
        assert _close(ra[k], rb[k]), (k, ra[k], rb[k])


def test_moved_paraboloid_field_equals_fresh_scene(gpu):
    rng = np.random.default_rng(4)
    n_helio = 30
    a0 = rng.uniform(-5e-3, 5e-3, (n_helio, 2))
    a1 = a0 + rng.uniform(-2e-3, 2e-3, (n_helio, 2))
    spec0 = S.config2(0, n_helio=n_helio, radius=200.0, aim_perturb=a0)
    spec1 = S.config2(0, n_helio=n_helio, radius=200.0, aim_perturb=a1)
    n = 1_000_000
    moved = L.Scene(spec0)
    e0 = moved.solve(n, seed=5)
    for a, b in zip(spec0.instances[:-1], spec1.instances[:-1]):
        moved.set_transform(a, b.transform)
    fresh = L.Scene(spec1)
    _compare(moved.solve(n, seed=6), fresh.solve(n, seed=6), spec0, spec1, -1)
    rcv0, rcv1 = spec0.instances[-1], spec1.instances[-1]
    shape0 = rcv0.object.sshapes[0][0]
    shape1 = rcv1.object.sshapes[0][0]
    m = moved.solve(n, seed=7).primitives(rcv0, shape0, S.FRONT, "absorbed_flux")[:, 0]
    f = fresh.solve(n, seed=7).primitives(rcv1, shape1, S.FRONT, "absorbed_flux")[:, 0]
    assert np.allclose(m, f, rtol=1e-10, atol=1e-9)
    # and back: the first transforms give the first result again
    for a in spec0.instances[:-1]:
        moved.set_transform(a, a.transform)
    _compare(moved.solve(n, seed=5), e0, spec0, spec0, -1)
    moved.close()
    fresh.close()


def test_moved_flat_heliostats_change_the_hierarchy(gpu):
    """Planar facet heliostats of the beam-down field, some moved far across
    the field (new top-level boxes, new tree)."""
    spec0 = S.config3(0, n_helio=300, r_min=50.0, r_max=300.0)
    spec1 = S.config3(0, n_helio=300, r_min=50.0, r_max=300.0)
    rng = np.random.default_rng(9)
    for inst in rng.choice(300, 40, replace=False):
        t = list(spec1.instances[inst].transform)
        t[9] = -t[9]
        t[10] = -t[10]
        spec1.instances[inst].transform = t
    n = 1_000_000
    moved = L.Scene(spec0)
    moved.solve(n, seed=1)
    for a, b in zip(spec0.instances, spec1.instances):
        moved.set_transform(a, b.transform)
    fresh = L.Scene(spec1)
    _compare(moved.solve(n, seed=2), fresh.solve(n, seed=2), spec0, spec1, -1)
    moved.close()
    fresh.close()
