"""Pending keys across solves. A path collects its receiver weights in
pending entries keyed by receiver and flushes them when it ends, clearing
their keys; a solve that drains therefore leaves every slot's keys clear,
and a context clears them itself only when its path state's layout changes
(width, form). Here the paths keep pending entries across several receiver
hits (a mirror bowl that is its own receiver on both sides), and a context
that changes width and form between solves must give, solve after solve,
what a fresh context gives for the same solve: the same paths, tallies
equal up to the order of the FP64 additions (1e-10 relative)."""
import numpy as np
import pytest

import extra_scenes as X
import ssol_ctypes as L

pytestmark = pytest.mark.gpu

# (form, wavefront width, realisations, seed): the width and the form
# change the state's layout, the repeats keep it
STEPS = [("fp64", 0, 300_000, 1), ("fp64", 0, 200_000, 2), ("fp64", 0, 200_000, 3),
         ("fp32", 0, 200_000, 4), ("fp64", 4096, 200_000, 5), ("fp64", 0, 200_000, 6)]


def _values(est, bowl):
    g = est.global_()
    out = [g[k][0] for k in sorted(g)]
    for side in (X.S.FRONT, X.S.BACK):
        r = est.receiver(bowl, side)
        out += [r["incoming_flux"][0], r["absorbed_flux"][0]]
    return np.array(out)


def _close(a, b, rel=1e-10):
    scale = np.maximum(np.abs(a), np.abs(b))
    assert np.all(np.abs(a - b) <= rel * scale + 1e-300), float(np.max(np.abs(a - b) / (scale + 1e-300)))


def test_reused_context_matches_fresh_ones(gpu):
    spec, h = X.hemisphere_deep_mirror()
    bowl = h["bowl"]
    lib = L.Scene(spec)
    try:
        for form, width, n, seed in STEPS:
            L.call("ssol_device_set_precision", lib.dev, 1 if form == "fp32" else 0)
            L.call("ssol_device_set_wavefront_size", lib.dev, width)
            est = lib.solve(n, seed=seed)
            fresh = L.Scene(spec, precision=form, path_slots=width)
            ref = fresh.solve(n, seed=seed)
            assert est.stats.segments == ref.stats.segments, (form, width, n)
            v, w = _values(est, bowl), _values(ref, bowl)
            _close(v, w)
            fresh.close()
    finally:
        lib.close()
