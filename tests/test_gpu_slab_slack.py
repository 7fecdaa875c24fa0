"""Conservative box tests. The trace kernels compute a plane distance as
p * (1/d) + (-o/d) with one FMA; the rounding of -o/d puts up to
|o/d| 2^-24 of error on it, which with origins 900 m from the world origin
exceeds what the box inflation absorbs at grazing box faces. The kernels
therefore accept a box when tnear - slack <= tfar, slack = scale *
max |o/d| (default 2^-21, ssolh_test_slab_slack sets it).

On the 1,000-paraboloid field (radius 900 m) the default is converged: a
512x larger slack traces the same segments and gives the same tallies to the
order of the FP64 additions; without slack (scale 0) rounding culls boxes
that rays cross and the absorbed flux moves by about 1e-5 relative (the
test only requires that it moves, which shows the comparison can see a
culled hit)."""
import ctypes as C

import numpy as np
import pytest

import ssol_ctypes as L
import ssol_scene as S

pytestmark = pytest.mark.gpu

PATHS = 20_000_000


def _set_slack(lib, scale):
    h = L.lib_hip()
    h.ssolh_test_slab_slack.argtypes = [C.c_void_p, C.c_float]
    h.ssolh_test_slab_slack.restype = C.c_int
    ctx = C.c_void_p()
    L.call("ssol_device_get_gpu_context", lib.dev, C.byref(ctx))
    assert h.ssolh_test_slab_slack(ctx, scale) == 0


def _run(lib, spec):
    est = lib.solve(PATHS, seed=2)
    rcv = spec.instances[-1]
    shape = rcv.object.sshapes[0][0]
    g = est.global_()
    out = np.concatenate([[g[k][0] for k in sorted(g)],
                          est.primitives(rcv, shape, S.FRONT, "absorbed_flux")[:, 0]])
    return est.stats.segments, out


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_default_slack_is_converged(gpu, precision):
    spec, _, _, _ = S.config5()
    lib = L.Scene(spec, precision=precision)
    lib.solve(100_000, seed=1)
    seg0, ref = _run(lib, spec)
    _set_slack(lib, 2.0 ** -12)
    seg1, big = _run(lib, spec)
    assert seg1 == seg0
    rel = 1e-6 if precision == "fp32" else 1e-10
    scale = np.maximum(np.abs(ref), np.abs(big))
    assert np.all(np.abs(big - ref) <= rel * scale + 1e-300), \
        float(np.max(np.abs(big - ref) / (scale + 1e-300)))
    if precision == "fp64":
        _set_slack(lib, 0.0)
        _, none = _run(lib, spec)
        moved = abs(none[0] - ref[0]) / abs(ref[0])   # absorbed_by_receivers
        print(f"absorbed without slack: {moved:.3g} relative", flush=True)
        assert moved > 1e-8
    lib.close()


def test_slack_rejects_out_of_range(gpu):
    lib = L.Scene(S.config1())
    lib.solve(1000, seed=1)
    h = L.lib_hip()
    h.ssolh_test_slab_slack.argtypes = [C.c_void_p, C.c_float]
    ctx = C.c_void_p()
    L.call("ssol_device_get_gpu_context", lib.dev, C.byref(ctx))
    assert h.ssolh_test_slab_slack(ctx, -1.0) != 0
    assert h.ssolh_test_slab_slack(ctx, float("nan")) != 0
    assert h.ssolh_test_slab_slack(ctx, 0.5) != 0
    lib.close()
