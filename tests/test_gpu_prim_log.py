"""Per-primitive tallies through the log of finished paths' entries: each
entry is appended to a log that is sorted by key and summed into the tally
after each wavefront, and whenever a host-driven iteration leaves the log
half full; entries that find it full go to per-lane atomics instead.
Capping the log (ssolh_test_prim_log) moves entries between these routes
without changing a path: the capped solves trace the same segments and
agree with the default one up to the order of the FP64 additions
(expectations 1e-10 relative, in both forms: the paths and their weights
are the same; variances and standard errors, which cancel, 1e-6).

Scene: a 200-heliostat flat field (beam-down) whose 64x64-cell receiver
keeps per-primitive tallies."""
import ctypes as C

import numpy as np
import pytest

import ssol_ctypes as L
import ssol_scene as S

pytestmark = pytest.mark.gpu

PATHS = 1_000_000


def _hip():
    h = L.lib_hip()
    h.ssolh_test_prim_log.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32),
                                      C.POINTER(C.c_uint64)]
    h.ssolh_test_prim_log.restype = C.c_int
    return h


def _ctx(lib):
    ctx = C.c_void_p()
    L.call("ssol_device_get_gpu_context", lib.dev, C.byref(ctx))
    assert ctx.value
    return ctx


def _log_stats(lib, cap):
    """The last solve's (sorted reductions, entries sent to atomics); sets
    the cap for the next solves"""
    f, o = C.c_uint32(), C.c_uint64()
    assert _hip().ssolh_test_prim_log(_ctx(lib), cap, C.byref(f), C.byref(o)) == 0
    return f.value, o.value


def _prims(spec, est):
    rcv = spec.instances[-1]
    shape = rcv.object.sshapes[0][0]
    means, spread = [], []
    for f in L.RECEIVER_FIELDS:
        p = est.primitives(rcv, shape, S.FRONT, f)
        means.append(p[:, 0])
        spread.append(p[:, 1:].ravel())
    return np.concatenate(means), np.concatenate(spread)


def _close(a, b, rel):
    scale = np.maximum(np.abs(a), np.abs(b))
    err = np.abs(a - b)
    assert np.all(err <= rel * scale + 1e-300), float(np.max(err / (scale + 1e-300)))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("graph", [0, 4])
def test_capped_log_changes_nothing(gpu, precision, graph):
    spec = S.config3(0, n_helio=200, r_min=50.0, r_max=250.0)
    lib = L.Scene(spec, precision=precision, graph_iterations=graph)
    ref = lib.solve(PATHS, seed=9)
    flushes, over = _log_stats(lib, 0)
    assert flushes >= 1 and over == 0         # default: the log takes every entry
    rm, rs = _prims(spec, ref)
    assert np.count_nonzero(rm) > 1000         # the receiver's cells are reached
    try:
        for cap in (50_000, 64, 1):
            _log_stats(lib, cap)
            est = lib.solve(PATHS, seed=9)
            flushes, over = _log_stats(lib, cap)
            print(f"cap {cap}: {flushes} reductions, {over} entries to atomics", flush=True)
            assert over > 0                    # the log filled: the atomics took the rest
            # host-driven: reduced once half full after an iteration, and
            # again at the end; graph-captured: at the end only
            assert flushes >= (2 if graph == 0 else 1)
            assert est.stats.segments == ref.stats.segments
            em, es = _prims(spec, est)
            _close(em, rm, 1e-10)
            _close(es, rs, 1e-6)
            g, h = est.global_(), ref.global_()
            for k in g:
                _close(np.array([g[k][0]]), np.array([h[k][0]]), 1e-10)
    finally:
        _log_stats(lib, 0)
        lib.close()
