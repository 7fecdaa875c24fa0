"""Traversal stack of the trace kernels. Each lane keeps the first entries of
its stack in LDS (16, 24 or 32) and the rest of the scene's bound (the TLAS
descent's pushes, the instance sentinel, the deepest BLAS descent's pushes
over the 4-wide nodes) in a device-memory overflow array. Capping the LDS
part (ssolh_test_trace_stack) moves entries to the overflow without changing
a single traversal: the capped solves trace the same segments and agree with
the default one up to the order of the FP64 tally additions.

Scenes: the 10,000-heliostat beam-down field (a TLAS over 10,003 instances
and a 64x64-cell receiver) and a 200-heliostat field whose receiver is a
256x256-cell mesh (131,072 triangles), whose BLAS alone is deeper than the
smallest LDS stack."""
import ctypes as C

import numpy as np
import pytest

import ssol_ctypes as L
import ssol_scene as S

pytestmark = pytest.mark.gpu


def _hip():
    h = L.lib_hip()
    h.ssolh_test_trace_stack.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32)]
    h.ssolh_test_trace_stack.restype = C.c_int
    h.ssolh_last_error.restype = C.c_char_p
    return h


def _ctx(scene):
    ctx = C.c_void_p()
    L.call("ssol_device_get_gpu_context", scene.dev, C.byref(ctx))
    assert ctx.value
    return ctx


def _outputs(lib, spec, est):
    rcv = spec.instances[-1]
    shape = rcv.object.sshapes[0][0]
    g = est.global_()
    return (np.array([g[k] for k in sorted(g)], dtype=np.float64).ravel(),
            est.primitives(rcv, shape, S.FRONT, "absorbed_flux").ravel().copy())


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("scene", ["beam_down_10k", "deep_receiver"])
def test_lds_stack_cap_changes_nothing(gpu, scene, precision):
    if scene == "beam_down_10k":
        spec, n = S.config3(), 4_000_000
    else:
        spec, n = S.config2(0, n_helio=200, radius=300.0, rcv_size=10.0, rcv_cells=256), 2_000_000
    lib = L.Scene(spec, precision=precision)
    hip = _hip()
    est = lib.solve(n, seed=3)
    ctx = _ctx(lib)
    bound = C.c_uint32()
    assert hip.ssolh_test_trace_stack(ctx, 0, C.byref(bound)) == 0
    print(f"{scene}: traversal stack bound {bound.value}", flush=True)
    assert bound.value > 16            # the cap of 16 sends entries to the overflow
    ref_seg = est.stats.segments
    ref = _outputs(lib, spec, est)
    assert ref[0].any() and ref[1].any()
    rel = 1e-6 if precision == "fp32" else 1e-10
    try:
        for cap in (16, 24):
            if bound.value <= cap:
                continue
            assert hip.ssolh_test_trace_stack(ctx, cap, None) == 0
            est = lib.solve(n, seed=3)
            assert est.stats.segments == ref_seg, cap
            for a, b in zip(_outputs(lib, spec, est), ref):
                scale = np.maximum(np.abs(a), np.abs(b))
                err = np.abs(a - b)
                assert np.all(err <= rel * scale + 1e-300), (cap, float(np.max(err / (scale + 1e-300))))
    finally:
        hip.ssolh_test_trace_stack(ctx, 0, None)
        lib.close()


def test_trace_stack_cap_rejects_other_sizes(gpu):
    spec = S.config2(0, n_helio=10, radius=100.0)
    lib = L.Scene(spec)
    lib.solve(1000, seed=1)
    hip, ctx = _hip(), _ctx(lib)
    assert hip.ssolh_test_trace_stack(ctx, 20, None) != 0
    assert b"cap 20" in hip.ssolh_last_error()
    lib.close()
