"""Moved instances: the TLAS refit on the GPU (ssolh_scene_move_instances,
reached through ssol_instance_set_transform and ssol_solve) and the TLAS
built on the GPU (ssolh_scene_build_tlas, a linear BVH) against a scene
built anew with the moved transforms (the reference rebuilds its views at
every solve, src/ssol_scene.c:346-433).

A refit keeps the tree and takes the instances' new world boxes, so the
closest hits, hence the traced segments and the tallies, are those of a
fresh build: equal segment counts, expectations equal up to the order of
the FP64 atomics (1e-10 relative; FP32 form 1e-6), variances and standard
errors to 1e-6. Aiming steps (small
rotations in place) stay on the refit path; a heliostat sent 5 km away
loosens the tree past the bound, and the solve builds the TLAS anew on the
GPU instead (no upload), with the same agreement. A GPU-built TLAS traces
what the uploaded one traces on the 10k-heliostat field, on C1 and on a
scene of one instance."""
import ctypes as C

import numpy as np
import pytest

import ssol_ctypes as L
import ssol_scene as S

pytestmark = pytest.mark.gpu

N_HELIO = 300
RADIUS = 400.0
PATHS = 2_000_000


def _hip():
    h = L.lib_hip()
    h.ssolh_test_scene_counters.argtypes = [C.c_void_p, C.POINTER(C.c_uint32),
                                            C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    h.ssolh_scene_build_tlas.argtypes = [C.c_void_p]
    h.ssolh_scene_build_tlas.restype = C.c_int
    h.ssolh_test_scene_counters.restype = C.c_int
    h.ssolh_scene_move_instances.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32,
                                             C.c_void_p, C.POINTER(C.c_double)]
    h.ssolh_scene_move_instances.restype = C.c_int
    h.ssolh_last_error.restype = C.c_char_p
    return h


def _ctx(lib):
    ctx = C.c_void_p()
    L.call("ssol_device_get_gpu_context", lib.dev, C.byref(ctx))
    return ctx


def _counters(lib, builds=False):
    u, r, b = C.c_uint32(), C.c_uint32(), C.c_uint32()
    assert _hip().ssolh_test_scene_counters(_ctx(lib), C.byref(u), C.byref(r), C.byref(b)) == 0
    return (u.value, r.value, b.value) if builds else (u.value, r.value)


def _field(aim):
    return S.config2(0, n_helio=N_HELIO, radius=RADIUS, aim_perturb=aim, rcv_size=8.0,
                     rcv_cells=16)


def _outputs(spec, est):
    """(expectations, their variances and standard errors): the global
    results, the receiver's per-triangle absorbed flux, the cos factor of
    20 heliostats"""
    rcv = spec.instances[-1]
    shape = rcv.object.sshapes[0][0]
    g = est.global_()
    glob = np.array([g[k] for k in sorted(g)], dtype=np.float64)
    prim = est.primitives(rcv, shape, S.FRONT, "absorbed_flux")
    cos = np.array([est.sampled(h)["cos_factor"] for h in spec.instances[:20]], dtype=np.float64)
    means = np.concatenate([glob[:, 0], prim[:, 0], cos[:, 0]])
    spread = np.concatenate([glob[:, 1:].ravel(), prim[:, 1:].ravel(), cos[:, 1:].ravel()])
    return means, spread


def _close(a, b, rel):
    scale = np.maximum(np.abs(a), np.abs(b))
    err = np.abs(a - b)
    assert np.all(err <= rel * scale + 1e-300), float(np.max(err / (scale + 1e-300)))


def _same_as_fresh(spec, est, spec_moved, precision, seed):
    """est: a solve of the scene of `spec` after its instances moved to
    those of `spec_moved`. Expectations agree to the order of the FP64
    additions; variances and standard errors (sum w2 / N - E^2, which
    cancels) to 1e-6."""
    fresh = L.Scene(spec_moved, precision=precision)
    ref = fresh.solve(PATHS, seed=seed)
    assert est.stats.segments == ref.stats.segments
    (ma, sa), (mb, sb) = _outputs(spec, est), _outputs(spec_moved, ref)
    _close(ma, mb, 1e-6 if precision == "fp32" else 1e-10)
    _close(sa, sb, 1e-6)
    fresh.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_aiming_steps_refit_equals_fresh_build(gpu, precision):
    rng = np.random.default_rng(5)
    aim0 = rng.uniform(-3e-3, 3e-3, size=(N_HELIO, 2))
    spec = _field(aim0)
    lib = L.Scene(spec, precision=precision)
    lib.solve(100_000, seed=1)
    up0, rf0 = _counters(lib)
    for step in range(1, 4):
        aim = aim0 + rng.uniform(-2e-3, 2e-3, size=aim0.shape)
        moved = _field(aim)
        for h, m in zip(spec.instances[:-1], moved.instances[:-1]):
            lib.set_transform(h, m.transform)
        est = lib.solve(PATHS, seed=10 + step)
        up, rf = _counters(lib)
        assert (up, rf) == (up0, rf0 + step), (up, rf)   # refit, no upload
        _same_as_fresh(spec, est, moved, precision, 10 + step)
    lib.close()


def test_far_move_rebuilds(gpu):
    aim0 = np.zeros((N_HELIO, 2))
    spec = _field(aim0)
    lib = L.Scene(spec)
    lib.solve(100_000, seed=1)
    up0, rf0 = _counters(lib)
    moved = _field(aim0)
    t = list(moved.instances[7].transform)
    t[9] += 5000.0                          # 5 km east
    moved.instances[7].transform = t
    lib.set_transform(spec.instances[7], t)
    b0 = _counters(lib, builds=True)[2]
    est = lib.solve(PATHS, seed=21)
    up, rf, b = _counters(lib, builds=True)
    # refit tried, tree too loose: built anew on the GPU, no upload
    assert (up, rf, b) == (up0, rf0 + 1, b0 + 1), (up, rf, b)
    _same_as_fresh(spec, est, moved, "fp64", 21)
    # and moved back: refit again (the rebuilt tree is the new reference)
    lib.set_transform(spec.instances[7], spec.instances[7].transform)
    est = lib.solve(PATHS, seed=22)
    assert _counters(lib)[1] == rf0 + 2
    _same_as_fresh(spec, est, _field(aim0), "fp64", 22)
    lib.close()


def test_move_rejects_other_counts(gpu):
    lib = L.Scene(_field(np.zeros((N_HELIO, 2))))
    lib.solve(10_000, seed=1)
    ctx = C.c_void_p()
    L.call("ssol_device_get_gpu_context", lib.dev, C.byref(ctx))
    hip = _hip()
    buf = (C.c_ubyte * (192 * 5))()
    r = hip.ssolh_scene_move_instances(ctx, 5, C.cast(buf, C.c_void_p), 0, None, None)
    assert r != 0 and b"instances" in hip.ssolh_last_error()
    assert hip.ssolh_scene_move_instances(ctx, N_HELIO + 1, None, 0, None, None) != 0
    lib.close()


def _one_instance():
    plate = S.Object([(S.square_mesh(2.0), S.matte(0.0), S.matte(0.0))])
    inst = S.Instance(plate, S.transform(None, (0.0, 0.0, 0.0)), receiver_mask=S.FRONT, sample=1)
    sun = S.Sun("directional", S.sun_dir_from(180.0, 60.0), 1000.0, S.mono_spectrum())
    return S.Scene([inst], sun, None, name="one_plate")


@pytest.mark.parametrize("scene,precision", [("config3", "fp64"), ("config3", "fp32"),
                                             ("config1", "fp64"), ("one", "fp64")])
def test_gpu_built_tlas_traces_like_the_uploaded_one(gpu, scene, precision):
    spec = {"config3": S.config3, "config1": S.config1, "one": _one_instance}[scene]()
    lib = L.Scene(spec, precision=precision)
    ref = lib.solve(PATHS, seed=5)
    up0, rf0, b0 = _counters(lib, builds=True)
    assert _hip().ssolh_scene_build_tlas(_ctx(lib)) == 0, _hip().ssolh_last_error()
    est = lib.solve(PATHS, seed=5)
    assert _counters(lib, builds=True) == (up0, rf0, b0 + 1)
    assert est.stats.segments == ref.stats.segments
    g, h = est.global_(), ref.global_()
    rel = 1e-6 if precision == "fp32" else 1e-10
    for k in g:
        _close(np.array([g[k][0]]), np.array([h[k][0]]), rel)
    if scene == "config3":
        rcv = spec.instances[-1]
        shape = rcv.object.sshapes[0][0]
        _close(est.primitives(rcv, shape, S.FRONT, "absorbed_flux")[:, 0],
               ref.primitives(rcv, shape, S.FRONT, "absorbed_flux")[:, 0], rel)
    lib.close()
