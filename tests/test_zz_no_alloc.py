"""ssolh_solve_noalloc (include/ssolh.h), the entry point that allocates
nothing, against the entry point it varies, ssolh_solve followed by
ssolh_tallies_download, through the C-ABI of libssol_hip.so, on the scene an
ssol_solve of the public API uploaded.

Realisation ids are static, so any two widths follow the same paths: the
traced segment counts are equal and the tallies agree up to the order of
the FP64 additions (1e-10 relative; FP32 form: its FP32 per-path weights,
1e-6). Two calls in a row, the second with fewer realisations, both match.
A width that does not fit the scratch array is refused with
SSOLH_ERR_LIMIT before any launch."""
import ctypes as C

import numpy as np
import pytest

import ssol_ctypes as L
import ssol_scene as S

pytestmark = pytest.mark.gpu

SSOLH_ERR_LIMIT = 4


class SolveArgs(C.Structure):
    _fields_ = [("realisations", C.c_uint64), ("max_failed", C.c_uint64),
                ("rng_key", C.c_uint64), ("rng_counter", C.c_uint64),
                ("precision", C.c_uint32), ("path_slots", C.c_uint32),
                ("pending_entries", C.c_uint32), ("graph_iterations", C.c_uint32)]


class Tallies(C.Structure):
    _fields_ = [("global_", (C.c_double * 2) * 6),
                ("receivers", C.POINTER(C.c_double)), ("sampled", C.POINTER(C.c_double)),
                ("sampled_count", C.POINTER(C.c_uint64)),
                ("sampled_x_receivers", C.POINTER(C.c_double)),
                ("primitives", C.POINTER(C.c_double)),
                ("realisation_count", C.c_uint64), ("failed_count", C.c_uint64),
                ("segments", C.c_uint64), ("pending_spills", C.c_uint64),
                ("aborted", C.c_uint32), ("iterations", C.c_uint32),
                ("wavefronts", C.c_uint32), ("refilled", C.c_uint32),
                ("kernel_ms", C.c_double), ("trace_ms", C.c_double), ("shade_ms", C.c_double)]


def _hip():
    h = L.lib_hip()
    h.ssolh_solve.argtypes = [C.c_void_p, C.POINTER(SolveArgs)]
    h.ssolh_solve.restype = C.c_int
    h.ssolh_tallies_download.argtypes = [C.c_void_p, C.POINTER(Tallies)]
    h.ssolh_tallies_download.restype = C.c_int
    h.ssolh_solve_noalloc.argtypes = [C.c_void_p, C.POINTER(SolveArgs), C.POINTER(Tallies)]
    h.ssolh_solve_noalloc.restype = C.c_int
    h.ssolh_last_error.restype = C.c_char_p
    return h


class Out:
    """Host arrays for one download, sized generously from the scene
    description (the download writes the scene's own layout into them)."""

    def __init__(self, spec, lib):
        ninst = len(spec.instances)
        nprim = sum(lib.shape_ntris(sh) for inst in spec.instances if inst.per_primitive
                    for sh, _, _ in inst.object.sshapes)
        self.rcv = np.zeros(ninst * 40)
        self.samp = np.zeros(ninst * 4)
        self.cnt = np.zeros(ninst, dtype=np.uint64)
        self.sxr = np.zeros(ninst * ninst * 40)
        self.prim = np.zeros(max(nprim, 1) * 40)
        self.t = Tallies()
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        self.t.receivers, self.t.sampled = dp(self.rcv), dp(self.samp)
        self.t.sampled_count = self.cnt.ctypes.data_as(C.POINTER(C.c_uint64))
        self.t.sampled_x_receivers, self.t.primitives = dp(self.sxr), dp(self.prim)

    def arrays(self):
        g = np.array([[self.t.global_[i][j] for j in range(2)] for i in range(6)])
        return [g, self.rcv, self.samp, self.sxr, self.prim]


def _ctx(scene):
    ctx = C.c_void_p()
    L.call("ssol_device_get_gpu_context", scene.dev, C.byref(ctx))
    assert ctx.value
    return ctx


def _same(a, b, rel):
    assert a.t.realisation_count == b.t.realisation_count
    assert a.t.failed_count == b.t.failed_count
    assert a.t.segments == b.t.segments          # the same paths, segment for segment
    assert np.array_equal(a.cnt, b.cnt)
    for x, y in zip(a.arrays(), b.arrays()):
        scale = np.maximum(np.abs(x), np.abs(y))
        assert np.all(np.abs(x - y) <= rel * scale + 1e-300), np.max(np.abs(x - y) / (scale + 1e-300))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_noalloc_matches_solve(gpu, precision):
    spec = S.config2(0, n_helio=20, radius=150.0, rcv_size=5.0, rcv_cells=8)
    lib = L.Scene(spec, precision=precision)
    lib.solve(1000, seed=1)                 # flattens and uploads the scene
    hip, ctx = _hip(), _ctx(lib)
    p = 1 if precision == "fp32" else 0
    rel = 1e-6 if precision == "fp32" else 1e-10
    for n in (600_000, 150_000):            # two calls, the second smaller
        a = SolveArgs(n, 2**62, 11, 1000, p, 0, 0, 0)
        ref = Out(spec, lib)
        assert hip.ssolh_solve(ctx, C.byref(a)) == 0, hip.ssolh_last_error()
        assert hip.ssolh_tallies_download(ctx, C.byref(ref.t)) == 0
        got = Out(spec, lib)
        assert hip.ssolh_solve_noalloc(ctx, C.byref(a), C.byref(got.t)) == 0, hip.ssolh_last_error()
        assert got.t.realisation_count == n
        assert got.t.global_[1][0] > 0      # absorbed by receivers
        _same(got, ref, rel)
    lib.close()


def test_noalloc_narrow_width_and_limits(gpu):
    spec = S.config2(0, n_helio=20, radius=150.0, rcv_size=5.0, rcv_cells=8)
    lib = L.Scene(spec)
    lib.solve(1000, seed=1)
    hip, ctx = _hip(), _ctx(lib)
    # an explicit width that fits: chunks of 8192 realisations
    a = SolveArgs(100_000, 2**62, 5, 0, 0, 8192, 0, 0)
    got = Out(spec, lib)
    assert hip.ssolh_solve_noalloc(ctx, C.byref(a), C.byref(got.t)) == 0, hip.ssolh_last_error()
    assert got.t.wavefronts == 13
    a.path_slots = 0
    ref = Out(spec, lib)
    assert hip.ssolh_solve(ctx, C.byref(a)) == 0
    assert hip.ssolh_tallies_download(ctx, C.byref(ref.t)) == 0
    _same(got, ref, 1e-10)
    # a width whose path state cannot fit in the 64 MiB array: refused
    a.path_slots = 1 << 20
    assert hip.ssolh_solve_noalloc(ctx, C.byref(a), C.byref(got.t)) == SSOLH_ERR_LIMIT
    assert b"scratch array" in hip.ssolh_last_error()
    # graph-captured iterations are not offered
    a.path_slots, a.graph_iterations = 0, 8
    assert hip.ssolh_solve_noalloc(ctx, C.byref(a), C.byref(got.t)) == 1
    lib.close()
