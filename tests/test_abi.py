"""The C-ABI libraries on the CPU: every entry point the headers declare is
exported, and the host-side contract that needs no GPU holds (argument
errors, data lookup, areas, RNG state, the failure without a device).

References: src/ssol.h (API), src/test_ssol_data.c:49-76 (data lookup),
src/test_ssol_solver3.c:161-162 (punched plane area), src/ssol_solver.c:
1160-1171 (ssol_solve argument errors)."""
import ctypes as C
import os
import re

import pytest

import ssol_ctypes as L
import ssol_scene as S
from conftest import ROOT

RES_OK, RES_BAD_ARG = 0, 1


def _declared(header, macro):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = "\n".join(l for l in text.splitlines() if not l.lstrip().startswith("#"))
    return sorted(set(re.findall(macro + r"\s+[\w\s\*]*?\b(\w+)\s*\(", text)))


@pytest.mark.parametrize("header,macro,lib", [
    ("ssol.h", "SSOL_API", "libssol.so"),
    ("ssp.h", "SSOL_API", "libssol.so"),
    ("ssolh.h", "SSOLH_API", "libssol_hip.so"),
])
def test_every_declared_symbol_is_exported(header, macro, lib):
    names = _declared(header, macro)
    assert len(names) > 10
    L.lib()
    so = C.CDLL(os.path.join(L.LIB_DIR, lib))
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


def test_core_entry_points_present():
    names = set(_declared("ssol.h", "SSOL_API"))
    for n in ("ssol_solve", "ssol_estimator_get_mc_global", "ssol_estimator_get_mc_receiver",
              "ssol_estimator_get_rng_state", "ssol_mc_receiver_get_mc_shape",
              "ssol_mc_shape_get_mc_primitive", "ssol_estimator_get_mc_sampled_x_receiver",
              "ssol_estimator_get_realisation_count", "ssol_estimator_get_failed_count",
              "ssol_estimator_get_sampled_area"):
        assert n in names


def test_data_lookup_reference_values():
    lib = L.lib()
    sc = L.Scene(S.config1())
    spec = S.Spectrum([float(i + 1) for i in range(10)], [float(i + 2) for i in range(10)])
    h = sc._spectrum(spec)
    d = L.Data()
    lib.ssol_data_set_real(C.byref(d), 1.25)
    get = lib.ssol_data_get_value
    get.restype = C.c_double
    get.argtypes = [C.POINTER(L.Data), C.c_double]
    assert get(C.byref(d), 1.0) == 1.25 and get(C.byref(d), 1.234) == 1.25
    lib.ssol_data_set_spectrum(C.byref(d), h)
    for i in range(10):
        assert get(C.byref(d), float(i + 1)) == float(i + 2)
    assert abs(get(C.byref(d), 1.5) - 2.5) <= 1e-6
    assert abs(get(C.byref(d), 1.25) - 2.25) <= 1e-6
    assert get(C.byref(d), 0.5) == 2 and get(C.byref(d), 0.1) == 2
    assert get(C.byref(d), 10) == 11 and get(C.byref(d), 10.1) == 11
    lib.ssol_data_clear(C.byref(d))
    sc.close()


def test_punched_plane_area():
    """A 20 x 20 m punched plane has area 400 to DBL_EPSILON."""
    import ref_scenes as R
    scene, _ = R.solver3()
    sc = L.Scene(scene)
    inst = scene.instances[0]
    area = C.c_double()
    assert L.lib().ssol_instance_get_area(sc.instances[id(inst)], C.byref(area)) == RES_OK
    assert abs(area.value - 400) <= 2.220446049250313e-16
    sc.close()


def test_solve_argument_errors():
    lib = L.lib()
    sc = L.Scene(S.config1())
    rng = L.Rng(0, 0)
    out = C.c_void_p()
    assert lib.ssol_solve(None, rng.h, 10, 0, None, C.byref(out)) == RES_BAD_ARG
    assert lib.ssol_solve(sc.scn, None, 10, 0, None, C.byref(out)) == RES_BAD_ARG
    assert lib.ssol_solve(sc.scn, rng.h, 0, 0, None, C.byref(out)) == RES_BAD_ARG
    assert lib.ssol_solve(sc.scn, rng.h, 10, 0, None, None) == RES_BAD_ARG
    assert lib.ssol_solve(sc.scn, rng.h, 2**63 + 5, 0, None, C.byref(out)) == RES_BAD_ARG
    assert lib.ssol_scene_detach_sun(sc.scn, sc.sun) == RES_OK
    assert lib.ssol_solve(sc.scn, rng.h, 10, 0, None, C.byref(out)) == RES_BAD_ARG
    assert not out.value
    sc.close()


def test_solve_fails_without_gpu():
    """No silent CPU path: without a visible device ssol_solve returns an error."""
    hip = L.lib_hip()
    hip.ssolh_device_count.restype = C.c_int
    if hip.ssolh_device_count() > 0:
        pytest.skip("a GPU is visible")
    sc = L.Scene(S.config1())
    with pytest.raises(L.SsolError):
        sc.solve(100)
    sc.close()


def test_rng_state_round_trip():
    lib = L.lib()
    lib.ssp_rng_get.restype = C.c_uint64
    lib.ssp_rng_canonical.restype = C.c_double
    a, b = L.Rng(1234, 0), L.Rng(0, 0)
    for _ in range(7):
        lib.ssp_rng_get(a.h)
    libc = C.CDLL(None)
    libc.tmpfile.restype = C.c_void_p
    libc.rewind.argtypes = [C.c_void_p]
    libc.fclose.argtypes = [C.c_void_p]
    f = libc.tmpfile()
    assert lib.ssp_rng_write(a.h, C.c_void_p(f)) == RES_OK
    libc.rewind(f)
    assert lib.ssp_rng_read(b.h, C.c_void_p(f)) == RES_OK
    libc.fclose(f)
    assert [lib.ssp_rng_get(a.h) for _ in range(9)] == [lib.ssp_rng_get(b.h) for _ in range(9)]
    # a fresh generator on the same seed replays the same numbers
    c, d = L.Rng(99, 0), L.Rng(99, 0)
    xs = [lib.ssp_rng_canonical(c.h) for _ in range(5)]
    assert xs == [lib.ssp_rng_canonical(d.h) for _ in range(5)]
    assert all(0 <= x < 1 for x in xs)
    t = C.c_int()
    assert lib.ssp_rng_get_type(c.h, C.byref(t)) == RES_OK and t.value == 4
    bad = C.c_void_p()
    assert lib.ssp_rng_create(None, 0, C.byref(bad)) == RES_BAD_ARG


def test_threefry_export_matches_oracle():
    import oracle_ctypes as O
    lib = L.lib()
    A = C.c_uint64 * 4
    for ctr, key in (([0, 0, 0, 0], [0, 0, 0, 0]), ([1, 2, 3, 4], [9, 8, 7, 6]),
                     ([2**64 - 1] * 4, [2**64 - 1] * 4)):
        out = A()
        lib.ssp_threefry4x64_20(A(*ctr), A(*key), out)
        assert list(out) == O.threefry(ctr, key)


def test_c_client_links_against_the_headers_and_libssol():
    """The C client (tests/c_client/config1_client.c) compiles against
    include/ssol.h and include/ssp.h alone and resolves every symbol from
    lib/libssol.so; without a GPU it builds its scene through the C API and
    stops at ssol_solve with an error (no CPU fallback)."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    client = os.path.join(root, "tests", "c_client", "config1_client")
    assert os.path.exists(client), "build the client with __graft_entry__.build()"
    ldd = subprocess.run(["ldd", client], capture_output=True, text=True).stdout
    assert "libssol.so =>" in ldd and "not found" not in ldd, ldd
    p = subprocess.run([client, "1000"], capture_output=True, text=True, timeout=120)
    if p.returncode != 0:
        assert p.returncode == 2 and "ssol_solve" in p.stderr, (p.returncode, p.stderr)
