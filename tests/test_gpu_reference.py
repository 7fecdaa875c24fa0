"""The GPU solver on the reference's known-answer solver tests (same values
and tolerances as src/test_ssol_solver*.c), then every global, receiver,
sampled and per-primitive estimate of those scenes against the oracle within
3 combined standard errors."""
import pytest

import ref_checks as K
import ref_scenes as R
import ssol_scene as S
from conftest import within
import ssol_ctypes as L

pytestmark = pytest.mark.gpu


def test_solver1_all(gpu):
    K.solver1_all(10000, "gpu")


def test_solver1_primary(gpu):
    K.solver1_primary(10000, "gpu")


def test_solver1_atmosphere_zero(gpu):
    K.solver1_primary(10000, "gpu", variant="atm0")


def test_solver1_atmosphere(gpu):
    K.solver1_atm(10000, "gpu")


def test_solver1_monochromatic_spectral_extinction(gpu):
    K.solver1_mono(10000, "gpu")


def test_solver2_mesh_heliostats(gpu):
    K.solver2(10000, "gpu")


def test_solver2b_punched_heliostats(gpu):
    K.solver2b(50000, "gpu")


def test_solver3_punched_plane(gpu):
    K.solver3(20000, "gpu")


def test_solver5_parabolic_cylinder(gpu):
    K.solver5(10000, "gpu")


def test_solver8_carved_parabolic_cylinder(gpu):
    K.solver8(100000, "gpu")


def test_solver4_paraboloid_focus(gpu):
    K.solver4(100000, "gpu")


def test_solver6_shadowing_and_rng_chaining(gpu):
    K.solver6_chain(10000, "gpu")


def test_solver7_beam_down(gpu):
    K.solver7(10000, "gpu")


def test_solver9_cube(gpu):
    K.solver9(100000, "gpu")


def test_solver10_pillbox(gpu):
    K.solver10(10000, 100000, "gpu", "pillbox")


def test_solver11_buie(gpu):
    K.solver10(10000, 100000, "gpu", "buie")


def test_solver12_gaussian(gpu):
    K.solver10(10000, 100000, "gpu", "gaussian")


def test_by_receiver_integration(gpu):
    K.by_receiver(10000, "gpu")


# Every estimate against the oracle ------------------------------------------------
def compare(scene, n, seed_gpu=11, seed_orc=12, k=3.0):
    g = R.Run("gpu", scene, n, seed=seed_gpu)
    o = R.Run("oracle", scene, n, seed=seed_orc)
    assert g.realisation_count == o.realisation_count == n
    bad = []

    def chk(name, a, b):
        ok, tol = within(a, b, k, floor=1e-9 * max(1.0, abs(b[0])))
        if not ok:
            bad.append((name, a, b, tol))
    for f, v in g.global_().items():
        chk(f"global.{f}", v, o.global_()[f])
    for inst in scene.instances:
        for side in (S.FRONT, S.BACK):
            if inst.receiver_mask & side:
                gv, ov = g.receiver(inst, side), o.receiver(inst, side)
                for f in L.RECEIVER_FIELDS:
                    chk(f"rcv[{scene.instances.index(inst)},{side}].{f}", gv[f], ov[f])
                if inst.per_primitive:
                    for s_idx, (shape, _, _) in enumerate(inst.object.sshapes):
                        ntris = g.lib.shape_ntris(shape)
                        for prim in range(ntris):
                            gp = g.primitive(inst, s_idx, prim, side)
                            op = o.primitive(inst, s_idx, prim, side)
                            for f in L.RECEIVER_FIELDS:
                                chk(f"prim[{prim},{side}].{f}", gp[f], op[f])
        if inst.sample:
            gs, os_ = g.sampled(inst), o.sampled(inst)
            assert gs["nb_samples"] > 0
            chk("sampled.cos", gs["cos_factor"], os_["cos_factor"])
            chk("sampled.shadowed", gs["shadowed"], os_["shadowed"])
            for rcv in scene.instances:
                for side in (S.FRONT, S.BACK):
                    if rcv.receiver_mask & side:
                        gx = g.sampled_x_receiver(inst, rcv, side)
                        ox = o.sampled_x_receiver(inst, rcv, side)
                        for f in L.RECEIVER_FIELDS:
                            chk(f"sxr.{f}", gx[f], ox[f])
    assert not bad, bad[:6]


@pytest.mark.parametrize("variant", ["all", "atm", "mono"])
def test_solver1_vs_oracle(gpu, variant):
    compare(R.solver1(variant)[0], 200000)


@pytest.mark.parametrize("builder", [R.solver4, R.solver6, R.solver7, R.solver9, R.solver3,
                                     R.solver5, R.solver8, lambda: R.solver2(True),
                                     lambda: R.solver2(False)])
def test_reference_scene_vs_oracle(gpu, builder):
    compare(builder()[0], 200000)


@pytest.mark.parametrize("kind", ["pillbox", "buie", "gaussian"])
def test_solver10_paraboloid_vs_oracle(gpu, kind):
    compare(R.solver10(kind, 2)[0], 200000)


def test_by_receiver_vs_oracle(gpu):
    compare(R.by_receiver(True)[0], 200000)
