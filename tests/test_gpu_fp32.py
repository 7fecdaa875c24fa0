"""The FP32 form (FP32 traversal, BSDFs and path state with per-instance
origin rebasing; FP64 quadric refinement and tallies) against the FP64
oracle on the reference's solver scenes and the benchmark fields.

Tolerance: |E_fp32 - E_oracle| <= 3 sqrt(SE_fp32^2 + SE_oracle^2)
+ 3e-5 |E_oracle| + 1e-6. The relative term bounds the single-precision
bias (float weights and positions: a few ulp of 2^-24 per bounce); the
reference's exact-equality checks (SE = 0 cases) are held to it instead of
1e-8."""
import math

import pytest

import ref_scenes as R
import ssol_ctypes as L
import ssol_scene as S

pytestmark = pytest.mark.gpu

REL = 3e-5


def _close(a, b):
    tol = 3.0 * math.hypot(a[2], b[2]) + REL * abs(b[0]) + 1e-6
    return abs(a[0] - b[0]) <= tol, tol


def compare32(scene, n, seed_gpu=21, seed_orc=22):
    g = R.Run("gpu32", scene, n, seed=seed_gpu)
    o = R.Run("oracle", scene, n, seed=seed_orc)
    assert g.realisation_count == o.realisation_count == n
    bad = []
    for f, v in g.global_().items():
        ok, tol = _close(v, o.global_()[f])
        if not ok:
            bad.append((f, v, o.global_()[f], tol))
    for inst in scene.instances:
        for side in (S.FRONT, S.BACK):
            if inst.receiver_mask & side:
                gv, ov = g.receiver(inst, side), o.receiver(inst, side)
                for f in L.RECEIVER_FIELDS:
                    ok, tol = _close(gv[f], ov[f])
                    if not ok:
                        bad.append((f, side, gv[f], ov[f], tol))
        if inst.sample:
            gs, os_ = g.sampled(inst), o.sampled(inst)
            for f in ("cos_factor", "shadowed"):
                ok, tol = _close(gs[f], os_[f])
                if not ok:
                    bad.append(("sampled." + f, gs[f], os_[f], tol))
    assert not bad, bad[:6]
    return g, o


@pytest.mark.parametrize("variant", ["all", "primary", "atm", "mono"])
def test_fp32_solver1(gpu, variant):
    compare32(R.solver1(variant)[0], 200000)


@pytest.mark.parametrize("builder", [R.solver3, R.solver4, R.solver5, R.solver6, R.solver7,
                                     R.solver8, R.solver9, lambda: R.solver2(True),
                                     lambda: R.solver2(False), lambda: R.by_receiver(True)])
def test_fp32_reference_scenes(gpu, builder):
    compare32(builder()[0], 200000)


@pytest.mark.parametrize("kind", ["pillbox", "buie", "gaussian"])
def test_fp32_solver10(gpu, kind):
    compare32(R.solver10(kind, 2)[0], 200000)


def test_fp32_paraboloid_field(gpu):
    compare32(S.config2(), 400000)


def test_fp32_beam_down_small_field(gpu):
    compare32(S.config3(n_helio=300), 1_000_000)


def test_fp32_beam_down_full_field(gpu):
    spec = S.config3()
    g = R.Run("gpu32", spec, 300_000, seed=5)
    o = R.Run("oracle", spec, 300_000, seed=6)
    for f in ("absorbed_by_receivers", "cos_factor", "shadowed", "missing",
              "extinguished_by_atmosphere", "other_absorbed"):
        ok, tol = _close(g.global_()[f], o.global_()[f])
        assert ok, (f, g.global_()[f], o.global_()[f], tol)


def test_fp32_fp64_same_streams_agree(gpu):
    """Same seed, both forms on the GPU: the FP32 form draws 24-bit uniforms
    from the same Threefry blocks, so the runs are correlated; their
    difference stays within the statistical bound."""
    spec = S.config3(n_helio=300)
    a = R.Run("gpu", spec, 2_000_000, seed=7)
    b = R.Run("gpu32", spec, 2_000_000, seed=7)
    ok, tol = _close(b.global_()["absorbed_by_receivers"], a.global_()["absorbed_by_receivers"])
    assert ok, (a.global_()["absorbed_by_receivers"], b.global_()["absorbed_by_receivers"], tol)
