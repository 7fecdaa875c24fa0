"""Assertions of the reference's known-answer solver tests, with the same
values and tolerances (eq_eps is an absolute tolerance). Each check takes a
ref_scenes.Run (GPU or oracle) so both solvers are held to the same bar.
File:line of each assertion group is given in its docstring."""
from __future__ import annotations

import math

import ref_scenes as R
import ssol_scene as S

DNI = 1000.0


def eq_eps(a, b, eps):
    return abs(a - b) <= eps


def solver1_all(n, backend, seed=0):
    """src/test_ssol_solver1.c:313-352"""
    scn, h = R.solver1("all")
    r = R.Run(backend, scn, n, seed=seed)
    assert r.realisation_count == n and r.failed_count == 0
    c = math.cos(math.pi / 4)
    m = 4 * DNI * c
    std = math.sqrt(((12 * DNI * c) ** 2 / 3 - (4 * DNI * c) ** 2) / n)
    g = r.global_()
    assert eq_eps(g["shadowed"][0], m, 2 * std), g["shadowed"]
    assert eq_eps(g["missing"][0], 2 * m, 2 * g["missing"][2]), g["missing"]
    rc = r.receiver(h["target"])
    assert eq_eps(rc["incoming_flux"][0], m, 3 * std), rc["incoming_flux"]
    assert eq_eps(rc["incoming_flux"][2], std, std * 1e-2), (rc["incoming_flux"], std)
    assert abs(r.sampled_area - 12) <= 12 * 2.2e-16 * 4
    return r


def solver1_primary(n, backend, variant="primary", seed=0):
    """src/test_ssol_solver1.c:354-400 (primary only, then zero extinction)"""
    scn, h = R.solver1(variant)
    r = R.Run(backend, scn, n, seed=seed)
    assert r.realisation_count == n
    c = math.cos(math.pi / 4)
    m = 4 * DNI * c
    g = r.global_()
    assert eq_eps(g["shadowed"][0], 0, 1e-4)
    assert eq_eps(g["missing"][0], m, 1e-4), g["missing"]
    assert eq_eps(g["cos_factor"][0], c, 1e-4)
    rc = r.receiver(h["target"])
    assert eq_eps(rc["incoming_flux"][0], m, 1e-8), rc["incoming_flux"]
    assert eq_eps(rc["incoming_flux"][2], 0, 1e-4)
    return r


def solver1_atm(n, backend, seed=0):
    """src/test_ssol_solver1.c:402-485 (reflectivity 0.87, extinction 0.03,
    per-primitive map summed over the target triangles)"""
    scn, h = R.solver1("atm")
    r = R.Run(backend, scn, n, seed=seed)
    assert r.realisation_count == n
    c = math.cos(math.pi / 4)
    m = 4 * DNI * c
    K = math.exp(-0.03 * 4 * math.sqrt(2))
    a_m = 0.87 * 4 * K * DNI * c
    g = r.global_()
    assert eq_eps(g["shadowed"][0], 0, 1e-4)
    total = sum(g[k][0] for k in ("missing", "shadowed", "absorbed_by_receivers",
                                   "extinguished_by_atmosphere", "other_absorbed"))
    assert eq_eps(total, m, 1e-4), (total, m)
    assert eq_eps(g["cos_factor"][0], c, 1e-4)
    rc = r.receiver(h["target"])
    assert eq_eps(rc["incoming_flux"][0], a_m, 1e-4), (rc["incoming_flux"], a_m)
    assert eq_eps(rc["incoming_flux"][2], 0, 1e-4)
    # per-primitive map x triangle area sums to the same value
    acc = 0.0
    for prim in range(2):
        p = r.primitive(h["target"], 0, prim)
        acc += p["incoming_flux"][0] * 2.0  # each triangle of the 2x2 square is 2 m2
    assert eq_eps(acc, a_m, 1e-4), (acc, a_m)
    return r


def solver1_mono(n, backend, seed=0):
    """src/test_ssol_solver1.c:491-528 (k = 0.1 + 0.21 * 0.1 = 0.121)"""
    scn, h = R.solver1("mono")
    r = R.Run(backend, scn, n, seed=seed)
    c = math.cos(math.pi / 4)
    m = 4 * math.exp(-0.121 * 4 * math.sqrt(2)) * DNI * c
    g = r.global_()
    assert eq_eps(g["shadowed"][0], 0, 1e-4)
    assert eq_eps(g["missing"][0], m, 1e-4), (g["missing"], m)
    assert eq_eps(g["cos_factor"][0], c, 1e-4)
    rc = r.receiver(h["target"])
    assert eq_eps(rc["incoming_flux"][0], m, 1e-4), (rc["incoming_flux"], m)
    assert eq_eps(rc["incoming_flux"][2], 0, 1e-4)
    return r


def solver2(n, backend, seed=0):
    """src/test_ssol_solver2.c:179-204 (mesh heliostats, punched secondary:
    the target gets the whole beam, SE 0)"""
    scn, h = R.solver2(True)
    r = R.Run(backend, scn, n, seed=seed)
    assert r.realisation_count == n
    c = math.cos(math.pi / 4)
    g = r.global_()
    assert eq_eps(g["shadowed"][0], 0, 1e-4)
    assert eq_eps(g["missing"][0], 4 * DNI * c, 1e-4), g["missing"]
    assert eq_eps(g["cos_factor"][0], c, 1e-4)
    with_back = r.receiver_or_none(h["heliostat1"], S.BACK)
    assert with_back is None
    r.receiver(h["secondary"])
    rc = r.receiver(h["target"])
    assert eq_eps(rc["incoming_flux"][0], 4 * DNI * c, 1e-8), rc["incoming_flux"]
    assert eq_eps(rc["incoming_flux"][2], 0, 1e-4)
    return r


def solver2b(n, backend, seed=0):
    """src/test_ssol_solver2b.c:183-206 (punched heliostats, half-size target)"""
    scn, h = R.solver2(False)
    r = R.Run(backend, scn, n, seed=seed)
    assert r.realisation_count == n
    c = math.cos(math.pi / 4)
    m = 2 * DNI * c
    std = math.sqrt(((4 * DNI * c) ** 2 / 2 - (2 * DNI * c) ** 2) / n)
    g = r.global_()
    assert eq_eps(g["shadowed"][0], 0, 1e-4)
    assert eq_eps(g["missing"][0], 4 * DNI * c, 1e-4), g["missing"]
    assert eq_eps(g["cos_factor"][0], c, 1e-4)
    rc = r.receiver(h["target"])
    assert eq_eps(rc["incoming_flux"][0], m, 2 * std), rc["incoming_flux"]
    assert eq_eps(rc["incoming_flux"][2], std, 1e-3), (rc["incoming_flux"], std)
    return r


def solver3(n, backend, seed=0):
    """src/test_ssol_solver3.c:138-163 (20x20 punched plane; area 400)"""
    scn, h = R.solver3()
    r = R.Run(backend, scn, n, seed=seed)
    assert r.realisation_count == n and r.failed_count == 0
    c = math.cos(math.pi / 4)
    m = 4 * DNI * c
    std = math.sqrt(((400 * DNI * c) ** 2 / 100 - (4 * DNI * c) ** 2) / n)
    g = r.global_()
    assert eq_eps(g["shadowed"][0], 0, 1e-4)
    assert eq_eps(g["missing"][0], 400 * DNI * c, 1e-4), g["missing"]
    assert eq_eps(g["cos_factor"][0], c, 1e-4)
    rc = r.receiver(h["target"])
    assert eq_eps(rc["incoming_flux"][0], m, 2 * std), rc["incoming_flux"]
    assert eq_eps(rc["incoming_flux"][2], std, 10), (rc["incoming_flux"], std)
    assert abs(r.sampled_area - 400) <= 2.220446049250313e-16 * 400
    return r


def solver5(n, backend, seed=0):
    """src/test_ssol_solver5.c:138-159 (parabolic cylinder focal line: the
    whole 400 m2 aperture flux, 1e-8)"""
    scn, h = R.solver5()
    r = R.Run(backend, scn, n, seed=seed)
    assert r.realisation_count == n
    g = r.global_()
    assert eq_eps(g["shadowed"][0], 0, 1e-4)
    assert eq_eps(g["missing"][0], 400 * DNI, 1e-4), g["missing"]
    rc = r.receiver(h["target"])
    assert eq_eps(rc["incoming_flux"][0], 400 * DNI, 1e-8), rc["incoming_flux"]
    assert eq_eps(rc["incoming_flux"][2], 0, 1e-4)
    return r


def solver8(n, backend, seed=0):
    """src/test_ssol_solver8.c:141-160 (y-carved parabolic cylinder under an
    oblique sun: projected area sqrt(2) * 10 * 4)"""
    scn, h = R.solver8()
    r = R.Run(backend, scn, n, seed=seed)
    assert r.realisation_count == n and r.failed_count == 0
    area = math.sqrt(2) * 10 * 4
    g = r.global_()
    assert eq_eps(g["shadowed"][0], 0, 1e-4)
    assert eq_eps(g["missing"][0], area * DNI, 3 * g["missing"][2]), g["missing"]
    rc = r.receiver(h["target"])
    assert eq_eps(rc["incoming_flux"][0], area * DNI, 2 * rc["incoming_flux"][2]), \
        rc["incoming_flux"]
    return r


def solver4(n, backend, seed=0):
    """src/test_ssol_solver4.c:146-172 (focus of a paraboloid: both stacked
    virtual targets receive the whole 400 m2 aperture flux, +-0.01 W)"""
    scn, h = R.solver4()
    r = R.Run(backend, scn, n, seed=seed)
    assert r.realisation_count == n and r.failed_count == 0
    g = r.global_()
    assert eq_eps(g["shadowed"][0], 0, 1e-4)
    assert eq_eps(g["missing"][0], 400 * DNI, 1e-2), g["missing"]
    for t in ("target1", "target2"):
        rc = r.receiver(h[t])
        assert eq_eps(rc["incoming_flux"][0], 400 * DNI, 1e-2), (t, rc["incoming_flux"])
        assert eq_eps(rc["incoming_flux"][2], 0, 1e-2), (t, rc["incoming_flux"])
    return r


def solver6(n, backend, seed=0):
    """src/test_ssol_solver6.c:182-235 (shadowing by matte targets seen
    through the reversed shadow ray; RNG-state chaining and merge)"""
    scn, h = R.solver6()
    r = R.Run(backend, scn, n, seed=seed)
    assert r.realisation_count == n
    g = r.global_()
    assert eq_eps(g["shadowed"][0], 100000, 2 * 100000 / math.sqrt(n)), g["shadowed"]
    assert eq_eps(g["missing"][0], 0, 0)
    rc = r.receiver(h["target1"])
    assert eq_eps(rc["incoming_flux"][0], 100000, 2 * 100000 / math.sqrt(n))
    assert rc["incoming_flux"][0] == rc["absorbed_flux"][0]
    rc = r.receiver(h["target2"], S.BACK)
    assert eq_eps(rc["incoming_flux"][0], 0, 1)
    assert rc["incoming_flux"][0] == rc["absorbed_flux"][0]
    return r, scn, h


def solver6_chain(n, backend, seed=0):
    """src/test_ssol_solver6.c:204-235: a second solve seeded from the first
    estimator's RNG state with 3x the realisations is independent."""
    r1, scn, h = solver6(n, backend, seed=seed)
    g1 = r1.global_()
    if backend == "oracle":
        # the oracle's stream of realisation i is (key, counter + i): the
        # continuation starts at counter + n
        r2 = R.Run(backend, scn, 3 * n, seed=seed, counter=n)
    else:
        lib = r1.lib
        e2 = lib.solve(3 * n, rng=r1.e.rng_state)
        r2 = r1.__class__.__new__(r1.__class__)
        r2.backend, r2.e, r2.lib = backend, e2, lib
    assert r2.realisation_count == 3 * n
    g2 = r2.global_()
    assert eq_eps(g2["shadowed"][0], 100000, 2 * 100000 / math.sqrt(n))
    assert eq_eps(g1["shadowed"][2] / math.sqrt(3), g2["shadowed"][2], 1e-1)
    assert g1["shadowed"][0] != g2["shadowed"][0]
    assert g1["shadowed"][2] != g2["shadowed"][2]
    sum_w = g1["shadowed"][0] * n + g2["shadowed"][0] * 3 * n
    V = g1["shadowed"][2] ** 2 * n
    sum_w2 = (V + g1["shadowed"][0] ** 2) * n
    V = g2["shadowed"][2] ** 2 * 3 * n
    sum_w2 += (V + g2["shadowed"][0] ** 2) * 3 * n
    E = sum_w / (4 * n)
    V = sum_w2 / (4 * n) - E * E
    SE = math.sqrt(V / (4 * n))
    assert eq_eps(E, 100000, 3 * SE)
    assert eq_eps(SE, g1["shadowed"][2] / 2, SE * 1e-2)


def solver7(n, backend, seed=0):
    """src/test_ssol_solver7.c:189-211 (beam-down: paraboloid + hyperboloid)"""
    scn, h = R.solver7()
    r = R.Run(backend, scn, n, seed=seed)
    total = 16 * DNI
    assert eq_eps(r.sampled_area * DNI, total, total * 1e-4), r.sampled_area
    rc = r.receiver(h["target"])
    assert eq_eps(rc["incoming_flux"][0], total, total * 1e-4), rc["incoming_flux"]
    assert eq_eps(rc["absorbed_flux"][0], 0.9 * total, 0.9 * total * 1e-4), rc["absorbed_flux"]
    assert eq_eps(rc["incoming_flux"][2], 0, 1e-2), rc["incoming_flux"]
    return r


def solver9(n, backend, seed=0):
    """src/test_ssol_solver9.c:135-154 (mirror cube under a virtual target)"""
    scn, h = R.solver9()
    r = R.Run(backend, scn, n, seed=seed)
    assert r.realisation_count == n and r.failed_count == 0
    g = r.global_()
    assert eq_eps(g["cos_factor"][0], 1 / 3, 3 * g["cos_factor"][2]), g["cos_factor"]
    assert eq_eps(g["shadowed"][0], DNI * 100, 3 * g["shadowed"][2]), g["shadowed"]
    assert eq_eps(g["missing"][0], DNI * 100, 3 * g["missing"][2]), g["missing"]
    rc = r.receiver(h["target"])
    assert eq_eps(rc["incoming_flux"][0], DNI * 60, 3 * rc["incoming_flux"][2]), rc["incoming_flux"]
    return r


def solver10(n1, n2, backend, sun_kind="pillbox", seed=0):
    """src/test_ssol_solver10.c:128-181 (and 11.c Buie, 12.c Gaussian):
    absorbed == DNI S exactly with SE == 0 and cos == 1 on a flat absorber;
    then a paraboloid absorbs DNI S within 3 SE."""
    scn, h = R.solver10(sun_kind, 1)
    r = R.Run(backend, scn, n1, seed=seed)
    assert r.realisation_count == n1 and r.failed_count == 0
    g = r.global_()
    assert g["cos_factor"][0] == 1 and g["cos_factor"][2] == 0, g["cos_factor"]
    assert g["absorbed_by_receivers"][0] == DNI * 100, g["absorbed_by_receivers"]
    assert g["absorbed_by_receivers"][2] == 0, g["absorbed_by_receivers"]
    scn2, h2 = R.solver10(sun_kind, 2)
    r2 = R.Run(backend, scn2, n2, seed=seed)
    assert r2.realisation_count == n2 and r2.failed_count == 0
    g2 = r2.global_()
    assert eq_eps(g2["absorbed_by_receivers"][0], DNI * 100,
                  3 * g2["absorbed_by_receivers"][2]), g2["absorbed_by_receivers"]
    return r, r2


def by_receiver(n, backend, seed=0):
    """src/test_ssol_by_receiver_integration.c:134-162"""
    s_dni_cos = 4 * DNI * math.cos(math.pi / 4)
    scn, h = R.by_receiver(False)
    r = R.Run(backend, scn, n, seed=seed)
    rc = r.receiver(h["target"])
    assert eq_eps(rc["incoming_flux"][0], s_dni_cos, rc["incoming_flux"][2] * 3), rc["incoming_flux"]
    scn, h = R.by_receiver(True)
    r = R.Run(backend, scn, 3 * n, seed=seed + 1)
    rc = r.receiver(h["target"])
    assert eq_eps(rc["incoming_flux"][0], s_dni_cos, rc["incoming_flux"][2] * 3), rc["incoming_flux"]
    sx = r.sampled_x_receiver(h["heliostat"], h["target"])
    assert eq_eps(sx["incoming_flux"][0], s_dni_cos, sx["incoming_flux"][2] * 3), sx["incoming_flux"]
    return r
