"""Scenes of the reference's known-answer solver tests, restated as
ssol_scene descriptions (same geometry, materials, suns and transforms), and
a uniform runner over the two solvers (the product's ssol_solve on the GPU
and the oracle).

The reference tests use d33_rotation_pitch for a rotation about X and
d33_rotation_yaw for a rotation about Y; the beam-down test's closed form
(src/test_ssol_solver7.c:189-211) fixes this reading of yaw.
"""
from __future__ import annotations

import math

import numpy as np

import ssol_scene as S

PI = math.pi


def spectrum3():
    """wavelengths {1,2,3}, intensities {1,0.8,1} (every solver test)."""
    return S.Spectrum([1.0, 2.0, 3.0], [1.0, 0.8, 1.0])


def t_pitch(a, tx=0.0, ty=0.0, tz=0.0):
    return S.transform(S.rot_x(a), (tx, ty, tz))


def t_yaw(a, tx=0.0, ty=0.0, tz=0.0):
    return S.transform(S.rot_y(a), (tx, ty, tz))


def t_move(tx=0.0, ty=0.0, tz=0.0):
    return S.transform(None, (tx, ty, tz))


# src/test_ssol_solver1.c -------------------------------------------------------
def solver1(variant: str):
    """variant: 'all' (every instance sampled), 'primary' (heliostat only),
    'atm0' (primary, zero extinction), 'atm' (reflectivity 0.87 heliostat,
    extinction 0.03, per-primitive target), 'mono' (wavelength 1.21,
    spectral extinction {1: 0.1, 2: 0.2})."""
    square = S.square_mesh(1.0)
    m_mtl = S.mirror(1.0)
    v_mtl = S.virtual()
    m_obj = S.Object([(square, m_mtl, m_mtl)])
    t_obj = S.Object([(square, v_mtl, v_mtl)])
    helio_obj = m_obj
    if variant == "atm":
        m2 = S.mirror(0.87)
        helio_obj = S.Object([(square, m2, m2)])
    helio = S.Instance(helio_obj, S.IDENTITY, receiver_mask=S.FRONT)
    secondary = S.Instance(m_obj, t_pitch(PI, 2, 0, 2), receiver_mask=S.FRONT)
    target = S.Instance(t_obj, t_move(4, 0, 0), receiver_mask=S.FRONT,
                        per_primitive=1 if variant == "atm" else 0)
    if variant != "all":
        secondary.sample = 0
        target.sample = 0
    spectrum = spectrum3()
    atm = None
    if variant == "atm0":
        atm = S.Data(0.0)
    elif variant == "atm":
        atm = S.Data(0.03)
    elif variant == "mono":
        spectrum = S.Spectrum([1.21], [1.0])
        atm = S.Data(0.0, S.Spectrum([1.0, 2.0], [0.1, 0.2]))
    sun = S.Sun("directional", (1, 0, -1), 1000.0, spectrum)
    scn = S.Scene([helio, secondary, target], sun, atm, name=f"solver1_{variant}")
    return scn, {"heliostat": helio, "secondary": secondary, "target": target, "square": square}


# src/test_ssol_solver2.c, src/test_ssol_solver2b.c ------------------------------
def solver2(mesh_heliostats: bool):
    """Two 1x2 m heliostats side by side, a flipped 2x2 m punched-plane
    secondary at (2, 0, 2) and a virtual target at x = 4. solver2: mesh
    heliostats and a 2x2 mesh target; solver2b (mesh_heliostats False):
    punched-plane heliostats and a 1x2 mesh target that sees half the beam."""
    m_mtl = S.mirror(1.0)
    v_mtl = S.virtual()
    rect = S.rect_mesh(np.float32(-0.5), np.float32(0.5), np.float32(-1), np.float32(1))
    if mesh_heliostats:
        h_shape = rect
        t_shape = S.square_mesh(1.0)
    else:
        h_shape = S.Punched(S.Q_PLANE, (0, 0), list(S.IDENTITY), [S.rect_carving(0.5, 1)])
        t_shape = rect
    sq = S.Punched(S.Q_PLANE, (0, 0), list(S.IDENTITY), [S.rect_carving(1, 1)])
    m_obj = S.Object([(h_shape, m_mtl, m_mtl)])
    h1 = S.Instance(m_obj, t_move(-0.5), receiver_mask=S.FRONT)
    h2 = S.Instance(m_obj, t_move(0.5), receiver_mask=S.FRONT)
    secondary = S.Instance(S.Object([(sq, m_mtl, m_mtl)]), t_pitch(PI, 2, 0, 2),
                           receiver_mask=S.FRONT, sample=0)
    target = S.Instance(S.Object([(t_shape, v_mtl, v_mtl)]), t_move(4),
                        receiver_mask=S.FRONT, sample=0)
    sun = S.Sun("directional", (1, 0, -1), 1000.0, spectrum3())
    name = "solver2" if mesh_heliostats else "solver2b"
    return (S.Scene([h1, h2, secondary, target], sun, name=name),
            {"heliostat1": h1, "secondary": secondary, "target": target})


# src/test_ssol_solver3.c -------------------------------------------------------
def solver3():
    """A 20x20 m punched-plane heliostat under a 2x2 m virtual target flipped
    at z = 2."""
    quad = S.Punched(S.Q_PLANE, (0, 0), list(S.IDENTITY), [S.rect_carving(10, 10)])
    m_mtl = S.mirror(1.0)
    v_mtl = S.virtual()
    helio = S.Instance(S.Object([(quad, m_mtl, m_mtl)]), S.IDENTITY, receiver_mask=S.FRONT)
    target = S.Instance(S.Object([(S.square_mesh(1.0), v_mtl, v_mtl)]), t_pitch(PI, 0, 0, 2),
                        receiver_mask=S.FRONT, sample=0)
    sun = S.Sun("directional", (1, 0, -1), 1000.0, spectrum3())
    return S.Scene([helio, target], sun, name="solver3"), {"heliostat": helio, "target": target}


# src/test_ssol_solver5.c -------------------------------------------------------
def solver5():
    """Parabolic cylinder (focal 10) carved to 20x20 m; a 20x0.2 m virtual
    target on the focal line."""
    f = 10.0
    quad = S.Punched(S.Q_PCYLINDER, (f, 0.0), list(S.IDENTITY), [S.rect_carving(10, 10)])
    m_mtl = S.mirror(1.0)
    v_mtl = S.virtual()
    helio = S.Instance(S.Object([(quad, m_mtl, m_mtl)]), S.IDENTITY)
    rect = S.rect_mesh(np.float32(-10), np.float32(10), np.float32(-0.1), np.float32(0.1))
    target = S.Instance(S.Object([(rect, v_mtl, v_mtl)]), t_pitch(PI, 0, 0, f),
                        receiver_mask=S.FRONT, sample=0)
    sun = S.Sun("directional", (0, 0, -1), 1000.0, spectrum3())
    return S.Scene([helio, target], sun, name="solver5"), {"target": target}


# src/test_ssol_solver8.c -------------------------------------------------------
def solver8():
    """Parabolic cylinder (focal 1) carved to x in [-5, 5], y in [0, 4]
    under an oblique sun (0, 1, -1); a 20x20 m virtual target flipped at
    z = 4."""
    quad = S.Punched(S.Q_PCYLINDER, (1.0, 0.0), list(S.IDENTITY),
                     [S.Carving(S.AND, np.array([[-5.0, 0.0], [-5.0, 4.0], [5.0, 4.0],
                                                 [5.0, 0.0]]))])
    m_mtl = S.mirror(1.0)
    v_mtl = S.virtual()
    helio = S.Instance(S.Object([(quad, m_mtl, m_mtl)]), S.IDENTITY)
    target = S.Instance(S.Object([(S.square_mesh(10.0), v_mtl, v_mtl)]), t_pitch(PI, 0, 0, 4),
                        receiver_mask=S.FRONT, sample=0)
    sun = S.Sun("directional", (0, 1, -1), 1000.0, spectrum3())
    return S.Scene([helio, target], sun, name="solver8"), {"target": target}


# src/test_ssol_solver4.c -------------------------------------------------------
def solver4():
    square = S.square_mesh(0.1)
    quad = S.Punched(S.Q_PARABOL, (10.0, 0.0), list(S.IDENTITY), [S.rect_carving(10, 10)],
                     slices_hint=100)
    m_mtl = S.mirror(1.0)
    v_mtl = S.virtual()
    helio = S.Instance(S.Object([(quad, m_mtl, m_mtl)]), S.IDENTITY)
    t_obj = S.Object([(square, v_mtl, v_mtl)])
    t1 = S.Instance(t_obj, t_pitch(PI, 0, 0, 10), receiver_mask=S.FRONT, sample=0)
    t2 = S.Instance(t_obj, t_pitch(PI, 0, 0, 10 + 1e-4), receiver_mask=S.FRONT, sample=0)
    sun = S.Sun("directional", (0, 0, -1), 1000.0, spectrum3())
    return S.Scene([helio, t1, t2], sun, name="solver4"), {"target1": t1, "target2": t2}


# src/test_ssol_solver6.c -------------------------------------------------------
def solver6():
    square = S.square_mesh(5.0)
    quad = S.Punched(S.Q_PLANE, (0, 0), list(S.IDENTITY), [S.rect_carving(5, 5)])
    m_mtl = S.mirror(1.0)
    bck = S.matte(0.0)
    v_mtl = S.virtual()
    h1 = S.Instance(S.Object([(quad, m_mtl, m_mtl)]), t_pitch(PI, -25, 0, 0))
    h2 = S.Instance(S.Object([(quad, m_mtl, m_mtl)]), t_pitch(PI, 25, 0, 0))
    tgt1 = S.Instance(S.Object([(square, bck, v_mtl)]), t_pitch(PI, -25, 0, 10),
                      receiver_mask=S.FRONT, sample=0)
    tgt2 = S.Instance(S.Object([(square, bck, v_mtl)]), t_move(25, 0, 10),
                      receiver_mask=S.BACK, sample=0)
    sun = S.Sun("directional", (0, 0, -1), 1000.0, spectrum3())
    return S.Scene([h1, h2, tgt1, tgt2], sun, name="solver6"), {"target1": tgt1, "target2": tgt2}


# src/test_ssol_solver7.c -------------------------------------------------------
def solver7():
    H = 10
    target_mesh = S.square_mesh(1.0)
    m_mtl = S.mirror(1.0)
    bck = S.matte(0.1)
    v_mtl = S.virtual()
    par = S.Punched(S.Q_PARABOL, (10 * math.sqrt(2) * H, 0.0), list(S.IDENTITY),
                    [S.rect_carving(2, 2)])
    helio = S.Instance(S.Object([(par, m_mtl, m_mtl)]), t_yaw(-0.25 * PI, 10 * H, 0, 0))
    hyp = S.Punched(S.Q_HYPERBOL, (1.0 * H, 9.0 * H), list(S.IDENTITY), [S.rect_carving(12, 12)])
    secondary = S.Instance(S.Object([(hyp, m_mtl, v_mtl)]), t_move(0, 0, 9 * H), sample=0)
    target = S.Instance(S.Object([(target_mesh, bck, v_mtl)]), S.IDENTITY,
                        receiver_mask=S.FRONT, sample=0)
    sun = S.Sun("directional", (1, 0, -1), 1000.0, spectrum3())
    return S.Scene([helio, secondary, target], sun, name="solver7"), {"target": target}


# src/test_ssol_solver9.c -------------------------------------------------------
def cube_mesh(half):
    h = float(np.float32(half))
    v = [[-h, -h, -h], [-h, -h, h], [-h, h, -h], [-h, h, h],
         [h, -h, -h], [h, -h, h], [h, h, -h], [h, h, h]]
    ids = [0, 6, 4, 0, 2, 6, 0, 3, 2, 0, 1, 3, 2, 7, 6, 2, 3, 7,
           4, 6, 7, 4, 7, 5, 0, 4, 5, 0, 5, 1, 1, 5, 7, 1, 7, 3]
    return S.Mesh(np.array(v, dtype=np.float32), np.array(ids, dtype=np.uint32))


def solver9():
    target_mesh = S.rect_mesh(np.float32(-3), np.float32(3), np.float32(-5), np.float32(5))
    cube = cube_mesh(5)
    m_mtl = S.mirror(1.0)
    v_mtl = S.virtual()
    helio = S.Instance(S.Object([(cube, m_mtl, m_mtl)]), S.IDENTITY)
    target = S.Instance(S.Object([(target_mesh, v_mtl, v_mtl)]), t_pitch(PI, 0, 0, 6),
                        receiver_mask=S.FRONT, sample=0)
    sun = S.Sun("directional", (0, 0, -1), 1000.0, spectrum3())
    return S.Scene([helio, target], sun, name="solver9"), {"target": target}


# src/test_ssol_solver10.c, 11.c, 12.c -------------------------------------------
def solver10(sun_kind="pillbox", part=1):
    square = S.square_mesh(5.0)
    m_mtl = S.matte(0.0)
    param = {"pillbox": 1.0, "buie": 0.1, "gaussian": 0.01}[sun_kind]
    sun = S.Sun(sun_kind, (0, 0, -1), 1000.0, spectrum3(), param=param)
    if part == 1:
        geom = S.Instance(S.Object([(square, m_mtl, m_mtl)]), t_move(-10, 0, 0),
                          receiver_mask=S.FRONT)
    else:
        par = S.Punched(S.Q_PARABOL, (10.0, 0.0), list(S.IDENTITY), [S.rect_carving(5, 5)])
        geom = S.Instance(S.Object([(par, m_mtl, m_mtl)]), t_move(10, 0, 0),
                          receiver_mask=S.FRONT)
    return S.Scene([geom], sun, name=f"solver10_{sun_kind}_{part}"), {"geom": geom}


# src/test_ssol_by_receiver_integration.c -----------------------------------------
def by_receiver(heliostat_receiver=False):
    square = S.square_mesh(1.0)
    quad = S.Punched(S.Q_PLANE, (0, 0), list(S.IDENTITY), [S.rect_carving(10, 10)])
    m_mtl = S.mirror(1.0)
    v_mtl = S.virtual()
    helio = S.Instance(S.Object([(quad, m_mtl, m_mtl)]), S.IDENTITY,
                       receiver_mask=S.FRONT if heliostat_receiver else 0)
    target = S.Instance(S.Object([(square, v_mtl, v_mtl)]), t_pitch(PI, 0, 0, 2),
                        receiver_mask=S.FRONT, sample=0)
    sun = S.Sun("directional", (1, 0, -1), 1000.0, spectrum3())
    return S.Scene([helio, target], sun, name="by_receiver"), {"heliostat": helio, "target": target}


# Uniform runner -----------------------------------------------------------------
class Run:
    """Results of one solve with the getters of the reference estimator."""

    def __init__(self, backend: str, scene, n: int, seed: int = 0, counter: int = 0,
                 max_failed: int = 0):
        self.backend = backend
        self.scene = scene
        if backend == "oracle":
            import oracle_ctypes as O
            self.o = O.Oracle(scene)
            self.res = self.o.solve(n, seed=seed, counter=counter, max_failed=max_failed)
            self.e = self.o
        else:
            import ssol_ctypes as L
            self.lib = L.Scene(scene, precision="fp32" if backend == "gpu32" else "fp64")
            self.e = self.lib.solve(n, seed=seed, counter=counter, max_failed=max_failed)
            self.res = self.e.res

    def global_(self):
        return self.e.global_()

    def receiver(self, inst, side=S.FRONT):
        return self.e.receiver(inst, side)

    def receiver_or_none(self, inst, side=S.FRONT):
        """None where the reference getter returns RES_BAD_ARG (a side that is
        not in the instance's receiver mask)."""
        try:
            return self.e.receiver(inst, side)
        except (ValueError, RuntimeError):
            return None

    def sampled(self, inst):
        return self.e.sampled(inst)

    def sampled_x_receiver(self, samp, rcv, side=S.FRONT):
        return self.e.sampled_x_receiver(samp, rcv, side)

    def primitive(self, inst, sshape_index, prim, side=S.FRONT):
        if self.backend == "oracle":
            return self.e.primitive(inst, sshape_index, prim, side)
        shape = inst.object.sshapes[sshape_index][0]
        return self.e.primitive(inst, shape, prim, side)

    @property
    def realisation_count(self):
        return self.e.realisation_count

    @property
    def failed_count(self):
        return self.e.failed_count

    @property
    def sampled_area(self):
        return self.e.sampled_area
