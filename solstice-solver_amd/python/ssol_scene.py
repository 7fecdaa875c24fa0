"""Scene descriptions for the Solstice flux integrator.

A scene here is plain data: the same information a client passes through the
``ssol_*`` scene API of the reference (src/ssol.h:527-1102): shapes (meshes or
punched quadric surfaces), materials, objects (lists of shaded shapes),
instances, a sun and an optional atmosphere. Two builders consume it:
``ssol_ctypes.build_scene`` drives the library's C API, the oracle's binding
builds the CPU restatement's own description. Keeping the description as data
lets both sides receive exactly the same scene.

Numbering follows the reference public header: side flags FRONT=1, BACK=2;
quadric types PLANE=0 .. HEMISPHERE=5; carving ops AND=0, SUB=1.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

FRONT, BACK = 1, 2
AND, SUB = 0, 1
Q_PLANE, Q_PARABOL, Q_PARABOL2F, Q_HYPERBOL, Q_PCYLINDER, Q_HEMISPHERE = range(6)
SIZE_MAX = 2**64 - 1


# --------------------------------------------------------------------------
# Transforms: 3x4 column-major arrays like the reference's
# --------------------------------------------------------------------------
def rot_x(a: float) -> np.ndarray:
    """Rotation about X (the reference tests' d33_rotation_pitch)."""
    c, s = math.cos(a), math.sin(a)
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]], dtype=np.float64)


def rot_y(a: float) -> np.ndarray:
    """Rotation about Y (the reference tests' d33_rotation_yaw)."""
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]], dtype=np.float64)


def rot_z(a: float) -> np.ndarray:
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]], dtype=np.float64)


def transform(R: Optional[np.ndarray] = None, t: Sequence[float] = (0, 0, 0)) -> List[float]:
    """3x4 column-major array from a 3x3 matrix and a translation."""
    R = np.eye(3) if R is None else np.asarray(R, dtype=np.float64)
    return list(R.T.reshape(-1)) + [float(x) for x in t]


IDENTITY = transform()


def frame_z_to(n: Sequence[float]) -> np.ndarray:
    """Rotation taking +Z onto the unit vector n (shortest arc)."""
    n = np.asarray(n, dtype=np.float64)
    n = n / np.linalg.norm(n)
    z = np.array([0.0, 0.0, 1.0])
    v = np.cross(z, n)
    c = float(np.dot(z, n))
    if np.linalg.norm(v) < 1e-15:
        return np.eye(3) if c > 0 else rot_x(math.pi)
    vx = np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])
    return np.eye(3) + vx + vx @ vx * (1 / (1 + c))


# --------------------------------------------------------------------------
# Scene elements
# --------------------------------------------------------------------------
@dataclass(eq=False)
class Spectrum:
    wavelengths: Sequence[float]
    values: Sequence[float]


@dataclass(eq=False)
class Data:
    real: float = 0.0
    spectrum: Optional[Spectrum] = None


@dataclass(eq=False)
class Medium:
    extinction: Data = field(default_factory=lambda: Data(0.0))
    refractive_index: Data = field(default_factory=lambda: Data(1.0))


VACUUM = Medium()


@dataclass(eq=False)
class Material:
    """kind: 'mirror', 'matte', 'dielectric', 'thin', 'virtual'."""
    kind: str
    reflectivity: float = 1.0
    roughness: float = 0.0
    distribution: int = 0          # 0 Beckmann, 1 pillbox
    normal: str = "Ns"             # the normal shader returns Ns or Ng
    medium_out: Medium = field(default_factory=Medium)
    medium_in: Medium = field(default_factory=Medium)
    slab: Medium = field(default_factory=Medium)
    thickness: float = 0.0

    TYPE = {"dielectric": 0, "matte": 1, "mirror": 2, "thin": 3, "virtual": 4}

    @property
    def type_id(self) -> int:
        return self.TYPE[self.kind]


def mirror(reflectivity=1.0, roughness=0.0, distribution=0) -> Material:
    return Material("mirror", reflectivity=reflectivity, roughness=roughness,
                    distribution=distribution)


def matte(reflectivity=0.0) -> Material:
    return Material("matte", reflectivity=reflectivity)


def virtual() -> Material:
    return Material("virtual")


def dielectric(outside: Medium, inside: Medium) -> Material:
    return Material("dielectric", medium_out=outside, medium_in=inside)


def thin_dielectric(outside: Medium, slab: Medium, thickness: float) -> Material:
    return Material("thin", medium_out=outside, medium_in=outside, slab=slab,
                    thickness=thickness)


@dataclass(eq=False)
class Mesh:
    positions: np.ndarray   # (n, 3) float32, object space
    indices: np.ndarray     # (m, 3) uint32

    def __post_init__(self):
        self.positions = np.ascontiguousarray(self.positions, dtype=np.float32).reshape(-1, 3)
        self.indices = np.ascontiguousarray(self.indices, dtype=np.uint32).reshape(-1, 3)


@dataclass(eq=False)
class Carving:
    op: int                 # AND / SUB
    xy: np.ndarray          # (n, 2) float64

    def __post_init__(self):
        self.xy = np.ascontiguousarray(self.xy, dtype=np.float64).reshape(-1, 2)


@dataclass(eq=False)
class Punched:
    quadric: int
    params: Tuple[float, float] = (0.0, 0.0)  # focal | fx,fy | img,real | radius
    transform: Sequence[float] = field(default_factory=lambda: list(IDENTITY))
    carvings: List[Carving] = field(default_factory=list)
    slices_hint: int = SIZE_MAX


@dataclass(eq=False)
class Object:
    sshapes: List[Tuple[object, Material, Material]]


@dataclass(eq=False)
class Instance:
    object: Object
    transform: Sequence[float] = field(default_factory=lambda: list(IDENTITY))
    receiver_mask: int = 0
    per_primitive: int = 0
    sample: int = 1


@dataclass(eq=False)
class Sun:
    """kind: 'directional', 'pillbox', 'gaussian', 'buie'."""
    kind: str
    direction: Sequence[float]
    dni: float
    spectrum: Spectrum
    param: float = 0.0

    TYPE = {"directional": 0, "pillbox": 1, "gaussian": 2, "buie": 3}

    @property
    def type_id(self) -> int:
        return self.TYPE[self.kind]


@dataclass(eq=False)
class Scene:
    instances: List[Instance]
    sun: Sun
    atmosphere: Optional[Data] = None
    name: str = ""


# --------------------------------------------------------------------------
# Geometry helpers
# --------------------------------------------------------------------------
def rect_mesh(x_min, x_max, y_min, y_max) -> Mesh:
    """Rectangle in z=0 facing +Z, vertices and triangles ordered like the
    reference's test fixture (src/test_ssol_rect_geometry.h:54-66)."""
    pos = [[x_min, y_min, 0], [x_max, y_min, 0], [x_max, y_max, 0], [x_min, y_max, 0]]
    return Mesh(np.array(pos, dtype=np.float32), np.array([[0, 2, 1], [2, 0, 3]]))


def square_mesh(half: float) -> Mesh:
    return rect_mesh(np.float32(-half), np.float32(half), np.float32(-half), np.float32(half))


def grid_mesh(size_x: float, size_y: float, nx: int, ny: int) -> Mesh:
    """nx x ny cells in z=0 centred on the origin, front face +Z, 2 triangles
    per cell (same winding as rect_mesh)."""
    xs = np.linspace(-size_x / 2, size_x / 2, nx + 1)
    ys = np.linspace(-size_y / 2, size_y / 2, ny + 1)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], axis=1)
    ids = []
    for i in range(nx):
        for j in range(ny):
            v0 = i * (ny + 1) + j          # (x0, y0)
            v1 = (i + 1) * (ny + 1) + j    # (x1, y0)
            v2 = (i + 1) * (ny + 1) + j + 1  # (x1, y1)
            v3 = i * (ny + 1) + j + 1      # (x0, y1)
            ids.append([v0, v2, v1])
            ids.append([v2, v0, v3])
    return Mesh(pos.astype(np.float32), np.array(ids, dtype=np.uint32))


def rect_carving(half_x: float, half_y: float, op: int = AND) -> Carving:
    """Rectangle polygon ordered like src/test_ssol_rect2D_geometry.h:52-57."""
    hx, hy = float(np.float32(half_x)), float(np.float32(half_y))
    return Carving(op, np.array([[-hx, -hy], [-hx, hy], [hx, hy], [hx, -hy]]))


def mono_spectrum(wl: float = 500.0) -> Spectrum:
    return Spectrum([wl], [1.0])


# --------------------------------------------------------------------------
# Workloads
# --------------------------------------------------------------------------
def config1() -> Scene:
    """One flat 10x10 m specular mirror (2 triangles, reflectivity 0.9) tilted
    45 deg toward a 5x5 m absorbing receiver plane 50 m away; pillbox sun,
    half-angle 4.65 mrad, DNI 1000, elevation 60 deg; no atmosphere."""
    el = math.radians(60)
    d = np.array([-math.cos(el), 0.0, -math.sin(el)])
    n = np.array([math.sin(math.pi / 4), 0.0, math.cos(math.pi / 4)])
    r = d - 2 * np.dot(d, n) * n
    mirror_obj = Object([(square_mesh(5.0), mirror(0.9), mirror(0.9))])
    rcv_obj = Object([(square_mesh(2.5), matte(0.0), matte(0.0))])
    helio = Instance(mirror_obj, transform(rot_y(math.pi / 4)), sample=1)
    theta = math.atan2(-r[0], -r[2])
    rcv = Instance(rcv_obj, transform(rot_y(theta), 50 * r), receiver_mask=FRONT,
                   per_primitive=0, sample=0)
    sun = Sun("pillbox", d, 1000.0, mono_spectrum(), param=4.65e-3)
    return Scene([helio, rcv], sun, None, name="config1")


CONFIG1_CLOSED_FORM = {
    "cos_factor": math.cos(math.radians(15)),
    "absorbed_by_receivers": 22500.0,
    "other_absorbed": (1 - 0.9) * 1000 * 100 * math.cos(math.radians(15)),
    "missing": 0.9 * 1000 * 100 * math.cos(math.radians(15)) - 22500.0,
    "shadowed": 0.0,
}


def _radial_staggered(n: int, r_min: float, r_max: float, rng: np.random.Generator):
    """Seeded radial-staggered positions: rings of growing radius, alternate
    rings shifted by half a spacing, jittered slightly."""
    pts = []
    ring = 0
    area = math.pi * (r_max ** 2 - r_min ** 2)
    spacing = math.sqrt(area / n) * 0.95
    r = r_min
    while len(pts) < n and r <= r_max * 1.5:
        m = max(6, int(2 * math.pi * r / spacing))
        off = 0.5 * (ring % 2)
        for k in range(m):
            a = 2 * math.pi * (k + off) / m
            rr = r + rng.uniform(-0.1, 0.1) * spacing
            pts.append((rr * math.cos(a), rr * math.sin(a)))
        ring += 1
        r += spacing
    pts = np.array(pts[:n])
    return pts


def _aim(pos, target, sun_dir):
    """Rotation taking +Z onto the bisector of -sun and the target direction."""
    to_t = np.asarray(target, dtype=np.float64) - np.asarray(pos, dtype=np.float64)
    to_t /= np.linalg.norm(to_t)
    s = -np.asarray(sun_dir, dtype=np.float64)
    s /= np.linalg.norm(s)
    n = s + to_t
    return frame_z_to(n / np.linalg.norm(n)), np.linalg.norm(
        np.asarray(target) - np.asarray(pos))


def sun_dir_from(azimuth_deg: float, elevation_deg: float) -> np.ndarray:
    """Direction the sunlight travels (downward)."""
    az, el = math.radians(azimuth_deg), math.radians(elevation_deg)
    return -np.array([math.cos(el) * math.cos(az), math.cos(el) * math.sin(az), math.sin(el)])


def config2(seed: int = 0, n_helio: int = 100, radius: float = 300.0,
            sun_dir=None, slope_error: float = 2e-3, sigma_sun: float = 2.51e-3,
            aim_perturb: Optional[np.ndarray] = None,
            slope_errors: Optional[Sequence[float]] = None,
            rcv_size: float = 10.0, rcv_cells: int = 32) -> Scene:
    """Radial-staggered field of paraboloid heliostats (seeded positions
    within `radius`), focal length = slant range, carved to 6x6 m, Beckmann
    glossy mirror of the given slope error (roughness = sqrt(2) * slope error,
    the RMS slope of an isotropic Gaussian slope distribution), reflectivity
    0.92; 10x10 m receiver of 32x32 cells (2048 triangles) facing down on a
    tower at 100 m; Gaussian sun sigma 2.51 mrad."""
    rng = np.random.default_rng(seed)
    pts = _radial_staggered(n_helio, 30.0, radius, rng)
    if sun_dir is None:
        sun_dir = sun_dir_from(180.0, 60.0)
    tower = np.array([0.0, 0.0, 100.0])
    rcv_mesh = grid_mesh(rcv_size, rcv_size, rcv_cells, rcv_cells)
    rcv_obj = Object([(rcv_mesh, matte(0.0), matte(0.0))])
    rcv = Instance(rcv_obj, transform(rot_x(math.pi), tower), receiver_mask=FRONT,
                   per_primitive=1, sample=0)
    insts = []
    for i, (x, y) in enumerate(pts):
        pos = np.array([x, y, 0.0])
        R, slant = _aim(pos, tower, sun_dir)
        if aim_perturb is not None:
            R = R @ rot_x(float(aim_perturb[i, 0])) @ rot_y(float(aim_perturb[i, 1]))
        se = slope_error if slope_errors is None else float(slope_errors[i])
        quad = Punched(Q_PARABOL, (slant, 0.0), list(IDENTITY), [rect_carving(3.0, 3.0)])
        mtl = mirror(0.92, roughness=math.sqrt(2) * se, distribution=0)
        obj = Object([(quad, mtl, mtl)])
        insts.append(Instance(obj, transform(R, pos), sample=1))
    sun = Sun("gaussian", sun_dir, 1000.0, mono_spectrum(), param=sigma_sun)
    return Scene(insts + [rcv], sun, None, name=f"config2_n{n_helio}")


def aim_axes(seed: int = 0, n_helio: int = 100, radius: float = 300.0, sun_dir=None,
             aim_perturb: Optional[np.ndarray] = None) -> np.ndarray:
    """World rotation axes [n][2][3] of the two aim angles of config2's
    heliostats, whose rotation is R rot_x(a0) rot_y(a1): d/da0 rotates about
    R e_x and d/da1 about R rot_x(a0) e_y."""
    rng = np.random.default_rng(seed)
    pts = _radial_staggered(n_helio, 30.0, radius, rng)
    if sun_dir is None:
        sun_dir = sun_dir_from(180.0, 60.0)
    tower = np.array([0.0, 0.0, 100.0])
    out = np.zeros((n_helio, 2, 3))
    for i, (x, y) in enumerate(pts):
        R, _ = _aim(np.array([x, y, 0.0]), tower, sun_dir)
        a0 = 0.0 if aim_perturb is None else float(aim_perturb[i, 0])
        out[i, 0] = R @ np.array([1.0, 0.0, 0.0])
        out[i, 1] = R @ rot_x(a0) @ np.array([0.0, 1.0, 0.0])
    return out


def config5(seed: int = 0, n_helio: int = 1000, radius: float = 900.0,
            aim_range: float = 5e-3, slope_range=(1e-3, 3e-3), **kw):
    """Differentiable-aiming field: config2's paraboloid heliostats with
    seeded uniform aim perturbations in [-aim_range, aim_range] (two angles
    each) and slope errors uniform in slope_range. Returns (scene, aim
    [n][2], slope errors [n], aim axes [n][2][3])."""
    rng = np.random.default_rng(seed + 5)
    aim = rng.uniform(-aim_range, aim_range, (n_helio, 2))
    slope = rng.uniform(slope_range[0], slope_range[1], n_helio)
    spec = config2(seed, n_helio=n_helio, radius=radius, aim_perturb=aim, slope_errors=slope, **kw)
    axes = aim_axes(seed, n_helio, radius, kw.get("sun_dir"), aim)
    return spec, aim, slope, axes


def config4(seed: int = 0, n_suns: int = 512, n_az: int = 32, n_el: int = 16,
            log_min: float = 5.0, log_max: float = 7.0):
    """Batched sweep: the 100-paraboloid field (aimed for its default sun)
    under n_az x n_el sun positions (azimuth 0-360 deg x elevation 5-85 deg)
    with seeded log-uniform realisation counts in [10^log_min, 10^log_max].
    Returns (scene, directions [n_suns][3], counts [n_suns])."""
    rng = np.random.default_rng(seed)
    az = np.linspace(0.0, 360.0, n_az, endpoint=False)
    el = np.linspace(5.0, 85.0, n_el)
    dirs = [sun_dir_from(a, e) for e in el for a in az][:n_suns]
    counts = np.floor(10.0 ** rng.uniform(log_min, log_max, len(dirs))).astype(np.int64)
    return config2(seed), np.array(dirs), counts


def config3(seed: int = 0, n_helio: int = 10000, r_min: float = 100.0,
            r_max: float = 1500.0, sun_dir=None) -> Scene:
    """Beam-down field of 10,000 flat heliostats, each an 8x8-facet 4x4 m
    mesh (128 triangles, one shared object), radial-staggered out to km
    scale; a hyperboloid secondary whose upper focus is the field's aim point
    and whose lower focus is the receiver on the ground; a 3 mm thin glass
    cover (n = 1.5) over a 64x64-cell absorbing receiver with per-primitive
    tallies; atmosphere extinction 1e-4 /m (0.1 /km)."""
    rng = np.random.default_rng(seed)
    pts = _radial_staggered(n_helio, r_min, r_max, rng)
    if sun_dir is None:
        sun_dir = sun_dir_from(180.0, 60.0)
    # Hyperboloid secondary: lower focus on the receiver (z = 0), upper focus
    # at the field's aim point (z = 600 m), vertex at z = 500 m, i.e. real
    # focal 500 m (vertex to lower focus) and image focal 100 m.
    f_upper = np.array([0.0, 0.0, 600.0])
    f_lower = np.array([0.0, 0.0, 0.0])
    H_vertex = 500.0
    helio_mesh = grid_mesh(4.0, 4.0, 8, 8)
    air_glass = mirror(0.93)
    helio_obj = Object([(helio_mesh, air_glass, air_glass)])
    insts = []
    for (x, y) in pts:
        pos = np.array([x, y, 0.0])
        R, _ = _aim(pos, f_upper, sun_dir)
        insts.append(Instance(helio_obj, transform(R, pos), sample=1))
    hyp = Punched(Q_HYPERBOL, (100.0, 500.0), list(IDENTITY),
                  [rect_carving(160.0, 160.0)])
    sec_mtl = mirror(0.95)
    sec_obj = Object([(hyp, sec_mtl, matte(0.0))])
    sec = Instance(sec_obj, transform(None, (0.0, 0.0, H_vertex)), sample=0)
    glass = thin_dielectric(VACUUM, Medium(Data(0.0), Data(1.5)), 3e-3)
    cover_obj = Object([(square_mesh(6.0), glass, glass)])
    cover = Instance(cover_obj, transform(None, (0.0, 0.0, 0.5)), sample=0)
    rcv_obj = Object([(grid_mesh(10.0, 10.0, 64, 64), matte(0.0), matte(0.0))])
    rcv = Instance(rcv_obj, transform(None, f_lower), receiver_mask=FRONT,
                   per_primitive=1, sample=0)
    sun = Sun("gaussian", sun_dir, 1000.0, mono_spectrum(), param=2.51e-3)
    return Scene(insts + [sec, cover, rcv], sun, Data(1e-4), name=f"config3_n{n_helio}")
