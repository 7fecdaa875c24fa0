"""ctypes mirror of the ssol C API (include/ssol.h, include/ssp.h).

``build_scene`` drives the library's scene-building entry points from a
``ssol_scene.Scene`` description exactly as a C client would (meshes through
``ssol_mesh_setup`` callbacks, punched surfaces through
``ssol_punched_surface_setup``, constant material shaders, instances,
sun, atmosphere). ``Scene.solve`` calls ``ssol_solve`` and wraps the returned
estimator; its getters mirror ``ssol_estimator_get_mc_global``,
``ssol_estimator_get_mc_receiver``, ``ssol_estimator_get_mc_sampled``,
``ssol_estimator_get_mc_sampled_x_receiver`` and
``ssol_mc_shape_get_mc_primitive``. The library runs ssol_solve on the GPU
only; without a usable gfx950 device ssol_solve fails.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, List

import numpy as np

import ssol_scene as S

_HERE = os.path.dirname(os.path.abspath(__file__))
# the built libraries; SSOL_LIB_DIR points at another build of them
LIB_DIR = os.environ.get("SSOL_LIB_DIR") or os.path.join(os.path.dirname(_HERE), "lib")
_LIB = None

RES = {0: "RES_OK", 1: "RES_BAD_ARG", 2: "RES_MEM_ERR", 3: "RES_IO_ERR",
       4: "RES_UNKNOWN_ERR", 5: "RES_BAD_OP", 6: "RES_EOF"}


class SsolError(RuntimeError):
    def __init__(self, fn, code):
        super().__init__(f"{fn} returned {RES.get(code, code)}")
        self.code = code


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(LIB_DIR, "libssol.so")
        if not os.path.exists(path):
            raise RuntimeError(f"libssol.so not built ({path}); run solstice-solver_amd/build.py")
        _LIB = C.CDLL(path)
        _declare(_LIB)
    return _LIB


def lib_hip():
    """The HIP kernel library (loaded as a dependency of libssol.so)."""
    lib()
    return C.CDLL(os.path.join(LIB_DIR, "libssol_hip.so"))


# Structures -------------------------------------------------------------------
class McResult(C.Structure):
    _fields_ = [("E", C.c_double), ("V", C.c_double), ("SE", C.c_double)]

    def tuple(self):
        return (self.E, self.V, self.SE)


GLOBAL_FIELDS = ("cos_factor", "absorbed_by_receivers", "shadowed", "missing",
                 "extinguished_by_atmosphere", "other_absorbed")
RECEIVER_FIELDS = ("incoming_flux", "incoming_if_no_atm_loss", "incoming_if_no_field_loss",
                   "incoming_lost_in_field", "incoming_lost_in_atmosphere", "absorbed_flux",
                   "absorbed_if_no_atm_loss", "absorbed_if_no_field_loss",
                   "absorbed_lost_in_field", "absorbed_lost_in_atmosphere")


class McGlobal(C.Structure):
    _fields_ = [(n, McResult) for n in GLOBAL_FIELDS]


class McReceiver(C.Structure):
    _fields_ = [(n, McResult) for n in RECEIVER_FIELDS] + [
        ("N__", C.c_size_t), ("mc__", C.c_void_p), ("instance__", C.c_void_p)]


class McShape(C.Structure):
    _fields_ = [("N__", C.c_size_t), ("mc__", C.c_void_p), ("shape__", C.c_void_p)]


class McSampled(C.Structure):
    _fields_ = [("cos_factor", McResult), ("shadowed", McResult), ("nb_samples", C.c_size_t)]


class McPrimitive(C.Structure):
    _fields_ = [(n, McResult) for n in RECEIVER_FIELDS]


class Data(C.Structure):
    class _V(C.Union):
        _fields_ = [("real", C.c_double), ("spectrum", C.c_void_p)]
    _fields_ = [("type", C.c_int), ("value", _V)]


class Medium(C.Structure):
    _fields_ = [("extinction", Data), ("refractive_index", Data)]


class Fragment(C.Structure):
    _fields_ = [("dir", C.c_double * 3), ("P", C.c_double * 3), ("Ng", C.c_double * 3),
                ("Ns", C.c_double * 3), ("uv", C.c_double * 2), ("dPdu", C.c_double * 3),
                ("dPdv", C.c_double * 3)]


SHADER = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.c_double, C.POINTER(Fragment),
                     C.POINTER(C.c_double))


class MirrorShader(C.Structure):
    _fields_ = [("normal", SHADER), ("reflectivity", SHADER), ("roughness", SHADER)]


class MatteShader(C.Structure):
    _fields_ = [("normal", SHADER), ("reflectivity", SHADER)]


class NormalShader(C.Structure):
    _fields_ = [("normal", SHADER)]


VERTEX_GET = C.CFUNCTYPE(None, C.c_uint, C.POINTER(C.c_float), C.c_void_p)
INDEX_GET = C.CFUNCTYPE(None, C.c_uint, C.POINTER(C.c_uint), C.c_void_p)
CARVING_GET = C.CFUNCTYPE(None, C.c_size_t, C.POINTER(C.c_double), C.c_void_p)
SPECTRUM_GET = C.CFUNCTYPE(None, C.c_size_t, C.POINTER(C.c_double), C.POINTER(C.c_double),
                           C.c_void_p)


class VertexData(C.Structure):
    _fields_ = [("usage", C.c_int), ("get", VERTEX_GET)]


class Quadric(C.Structure):
    class _D(C.Union):
        _fields_ = [("v", C.c_double * 2)]
    _fields_ = [("type", C.c_int), ("data", _D), ("transform", C.c_double * 12),
                ("slices_count_hint", C.c_size_t)]


class Carving(C.Structure):
    _fields_ = [("get", CARVING_GET), ("nb_vertices", C.c_size_t), ("operation", C.c_int),
                ("context", C.c_void_p)]


class PunchedSurface(C.Structure):
    _fields_ = [("quadric", C.POINTER(Quadric)), ("carvings", C.POINTER(Carving)),
                ("nb_carvings", C.c_size_t)]


class AutotuneResult(C.Structure):
    _fields_ = [("path_slots", C.c_uint), ("graph_iterations", C.c_uint),
                ("paths_per_second", C.c_double), ("candidates", C.c_uint)]


class SolveStats(C.Structure):
    _fields_ = [("segments", C.c_uint64), ("paths", C.c_uint64), ("pending_spills", C.c_uint64),
                ("kernel_ms", C.c_double), ("trace_ms", C.c_double), ("shade_ms", C.c_double),
                ("iterations", C.c_uint), ("wavefronts", C.c_uint), ("refilled", C.c_uint)]


class PathTracker(C.Structure):
    """struct ssol_path_tracker: a negative length takes the default"""
    _fields_ = [("sun_ray_length", C.c_double), ("infinite_ray_length", C.c_double)]


class Path(C.Structure):
    _fields_ = [("path__", C.c_void_p)]


class PathVertex(C.Structure):
    _fields_ = [("pos", C.c_double * 3), ("weight", C.c_double)]


PATH_TYPES = ("missing", "shadow", "success", "error")  # enum ssol_path_type

VP = C.c_void_p
PVP = C.POINTER(C.c_void_p)


def _declare(L):
    sigs = {
        "ssol_device_create": [VP, VP, C.c_uint, C.c_int, PVP],
        "ssol_camera_create": [VP, PVP],
        "ssol_camera_ref_put": [VP],
        "ssol_camera_set_proj_ratio": [VP, C.c_double],
        "ssol_camera_set_fov": [VP, C.c_double],
        "ssol_camera_look_at": [VP, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                C.POINTER(C.c_double)],
        "ssol_image_create": [VP, PVP],
        "ssol_image_ref_put": [VP],
        "ssol_image_setup": [VP, C.c_size_t, C.c_size_t, C.c_int],
        "ssol_image_get_layout": [VP, VP],
        "ssol_image_map": [VP, C.POINTER(VP)],
        "ssol_image_unmap": [VP],
        "ssol_image_sample": [VP, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double),
                              C.POINTER(C.c_double)],
        "ssol_image_write": [VP, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t), C.c_int, VP],
        "ssol_draw_draft": [VP, VP, C.c_size_t, C.c_size_t, C.c_size_t, VP, VP],
        "ssol_draw_pt": [VP, VP, C.c_size_t, C.c_size_t, C.c_size_t, C.POINTER(C.c_double),
                         VP, VP],
        "ssol_device_ref_put": [VP],
        "ssol_device_set_precision": [VP, C.c_int],
        "ssol_device_set_wavefront_size": [VP, C.c_size_t],
        "ssol_device_get_gpu_context": [VP, C.POINTER(VP)],
        "ssol_device_set_graph_iterations": [VP, C.c_uint],
        "ssol_device_autotune": [VP, VP, C.c_size_t, C.POINTER(AutotuneResult)],
        "ssol_device_autotune_batch": [VP, VP, C.c_size_t, C.POINTER(VP), C.POINTER(C.c_size_t),
                                       C.POINTER(AutotuneResult)],
        "ssol_solve_adjoint": [VP, VP, C.c_size_t, C.c_size_t, C.POINTER(VP),
                               C.POINTER(C.c_double), C.c_double, VP, VP,
                               C.POINTER(C.c_double), C.POINTER(C.c_double)],
        "ssol_solve_sun_batch": [VP, VP, C.c_size_t, C.POINTER(VP), C.POINTER(C.c_size_t),
                                 C.POINTER(McGlobal), C.POINTER(C.c_size_t)],
        "ssol_scene_create": [VP, PVP],
        "ssol_scene_ref_put": [VP],
        "ssol_scene_attach_instance": [VP, VP],
        "ssol_scene_detach_instance": [VP, VP],
        "ssol_scene_attach_sun": [VP, VP],
        "ssol_scene_detach_sun": [VP, VP],
        "ssol_scene_attach_atmosphere": [VP, VP],
        "ssol_scene_detach_atmosphere": [VP, VP],
        "ssol_shape_create_mesh": [VP, PVP],
        "ssol_shape_create_punched_surface": [VP, PVP],
        "ssol_shape_ref_put": [VP],
        "ssol_shape_get_triangles_count": [VP, C.POINTER(C.c_uint)],
        "ssol_shape_get_vertices_count": [VP, C.POINTER(C.c_uint)],
        "ssol_shape_get_vertex_attrib": [VP, C.c_uint, C.c_int, C.c_double * 3],
        "ssol_shape_get_triangle_indices": [VP, C.c_uint, C.c_uint * 3],
        "ssol_mesh_setup": [VP, C.c_uint, INDEX_GET, C.c_uint, C.POINTER(VertexData), C.c_uint, VP],
        "ssol_punched_surface_setup": [VP, C.POINTER(PunchedSurface)],
        "ssol_material_create_mirror": [VP, PVP],
        "ssol_material_create_matte": [VP, PVP],
        "ssol_material_create_virtual": [VP, PVP],
        "ssol_material_create_dielectric": [VP, PVP],
        "ssol_material_create_thin_dielectric": [VP, PVP],
        "ssol_material_ref_put": [VP],
        "ssol_mirror_setup": [VP, C.POINTER(MirrorShader), C.c_int],
        "ssol_matte_setup": [VP, C.POINTER(MatteShader)],
        "ssol_dielectric_setup": [VP, C.POINTER(NormalShader), C.POINTER(Medium), C.POINTER(Medium)],
        "ssol_thin_dielectric_setup": [VP, C.POINTER(NormalShader), C.POINTER(Medium),
                                       C.POINTER(Medium), C.c_double],
        "ssol_object_create": [VP, PVP],
        "ssol_object_ref_put": [VP],
        "ssol_object_add_shaded_shape": [VP, VP, VP, VP],
        "ssol_object_instantiate": [VP, PVP],
        "ssol_instance_ref_put": [VP],
        "ssol_instance_set_transform": [VP, C.c_double * 12],
        "ssol_instance_set_receiver": [VP, C.c_int, C.c_int],
        "ssol_instance_sample": [VP, C.c_int],
        "ssol_spectrum_create": [VP, PVP],
        "ssol_spectrum_ref_put": [VP],
        "ssol_spectrum_setup": [VP, SPECTRUM_GET, C.c_size_t, VP],
        "ssol_sun_create_directional": [VP, PVP],
        "ssol_sun_create_pillbox": [VP, PVP],
        "ssol_sun_create_gaussian": [VP, PVP],
        "ssol_sun_create_buie": [VP, PVP],
        "ssol_sun_ref_put": [VP],
        "ssol_sun_set_direction": [VP, C.c_double * 3],
        "ssol_sun_set_dni": [VP, C.c_double],
        "ssol_sun_set_spectrum": [VP, VP],
        "ssol_sun_pillbox_set_half_angle": [VP, C.c_double],
        "ssol_sun_gaussian_set_std_dev": [VP, C.c_double],
        "ssol_sun_set_buie_param": [VP, C.c_double],
        "ssol_atmosphere_create": [VP, PVP],
        "ssol_atmosphere_ref_put": [VP],
        "ssol_atmosphere_set_extinction": [VP, C.POINTER(Data)],
        "ssol_solve": [VP, VP, C.c_size_t, C.c_size_t, VP, PVP],
        "ssol_estimator_ref_put": [VP],
        "ssol_estimator_get_mc_global": [VP, C.POINTER(McGlobal)],
        "ssol_estimator_get_mc_receiver": [VP, VP, C.c_int, C.POINTER(McReceiver)],
        "ssol_estimator_get_mc_sampled": [VP, VP, C.POINTER(McSampled)],
        "ssol_estimator_get_mc_sampled_x_receiver": [VP, VP, VP, C.c_int, C.POINTER(McReceiver)],
        "ssol_estimator_get_realisation_count": [VP, C.POINTER(C.c_size_t)],
        "ssol_estimator_get_failed_count": [VP, C.POINTER(C.c_size_t)],
        "ssol_estimator_get_sampled_area": [VP, C.POINTER(C.c_double)],
        "ssol_estimator_get_sampled_count": [VP, C.POINTER(C.c_size_t)],
        "ssol_estimator_get_rng_state": [VP, PVP],
        "ssol_estimator_get_solve_stats": [VP, C.POINTER(SolveStats)],
        "ssol_estimator_get_tracked_paths_count": [VP, C.POINTER(C.c_size_t)],
        "ssol_estimator_get_tracked_path": [VP, C.c_size_t, C.POINTER(Path)],
        "ssol_path_get_vertices_count": [C.POINTER(Path), C.POINTER(C.c_size_t)],
        "ssol_path_get_vertex": [C.POINTER(Path), C.c_size_t, C.POINTER(PathVertex)],
        "ssol_path_get_type": [C.POINTER(Path), C.POINTER(C.c_int)],
        "ssol_mc_receiver_get_mc_shape": [C.POINTER(McReceiver), VP, C.POINTER(McShape)],
        "ssol_mc_shape_get_mc_primitive": [C.POINTER(McShape), C.c_uint, C.POINTER(McPrimitive)],
        "ssp_rng_create": [VP, C.c_int, PVP],
        "ssp_rng_ref_put": [VP],
        "ssp_rng_get_counter_state": [VP, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)],
        "ssp_rng_set_counter_state": [VP, C.c_uint64, C.c_uint64],
    }
    for name, args in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.ssol_data_set_real.restype = C.c_void_p
    L.ssol_data_set_real.argtypes = [C.POINTER(Data), C.c_double]
    L.ssol_data_set_spectrum.restype = C.c_void_p
    L.ssol_data_set_spectrum.argtypes = [C.POINTER(Data), VP]
    L.ssol_data_clear.restype = C.c_void_p
    L.ssol_data_clear.argtypes = [C.POINTER(Data)]
    L.ssp_threefry4x64_20.restype = None
    L.ssp_threefry4x64_20.argtypes = [C.c_uint64 * 4, C.c_uint64 * 4, C.c_uint64 * 4]


def call(name, *args):
    r = getattr(lib(), name)(*args)
    if r != 0:
        raise SsolError(name, r)
    return r


def new(name, *args):
    h = C.c_void_p()
    call(name, *args, C.byref(h))
    return h


SSP_RNG_THREEFRY = 4


class Rng:
    """ssp_rng of type SSP_RNG_THREEFRY with an explicit (key, counter)."""

    def __init__(self, key: int = 0, counter: int = 0):
        self.h = new("ssp_rng_create", None, SSP_RNG_THREEFRY)
        call("ssp_rng_set_counter_state", self.h, key, counter)

    def __del__(self):
        if getattr(self, "h", None):
            lib().ssp_rng_ref_put(self.h)
            self.h = None


def rng_state(handle):
    k, c = C.c_uint64(), C.c_uint64()
    call("ssp_rng_get_counter_state", handle, C.byref(k), C.byref(c))
    return k.value, c.value


# Scene construction --------------------------------------------------------------
def _const_shader(value):
    def fn(dev, buf, wl, frag, out):
        out[0] = value
    return SHADER(fn)


def _normal_shader(kind):
    def fn(dev, buf, wl, frag, out):
        src = frag.contents.Ng if kind == "Ng" else frag.contents.Ns
        out[0], out[1], out[2] = src[0], src[1], src[2]
    return SHADER(fn)


class Scene:
    """A scene built through the C API from a ssol_scene.Scene description."""

    def __init__(self, spec: S.Scene, precision: str = "fp64", path_slots: int = 0,
                 graph_iterations: int = 0):
        self.spec = spec
        self._keep: List[object] = []
        self._handles: List[tuple] = []  # (release fn name, handle)
        self.dev = new("ssol_device_create", None, None, 0xFFFFFFFF, 0)
        if precision == "fp32":
            call("ssol_device_set_precision", self.dev, 1)
        if path_slots:
            call("ssol_device_set_wavefront_size", self.dev, path_slots)
        if graph_iterations:
            call("ssol_device_set_graph_iterations", self.dev, graph_iterations)
        self.scn = new("ssol_scene_create", self.dev)
        self._spectra: Dict[int, C.c_void_p] = {}
        self._materials: Dict[int, C.c_void_p] = {}
        self.shapes: Dict[int, C.c_void_p] = {}
        self._objects: Dict[int, C.c_void_p] = {}
        self.instances: Dict[int, C.c_void_p] = {}
        for inst in spec.instances:
            h = new("ssol_object_instantiate", self._object(inst.object))
            self._handles.append(("ssol_instance_ref_put", h))
            call("ssol_instance_set_transform", h, (C.c_double * 12)(*inst.transform))
            call("ssol_instance_set_receiver", h, inst.receiver_mask, inst.per_primitive)
            call("ssol_instance_sample", h, inst.sample)
            call("ssol_scene_attach_instance", self.scn, h)
            self.instances[id(inst)] = h
        self.sun = self._make_sun(spec.sun)
        call("ssol_scene_attach_sun", self.scn, self.sun)
        if spec.atmosphere is not None:
            self.atm = new("ssol_atmosphere_create", self.dev)
            self._handles.append(("ssol_atmosphere_ref_put", self.atm))
            d = self._data(spec.atmosphere)
            call("ssol_atmosphere_set_extinction", self.atm, C.byref(d))
            lib().ssol_data_clear(C.byref(d))
            call("ssol_scene_attach_atmosphere", self.scn, self.atm)

    def _make_sun(self, sun: S.Sun, direction=None):
        ctor = {"directional": "ssol_sun_create_directional", "pillbox": "ssol_sun_create_pillbox",
                "gaussian": "ssol_sun_create_gaussian", "buie": "ssol_sun_create_buie"}[sun.kind]
        h = new(ctor, self.dev)
        self._handles.append(("ssol_sun_ref_put", h))
        d = sun.direction if direction is None else direction
        call("ssol_sun_set_direction", h, (C.c_double * 3)(*d))
        call("ssol_sun_set_dni", h, sun.dni)
        call("ssol_sun_set_spectrum", h, self._spectrum(sun.spectrum))
        if sun.kind == "pillbox":
            call("ssol_sun_pillbox_set_half_angle", h, sun.param)
        elif sun.kind == "gaussian":
            call("ssol_sun_gaussian_set_std_dev", h, sun.param)
        elif sun.kind == "buie":
            call("ssol_sun_set_buie_param", h, sun.param)
        return h

    def solve_sun_batch(self, directions, counts, seed: int = 0, counter: int = 0):
        """ssol_solve_sun_batch over suns like the scene's, pointing along
        `directions`; returns a list of global-estimate dicts and counts."""
        suns = [self._make_sun(self.spec.sun, d) for d in directions]
        n = len(suns)
        arr = (VP * n)(*[h.value if isinstance(h, C.c_void_p) else h for h in suns])
        cnt = (C.c_size_t * n)(*[int(x) for x in counts])
        out = (McGlobal * n)()
        got = (C.c_size_t * n)()
        r = Rng(seed, counter)
        call("ssol_solve_sun_batch", self.scn, r.h, n, arr, cnt, out, got)
        res = [{f: getattr(out[j], f).tuple() for f in GLOBAL_FIELDS} for j in range(n)]
        for h in suns:
            lib().ssol_sun_ref_put(h)
            self._handles.remove(("ssol_sun_ref_put", h))
        return res, list(got)

    def autotune_batch(self, directions, counts) -> AutotuneResult:
        """ssol_device_autotune_batch on a probe batch of suns like the
        scene's, pointing along `directions`; the pick stays on the device."""
        suns = [self._make_sun(self.spec.sun, d) for d in directions]
        n = len(suns)
        arr = (VP * n)(*[h.value if isinstance(h, C.c_void_p) else h for h in suns])
        cnt = (C.c_size_t * n)(*[int(x) for x in counts])
        out = AutotuneResult()
        call("ssol_device_autotune_batch", self.dev, self.scn, n, arr, cnt, C.byref(out))
        for h in suns:
            lib().ssol_sun_ref_put(h)
            self._handles.remove(("ssol_sun_ref_put", h))
        return out

    def draw(self, width, height, spp, position, target, up, fov=None, proj_ratio=None,
             sky_up=None):
        """ssol_draw_draft (sky_up None) or ssol_draw_pt into an ssol_image
        through ssol_image_write; returns the image as [height][width][3]"""
        cam = new("ssol_camera_create", self.dev)
        img = new("ssol_image_create", self.dev)
        try:
            d3 = lambda v: (C.c_double * 3)(*[float(x) for x in v])  # noqa: E731
            if fov is not None:
                call("ssol_camera_set_fov", cam, fov)
            if proj_ratio is not None:
                call("ssol_camera_set_proj_ratio", cam, proj_ratio)
            call("ssol_camera_look_at", cam, d3(position), d3(target), d3(up))
            call("ssol_image_setup", img, width, height, 0)
            writer = C.cast(lib().ssol_image_write, VP)
            if sky_up is None:
                call("ssol_draw_draft", self.scn, cam, width, height, spp, writer, img)
            else:
                call("ssol_draw_pt", self.scn, cam, width, height, spp, d3(sky_up), writer, img)
            mem = C.c_void_p()
            call("ssol_image_map", img, C.byref(mem))
            buf = C.string_at(mem.value, width * height * 24)
            call("ssol_image_unmap", img)
            return np.frombuffer(buf, dtype=np.float64).reshape(height, width, 3).copy()
        finally:
            lib().ssol_image_ref_put(img)
            lib().ssol_camera_ref_put(cam)

    def set_transform(self, inst, transform):
        """ssol_instance_set_transform on the instance built from spec `inst`
        (the next solve moves it without rebuilding the objects)"""
        call("ssol_instance_set_transform", self.instances[id(inst)],
             (C.c_double * 12)(*transform))

    def solve_adjoint(self, n: int, heliostats, axes, d_absorbed: float = 0.0,
                      map_receiver=None, map_shape=None, d_map=None, seed: int = 0,
                      counter: int = 0) -> np.ndarray:
        """ssol_solve_adjoint: gradient [len(heliostats)][3] (two aim angles
        about `axes` [p][2][3], slope error) of d_absorbed * absorbed +
        d_map . per-primitive absorbed flux map (W/m2) of map_shape on
        map_receiver, from the realisations of solve(n, seed, counter)."""
        hs = (VP * len(heliostats))(*[self.instances[id(h)].value for h in heliostats])
        ax = np.ascontiguousarray(axes, dtype=np.float64).reshape(-1)
        assert ax.size == 6 * len(heliostats)
        grad = np.zeros((len(heliostats), 3))
        dm = None
        if d_map is not None:
            dm = np.ascontiguousarray(d_map, dtype=np.float64)
        r = Rng(seed, counter)
        call("ssol_solve_adjoint", self.scn, r.h, n, len(heliostats), hs,
             ax.ctypes.data_as(C.POINTER(C.c_double)), float(d_absorbed),
             self.instances[id(map_receiver)] if map_receiver is not None else None,
             self.shapes[id(map_shape)] if map_shape is not None else None,
             dm.ctypes.data_as(C.POINTER(C.c_double)) if dm is not None else None,
             grad.ctypes.data_as(C.POINTER(C.c_double)))
        return grad

    def autotune(self, probe: int) -> AutotuneResult:
        out = AutotuneResult()
        call("ssol_device_autotune", self.dev, self.scn, probe, C.byref(out))
        return out

    # element builders
    def _spectrum(self, sp: S.Spectrum):
        if id(sp) in self._spectra:
            return self._spectra[id(sp)]
        h = new("ssol_spectrum_create", self.dev)
        self._handles.append(("ssol_spectrum_ref_put", h))
        wl = list(sp.wavelengths)
        va = list(sp.values)

        def get(i, w, v, ctx):
            w[0] = wl[i]
            v[0] = va[i]
        cb = SPECTRUM_GET(get)
        call("ssol_spectrum_setup", h, cb, len(wl), None)
        self._spectra[id(sp)] = h
        return h

    def _data(self, d: S.Data) -> Data:
        out = Data()
        out.type = 0
        out.value.real = 0.0
        if d.spectrum is None:
            lib().ssol_data_set_real(C.byref(out), d.real)
        else:
            lib().ssol_data_set_spectrum(C.byref(out), self._spectrum(d.spectrum))
        return out

    def _medium(self, m: S.Medium) -> Medium:
        out = Medium()
        out.extinction = self._data(m.extinction)
        out.refractive_index = self._data(m.refractive_index)
        self._keep.append(out)
        return out

    def _material(self, m: S.Material):
        if id(m) in self._materials:
            return self._materials[id(m)]
        ctor = {"mirror": "ssol_material_create_mirror", "matte": "ssol_material_create_matte",
                "virtual": "ssol_material_create_virtual",
                "dielectric": "ssol_material_create_dielectric",
                "thin": "ssol_material_create_thin_dielectric"}[m.kind]
        h = new(ctor, self.dev)
        self._handles.append(("ssol_material_ref_put", h))
        nrm = _normal_shader(m.normal)
        self._keep.append(nrm)
        if m.kind == "mirror":
            refl, rough = _const_shader(m.reflectivity), _const_shader(m.roughness)
            sh = MirrorShader(nrm, refl, rough)
            self._keep += [sh, refl, rough]
            call("ssol_mirror_setup", h, C.byref(sh), m.distribution)
        elif m.kind == "matte":
            refl = _const_shader(m.reflectivity)
            sh = MatteShader(nrm, refl)
            self._keep += [sh, refl]
            call("ssol_matte_setup", h, C.byref(sh))
        elif m.kind == "dielectric":
            sh = NormalShader(nrm)
            self._keep.append(sh)
            call("ssol_dielectric_setup", h, C.byref(sh), C.byref(self._medium(m.medium_out)),
                 C.byref(self._medium(m.medium_in)))
        elif m.kind == "thin":
            sh = NormalShader(nrm)
            self._keep.append(sh)
            call("ssol_thin_dielectric_setup", h, C.byref(sh), C.byref(self._medium(m.medium_out)),
                 C.byref(self._medium(m.slab)), m.thickness)
        self._materials[id(m)] = h
        return h

    def _shape(self, s):
        if id(s) in self.shapes:
            return self.shapes[id(s)]
        if isinstance(s, S.Mesh):
            h = new("ssol_shape_create_mesh", self.dev)
            pos = np.ascontiguousarray(s.positions, dtype=np.float32)
            ids = np.ascontiguousarray(s.indices, dtype=np.uint32)

            def get_ids(i, out, ctx):
                out[0], out[1], out[2] = int(ids[i, 0]), int(ids[i, 1]), int(ids[i, 2])

            def get_pos(i, out, ctx):
                out[0], out[1], out[2] = float(pos[i, 0]), float(pos[i, 1]), float(pos[i, 2])
            cb_ids, cb_pos = INDEX_GET(get_ids), VERTEX_GET(get_pos)
            attr = (VertexData * 1)(VertexData(0, cb_pos))
            call("ssol_mesh_setup", h, len(ids), cb_ids, len(pos), attr, 1, None)
        else:
            h = new("ssol_shape_create_punched_surface", self.dev)
            q = Quadric()
            q.type = s.quadric
            q.data.v[0], q.data.v[1] = s.params
            q.transform = (C.c_double * 12)(*s.transform)
            q.slices_count_hint = s.slices_hint
            carv = (Carving * max(1, len(s.carvings)))()
            cbs = []
            for i, c in enumerate(s.carvings):
                xy = np.ascontiguousarray(c.xy, dtype=np.float64)

                def get(iv, out, ctx, xy=xy):
                    out[0], out[1] = float(xy[iv, 0]), float(xy[iv, 1])
                cb = CARVING_GET(get)
                cbs.append(cb)
                carv[i] = Carving(cb, len(xy), c.op, None)
            ps = PunchedSurface(C.pointer(q), carv, len(s.carvings))
            call("ssol_punched_surface_setup", h, C.byref(ps))
        self._handles.append(("ssol_shape_ref_put", h))
        self.shapes[id(s)] = h
        return h

    def _object(self, o: S.Object):
        if id(o) in self._objects:
            return self._objects[id(o)]
        h = new("ssol_object_create", self.dev)
        self._handles.append(("ssol_object_ref_put", h))
        for shape, mf, mb in o.sshapes:
            call("ssol_object_add_shaded_shape", h, self._shape(shape), self._material(mf),
                 self._material(mb))
        self._objects[id(o)] = h
        return h

    def shape_ntris(self, shape) -> int:
        n = C.c_uint()
        call("ssol_shape_get_triangles_count", self.shapes[id(shape)], C.byref(n))
        return n.value

    def solve(self, n: int, seed: int = 0, counter: int = 0, max_failed: int = 0,
              rng=None, tracker=None) -> "Estimator":
        """tracker: None, or (sun_ray_length, infinite_ray_length) to record
        every path (negative lengths take the default)"""
        r = rng if rng is not None else Rng(seed, counter)
        h = C.c_void_p()
        trk = None
        if tracker is not None:
            t = PathTracker(*tracker)
            trk = C.cast(C.pointer(t), C.c_void_p)
        res = lib().ssol_solve(self.scn, r.h if isinstance(r, Rng) else r, n, max_failed, trk,
                               C.byref(h))
        if res not in (0, 5) or not h.value:
            raise SsolError("ssol_solve", res)
        return Estimator(self, h, res)

    def close(self):
        if getattr(self, "scn", None):
            lib().ssol_scene_ref_put(self.scn)
            self.scn = None
            for fn, h in reversed(self._handles):
                getattr(lib(), fn)(h)
            self._handles = []
            lib().ssol_device_ref_put(self.dev)
            self.dev = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Estimator:
    def __init__(self, scene: Scene, h, res: int):
        self.scene = scene
        self.h = h
        self.res = res

    def __del__(self):
        if getattr(self, "h", None):
            lib().ssol_estimator_ref_put(self.h)
            self.h = None

    def global_(self):
        g = McGlobal()
        call("ssol_estimator_get_mc_global", self.h, C.byref(g))
        return {n: getattr(g, n).tuple() for n in GLOBAL_FIELDS}

    def receiver_struct(self, inst, side) -> McReceiver:
        r = McReceiver()
        call("ssol_estimator_get_mc_receiver", self.h, self.scene.instances[id(inst)], side,
             C.byref(r))
        return r

    def receiver(self, inst, side):
        r = self.receiver_struct(inst, side)
        return {n: getattr(r, n).tuple() for n in RECEIVER_FIELDS}

    def sampled(self, inst):
        s = McSampled()
        call("ssol_estimator_get_mc_sampled", self.h, self.scene.instances[id(inst)], C.byref(s))
        return {"cos_factor": s.cos_factor.tuple(), "shadowed": s.shadowed.tuple(),
                "nb_samples": s.nb_samples}

    def sampled_x_receiver(self, samp, rcv, side):
        r = McReceiver()
        call("ssol_estimator_get_mc_sampled_x_receiver", self.h, self.scene.instances[id(samp)],
             self.scene.instances[id(rcv)], side, C.byref(r))
        return {n: getattr(r, n).tuple() for n in RECEIVER_FIELDS}

    def primitive(self, inst, shape, prim, side):
        r = self.receiver_struct(inst, side)
        m = McShape()
        call("ssol_mc_receiver_get_mc_shape", C.byref(r), self.scene.shapes[id(shape)], C.byref(m))
        p = McPrimitive()
        call("ssol_mc_shape_get_mc_primitive", C.byref(m), prim, C.byref(p))
        return {n: getattr(p, n).tuple() for n in RECEIVER_FIELDS}

    def primitives(self, inst, shape, side, field="incoming_flux"):
        """E of one field for every triangle of `shape` (W/m2)."""
        r = self.receiver_struct(inst, side)
        m = McShape()
        call("ssol_mc_receiver_get_mc_shape", C.byref(r), self.scene.shapes[id(shape)], C.byref(m))
        n = self.scene.shape_ntris(shape)
        out = np.zeros((n, 3))
        p = McPrimitive()
        for i in range(n):
            call("ssol_mc_shape_get_mc_primitive", C.byref(m), i, C.byref(p))
            out[i] = getattr(p, field).tuple()
        return out

    @property
    def realisation_count(self):
        n = C.c_size_t()
        call("ssol_estimator_get_realisation_count", self.h, C.byref(n))
        return n.value

    @property
    def failed_count(self):
        n = C.c_size_t()
        call("ssol_estimator_get_failed_count", self.h, C.byref(n))
        return n.value

    @property
    def sampled_area(self):
        a = C.c_double()
        call("ssol_estimator_get_sampled_area", self.h, C.byref(a))
        return a.value

    @property
    def rng_state(self):
        h = C.c_void_p()
        call("ssol_estimator_get_rng_state", self.h, C.byref(h))
        return h

    def tracked_paths(self):
        """[(type name, vertices [n][4] = x, y, z, weight)] of a solve with a
        path tracker, through ssol_estimator_get_tracked_path and the
        ssol_path getters"""
        n = C.c_size_t()
        call("ssol_estimator_get_tracked_paths_count", self.h, C.byref(n))
        out = []
        path, nv, ty, v = Path(), C.c_size_t(), C.c_int(), PathVertex()
        for i in range(n.value):
            call("ssol_estimator_get_tracked_path", self.h, i, C.byref(path))
            call("ssol_path_get_vertices_count", C.byref(path), C.byref(nv))
            call("ssol_path_get_type", C.byref(path), C.byref(ty))
            a = np.zeros((nv.value, 4))
            for k in range(nv.value):
                call("ssol_path_get_vertex", C.byref(path), k, C.byref(v))
                a[k, :3] = v.pos[:]
                a[k, 3] = v.weight
            out.append((PATH_TYPES[ty.value], a))
        return out

    @property
    def stats(self) -> SolveStats:
        s = SolveStats()
        call("ssol_estimator_get_solve_stats", self.h, C.byref(s))
        return s
