"""ORACLE -- test infrastructure only.

ctypes binding of the oracle (liboracle.so, the CPU FP64 restatement of the
reference flux integrator). Only the tests, the smoke check and the
benchmark's CPU-baseline leg use it.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, List

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library not built: {path} (run make -C oracle)")
        _LIB = C.CDLL(path)
        _declare(_LIB)
    return _LIB


class OData(C.Structure):
    _fields_ = [("real", C.c_double), ("spectrum", C.c_int)]


class OMedium(C.Structure):
    _fields_ = [("extinction", OData), ("refractive_index", OData)]


class OSpectrum(C.Structure):
    _fields_ = [("n", C.c_size_t), ("wavelengths", C.POINTER(C.c_double)),
                ("values", C.POINTER(C.c_double))]


class OMaterial(C.Structure):
    _fields_ = [("type", C.c_int), ("distribution", C.c_int), ("normal_from_ng", C.c_int),
                ("reflectivity", C.c_double), ("roughness", C.c_double),
                ("medium_out", OMedium), ("medium_in", OMedium), ("slab", OMedium),
                ("thickness", C.c_double)]


class OCarving(C.Structure):
    _fields_ = [("operation", C.c_int), ("nverts", C.c_size_t), ("xy", C.POINTER(C.c_double))]


class OShape(C.Structure):
    _fields_ = [("type", C.c_int), ("nverts", C.c_uint), ("ntris", C.c_uint),
                ("positions", C.POINTER(C.c_float)), ("indices", C.POINTER(C.c_uint)),
                ("quadric", C.c_int), ("qdata", C.c_double * 2),
                ("qtransform", C.c_double * 12), ("slices_hint", C.c_size_t),
                ("ncarvings", C.c_size_t), ("carvings", C.POINTER(OCarving))]


class OSShape(C.Structure):
    _fields_ = [("shape", C.c_int), ("mtl_front", C.c_int), ("mtl_back", C.c_int)]


class OObject(C.Structure):
    _fields_ = [("nsshapes", C.c_size_t), ("sshapes", C.POINTER(OSShape))]


class OInstance(C.Structure):
    _fields_ = [("object", C.c_int), ("transform", C.c_double * 12),
                ("receiver_mask", C.c_int), ("per_primitive", C.c_int), ("sample", C.c_int)]


class OSun(C.Structure):
    _fields_ = [("type", C.c_int), ("dir", C.c_double * 3), ("dni", C.c_double),
                ("param", C.c_double), ("spectrum", C.c_int)]


class OScene(C.Structure):
    _fields_ = [("nspectra", C.c_size_t), ("spectra", C.POINTER(OSpectrum)),
                ("nmaterials", C.c_size_t), ("materials", C.POINTER(OMaterial)),
                ("nshapes", C.c_size_t), ("shapes", C.POINTER(OShape)),
                ("nobjects", C.c_size_t), ("objects", C.POINTER(OObject)),
                ("ninstances", C.c_size_t), ("instances", C.POINTER(OInstance)),
                ("sun", OSun), ("has_atmosphere", C.c_int),
                ("atmosphere_extinction", OData),
                ("track_sampled_x_receiver", C.c_int)]


R3 = C.c_double * 3
R10x3 = (C.c_double * 3) * 10


def _declare(L):
    L.oracle_create.restype = C.c_void_p
    L.oracle_create.argtypes = [C.POINTER(OScene), C.POINTER(C.c_int)]
    L.oracle_destroy.argtypes = [C.c_void_p]
    L.oracle_solve.restype = C.c_int
    L.oracle_solve.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_int]
    L.oracle_solve_tracked.restype = C.c_int
    L.oracle_solve_tracked.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64,
                                       C.c_uint64, C.c_double, C.c_double]
    L.oracle_tracked_count.restype = C.c_size_t
    L.oracle_tracked_count.argtypes = [C.c_void_p]
    L.oracle_tracked_path.restype = C.c_int
    L.oracle_tracked_path.argtypes = [C.c_void_p, C.c_size_t, C.POINTER(C.c_int),
                                      C.POINTER(C.c_size_t), C.POINTER(C.POINTER(C.c_double))]
    L.oracle_tracked_extend.restype = C.c_double
    L.oracle_tracked_extend.argtypes = [C.c_void_p]
    L.oracle_get_global.argtypes = [C.c_void_p, (C.c_double * 3) * 6]
    L.oracle_get_receiver.argtypes = [C.c_void_p, C.c_size_t, C.c_int, R10x3]
    L.oracle_get_sampled.argtypes = [C.c_void_p, C.c_size_t, (C.c_double * 3) * 2,
                                     C.POINTER(C.c_uint64)]
    L.oracle_get_sampled_x_receiver.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_int, R10x3]
    L.oracle_get_primitive.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_uint, C.c_int, R10x3]
    for f in ("oracle_realisation_count", "oracle_failed_count", "oracle_segments"):
        getattr(L, f).restype = C.c_uint64
        getattr(L, f).argtypes = [C.c_void_p]
    for f in ("oracle_sampled_area", "oracle_sampled_area_proxy"):
        getattr(L, f).restype = C.c_double
        getattr(L, f).argtypes = [C.c_void_p]
    L.oracle_shape_ntris.restype = C.c_uint
    L.oracle_shape_ntris.argtypes = [C.c_void_p, C.c_size_t]
    L.oracle_shape_rt_area.restype = C.c_double
    L.oracle_shape_rt_area.argtypes = [C.c_void_p, C.c_size_t]
    L.oracle_shape_samp_area.restype = C.c_double
    L.oracle_shape_samp_area.argtypes = [C.c_void_p, C.c_size_t]
    L.oracle_shape_rt_vertex_count.restype = C.c_uint
    L.oracle_shape_rt_vertex_count.argtypes = [C.c_void_p, C.c_size_t]
    L.oracle_shape_rt_vertex.argtypes = [C.c_void_p, C.c_size_t, C.c_uint, R3]
    L.oracle_threefry4x64_20.argtypes = [C.c_uint64 * 4, C.c_uint64 * 4, C.c_uint64 * 4]
    L.oracle_solve_second.restype = C.c_int
    L.oracle_solve_second.argtypes = [C.c_double] * 4 + [C.c_double * 2]
    L.oracle_quadric_intersect_local.restype = C.c_int
    L.oracle_quadric_intersect_local.argtypes = [C.c_int, C.c_double * 2, R3, R3, C.c_double,
                                                 C.POINTER(C.c_double), R3]
    L.oracle_quadric_z.restype = C.c_double
    L.oracle_quadric_z.argtypes = [C.c_int, C.c_double * 2, C.c_double * 2]
    L.oracle_basis.argtypes = [R3, C.c_double * 9]
    L.oracle_sun_sample.restype = C.c_int
    L.oracle_sun_sample.argtypes = [C.c_int, C.c_double, R3, C.POINTER(C.c_double), C.c_int, R3]
    L.oracle_buie_params.argtypes = [C.c_double, C.c_double * 8]
    L.oracle_buie_pdf_theta.restype = C.c_double
    L.oracle_buie_pdf_theta.argtypes = [C.c_double, C.c_double]
    L.oracle_spectrum_interpolate.restype = C.c_double
    L.oracle_spectrum_interpolate.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double),
                                              C.c_size_t, C.c_double]
    L.oracle_piecewise_linear_sample.restype = C.c_double
    L.oracle_piecewise_linear_sample.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double),
                                                 C.c_size_t, C.c_double]
    L.oracle_bsdf_sample.restype = C.c_double
    L.oracle_bsdf_sample.argtypes = [C.POINTER(OMaterial), C.c_double, C.c_double, C.c_double,
                                     C.c_double, R3, R3, C.POINTER(C.c_double), C.c_int, R3,
                                     C.POINTER(C.c_int)]


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    """The oracle over one scene description (ssol_scene.Scene)."""

    def __init__(self, scene, track_sampled_x_receiver: bool = True):
        L = lib()
        self._keep: List[object] = []
        self.scene = scene
        spectra: Dict[int, int] = {}
        spec_list = []

        def spec_id(sp):
            if sp is None:
                return -1
            if id(sp) not in spectra:
                wl = np.ascontiguousarray(sp.wavelengths, dtype=np.float64)
                va = np.ascontiguousarray(sp.values, dtype=np.float64)
                self._keep += [wl, va]
                spectra[id(sp)] = len(spec_list)
                spec_list.append(OSpectrum(len(wl), _dptr(wl), _dptr(va)))
            return spectra[id(sp)]

        def odata(d):
            return OData(float(d.real), spec_id(d.spectrum))

        def omedium(m):
            return OMedium(odata(m.extinction), odata(m.refractive_index))

        materials: Dict[int, int] = {}
        mat_list = []

        def mat_id(m):
            if id(m) not in materials:
                materials[id(m)] = len(mat_list)
                mat_list.append(OMaterial(m.type_id, m.distribution, 1 if m.normal == "Ng" else 0,
                                          m.reflectivity, m.roughness, omedium(m.medium_out),
                                          omedium(m.medium_in), omedium(m.slab), m.thickness))
            return materials[id(m)]

        shapes: Dict[int, int] = {}
        shape_list = []

        def shape_id(s):
            if id(s) in shapes:
                return shapes[id(s)]
            from ssol_scene import Mesh
            if isinstance(s, Mesh):
                pos = np.ascontiguousarray(s.positions, dtype=np.float32)
                ids = np.ascontiguousarray(s.indices, dtype=np.uint32)
                self._keep += [pos, ids]
                sh = OShape(0, len(pos), len(ids), pos.ctypes.data_as(C.POINTER(C.c_float)),
                            ids.ctypes.data_as(C.POINTER(C.c_uint)))
            else:
                carv = (OCarving * max(1, len(s.carvings)))()
                for i, c in enumerate(s.carvings):
                    xy = np.ascontiguousarray(c.xy, dtype=np.float64)
                    self._keep.append(xy)
                    carv[i] = OCarving(c.op, len(xy), _dptr(xy))
                self._keep.append(carv)
                sh = OShape(1, 0, 0, None, None, s.quadric, (C.c_double * 2)(*s.params),
                            (C.c_double * 12)(*s.transform), s.slices_hint, len(s.carvings), carv)
            shapes[id(s)] = len(shape_list)
            shape_list.append(sh)
            return shapes[id(s)]

        objects: Dict[int, int] = {}
        obj_list = []

        def obj_id(o):
            if id(o) not in objects:
                arr = (OSShape * len(o.sshapes))()
                for i, (s, mf, mb) in enumerate(o.sshapes):
                    arr[i] = OSShape(shape_id(s), mat_id(mf), mat_id(mb))
                self._keep.append(arr)
                objects[id(o)] = len(obj_list)
                obj_list.append(OObject(len(o.sshapes), arr))
            return objects[id(o)]

        insts = (OInstance * len(scene.instances))()
        for i, inst in enumerate(scene.instances):
            insts[i] = OInstance(obj_id(inst.object), (C.c_double * 12)(*inst.transform),
                                 inst.receiver_mask, inst.per_primitive, inst.sample)
        sun = scene.sun
        osun = OSun(sun.type_id, R3(*sun.direction), sun.dni, sun.param, spec_id(sun.spectrum))
        atm = odata(scene.atmosphere) if scene.atmosphere is not None else OData(0.0, -1)

        def arr(T, lst):
            a = (T * max(1, len(lst)))(*lst)
            self._keep.append(a)
            return a

        self._scn = OScene(len(spec_list), arr(OSpectrum, spec_list),
                           len(mat_list), arr(OMaterial, mat_list),
                           len(shape_list), arr(OShape, shape_list),
                           len(obj_list), arr(OObject, obj_list),
                           len(scene.instances), insts, osun,
                           1 if scene.atmosphere is not None else 0, atm,
                           1 if track_sampled_x_receiver else 0)
        self._keep.append(insts)
        self.shape_index = shapes
        self.instance_index = {id(inst): i for i, inst in enumerate(scene.instances)}
        err = C.c_int(0)
        self._h = L.oracle_create(C.byref(self._scn), C.byref(err))
        if not self._h:
            raise ValueError(f"oracle_create failed ({err.value})")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().oracle_destroy(h)
            self._h = None

    def solve(self, n: int, seed: int = 0, counter: int = 0, max_failed: int = 0,
              nthreads: int = 0) -> int:
        return lib().oracle_solve(self._h, seed, counter, n, max_failed, nthreads)

    def solve_tracked(self, n: int, seed: int = 0, counter: int = 0, max_failed: int = 0,
                      sun_ray_length: float = -1.0, infinite_ray_length: float = -1.0) -> int:
        """Single-threaded solve recording every path (reference path tracker)."""
        return lib().oracle_solve_tracked(self._h, seed, counter, n, max_failed,
                                          sun_ray_length, infinite_ray_length)

    def tracked_paths(self):
        """[(type, vertices [n][4] = x, y, z, weight)] in realisation order."""
        out = []
        for i in range(lib().oracle_tracked_count(self._h)):
            ty, n, v = C.c_int(0), C.c_size_t(0), C.POINTER(C.c_double)()
            assert lib().oracle_tracked_path(self._h, i, C.byref(ty), C.byref(n), C.byref(v)) == 0
            a = np.ctypeslib.as_array(v, shape=(n.value * 4,)).reshape(n.value, 4).copy() \
                if n.value else np.zeros((0, 4))
            out.append((ty.value, a))
        return out

    def tracked_extend(self) -> float:
        return lib().oracle_tracked_extend(self._h)

    @staticmethod
    def _res(a):
        return [tuple(a[i]) for i in range(len(a))]

    GLOBAL = ("cos_factor", "absorbed_by_receivers", "shadowed", "missing",
              "extinguished_by_atmosphere", "other_absorbed")
    RECEIVER = ("incoming_flux", "incoming_if_no_atm_loss", "incoming_if_no_field_loss",
                "incoming_lost_in_field", "incoming_lost_in_atmosphere", "absorbed_flux",
                "absorbed_if_no_atm_loss", "absorbed_if_no_field_loss",
                "absorbed_lost_in_field", "absorbed_lost_in_atmosphere")

    def global_(self):
        a = ((C.c_double * 3) * 6)()
        lib().oracle_get_global(self._h, a)
        return dict(zip(self.GLOBAL, self._res(a)))

    def receiver(self, inst, side):
        a = R10x3()
        if lib().oracle_get_receiver(self._h, self.instance_index[id(inst)], side, a):
            raise ValueError("not a receiver side")
        return dict(zip(self.RECEIVER, self._res(a)))

    def sampled(self, inst):
        a = ((C.c_double * 3) * 2)()
        nb = C.c_uint64(0)
        if lib().oracle_get_sampled(self._h, self.instance_index[id(inst)], a, C.byref(nb)):
            raise ValueError("not sampled")
        return {"cos_factor": tuple(a[0]), "shadowed": tuple(a[1]), "nb_samples": nb.value}

    def sampled_x_receiver(self, samp, rcv, side):
        a = R10x3()
        if lib().oracle_get_sampled_x_receiver(self._h, self.instance_index[id(samp)],
                                               self.instance_index[id(rcv)], side, a):
            raise ValueError("no sampled x receiver tally")
        return dict(zip(self.RECEIVER, self._res(a)))

    def primitive(self, inst, sshape, prim, side):
        a = R10x3()
        if lib().oracle_get_primitive(self._h, self.instance_index[id(inst)], sshape, prim,
                                      side, a):
            raise ValueError("no per-primitive tally")
        return dict(zip(self.RECEIVER, self._res(a)))

    @property
    def realisation_count(self):
        return lib().oracle_realisation_count(self._h)

    @property
    def failed_count(self):
        return lib().oracle_failed_count(self._h)

    @property
    def segments(self):
        return lib().oracle_segments(self._h)

    @property
    def sampled_area(self):
        return lib().oracle_sampled_area(self._h)

    @property
    def sampled_area_proxy(self):
        return lib().oracle_sampled_area_proxy(self._h)

    def shape_ntris(self, shape):
        return lib().oracle_shape_ntris(self._h, self.shape_index[id(shape)])

    def shape_rt_area(self, shape):
        return lib().oracle_shape_rt_area(self._h, self.shape_index[id(shape)])

    def shape_rt_vertices(self, shape):
        L = lib()
        s = self.shape_index[id(shape)]
        n = L.oracle_shape_rt_vertex_count(self._h, s)
        out = np.zeros((n, 3))
        v = R3()
        for i in range(n):
            L.oracle_shape_rt_vertex(self._h, s, i, v)
            out[i] = tuple(v)
        return out


# Deterministic kernels -------------------------------------------------------
def threefry(ctr, key):
    A = C.c_uint64 * 4
    out = A()
    lib().oracle_threefry4x64_20(A(*ctr), A(*key), out)
    return list(out)


def solve_second(a, b, c, hint):
    d = (C.c_double * 2)()
    n = lib().oracle_solve_second(a, b, c, hint, d)
    return n, list(d)


def quadric_intersect(quadric, qdata, org, dir, hint):
    t = C.c_double(0)
    g = R3()
    ok = lib().oracle_quadric_intersect_local(quadric, (C.c_double * 2)(*qdata), R3(*org),
                                               R3(*dir), hint, C.byref(t), g)
    return ok, t.value, np.array(tuple(g))


def quadric_z(quadric, qdata, xy):
    return lib().oracle_quadric_z(quadric, (C.c_double * 2)(*qdata), (C.c_double * 2)(*xy))


def basis(d):
    m = (C.c_double * 9)()
    lib().oracle_basis(R3(*d), m)
    return np.array(list(m)).reshape(3, 3).T  # columns x, y, z


def sun_sample(kind, param, d, u):
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = R3()
    used = lib().oracle_sun_sample(kind, param, R3(*d), _dptr(u), len(u), out)
    return np.array(tuple(out)), used


def buie_params(csr):
    out = (C.c_double * 8)()
    lib().oracle_buie_params(csr, out)
    return list(out)


def buie_pdf_theta(csr, theta):
    return lib().oracle_buie_pdf_theta(csr, theta)


def spectrum_interpolate(wl, val, w):
    wl = np.ascontiguousarray(wl, dtype=np.float64)
    val = np.ascontiguousarray(val, dtype=np.float64)
    return lib().oracle_spectrum_interpolate(_dptr(wl), _dptr(val), len(wl), w)


def piecewise_linear_sample(wl, val, u):
    wl = np.ascontiguousarray(wl, dtype=np.float64)
    val = np.ascontiguousarray(val, dtype=np.float64)
    return lib().oracle_piecewise_linear_sample(_dptr(wl), _dptr(val), len(wl), u)


def bsdf_sample(material, wi, N, u, eta_out=1.0, eta_in=1.0, slab_eta=1.0, slab_abs=0.0):
    m = OMaterial(material.type_id, material.distribution, 0, material.reflectivity,
                  material.roughness, OMedium(), OMedium(), OMedium(), material.thickness)
    u = np.ascontiguousarray(u, dtype=np.float64)
    wo = R3()
    t = C.c_int(0)
    w = lib().oracle_bsdf_sample(C.byref(m), eta_out, eta_in, slab_eta, slab_abs, R3(*wi),
                                 R3(*N), _dptr(u), len(u), wo, C.byref(t))
    return w, np.array(tuple(wo)), t.value
