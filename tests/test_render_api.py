"""Camera and image API of the renderer (reference src/ssol_camera.c,
src/ssol_image.c, whose behaviour the tests restate): argument checks, the
camera frame that look_at / set_fov / set_proj_ratio produce (read back
through a draft render's first-ray geometry is a GPU test; here the image
side), the image layout, ssol_image_write and nearest / bilinear sampling
with clamped and repeated addressing, against numpy. CPU only: no solve."""
import ctypes as C

import numpy as np
import pytest

import ssol_ctypes as L

RES_OK, RES_BAD_ARG = 0, 1
DOUBLE3, FILTER_LINEAR, FILTER_NEAREST, CLAMP, REPEAT = 0, 0, 1, 0, 1


class Layout(C.Structure):
    _fields_ = [("row_pitch", C.c_size_t), ("offset", C.c_size_t), ("size", C.c_size_t),
                ("width", C.c_size_t), ("height", C.c_size_t), ("pixel_format", C.c_int)]


@pytest.fixture
def dev():
    d = L.new("ssol_device_create", None, None, 0xFFFFFFFF, 0)
    yield d
    L.lib().ssol_device_ref_put(d)


def d3(*v):
    return (C.c_double * 3)(*v)


def test_camera_arguments(dev):
    lib = L.lib()
    cam = L.new("ssol_camera_create", dev)
    assert lib.ssol_camera_create(None, C.byref(C.c_void_p())) == RES_BAD_ARG
    assert lib.ssol_camera_set_fov(cam, 0.0) == RES_BAD_ARG
    assert lib.ssol_camera_set_fov(cam, 1.0) == RES_OK
    assert lib.ssol_camera_set_proj_ratio(cam, -1.0) == RES_BAD_ARG
    assert lib.ssol_camera_set_proj_ratio(cam, 4.0 / 3.0) == RES_OK
    # degenerate frames: target on the position, up along the view
    assert lib.ssol_camera_look_at(cam, d3(0, 0, 0), d3(0, 0, 0), d3(0, 1, 0)) == RES_BAD_ARG
    assert lib.ssol_camera_look_at(cam, d3(0, 0, 0), d3(0, 1, 0), d3(0, 1, 0)) == RES_BAD_ARG
    assert lib.ssol_camera_look_at(cam, d3(1, 2, 3), d3(0, 0, 0), d3(0, 0, 1)) == RES_OK
    assert lib.ssol_camera_ref_put(cam) == RES_OK


def _image(dev, w, h):
    img = L.new("ssol_image_create", dev)
    assert L.lib().ssol_image_setup(img, w, h, DOUBLE3) == RES_OK
    return img


def test_image_layout_write_and_map(dev):
    lib = L.lib()
    img = L.new("ssol_image_create", dev)
    assert lib.ssol_image_setup(img, 0, 4, DOUBLE3) == RES_BAD_ARG
    assert lib.ssol_image_setup(img, 5, 4, 1) == RES_BAD_ARG
    assert lib.ssol_image_setup(img, 5, 4, DOUBLE3) == RES_OK
    lay = Layout()
    assert lib.ssol_image_get_layout(img, C.byref(lay)) == RES_OK
    assert (lay.row_pitch, lay.offset, lay.size, lay.width, lay.height, lay.pixel_format) == \
        (5 * 24, 0, 20, 5, 4, DOUBLE3)
    tile = np.arange(2 * 3 * 3, dtype=np.float64).reshape(2, 3, 3)
    org, size = (C.c_size_t * 2)(1, 2), (C.c_size_t * 2)(3, 2)
    ptr = tile.ctypes.data_as(C.c_void_p)
    assert lib.ssol_image_write(img, org, size, DOUBLE3, ptr) == RES_OK
    assert lib.ssol_image_write(img, (C.c_size_t * 2)(3, 2), size, DOUBLE3, ptr) == RES_BAD_ARG
    assert lib.ssol_image_write(img, org, size, 1, ptr) == RES_BAD_ARG
    mem = C.c_void_p()
    assert lib.ssol_image_map(img, C.byref(mem)) == RES_OK
    a = np.frombuffer(C.string_at(mem.value, 5 * 4 * 24), dtype=np.float64).reshape(4, 5, 3)
    want = np.zeros((4, 5, 3))
    want[2:4, 1:4] = tile
    assert np.array_equal(a, want)
    assert lib.ssol_image_unmap(img) == RES_OK
    lib.ssol_image_ref_put(img)


def _map(u, mode):
    if mode == CLAMP:
        return min(max(u, 0.0), np.nextafter(1.0, 0.0))
    f = np.modf(u)[0]
    return f + 1.0 if f < 0 else f


def _sample(px, filt, au, av, uv):
    h, w, _ = px.shape
    x0, y0 = _map(uv[0], au) * w, _map(uv[1], av) * h
    at = lambda x, y: px[min(int(y), h - 1), min(int(x), w - 1)]  # noqa: E731
    if filt == FILTER_NEAREST:
        return at(x0, y0)
    x1, y1 = _map(uv[0] + 1.0 / w, au) * w, _map(uv[1] + 1.0 / h, av) * h
    s, t = np.modf(x0)[0], np.modf(y0)[0]
    return (1 - s) * ((1 - t) * at(x0, y0) + t * at(x0, y1)) + \
        s * ((1 - t) * at(x1, y0) + t * at(x1, y1))


def test_image_sampling_against_numpy(dev):
    lib = L.lib()
    w, h = 7, 5
    img = _image(dev, w, h)
    px = np.random.default_rng(3).random((h, w, 3))
    assert lib.ssol_image_write(img, (C.c_size_t * 2)(0, 0), (C.c_size_t * 2)(w, h), DOUBLE3,
                                px.ctypes.data_as(C.c_void_p)) == RES_OK
    out = (C.c_double * 3)()
    for uv in ((0.0, 0.0), (0.31, 0.77), (0.999, 0.5), (1.25, -0.4), (-2.1, 3.7), (0.5, 1.0)):
        for filt in (FILTER_NEAREST, FILTER_LINEAR):
            for au in (CLAMP, REPEAT):
                for av in (CLAMP, REPEAT):
                    assert lib.ssol_image_sample(img, filt, au, av, (C.c_double * 2)(*uv),
                                                 out) == RES_OK
                    assert np.allclose(list(out), _sample(px, filt, au, av, uv),
                                       rtol=1e-12, atol=1e-15), (uv, filt, au, av)
    lib.ssol_image_ref_put(img)
