"""Renderer on the solver's kernels (ssol_draw_draft, ssol_draw_pt; reference
src/ssol_draw_draft.c, src/ssol_draw_pt.c:126-285).

- Draft: every pixel of a camera looking down on a large plate is the mean,
  over the draft's pixel offsets (read back from the public ssp_rng API the
  host draws them from), of the cosine between the plate's normal and the
  camera ray that camera_ray builds (FP32 frame, as the reference): checked
  per pixel to 2e-6 against a numpy restatement.
- Path tracing: a sunlit Lambertian plate (rho 0.5, sun 60 degrees up) seen
  from above. Each sample gathers the sun term rho cos(theta_s) / pi (the
  BSDF value times the cosine, unit sun) and, after one bounce that
  multiplies the throughput by |cos| x rho and escapes upwards, 0.1 of it:
  E = rho cos(theta_s) / pi + 0.1 rho E[cos] = rho cos(theta_s) / pi + rho / 15
  for cosine-weighted bounces. The image mean is checked to 5 standard
  errors of the per-pixel spread."""
import ctypes as C
import math

import numpy as np
import pytest

import ssol_ctypes as L
import ssol_scene as S

pytestmark = pytest.mark.gpu

RHO = 0.5


def _plate_scene():
    plate = S.Object([(S.square_mesh(100.0), S.matte(RHO), S.matte(RHO))])
    inst = S.Instance(plate, S.transform(None, (0.0, 0.0, 0.0)), receiver_mask=S.FRONT,
                      sample=1)
    sun = S.Sun("directional", S.sun_dir_from(180.0, 60.0), 1000.0, S.mono_spectrum())
    return S.Scene([inst], sun, None, name="plate")


def _draft_offsets(spp):
    lib = L.lib()
    lib.ssp_rng_canonical_float.restype = C.c_float
    lib.ssp_rng_canonical_float.argtypes = [C.c_void_p]
    r = L.new("ssp_rng_create", None, L.SSP_RNG_THREEFRY)
    off = np.array([lib.ssp_rng_canonical_float(r) for _ in range(2 * spp)], dtype=np.float32)
    lib.ssp_rng_ref_put(r)
    return off.reshape(spp, 2)


def _frame(pos, tgt, up, fov):
    f = lambda v: np.asarray(v, dtype=np.float32)  # noqa: E731
    n = lambda v: (v / np.float32(np.sqrt(np.float32(np.dot(v, v))))).astype(np.float32)  # noqa: E731
    z = n(f(tgt) - f(pos))
    x = n(np.cross(z, f(up)).astype(np.float32))
    y = n(np.cross(z, x).astype(np.float32))
    depth = np.float32(1.0 / math.tan(fov * 0.5))
    return x, y, (z * depth).astype(np.float32)


def test_draft_plate_cosines(gpu):
    spec = _plate_scene()
    lib = L.Scene(spec)
    W, H, spp, fov = 48, 32, 4, math.radians(60.0)
    pos, tgt, up = (0.0, 0.0, 10.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0)
    img = lib.draw(W, H, spp, pos, tgt, up, fov=fov)
    lib.close()
    assert np.array_equal(img[..., 0], img[..., 1]) and np.array_equal(img[..., 0], img[..., 2])
    ax, ay, az = _frame(pos, tgt, up, fov)
    off = _draft_offsets(spp)
    want = np.zeros((H, W))
    for k in range(spp):
        sx = ((np.arange(W, dtype=np.float32) + off[k, 0]) / np.float32(W)).astype(np.float32)
        sy = ((np.arange(H, dtype=np.float32) + off[k, 1]) / np.float32(H)).astype(np.float32)
        d = (ax[None, None, :] * (sx[None, :, None] * 2 - 1) +
             ay[None, None, :] * (sy[:, None, None] * 2 - 1) + az[None, None, :]).astype(np.float32)
        d = d / np.sqrt((d * d).sum(-1, keepdims=True)).astype(np.float32)
        want += np.abs(d[..., 2].astype(np.float64))
    want /= spp
    err = np.abs(img[..., 0] - want)
    assert np.max(err) <= 2e-6, float(np.max(err))


def test_path_traced_lambertian_plate_closed_form(gpu):
    spec = _plate_scene()
    lib = L.Scene(spec)
    W, H, spp = 64, 64, 32
    img = lib.draw(W, H, spp, (0.0, 0.0, 10.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0),
                   fov=math.radians(60.0), sky_up=(0.0, 0.0, 1.0))
    lib.close()
    v = img[..., 0].ravel()
    cos_s = math.sin(math.radians(60.0))
    expect = RHO * cos_s / math.pi + RHO / 15.0
    se = v.std(ddof=1) / math.sqrt(v.size)
    print(f"mean {v.mean():.6f} expected {expect:.6f} se {se:.2e}", flush=True)
    assert abs(v.mean() - expect) <= 5 * se + 1e-7
    # the sun term alone is the floor of every sample
    assert v.min() >= RHO * cos_s / math.pi - 1e-9


def test_draw_arguments(gpu):
    lib = L.Scene(_plate_scene())
    cam = L.new("ssol_camera_create", lib.dev)
    img = L.new("ssol_image_create", lib.dev)
    L.call("ssol_image_setup", img, 8, 8, 0)
    w = C.cast(L.lib().ssol_image_write, C.c_void_p)
    up = (C.c_double * 3)(0, 0, 1)
    assert L.lib().ssol_draw_draft(lib.scn, cam, 8, 8, 0, w, img) != 0
    assert L.lib().ssol_draw_draft(lib.scn, cam, 8, 8, 1, None, img) != 0
    assert L.lib().ssol_draw_pt(lib.scn, cam, 8, 8, 1, None, w, img) != 0
    assert L.lib().ssol_draw_pt(lib.scn, None, 8, 8, 1, up, w, img) != 0
    # a writer that refuses: the error comes back
    assert L.lib().ssol_draw_draft(lib.scn, cam, 16, 8, 1, w, img) != 0  # image too small
    L.lib().ssol_image_ref_put(img)
    L.lib().ssol_camera_ref_put(cam)
    lib.close()
