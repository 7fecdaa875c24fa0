// Rendering API: pinhole camera, images and the draft / path-traced draws
// (reference src/ssol_camera.c, src/ssol_image.c, src/ssol_draw.c,
// src/ssol_draw_draft.c, src/ssol_draw_pt.c). The pixels come from
// ssolh_render on the scene ssol_solve would upload; the host hands them to
// the caller's writer in 32x32 tiles, as the reference's tiled loop does.
#include <math.h>
#include <string.h>

#include <cmath>

#include <algorithm>
#include <vector>

#include "ssol_internal.hpp"

struct ssol_camera {
  int ref = 1;
  ssol_device* dev = nullptr;
  float axis_x[3] = {1, 0, 0}, axis_y[3] = {0, 1, 0}, axis_z[3] = {0, 0, -1};
  float position[3] = {0, 0, 0};
  float fov_x = (float)(M_PI / 2);
  float rcp_proj_ratio = 1.f;  // height / width
};

struct ssol_image {
  int ref = 1;
  ssol_device* dev = nullptr;
  std::vector<char> mem;
  size_t pitch = 0, width = 0, height = 0;
  enum ssol_pixel_format format = SSOL_PIXEL_FORMATS_COUNT__;
};

namespace {

constexpr size_t TILE = 32;

float f3_len(const float v[3]) { return sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

// v / |v| in place; false if |v| is 0 or not finite
bool f3_normalize(float v[3]) {
  const float l = f3_len(v);
  if (!(l > 0) || !std::isfinite(l)) return false;
  for (int k = 0; k < 3; ++k) v[k] /= l;
  return true;
}

void f3_cross(const float a[3], const float b[3], float o[3]) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

// Texel address in [0, 1) (reference map_address, src/ssol_image.c)
double map_address(double a, enum ssol_address_mode mode) {
  if (mode == SSOL_ADDRESS_CLAMP) return std::min(std::max(a, 0.0), nextafter(1.0, 0.0));
  double ip;
  double f = modf(a, &ip);
  if (f < 0) f += 1.0;
  return f;
}

const double* texel(const ssol_image* img, size_t x, size_t y) {
  x = std::min(x, img->width - 1);
  y = std::min(y, img->height - 1);
  return (const double*)(img->mem.data() + y * img->pitch + x * 3 * sizeof(double));
}

// Renders through ssolh_render and hands the image to `writer` tile by tile
res_T draw(ssol_scene* scn, ssol_camera* cam, size_t width, size_t height, size_t spp,
           uint32_t mode, const double up[3], ssol_write_pixels_T writer, void* data,
           const char* who) {
  if (!scn || !cam || !width || !height || !spp || !writer) return RES_BAD_ARG;
  if (width > (1u << 16) || height > (1u << 16) || spp > (1u << 20)) return RES_BAD_ARG;
  res_T r = ssol::scene_for_gpu(scn, who);
  if (r != RES_OK) return r;
  ssolh_render_args a{};
  a.width = (uint32_t)width;
  a.height = (uint32_t)height;
  a.spp = (uint32_t)spp;
  a.mode = mode;
  for (int k = 0; k < 3; ++k) {
    a.position[k] = cam->position[k];
    a.axis_x[k] = cam->axis_x[k];
    a.axis_y[k] = cam->axis_y[k];
    a.axis_z[k] = cam->axis_z[k];
    a.up[k] = up ? up[k] : 0.0;
  }
  // draft: pixel offsets shared by every pixel, from a Threefry generator
  // (reference: ssp_rng_canonical_float of a fresh generator); path
  // tracing: per-sample streams keyed by the fixed key below
  std::vector<float> draft;
  if (mode == 0) {
    ssp_rng* rng = nullptr;
    if (ssp_rng_create(nullptr, SSP_RNG_THREEFRY, &rng) != RES_OK) return RES_MEM_ERR;
    draft.resize(2 * spp);
    for (float& x : draft) x = ssp_rng_canonical_float(rng);
    ssp_rng_ref_put(rng);
    a.draft_samples = draft.data();
  }
  a.rng_key = 0x7265656e64657273ull;
  std::vector<double> pix(width * height);
  const int e = ssolh_render(scn->dev->gpu, &a, pix.data());
  if (e) {
    ssol::log_error(scn->dev, "%s: GPU render failed: %s\n", who, ssolh_last_error());
    return e == SSOLH_ERR_ARG ? RES_BAD_ARG : RES_UNKNOWN_ERR;
  }
  std::vector<double> tile(TILE * TILE * 3);
  for (size_t ty = 0; ty < height; ty += TILE)
    for (size_t tx = 0; tx < width; tx += TILE) {
      const size_t origin[2] = {tx, ty};
      const size_t size[2] = {std::min(TILE, width - tx), std::min(TILE, height - ty)};
      for (size_t y = 0; y < size[1]; ++y)
        for (size_t x = 0; x < size[0]; ++x) {
          const double v = pix[(ty + y) * width + tx + x];
          double* o = tile.data() + (y * size[0] + x) * 3;
          o[0] = o[1] = o[2] = v;
        }
      r = writer(data, origin, size, SSOL_PIXEL_DOUBLE3, tile.data());
      if (r != RES_OK) return r;
    }
  return RES_OK;
}

}  // namespace

extern "C" {

// Camera (reference src/ssol_camera.c) ---------------------------------------
res_T ssol_camera_create(struct ssol_device* dev, struct ssol_camera** out) {
  if (!dev || !out) return RES_BAD_ARG;
  ssol_camera* cam = new ssol_camera();
  cam->dev = dev;
  ssol::ref_get(dev);
  const double pos[3] = {0, 0, 0}, tgt[3] = {0, 0, -1}, up[3] = {0, 1, 0};
  ssol_camera_look_at(cam, pos, tgt, up);
  *out = cam;
  return RES_OK;
}

res_T ssol_camera_ref_get(struct ssol_camera* cam) {
  if (!cam) return RES_BAD_ARG;
  ++cam->ref;
  return RES_OK;
}

res_T ssol_camera_ref_put(struct ssol_camera* cam) {
  if (!cam) return RES_BAD_ARG;
  if (--cam->ref == 0) {
    ssol_device* d = cam->dev;
    delete cam;
    ssol::device_ref_put_(d);
  }
  return RES_OK;
}

res_T ssol_camera_set_proj_ratio(struct ssol_camera* cam, const double ratio) {
  if (!cam || !(ratio > 0)) return RES_BAD_ARG;
  float y[3] = {cam->axis_y[0], cam->axis_y[1], cam->axis_y[2]};
  if (!f3_normalize(y)) return RES_BAD_ARG;
  cam->rcp_proj_ratio = (float)(1.0 / ratio);
  for (int k = 0; k < 3; ++k) cam->axis_y[k] = y[k] * cam->rcp_proj_ratio;
  return RES_OK;
}

res_T ssol_camera_set_fov(struct ssol_camera* cam, const double fov_x) {
  if (!cam || !((float)fov_x > 0)) return RES_BAD_ARG;
  float z[3] = {cam->axis_z[0], cam->axis_z[1], cam->axis_z[2]};
  if (!f3_normalize(z)) return RES_BAD_ARG;
  const float depth = (float)(1.0 / tan(fov_x * 0.5));
  for (int k = 0; k < 3; ++k) cam->axis_z[k] = z[k] * depth;
  cam->fov_x = (float)fov_x;
  return RES_OK;
}

res_T ssol_camera_look_at(struct ssol_camera* cam, const double pos[3], const double tgt[3],
                          const double up[3]) {
  if (!cam || !pos || !tgt || !up) return RES_BAD_ARG;
  float p[3], t[3], u[3], x[3], y[3], z[3];
  for (int k = 0; k < 3; ++k) { p[k] = (float)pos[k]; t[k] = (float)tgt[k]; u[k] = (float)up[k]; }
  for (int k = 0; k < 3; ++k) z[k] = t[k] - p[k];
  if (!f3_normalize(z)) return RES_BAD_ARG;
  f3_cross(z, u, x);
  if (!f3_normalize(x)) return RES_BAD_ARG;
  f3_cross(z, x, y);
  if (!f3_normalize(y)) return RES_BAD_ARG;
  const float depth = (float)(1.0 / tan(cam->fov_x * 0.5));
  for (int k = 0; k < 3; ++k) {
    cam->axis_x[k] = x[k];
    cam->axis_y[k] = y[k] * cam->rcp_proj_ratio;
    cam->axis_z[k] = z[k] * depth;
    cam->position[k] = p[k];
  }
  return RES_OK;
}

// Image (reference src/ssol_image.c) -------------------------------------------
res_T ssol_image_create(struct ssol_device* dev, struct ssol_image** out) {
  if (!dev || !out) return RES_BAD_ARG;
  ssol_image* img = new ssol_image();
  img->dev = dev;
  ssol::ref_get(dev);
  *out = img;
  return RES_OK;
}

res_T ssol_image_ref_get(struct ssol_image* img) {
  if (!img) return RES_BAD_ARG;
  ++img->ref;
  return RES_OK;
}

res_T ssol_image_ref_put(struct ssol_image* img) {
  if (!img) return RES_BAD_ARG;
  if (--img->ref == 0) {
    ssol_device* d = img->dev;
    delete img;
    ssol::device_ref_put_(d);
  }
  return RES_OK;
}

res_T ssol_image_setup(struct ssol_image* img, const size_t width, const size_t height,
                       const enum ssol_pixel_format fmt) {
  if (!img || !width || !height || (unsigned)fmt >= (unsigned)SSOL_PIXEL_FORMATS_COUNT__)
    return RES_BAD_ARG;
  const size_t pitch = width * ssol_sizeof_pixel_format(fmt);
  img->mem.assign(pitch * height, 0);
  img->pitch = pitch;
  img->width = width;
  img->height = height;
  img->format = fmt;
  return RES_OK;
}

res_T ssol_image_get_layout(const struct ssol_image* img, struct ssol_image_layout* layout) {
  if (!img || !layout) return RES_BAD_ARG;
  layout->row_pitch = img->pitch;
  layout->offset = 0;
  layout->size = img->width * img->height;
  layout->width = img->width;
  layout->height = img->height;
  layout->pixel_format = img->format;
  return RES_OK;
}

res_T ssol_image_map(const struct ssol_image* img, char** mem) {
  if (!img || !mem) return RES_BAD_ARG;
  *mem = img->mem.empty() ? nullptr : const_cast<char*>(img->mem.data());
  return RES_OK;
}

res_T ssol_image_unmap(const struct ssol_image* img) { return img ? RES_OK : RES_BAD_ARG; }

res_T ssol_image_sample(const struct ssol_image* img, const enum ssol_filter_mode filter,
                        const enum ssol_address_mode au, const enum ssol_address_mode av,
                        const double uv[2], void* val) {
  if (!img || !uv || !val || img->format != SSOL_PIXEL_DOUBLE3 || img->mem.empty())
    return RES_BAD_ARG;
  double* out = (double*)val;
  const double x0 = map_address(uv[0], au) * (double)img->width;
  const double y0 = map_address(uv[1], av) * (double)img->height;
  if (filter == SSOL_FILTER_NEAREST) {
    const double* z = texel(img, (size_t)x0, (size_t)y0);
    for (int k = 0; k < 3; ++k) out[k] = z[k];
    return RES_OK;
  }
  // bilinear between the texel and its +x, +y neighbours (addressed like uv)
  const double x1 = map_address(uv[0] + 1.0 / (double)img->width, au) * (double)img->width;
  const double y1 = map_address(uv[1] + 1.0 / (double)img->height, av) * (double)img->height;
  const double* z00 = texel(img, (size_t)x0, (size_t)y0);
  const double* z01 = texel(img, (size_t)x0, (size_t)y1);
  const double* z10 = texel(img, (size_t)x1, (size_t)y0);
  const double* z11 = texel(img, (size_t)x1, (size_t)y1);
  double ip;
  const double s = modf(x0, &ip), t = modf(y0, &ip);
  for (int k = 0; k < 3; ++k)
    out[k] = (1 - s) * ((1 - t) * z00[k] + t * z01[k]) + s * ((1 - t) * z10[k] + t * z11[k]);
  return RES_OK;
}

res_T ssol_image_write(void* image, const size_t origin[2], const size_t size[2],
                       const enum ssol_pixel_format fmt, const void* pixels) {
  ssol_image* img = (ssol_image*)image;
  if (!img || !origin || !size || !pixels) return RES_BAD_ARG;
  if (fmt != img->format || img->mem.empty()) return RES_BAD_ARG;
  if (origin[0] + size[0] > img->width || origin[1] + size[1] > img->height) return RES_BAD_ARG;
  const size_t bpp = ssol_sizeof_pixel_format(fmt), row = size[0] * bpp;
  const char* src = (const char*)pixels;
  for (size_t y = 0; y < size[1]; ++y, src += row)
    memcpy(img->mem.data() + (origin[1] + y) * img->pitch + origin[0] * bpp, src, row);
  return RES_OK;
}

// Draws -----------------------------------------------------------------------
res_T ssol_draw_draft(struct ssol_scene* scn, struct ssol_camera* cam, const size_t width,
                      const size_t height, const size_t spp, ssol_write_pixels_T writer,
                      void* data) {
  return draw(scn, cam, width, height, spp, 0, nullptr, writer, data, "ssol_draw_draft");
}

res_T ssol_draw_pt(struct ssol_scene* scn, struct ssol_camera* cam, const size_t width,
                   const size_t height, const size_t spp, const double up[3],
                   ssol_write_pixels_T writer, void* data) {
  if (!up) return RES_BAD_ARG;
  return draw(scn, cam, width, height, spp, 1, up, writer, data, "ssol_draw_pt");
}

}  // extern "C"
