// Scene-building API: device, spectra, data, parameter buffers, suns,
// atmospheres, materials, objects, instances and scenes. Argument checks and
// ownership follow the reference (src/ssol_device.c, ssol_spectrum.c,
// ssol_data.c, ssol_param_buffer.c, ssol_sun.c, ssol_atmosphere.c,
// ssol_material.c:301-548, ssol_object.c, ssol_instance.c, ssol_scene.c:1-330).
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>

#include "ssol_internal.hpp"

namespace ssol {

void log_error(const ssol_device* dev, const char* fmt, ...) {
  if (dev && !dev->verbose) return;
  va_list ap;
  va_start(ap, fmt);
  fprintf(stderr, "ssol error: ");
  vfprintf(stderr, fmt, ap);
  va_end(ap);
}

void log_warning(const ssol_device* dev, const char* fmt, ...) {
  if (!dev || !dev->verbose) return;
  va_list ap;
  va_start(ap, fmt);
  fprintf(stderr, "ssol warning: ");
  vfprintf(stderr, fmt, ap);
  va_end(ap);
}

uint64_t fnv64(const void* data, size_t n) {
  const unsigned char* p = (const unsigned char*)data;
  uint64_t h = 0xcbf29ce484222325ull;
  for (size_t i = 0; i < n; ++i) {
    h ^= p[i];
    h *= 0x100000001b3ull;
  }
  return h;
}

uint64_t next_gen() {
  static uint64_t g = 0;
  return ++g;
}

void device_ref_put_(ssol_device* d) {
  if (--d->ref == 0) {
    solve_cache_free(d);
    if (d->gpu) ssolh_destroy(d->gpu);
    delete d;
  }
}

}  // namespace ssol

using namespace ssol;

namespace {

uint32_t g_next_uid = 1;

template <typename T>
T* new_obj(ssol_device* dev) {
  T* o = new T();
  o->dev = dev;
  ref_get(dev);
  return o;
}

void spectrum_put(ssol_spectrum* s) {
  if (--s->ref == 0) {
    ssol_device* d = s->dev;
    delete s;
    device_ref_put_(d);
  }
}

void medium_release(ssol_medium* m) {
  ssol_data_clear(&m->extinction);
  ssol_data_clear(&m->refractive_index);
}

bool spectrum_check(const ssol_spectrum* s, double lower, double upper) {
  if (!s || s->intensities.empty() || s->intensities.size() != s->wavelengths.size())
    return false;
  double cur = 0;
  for (size_t i = 0; i < s->wavelengths.size(); ++i) {
    const double wl = s->wavelengths[i], v = s->intensities[i];
    if (v < lower || v > upper || wl <= 0 || wl <= cur) return false;
    cur = wl;
  }
  return true;
}

bool check_medium(const ssol_medium* m) {
  if (!m) return false;
  if (m->extinction.type == SSOL_DATA_REAL) {
    if (m->extinction.value.real < 0) return false;
  } else if (!spectrum_check(m->extinction.value.spectrum, 0, 1.7976931348623157e308)) {
    return false;
  }
  if (m->refractive_index.type == SSOL_DATA_REAL) {
    if (m->refractive_index.value.real <= 0) return false;
  } else if (!spectrum_check(m->refractive_index.value.spectrum, 2.220446049250313e-16,
                             1.7976931348623157e308)) {
    return false;
  }
  return true;
}

}  // namespace

extern "C" {

// Device ----------------------------------------------------------------------
res_T ssol_device_create(struct logger*, struct mem_allocator*, const unsigned nthreads_hint,
                         const int verbose, struct ssol_device** out) {
  if (!out) return RES_BAD_ARG;
  ssol_device* d = new ssol_device();
  d->verbose = verbose;
  d->nthreads = nthreads_hint ? nthreads_hint : 1;
  *out = d;
  return RES_OK;
}

res_T ssol_device_ref_get(struct ssol_device* d) {
  if (!d) return RES_BAD_ARG;
  ref_get(d);
  return RES_OK;
}

res_T ssol_device_ref_put(struct ssol_device* d) {
  if (!d) return RES_BAD_ARG;
  device_ref_put_(d);
  return RES_OK;
}

res_T ssol_device_set_precision(struct ssol_device* d, const enum ssol_precision p) {
  if (!d || (p != SSOL_PRECISION_FP64 && p != SSOL_PRECISION_FP32)) return RES_BAD_ARG;
  d->precision = p;
  return RES_OK;
}

res_T ssol_device_set_wavefront_size(struct ssol_device* d, const size_t slots) {
  if (!d) return RES_BAD_ARG;
  d->path_slots = slots;
  return RES_OK;
}

res_T ssol_device_get_gpu_context(const struct ssol_device* d, void** ctx) {
  if (!d || !ctx) return RES_BAD_ARG;
  *ctx = d->gpu;
  return d->gpu ? RES_OK : RES_BAD_OP;
}

// Spectrum (src/ssol_spectrum.c) -----------------------------------------------
res_T ssol_spectrum_create(struct ssol_device* dev, struct ssol_spectrum** out) {
  if (!dev || !out) return RES_BAD_ARG;
  *out = new_obj<ssol_spectrum>(dev);
  return RES_OK;
}

res_T ssol_spectrum_ref_get(struct ssol_spectrum* s) {
  if (!s) return RES_BAD_ARG;
  ref_get(s);
  return RES_OK;
}

res_T ssol_spectrum_ref_put(struct ssol_spectrum* s) {
  if (!s) return RES_BAD_ARG;
  spectrum_put(s);
  return RES_OK;
}

res_T ssol_spectrum_setup(struct ssol_spectrum* s,
                          void (*get)(const size_t, double*, double*, void*),
                          const size_t n, void* ctx) {
  if (!s || !n || !get) return RES_BAD_ARG;
  s->wavelengths.assign(n, 0);
  s->intensities.assign(n, 0);
  double cur = 0;
  for (size_t i = 0; i < n; ++i) {
    get(i, &s->wavelengths[i], &s->intensities[i], ctx);
    if (s->wavelengths[i] <= cur || s->intensities[i] < 0) {
      s->wavelengths.clear();
      s->intensities.clear();
      s->checksum[0] = s->checksum[1] = 0;
      return RES_BAD_ARG;
    }
    cur = s->wavelengths[i];
  }
  s->gen = next_gen();
  s->checksum[0] = fnv64(s->wavelengths.data(), n * sizeof(double));
  s->checksum[1] = fnv64(s->intensities.data(), n * sizeof(double));
  return RES_OK;
}

// Data (src/ssol_data.c) ---------------------------------------------------------
struct ssol_data* ssol_data_set_real(struct ssol_data* d, const double real) {
  ssol_data_clear(d);
  d->type = SSOL_DATA_REAL;
  d->value.real = real;
  return d;
}

struct ssol_data* ssol_data_set_spectrum(struct ssol_data* d, struct ssol_spectrum* s) {
  if (d->type == SSOL_DATA_SPECTRUM && d->value.spectrum == s) return d;
  ssol_data_clear(d);
  d->type = SSOL_DATA_SPECTRUM;
  d->value.spectrum = s;
  ref_get(s);
  return d;
}

struct ssol_data* ssol_data_clear(struct ssol_data* d) {
  if (d->type != SSOL_DATA_SPECTRUM) return d;
  if (d->value.spectrum) spectrum_put(d->value.spectrum);
  *d = SSOL_DATA_NULL;
  return d;
}

struct ssol_data* ssol_data_copy(struct ssol_data* dst, const struct ssol_data* src) {
  if (dst == src) return dst;
  if (src->type == SSOL_DATA_SPECTRUM) {
    ssol_data_set_spectrum(dst, src->value.spectrum);
  } else {
    ssol_data_clear(dst);
    dst->type = SSOL_DATA_REAL;
    dst->value.real = src->value.real;
  }
  return dst;
}

double ssol_data_get_value(const struct ssol_data* d, const double w) {
  if (d->type == SSOL_DATA_REAL) return d->value.real;
  const ssol_spectrum* s = d->value.spectrum;
  const size_t n = s->wavelengths.size();
  const double* wl = s->wavelengths.data();
  const double* va = s->intensities.data();
  const size_t i = (size_t)(std::lower_bound(wl, wl + n, w) - wl);
  if (i == n) return va[n - 1];
  if (i == 0) return va[0];
  return va[i - 1] + (w - wl[i - 1]) * ((va[i] - va[i - 1]) / (wl[i] - wl[i - 1]));
}

// Param buffer (src/ssol_param_buffer.c) ----------------------------------------
res_T ssol_param_buffer_create(struct ssol_device* dev, const size_t capacity,
                               struct ssol_param_buffer** out) {
  if (!dev || !out || !capacity) return RES_BAD_ARG;
  ssol_param_buffer* b = new_obj<ssol_param_buffer>(dev);
  b->mem.assign(capacity + 64, 0);
  *out = b;
  return RES_OK;
}

res_T ssol_param_buffer_ref_get(struct ssol_param_buffer* b) {
  if (!b) return RES_BAD_ARG;
  ref_get(b);
  return RES_OK;
}

res_T ssol_param_buffer_clear(struct ssol_param_buffer* b) {
  if (!b) return RES_BAD_ARG;
  for (auto& r : b->releases) if (r.fn) r.fn(r.p);
  b->releases.clear();
  b->used = 0;
  return RES_OK;
}

res_T ssol_param_buffer_ref_put(struct ssol_param_buffer* b) {
  if (!b) return RES_BAD_ARG;
  if (--b->ref == 0) {
    ssol_param_buffer_clear(b);
    ssol_device* d = b->dev;
    delete b;
    device_ref_put_(d);
  }
  return RES_OK;
}

void* ssol_param_buffer_allocate(struct ssol_param_buffer* b, const size_t size,
                                 const size_t align, void (*release)(void*)) {
  if (!b || !size || !align || (align & (align - 1)) || align > 64) return nullptr;
  uintptr_t base = (uintptr_t)b->mem.data();
  uintptr_t p = (base + b->used + align - 1) & ~(uintptr_t)(align - 1);
  if (p + size > base + b->mem.size()) return nullptr;
  b->used = (size_t)(p + size - base);
  b->releases.push_back({(void*)p, release});
  return (void*)p;
}

void* ssol_param_buffer_get(struct ssol_param_buffer* b) {
  if (!b || b->releases.empty()) return nullptr;
  return b->releases.front().p;
}

// Sun (src/ssol_sun.c) ---------------------------------------------------------
static res_T sun_create(ssol_device* dev, ssol_sun** out, sun_kind k) {
  if (!dev || !out) return RES_BAD_ARG;
  ssol_sun* s = new_obj<ssol_sun>(dev);
  s->kind = k;
  *out = s;
  return RES_OK;
}

res_T ssol_sun_create_directional(struct ssol_device* d, struct ssol_sun** s) {
  return sun_create(d, s, sun_kind::directional);
}
res_T ssol_sun_create_pillbox(struct ssol_device* d, struct ssol_sun** s) {
  return sun_create(d, s, sun_kind::pillbox);
}
res_T ssol_sun_create_gaussian(struct ssol_device* d, struct ssol_sun** s) {
  return sun_create(d, s, sun_kind::gaussian);
}
res_T ssol_sun_create_buie(struct ssol_device* d, struct ssol_sun** s) {
  return sun_create(d, s, sun_kind::buie);
}

res_T ssol_sun_ref_get(struct ssol_sun* s) {
  if (!s) return RES_BAD_ARG;
  ref_get(s);
  return RES_OK;
}

res_T ssol_sun_ref_put(struct ssol_sun* s) {
  if (!s) return RES_BAD_ARG;
  if (--s->ref == 0) {
    if (s->spectrum) spectrum_put(s->spectrum);
    ssol_device* d = s->dev;
    delete s;
    device_ref_put_(d);
  }
  return RES_OK;
}

res_T ssol_sun_set_direction(struct ssol_sun* s, const double dir[3]) {
  if (s) s->gen = next_gen();
  if (!s || !dir) return RES_BAD_ARG;
  const double n = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
  if (n == 0) return RES_BAD_ARG;
  for (int i = 0; i < 3; ++i) s->direction[i] = dir[i] / n;
  return RES_OK;
}

res_T ssol_sun_get_direction(const struct ssol_sun* s, double dir[3]) {
  if (!s || !dir) return RES_BAD_ARG;
  for (int i = 0; i < 3; ++i) dir[i] = s->direction[i];
  return RES_OK;
}

res_T ssol_sun_set_dni(struct ssol_sun* s, const double dni) {
  if (s) s->gen = next_gen();
  if (!s || dni <= 0) return RES_BAD_ARG;
  s->dni = dni;
  return RES_OK;
}

res_T ssol_sun_get_dni(const struct ssol_sun* s, double* dni) {
  if (!s || !dni) return RES_BAD_ARG;
  *dni = s->dni;
  return RES_OK;
}

res_T ssol_sun_set_spectrum(struct ssol_sun* s, struct ssol_spectrum* sp) {
  if (s) s->gen = next_gen();
  if (!s || !sp) return RES_BAD_ARG;
  if (sp == s->spectrum) return RES_OK;
  if (s->spectrum) spectrum_put(s->spectrum);
  ref_get(sp);
  s->spectrum = sp;
  return RES_OK;
}

res_T ssol_sun_pillbox_set_half_angle(struct ssol_sun* s, const double a) {
  if (s) s->gen = next_gen();
  if (!s || a <= 0 || a > M_PI * 0.5 || s->kind != sun_kind::pillbox) return RES_BAD_ARG;
  s->param = a;
  return RES_OK;
}

res_T ssol_sun_gaussian_set_std_dev(struct ssol_sun* s, const double sd) {
  if (s) s->gen = next_gen();
  if (!s || sd <= 0 || s->kind != sun_kind::gaussian) return RES_BAD_ARG;
  s->param = sd;
  return RES_OK;
}

res_T ssol_sun_set_buie_param(struct ssol_sun* s, const double r) {
  if (s) s->gen = next_gen();
  if (!s || r <= 0 || r >= 1 || s->kind != sun_kind::buie) return RES_BAD_ARG;
  s->param = r;
  return RES_OK;
}

// Atmosphere (src/ssol_atmosphere.c) -------------------------------------------
res_T ssol_atmosphere_create(struct ssol_device* dev, struct ssol_atmosphere** out) {
  if (!dev || !out) return RES_BAD_ARG;
  *out = new_obj<ssol_atmosphere>(dev);
  return RES_OK;
}

res_T ssol_atmosphere_ref_get(struct ssol_atmosphere* a) {
  if (!a) return RES_BAD_ARG;
  ref_get(a);
  return RES_OK;
}

res_T ssol_atmosphere_ref_put(struct ssol_atmosphere* a) {
  if (!a) return RES_BAD_ARG;
  if (--a->ref == 0) {
    ssol_data_clear(&a->extinction);
    ssol_device* d = a->dev;
    delete a;
    device_ref_put_(d);
  }
  return RES_OK;
}

res_T ssol_atmosphere_set_extinction(struct ssol_atmosphere* a, struct ssol_data* ext) {
  if (!a || !ext) return RES_BAD_ARG;
  if (ext->type == SSOL_DATA_REAL) {
    if (ext->value.real < 0 || ext->value.real > 1) return RES_BAD_ARG;
  } else if (!spectrum_check(ext->value.spectrum, 0, 1)) {
    return RES_BAD_ARG;
  }
  ssol_data_copy(&a->extinction, ext);
  a->gen = next_gen();
  return RES_OK;
}

// Material (src/ssol_material.c:301-548) ---------------------------------------
static void default_normal(struct ssol_device*, struct ssol_param_buffer*, const double,
                           const struct ssol_surface_fragment* f, double* val) {
  val[0] = f->Ns[0];
  val[1] = f->Ns[1];
  val[2] = f->Ns[2];
}

static res_T material_create(ssol_device* dev, ssol_material** out, ssol_material_type t) {
  if (!dev || !out) return RES_BAD_ARG;
  ssol_material* m = new_obj<ssol_material>(dev);
  m->type = t;
  m->normal = default_normal;
  *out = m;
  return RES_OK;
}

res_T ssol_material_create_dielectric(struct ssol_device* d, struct ssol_material** m) {
  return material_create(d, m, SSOL_MATERIAL_DIELECTRIC);
}
res_T ssol_material_create_mirror(struct ssol_device* d, struct ssol_material** m) {
  return material_create(d, m, SSOL_MATERIAL_MIRROR);
}
res_T ssol_material_create_matte(struct ssol_device* d, struct ssol_material** m) {
  return material_create(d, m, SSOL_MATERIAL_MATTE);
}
res_T ssol_material_create_virtual(struct ssol_device* d, struct ssol_material** m) {
  return material_create(d, m, SSOL_MATERIAL_VIRTUAL);
}
res_T ssol_material_create_thin_dielectric(struct ssol_device* d, struct ssol_material** m) {
  return material_create(d, m, SSOL_MATERIAL_THIN_DIELECTRIC);
}

res_T ssol_material_get_type(const struct ssol_material* m, enum ssol_material_type* t) {
  if (!m || !t) return RES_BAD_ARG;
  *t = m->type;
  return RES_OK;
}

res_T ssol_material_ref_get(struct ssol_material* m) {
  if (!m) return RES_BAD_ARG;
  ref_get(m);
  return RES_OK;
}

res_T ssol_material_ref_put(struct ssol_material* m) {
  if (!m) return RES_BAD_ARG;
  if (--m->ref == 0) {
    if (m->buf) ssol_param_buffer_ref_put(m->buf);
    medium_release(&m->in_medium);
    medium_release(&m->out_medium);
    medium_release(&m->slab_medium);
    ssol_device* d = m->dev;
    delete m;
    device_ref_put_(d);
  }
  return RES_OK;
}

res_T ssol_material_set_param_buffer(struct ssol_material* m, struct ssol_param_buffer* b) {
  if (m) m->gen = next_gen();
  if (!m || !b) return RES_BAD_ARG;
  ref_get(b);
  if (m->buf) ssol_param_buffer_ref_put(m->buf);
  m->buf = b;
  return RES_OK;
}

res_T ssol_dielectric_setup(struct ssol_material* m, const struct ssol_dielectric_shader* sh,
                            const struct ssol_medium* out_med, const struct ssol_medium* in_med) {
  if (m) m->gen = next_gen();
  if (!m || m->type != SSOL_MATERIAL_DIELECTRIC || !sh || !sh->normal || !check_medium(out_med) ||
      !check_medium(in_med))
    return RES_BAD_ARG;
  ssol_medium_copy(&m->out_medium, out_med);
  ssol_medium_copy(&m->in_medium, in_med);
  m->normal = sh->normal;
  return RES_OK;
}

res_T ssol_mirror_setup(struct ssol_material* m, const struct ssol_mirror_shader* sh,
                        const enum ssol_microfacet_distribution distrib) {
  if (m) m->gen = next_gen();
  if (!m || m->type != SSOL_MATERIAL_MIRROR || !sh || !sh->normal || !sh->reflectivity ||
      !sh->roughness || (unsigned)distrib >= SSOL_MICROFACET_DISTRIBUTIONS_COUNT__)
    return RES_BAD_ARG;
  m->normal = sh->normal;
  m->reflectivity = sh->reflectivity;
  m->roughness = sh->roughness;
  m->distrib = distrib;
  return RES_OK;
}

res_T ssol_matte_setup(struct ssol_material* m, const struct ssol_matte_shader* sh) {
  if (m) m->gen = next_gen();
  if (!m || m->type != SSOL_MATERIAL_MATTE || !sh || !sh->normal || !sh->reflectivity)
    return RES_BAD_ARG;
  m->normal = sh->normal;
  m->reflectivity = sh->reflectivity;
  return RES_OK;
}

res_T ssol_thin_dielectric_setup(struct ssol_material* m,
                                 const struct ssol_thin_dielectric_shader* sh,
                                 const struct ssol_medium* out_med,
                                 const struct ssol_medium* slab, const double thickness) {
  if (m) m->gen = next_gen();
  if (!m || m->type != SSOL_MATERIAL_THIN_DIELECTRIC || !sh || !sh->normal ||
      !check_medium(out_med) || !check_medium(slab) || thickness < 0)
    return RES_BAD_ARG;
  ssol_medium_copy(&m->slab_medium, slab);
  m->thickness = thickness;
  ssol_medium_copy(&m->out_medium, out_med);
  ssol_medium_copy(&m->in_medium, out_med);
  m->normal = sh->normal;
  return RES_OK;
}

// Object (src/ssol_object.c) ---------------------------------------------------
res_T ssol_object_create(struct ssol_device* dev, struct ssol_object** out) {
  if (!dev || !out) return RES_BAD_ARG;
  *out = new_obj<ssol_object>(dev);
  return RES_OK;
}

res_T ssol_object_ref_get(struct ssol_object* o) {
  if (!o) return RES_BAD_ARG;
  ref_get(o);
  return RES_OK;
}

res_T ssol_object_clear(struct ssol_object* o) {
  if (!o) return RES_BAD_ARG;
  for (auto& s : o->sshapes) {
    ssol_shape_ref_put(s.shape);
    ssol_material_ref_put(s.front);
    ssol_material_ref_put(s.back);
  }
  o->sshapes.clear();
  o->rt_area = o->samp_area = 0;
  return RES_OK;
}

res_T ssol_object_ref_put(struct ssol_object* o) {
  if (!o) return RES_BAD_ARG;
  if (--o->ref == 0) {
    ssol_object_clear(o);
    ssol_device* d = o->dev;
    delete o;
    device_ref_put_(d);
  }
  return RES_OK;
}

res_T ssol_object_add_shaded_shape(struct ssol_object* o, struct ssol_shape* s,
                                   struct ssol_material* f, struct ssol_material* b) {
  if (!o || !s || !f || !b) return RES_BAD_ARG;
  for (auto& ss : o->sshapes) {
    if (ss.shape == s) {
      log_warning(o->dev, "ssol_object_add_shaded_shape: the object already own the shape.\n");
      return RES_OK;
    }
  }
  ref_get(s);
  ref_get(f);
  ref_get(b);
  o->sshapes.push_back({s, f, b});
  o->rt_area += s->rt_area;
  o->samp_area += s->samp_area;
  return RES_OK;
}

res_T ssol_object_get_area(const struct ssol_object* o, double* area) {
  if (!o || !area) return RES_BAD_ARG;
  *area = o->rt_area;
  return RES_OK;
}

// Instance (src/ssol_instance.c) -------------------------------------------------
res_T ssol_object_instantiate(struct ssol_object* o, struct ssol_instance** out) {
  if (!o || !out) return RES_BAD_ARG;
  ssol_instance* i = new_obj<ssol_instance>(o->dev);
  ref_get(o);
  i->object = o;
  i->id = g_next_uid++;
  *out = i;
  return RES_OK;
}

res_T ssol_instance_ref_get(struct ssol_instance* i) {
  if (!i) return RES_BAD_ARG;
  ref_get(i);
  return RES_OK;
}

res_T ssol_instance_ref_put(struct ssol_instance* i) {
  if (!i) return RES_BAD_ARG;
  if (--i->ref == 0) {
    ssol_object_ref_put(i->object);
    ssol_device* d = i->dev;
    delete i;
    device_ref_put_(d);
  }
  return RES_OK;
}

res_T ssol_instance_set_transform(struct ssol_instance* i, const double t[12]) {
  if (!i || !t) return RES_BAD_ARG;
  memcpy(i->transform, t, sizeof(i->transform));
  return RES_OK;
}

res_T ssol_instance_set_receiver(struct ssol_instance* i, const int mask, const int per_prim) {
  if (!i) return RES_BAD_ARG;
  i->receiver_mask = mask;
  i->per_primitive = per_prim;
  return RES_OK;
}

res_T ssol_instance_is_receiver(struct ssol_instance* i, int* mask, int* per_prim) {
  if (!i || !mask || !per_prim) return RES_BAD_ARG;
  *mask = i->receiver_mask;
  *per_prim = i->per_primitive;
  return RES_OK;
}

res_T ssol_instance_sample(struct ssol_instance* i, const int sample) {
  if (!i) return RES_BAD_ARG;
  i->sample = sample;
  return RES_OK;
}

res_T ssol_instance_get_id(const struct ssol_instance* i, uint32_t* id) {
  if (!i || !id) return RES_BAD_ARG;
  *id = i->id;
  return RES_OK;
}

res_T ssol_instance_get_area(const struct ssol_instance* i, double* area) {
  if (!i || !area) return RES_BAD_ARG;
  *area = i->object->rt_area;
  return RES_OK;
}

res_T ssol_instance_get_shaded_shapes_count(const struct ssol_instance* i, size_t* n) {
  if (!i || !n) return RES_BAD_ARG;
  *n = i->object->sshapes.size();
  return RES_OK;
}

static void invtrans3(double out[9], const double m[9]) {
  // inverse transpose of a column-major 3x3
  double inv[9];
  inv[0] = m[4] * m[8] - m[7] * m[5];
  inv[1] = m[7] * m[2] - m[1] * m[8];
  inv[2] = m[1] * m[5] - m[4] * m[2];
  inv[3] = m[6] * m[5] - m[3] * m[8];
  inv[4] = m[0] * m[8] - m[6] * m[2];
  inv[5] = m[3] * m[2] - m[0] * m[5];
  inv[6] = m[3] * m[7] - m[6] * m[4];
  inv[7] = m[6] * m[1] - m[0] * m[7];
  inv[8] = m[0] * m[4] - m[3] * m[1];
  const double det = m[0] * inv[0] + m[3] * inv[1] + m[6] * inv[2];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) out[c * 3 + r] = inv[r * 3 + c] / det;
}

res_T ssol_instance_get_shaded_shape(const struct ssol_instance* i, const size_t k,
                                     struct ssol_instantiated_shaded_shape* out) {
  if (!i || !out || k >= i->object->sshapes.size()) return RES_BAD_ARG;
  const shaded_shape& s = i->object->sshapes[k];
  out->shape = s.shape;
  out->mtl_front = s.front;
  out->mtl_back = s.back;
  memcpy(out->R__, i->transform, 9 * sizeof(double));
  memcpy(out->T__, i->transform + 9, 3 * sizeof(double));
  invtrans3(out->R_invtrans__, out->R__);
  return RES_OK;
}

res_T ssol_instantiated_shaded_shape_get_vertex_attrib(
    const struct ssol_instantiated_shaded_shape* s, const unsigned iv,
    const enum ssol_attrib_usage usage, double value[]) {
  if (!s || (unsigned)usage >= SSOL_ATTRIBS_COUNT__ || !value) return RES_BAD_ARG;
  res_T r = ssol_shape_get_vertex_attrib(s->shape, iv, usage, value);
  if (r != RES_OK) return r;
  if (usage == SSOL_NORMAL) {
    double v[3] = {value[0], value[1], value[2]};
    for (int k = 0; k < 3; ++k)
      value[k] = s->R_invtrans__[k] * v[0] + s->R_invtrans__[3 + k] * v[1] + s->R_invtrans__[6 + k] * v[2];
  } else if (usage == SSOL_POSITION) {
    double v[3] = {value[0], value[1], value[2]};
    for (int k = 0; k < 3; ++k)
      value[k] = s->R__[k] * v[0] + s->R__[3 + k] * v[1] + s->R__[6 + k] * v[2] + s->T__[k];
  }
  return RES_OK;
}

// Scene (src/ssol_scene.c:1-330) -----------------------------------------------
res_T ssol_scene_create(struct ssol_device* dev, struct ssol_scene** out) {
  if (!dev || !out) return RES_BAD_ARG;
  *out = new_obj<ssol_scene>(dev);
  return RES_OK;
}

res_T ssol_scene_ref_get(struct ssol_scene* s) {
  if (!s) return RES_BAD_ARG;
  ref_get(s);
  return RES_OK;
}

res_T ssol_scene_clear(struct ssol_scene* s) {
  if (!s) return RES_BAD_ARG;
  for (ssol_instance* i : s->instances) ssol_instance_ref_put(i);
  s->instances.clear();
  if (s->sun) ssol_scene_detach_sun(s, s->sun);
  if (s->atmosphere) ssol_scene_detach_atmosphere(s, s->atmosphere);
  return RES_OK;
}

res_T ssol_scene_ref_put(struct ssol_scene* s) {
  if (!s) return RES_BAD_ARG;
  if (--s->ref == 0) {
    ssol_scene_clear(s);
    ssol_device* d = s->dev;
    delete s;
    device_ref_put_(d);
  }
  return RES_OK;
}

res_T ssol_scene_attach_instance(struct ssol_scene* s, struct ssol_instance* i) {
  if (!s || !i) return RES_BAD_ARG;
  if (std::find(s->instances.begin(), s->instances.end(), i) != s->instances.end()) return RES_OK;
  ref_get(i);
  s->instances.push_back(i);
  return RES_OK;
}

res_T ssol_scene_detach_instance(struct ssol_scene* s, struct ssol_instance* i) {
  if (!s || !i) return RES_BAD_ARG;
  auto it = std::find(s->instances.begin(), s->instances.end(), i);
  if (it == s->instances.end()) return RES_BAD_ARG;
  s->instances.erase(it);
  ssol_instance_ref_put(i);
  return RES_OK;
}

res_T ssol_scene_compute_aabb(const struct ssol_scene* s, float lo[3], float hi[3]) {
  if (!s || !lo || !hi) return RES_BAD_ARG;
  for (int k = 0; k < 3; ++k) { lo[k] = 3.4e38f; hi[k] = -3.4e38f; }
  for (const ssol_instance* i : s->instances) {
    for (const shaded_shape& ss : i->object->sshapes) {
      const std::vector<float>& p = ss.shape->pos;
      for (size_t v = 0; v + 2 < p.size(); v += 3) {
        for (int k = 0; k < 3; ++k) {
          const double w = i->transform[k] * p[v] + i->transform[3 + k] * p[v + 1] +
                           i->transform[6 + k] * p[v + 2] + i->transform[9 + k];
          lo[k] = std::min(lo[k], (float)w);
          hi[k] = std::max(hi[k], (float)w);
        }
      }
    }
  }
  return RES_OK;
}

res_T ssol_scene_attach_sun(struct ssol_scene* s, struct ssol_sun* sun) {
  if (!s || !sun) return RES_BAD_ARG;
  if (sun->scene || s->sun) {
    if (sun->scene != s || s->sun != sun) return RES_BAD_ARG;
    return RES_OK;
  }
  ref_get(sun);
  s->sun = sun;
  sun->scene = s;
  return RES_OK;
}

res_T ssol_scene_detach_sun(struct ssol_scene* s, struct ssol_sun* sun) {
  if (!s || !sun || !s->sun || sun->scene != s) return RES_BAD_ARG;
  sun->scene = nullptr;
  s->sun = nullptr;
  ssol_sun_ref_put(sun);
  return RES_OK;
}

res_T ssol_scene_attach_atmosphere(struct ssol_scene* s, struct ssol_atmosphere* a) {
  if (!s || !a) return RES_BAD_ARG;
  if (a->scene || s->atmosphere) {
    if (a->scene != s || s->atmosphere != a) return RES_BAD_ARG;
    return RES_OK;
  }
  ref_get(a);
  s->atmosphere = a;
  a->scene = s;
  return RES_OK;
}

res_T ssol_scene_detach_atmosphere(struct ssol_scene* s, struct ssol_atmosphere* a) {
  if (!s || !a || !s->atmosphere || a->scene != s) return RES_BAD_ARG;
  a->scene = nullptr;
  s->atmosphere = nullptr;
  ssol_atmosphere_ref_put(a);
  return RES_OK;
}

res_T ssol_scene_for_each_instance(struct ssol_scene* s,
                                   res_T (*func)(struct ssol_instance*, void*), void* ctx) {
  if (!s || !func) return RES_BAD_ARG;
  std::vector<ssol_instance*> copy = s->instances;
  for (ssol_instance* i : copy) {
    res_T r = func(i, ctx);
    if (r != RES_OK) return r;
  }
  return RES_OK;
}

// Shape basics (src/ssol_shape.c:1527-1609, 1701-1769) -------------------------
static res_T shape_create(ssol_device* dev, ssol_shape** out, shape_kind k) {
  if (!dev || !out) return RES_BAD_ARG;
  ssol_shape* s = new_obj<ssol_shape>(dev);
  s->kind = k;
  s->uid = g_next_uid++;
  *out = s;
  return RES_OK;
}

res_T ssol_shape_create_mesh(struct ssol_device* d, struct ssol_shape** s) {
  return shape_create(d, s, shape_kind::mesh);
}

res_T ssol_shape_create_punched_surface(struct ssol_device* d, struct ssol_shape** s) {
  return shape_create(d, s, shape_kind::punched);
}

res_T ssol_shape_ref_get(struct ssol_shape* s) {
  if (!s) return RES_BAD_ARG;
  ref_get(s);
  return RES_OK;
}

res_T ssol_shape_ref_put(struct ssol_shape* s) {
  if (!s) return RES_BAD_ARG;
  if (--s->ref == 0) {
    ssol_device* d = s->dev;
    delete s;
    device_ref_put_(d);
  }
  return RES_OK;
}

res_T ssol_shape_get_vertices_count(const struct ssol_shape* s, unsigned* n) {
  if (!s || !n) return RES_BAD_ARG;
  *n = (unsigned)(s->pos.size() / 3);
  return RES_OK;
}

res_T ssol_shape_get_vertex_attrib(const struct ssol_shape* s, const unsigned iv,
                                   const enum ssol_attrib_usage usage, double value[]) {
  if (!s || (unsigned)usage >= SSOL_ATTRIBS_COUNT__ || !value) return RES_BAD_ARG;
  const size_t nv = s->pos.size() / 3;
  if (iv >= nv) return RES_BAD_ARG;
  value[0] = value[1] = value[2] = 1;
  if (usage == SSOL_POSITION) {
    for (int k = 0; k < 3; ++k) value[k] = s->pos[3 * iv + k];
  } else if (usage == SSOL_NORMAL) {
    if (s->normals.empty()) return RES_BAD_ARG;
    for (int k = 0; k < 3; ++k) value[k] = s->normals[3 * iv + k];
  } else {
    if (s->uvs.empty()) return RES_BAD_ARG;
    for (int k = 0; k < 2; ++k) value[k] = s->uvs[2 * iv + k];
  }
  // punched-surface vertices are stored in object space already (the
  // reference stores quadric-space data and applies the quadric transform
  // in this getter, src/ssol_shape.c:1580-1590)
  return RES_OK;
}

res_T ssol_shape_get_triangles_count(const struct ssol_shape* s, unsigned* n) {
  if (!s || !n) return RES_BAD_ARG;
  *n = (unsigned)(s->ids.size() / 3);
  return RES_OK;
}

res_T ssol_shape_get_triangle_indices(const struct ssol_shape* s, const unsigned it,
                                      unsigned ids[3]) {
  if (!s || !ids || it >= s->ids.size() / 3) return RES_BAD_ARG;
  for (int k = 0; k < 3; ++k) ids[k] = s->ids[3 * it + k];
  return RES_OK;
}

static double tri_area(const float* a, const float* b, const float* c) {
  const double e0[3] = {(double)b[0] - a[0], (double)b[1] - a[1], (double)b[2] - a[2]};
  const double e1[3] = {(double)c[0] - a[0], (double)c[1] - a[1], (double)c[2] - a[2]};
  const double n[3] = {e0[1] * e1[2] - e0[2] * e1[1], e0[2] * e1[0] - e0[0] * e1[2],
                       e0[0] * e1[1] - e0[1] * e1[0]};
  return std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
}

res_T ssol_mesh_setup(struct ssol_shape* s, const unsigned ntris,
                      void (*get_indices)(const unsigned, unsigned[3], void*),
                      const unsigned nverts, const struct ssol_vertex_data attribs[],
                      const unsigned nattribs, void* data) {
  if (s) s->gen = next_gen();
  if (!s || s->kind != shape_kind::mesh || !get_indices || !ntris || !nverts || !attribs ||
      !nattribs)
    return RES_BAD_ARG;
  if (nattribs > SSOL_ATTRIBS_COUNT__) return RES_MEM_ERR;
  const ssol_vertex_data* pos_attr = nullptr;
  const ssol_vertex_data* nrm_attr = nullptr;
  const ssol_vertex_data* uv_attr = nullptr;
  for (unsigned i = 0; i < nattribs; ++i) {
    if (!attribs[i].get) return RES_BAD_ARG;
    switch (attribs[i].usage) {
      case SSOL_POSITION: pos_attr = attribs + i; break;
      case SSOL_NORMAL: nrm_attr = attribs + i; break;
      case SSOL_TEXCOORD: uv_attr = attribs + i; break;
      default: return RES_BAD_ARG;
    }
  }
  if (!pos_attr) return RES_BAD_ARG;
  std::vector<float> pos(3 * (size_t)nverts), nrm, uv;
  std::vector<unsigned> ids(3 * (size_t)ntris);
  for (unsigned t = 0; t < ntris; ++t) {
    get_indices(t, &ids[3 * t], data);
    for (int k = 0; k < 3; ++k)
      if (ids[3 * t + k] >= nverts) return RES_BAD_ARG;
  }
  for (unsigned v = 0; v < nverts; ++v) pos_attr->get(v, &pos[3 * v], data);
  if (nrm_attr) {
    nrm.resize(3 * (size_t)nverts);
    for (unsigned v = 0; v < nverts; ++v) nrm_attr->get(v, &nrm[3 * v], data);
  }
  if (uv_attr) {
    uv.resize(2 * (size_t)nverts);
    for (unsigned v = 0; v < nverts; ++v) uv_attr->get(v, &uv[2 * v], data);
  }
  double area = 0;
  for (unsigned t = 0; t < ntris; ++t)
    area += tri_area(&pos[3 * ids[3 * t]], &pos[3 * ids[3 * t + 1]], &pos[3 * ids[3 * t + 2]]);
  s->pos.swap(pos);
  s->ids.swap(ids);
  s->normals.swap(nrm);
  s->uvs.swap(uv);
  s->samp_pos.clear();
  s->rt_area = s->samp_area = area * 0.5;
  s->ready = true;
  return RES_OK;
}

res_T ssol_punched_surface_setup(struct ssol_shape* s, const struct ssol_punched_surface* ps) {
  if (s) s->gen = next_gen();
  if (!s || s->kind != shape_kind::punched) return RES_BAD_ARG;
  return punched_setup(s, ps);
}

}  // extern "C"
