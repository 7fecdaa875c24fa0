/* A C client of the drop-in API: it includes only include/ssol.h and
 * include/ssp.h and links -lssol, as a program written against the
 * reference's header would (src/ssol.h:1107-1223).
 *
 * Scene: one 10x10 m specular mirror (reflectivity from a parameter
 * buffer) tilted 45 degrees about Y, a 5x5 m absorbing receiver 50 m away on
 * the reflected axis, pillbox sun of half-angle 4.65 mrad at elevation 60
 * degrees, DNI 1000. Closed forms: cos factor cos 15 deg exactly; absorbed =
 * DNI rho A_receiver (the receiver lies inside the reflected beam's umbra);
 * other absorbed = (1 - rho) DNI 100 cos 15; missing = the rest.
 *
 * Checks, in order:
 *  1. the closed forms with rho = 0.9, and the receiver and per-primitive
 *     getters against the global estimate;
 *  2. rho rewritten in the parameter buffer (no setter call) between two
 *     solves: the second solve sees rho = 0.5;
 *  3. the estimator's RNG state written to a file with ssp_rng_write and
 *     read into a fresh generator with ssp_rng_read seeds a solve of the
 *     same realisations as a solve seeded from the borrowed state itself
 *     (src/ssol_estimator.c:368-418): equal estimates up to the order of
 *     the GPU's FP64 atomic additions.
 * Exit status 0 when every check holds; each failure prints one line. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ssol.h"
#include "ssp.h"

#define DNI 1000.0
static int nfail = 0;

#define CHECK(cond, ...) do { if(!(cond)) { ++nfail;                           \
  fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__);                         \
  fprintf(stderr, __VA_ARGS__); fputc('\n', stderr); } } while(0)
#define CALL(x) do { res_T r_ = (x); if(r_ != RES_OK) {                        \
  fprintf(stderr, "FAIL %s:%d: %s returned %d\n", __FILE__, __LINE__, #x, r_); \
  exit(2); } } while(0)

struct mirror_params { double reflectivity, roughness; };

static void get_ids(const unsigned itri, unsigned ids[3], void* ctx)
{
  static const unsigned tris[2][3] = { {0, 2, 1}, {2, 0, 3} };
  (void)ctx;
  ids[0] = tris[itri][0]; ids[1] = tris[itri][1]; ids[2] = tris[itri][2];
}

static void get_pos(const unsigned ivert, float pos[], void* ctx)
{
  const float h = *(const float*)ctx;
  const float xy[4][2] = { {-h, -h}, {h, -h}, {h, h}, {-h, h} };
  pos[0] = xy[ivert][0]; pos[1] = xy[ivert][1]; pos[2] = 0;
}

static void get_wl(const size_t i, double* wl, double* val, void* ctx)
{
  (void)i; (void)ctx;
  *wl = 500; *val = 1;
}

static void normal_ns(struct ssol_device* dev, struct ssol_param_buffer* buf,
  const double wl, const struct ssol_surface_fragment* f, double* val)
{
  (void)dev; (void)buf; (void)wl;
  val[0] = f->Ns[0]; val[1] = f->Ns[1]; val[2] = f->Ns[2];
}

static void mirror_reflectivity(struct ssol_device* dev, struct ssol_param_buffer* buf,
  const double wl, const struct ssol_surface_fragment* f, double* val)
{
  (void)dev; (void)wl; (void)f;
  *val = ((const struct mirror_params*)ssol_param_buffer_get(buf))->reflectivity;
}

static void mirror_roughness(struct ssol_device* dev, struct ssol_param_buffer* buf,
  const double wl, const struct ssol_surface_fragment* f, double* val)
{
  (void)dev; (void)wl; (void)f;
  *val = ((const struct mirror_params*)ssol_param_buffer_get(buf))->roughness;
}

static void zero(struct ssol_device* dev, struct ssol_param_buffer* buf,
  const double wl, const struct ssol_surface_fragment* f, double* val)
{
  (void)dev; (void)buf; (void)wl; (void)f;
  *val = 0;
}

/* 3x4 column-major rotation about Y by a, then translation t */
static void yaw(double a, const double t[3], double m[12])
{
  const double c = cos(a), s = sin(a);
  const double v[12] = { c, 0, -s,  0, 1, 0,  s, 0, c,  t[0], t[1], t[2] };
  memcpy(m, v, sizeof(v));
}

static int close_to(double a, double b, double tol) { return fabs(a - b) <= tol; }

static int same_result(const struct ssol_mc_result* a, const struct ssol_mc_result* b)
{
  return close_to(a->E, b->E, 1e-12 * fabs(b->E) + 1e-300)
      && close_to(a->V, b->V, 1e-9 * fabs(b->V) + 1e-9 * b->E * b->E)
      && close_to(a->SE, b->SE, 1e-9 * fabs(b->SE) + 1e-9 * fabs(b->E));
}

static int same_global(const struct ssol_mc_global* a, const struct ssol_mc_global* b)
{
  return same_result(&a->cos_factor, &b->cos_factor)
      && same_result(&a->absorbed_by_receivers, &b->absorbed_by_receivers)
      && same_result(&a->shadowed, &b->shadowed)
      && same_result(&a->missing, &b->missing)
      && same_result(&a->extinguished_by_atmosphere, &b->extinguished_by_atmosphere)
      && same_result(&a->other_absorbed, &b->other_absorbed);
}

static struct ssol_shape* square(struct ssol_device* dev, float* half)
{
  struct ssol_shape* shape;
  struct ssol_vertex_data attr = SSOL_VERTEX_DATA_NULL;
  attr.usage = SSOL_POSITION;
  attr.get = get_pos;
  CALL(ssol_shape_create_mesh(dev, &shape));
  CALL(ssol_mesh_setup(shape, 2, get_ids, 4, &attr, 1, half));
  return shape;
}

int main(int argc, char** argv)
{
  struct ssol_device* dev;
  struct ssol_scene* scn;
  struct ssol_spectrum* spectrum;
  struct ssol_sun* sun;
  struct ssol_shape *m_shape, *r_shape;
  struct ssol_material *m_mtl, *r_mtl;
  struct ssol_param_buffer* buf;
  struct mirror_params* prm;
  struct ssol_object *m_obj, *r_obj;
  struct ssol_instance *helio, *rcv;
  struct ssp_rng *rng, *copy;
  const struct ssp_rng* next;
  struct ssol_estimator *est, *est2, *e_borrowed, *e_read;
  struct ssol_mc_global g, g_borrowed, g_read;
  struct ssol_mc_receiver mr;
  struct ssol_mc_shape ms;
  struct ssol_mc_primitive mp;
  struct ssol_mirror_shader msh = SSOL_MIRROR_SHADER_NULL;
  struct ssol_matte_shader rsh = SSOL_MATTE_SHADER_NULL;
  const double el = 60 * M_PI / 180, c15 = cos(15 * M_PI / 180);
  const double d[3] = { -cos(el), 0, -sin(el) };
  const double n[3] = { sin(M_PI / 4), 0, cos(M_PI / 4) };
  double r[3], t[12], zero3[3] = { 0, 0, 0 }, prim_sum = 0, area;
  const size_t N = argc > 1 ? (size_t)strtoull(argv[1], NULL, 10) : 1000000;
  size_t count;
  float m_half = 5, r_half = 2.5f;
  FILE* f;
  unsigned i;
  int k;

  CALL(ssol_device_create(NULL, NULL, SSOL_NTHREADS_DEFAULT, 0, &dev));
  CALL(ssol_scene_create(dev, &scn));
  CALL(ssol_spectrum_create(dev, &spectrum));
  CALL(ssol_spectrum_setup(spectrum, get_wl, 1, NULL));
  CALL(ssol_sun_create_pillbox(dev, &sun));
  CALL(ssol_sun_set_direction(sun, d));
  CALL(ssol_sun_set_dni(sun, DNI));
  CALL(ssol_sun_set_spectrum(sun, spectrum));
  CALL(ssol_sun_pillbox_set_half_angle(sun, 4.65e-3));
  CALL(ssol_scene_attach_sun(scn, sun));

  /* mirror: reflectivity and roughness live in a parameter buffer */
  m_shape = square(dev, &m_half);
  CALL(ssol_param_buffer_create(dev, 64, &buf));
  prm = ssol_param_buffer_allocate(buf, sizeof(*prm), 16, NULL);
  CHECK(prm != NULL, "ssol_param_buffer_allocate");
  prm->reflectivity = 0.9;
  prm->roughness = 0;
  CALL(ssol_material_create_mirror(dev, &m_mtl));
  msh.normal = normal_ns;
  msh.reflectivity = mirror_reflectivity;
  msh.roughness = mirror_roughness;
  CALL(ssol_mirror_setup(m_mtl, &msh, SSOL_MICROFACET_BECKMANN));
  CALL(ssol_material_set_param_buffer(m_mtl, buf));
  CALL(ssol_object_create(dev, &m_obj));
  CALL(ssol_object_add_shaded_shape(m_obj, m_shape, m_mtl, m_mtl));
  CALL(ssol_object_instantiate(m_obj, &helio));
  yaw(M_PI / 4, zero3, t);
  CALL(ssol_instance_set_transform(helio, t));
  CALL(ssol_scene_attach_instance(scn, helio));

  /* receiver: absorbing, per-primitive tallies, not sampled */
  r_shape = square(dev, &r_half);
  CALL(ssol_material_create_matte(dev, &r_mtl));
  rsh.normal = normal_ns;
  rsh.reflectivity = zero;
  CALL(ssol_matte_setup(r_mtl, &rsh));
  CALL(ssol_object_create(dev, &r_obj));
  CALL(ssol_object_add_shaded_shape(r_obj, r_shape, r_mtl, r_mtl));
  CALL(ssol_object_instantiate(r_obj, &rcv));
  for(k = 0; k < 3; ++k) {
    const double dn = d[0]*n[0] + d[1]*n[1] + d[2]*n[2];
    r[k] = d[k] - 2 * dn * n[k];
  }
  {
    const double pos[3] = { 50 * r[0], 50 * r[1], 50 * r[2] };
    yaw(atan2(-r[0], -r[2]), pos, t);
  }
  CALL(ssol_instance_set_transform(rcv, t));
  CALL(ssol_instance_set_receiver(rcv, SSOL_FRONT, 1));
  CALL(ssol_instance_sample(rcv, 0));
  CALL(ssol_scene_attach_instance(scn, rcv));

  CALL(ssp_rng_create(NULL, SSP_RNG_THREEFRY, &rng));
  CALL(ssp_rng_set(rng, 0));

  /* 1. closed forms at rho = 0.9 */
  CALL(ssol_solve(scn, rng, N, 0, NULL, &est));
  CALL(ssol_estimator_get_mc_global(est, &g));
  CALL(ssol_estimator_get_realisation_count(est, &count));
  CHECK(count == N, "realisation count %zu", count);
  CHECK(close_to(g.cos_factor.E, c15, 1e-9) && g.cos_factor.SE < 1e-9,
    "cos factor %.12g +- %g, want %.12g", g.cos_factor.E, g.cos_factor.SE, c15);
  CHECK(g.shadowed.E == 0, "shadowed %g", g.shadowed.E);
  CHECK(close_to(g.absorbed_by_receivers.E, 22500,
    4 * g.absorbed_by_receivers.SE + 2e-5 * 22500),
    "absorbed %.3f +- %.3f, want 22500", g.absorbed_by_receivers.E, g.absorbed_by_receivers.SE);
  CHECK(close_to(g.other_absorbed.E, 0.1 * DNI * 100 * c15, 4 * g.other_absorbed.SE + 1e-6 * DNI),
    "other absorbed %.4f +- %.4f, want %.4f", g.other_absorbed.E, g.other_absorbed.SE,
    0.1 * DNI * 100 * c15);
  CHECK(close_to(g.missing.E, 0.9 * DNI * 100 * c15 - 22500,
    4 * g.missing.SE + 2e-5 * 22500), "missing %.3f", g.missing.E);
  CALL(ssol_estimator_get_mc_receiver(est, rcv, SSOL_FRONT, &mr));
  CHECK(close_to(mr.absorbed_flux.E, g.absorbed_by_receivers.E, 1e-9 * 22500),
    "receiver absorbed %.6f vs global %.6f", mr.absorbed_flux.E, g.absorbed_by_receivers.E);
  CHECK(close_to(mr.incoming_flux.E, mr.absorbed_flux.E, 1e-9 * 22500),
    "absorbing receiver: incoming %.6f vs absorbed %.6f", mr.incoming_flux.E, mr.absorbed_flux.E);
  CHECK(ssol_estimator_get_mc_receiver(est, rcv, SSOL_BACK, &mr) == RES_BAD_ARG,
    "a side outside the receiver mask is RES_BAD_ARG");
  CALL(ssol_estimator_get_mc_receiver(est, rcv, SSOL_FRONT, &mr));
  CALL(ssol_mc_receiver_get_mc_shape(&mr, r_shape, &ms));
  area = 5.0 * 5.0 / 2; /* each of the 2 triangles of the 5x5 m square */
  for(i = 0; i < 2; ++i) {
    CALL(ssol_mc_shape_get_mc_primitive(&ms, i, &mp));
    prim_sum += mp.absorbed_flux.E * area; /* W/m2 x m2 */
  }
  CHECK(close_to(prim_sum, g.absorbed_by_receivers.E, 1e-6 * 22500),
    "per-primitive flux x area %.6f vs absorbed %.6f", prim_sum, g.absorbed_by_receivers.E);
  CHECK(ssol_mc_shape_get_mc_primitive(&ms, 2, &mp) == RES_BAD_ARG,
    "primitive index past the shape is RES_BAD_ARG");

  /* 2. a parameter rewritten between solves is seen by the next solve */
  prm->reflectivity = 0.5;
  CALL(ssol_solve(scn, rng, N, 0, NULL, &est2));
  CALL(ssol_estimator_get_mc_global(est2, &g));
  CHECK(close_to(g.absorbed_by_receivers.E, 12500,
    4 * g.absorbed_by_receivers.SE + 2e-5 * 12500),
    "after rho 0.9 -> 0.5 in the parameter buffer: absorbed %.3f, want 12500",
    g.absorbed_by_receivers.E);
  CHECK(close_to(g.other_absorbed.E, 0.5 * DNI * 100 * c15, 4 * g.other_absorbed.SE + 1e-6 * DNI),
    "after rho 0.9 -> 0.5: other absorbed %.4f +- %.4f", g.other_absorbed.E, g.other_absorbed.SE);
  prm->reflectivity = 0.9;

  /* 3. the RNG state round trip through a file */
  CALL(ssol_estimator_get_rng_state(est, &next));
  f = tmpfile();
  CHECK(f != NULL, "tmpfile");
  CALL(ssp_rng_write(next, f));
  rewind(f);
  CALL(ssp_rng_create(NULL, SSP_RNG_THREEFRY, &copy));
  CALL(ssp_rng_read(copy, f));
  fclose(f);
  CALL(ssol_solve(scn, next, N / 4, 0, NULL, &e_borrowed));
  CALL(ssol_solve(scn, copy, N / 4, 0, NULL, &e_read));
  CALL(ssol_estimator_get_mc_global(e_borrowed, &g_borrowed));
  CALL(ssol_estimator_get_mc_global(e_read, &g_read));
  /* the same realisations: equal up to the order of the GPU's FP64 atomic
   * additions */
  CHECK(same_global(&g_borrowed, &g_read),
    "solve seeded from the state read back differs from the borrowed state's: %.17g vs %.17g",
    g_read.absorbed_by_receivers.E, g_borrowed.absorbed_by_receivers.E);
  CHECK(g_borrowed.absorbed_by_receivers.E != g.absorbed_by_receivers.E,
    "the chained solve repeated the first solve's streams");

  CALL(ssol_estimator_ref_put(e_read));
  CALL(ssol_estimator_ref_put(e_borrowed));
  CALL(ssol_estimator_ref_put(est2));
  CALL(ssol_estimator_ref_put(est));
  CALL(ssp_rng_ref_put(copy));
  CALL(ssp_rng_ref_put(rng));
  CALL(ssol_instance_ref_put(rcv));
  CALL(ssol_instance_ref_put(helio));
  CALL(ssol_object_ref_put(r_obj));
  CALL(ssol_object_ref_put(m_obj));
  CALL(ssol_material_ref_put(r_mtl));
  CALL(ssol_material_ref_put(m_mtl));
  CALL(ssol_param_buffer_ref_put(buf));
  CALL(ssol_shape_ref_put(r_shape));
  CALL(ssol_shape_ref_put(m_shape));
  CALL(ssol_sun_ref_put(sun));
  CALL(ssol_spectrum_ref_put(spectrum));
  CALL(ssol_scene_ref_put(scn));
  CALL(ssol_device_ref_put(dev));
  printf("%s: %d failed checks\n", nfail ? "FAILED" : "ok", nfail);
  return nfail ? 1 : 0;
}
