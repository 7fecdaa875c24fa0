/* Solstice flux integrator -- public C API of the MI355X build.
 *
 * This header is the drop-in surface. Every type, enumerator and entry point
 * keeps the name, argument order and meaning of the reference library's
 * public header (src/ssol.h, solstice-solver 0.9). A client written against
 * the reference compiles against this header unchanged for everything on the
 * flux-integration path: device, scene, shapes, materials, objects,
 * instances, sun, atmosphere, spectra, ssol_solve and the estimator getters.
 *
 * Differences, all deliberate:
 *  - The rendering API (camera, image, ssol_draw_draft, ssol_draw_pt) runs
 *    on the solver's GPU kernels; like ssol_solve it needs the scene's sun
 *    (its spectrum selects the materials' baked values).
 *  - `struct logger` and `struct mem_allocator` stay opaque. They are
 *    accepted by ssol_device_create and ignored: the library allocates with
 *    the C allocator and logs to stderr when `verbose` is set.
 *  - Material shaders are host callbacks. ssol_solve evaluates ("bakes") them
 *    once per material on probe fragments and requires them to be constant
 *    in wavelength and position; the normal shader must return the
 *    fragment's shading or geometry normal. A shader that breaks this makes
 *    ssol_solve fail with RES_BAD_ARG and a log line.
 *  - res_T codes follow RSys 0.10 (RES_OK = 0, then BAD_ARG, MEM_ERR,
 *    IO_ERR, UNKNOWN_ERR, BAD_OP, EOF).
 *
 * ssol_solve runs the realisations on the GPU through the C-ABI declared in
 * ssolh.h; there is no CPU fallback. */
#ifndef SSOL_H
#define SSOL_H

#include <stddef.h>
#include <stdint.h>

#if defined(SSOL_SHARED_BUILD)
  #define SSOL_API extern __attribute__((visibility("default")))
#else
  #define SSOL_API extern
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* RSys result codes (reference: rsys/rsys.h, enum res_type) */
#ifndef SSOL_RES_T_DEFINED
#define SSOL_RES_T_DEFINED
typedef int res_T;
enum ssol_res_code {
  RES_OK = 0,
  RES_BAD_ARG,
  RES_MEM_ERR,
  RES_IO_ERR,
  RES_UNKNOWN_ERR,
  RES_BAD_OP,
  RES_EOF
};
#endif

#define SSOL_NTHREADS_DEFAULT (~0u)

struct logger;
struct mem_allocator;
struct ssp_rng;

struct ssol_atmosphere;
struct ssol_camera;
struct ssol_image;
struct ssol_device;
struct ssol_material;
struct ssol_object;
struct ssol_instance;
struct ssol_param_buffer;
struct ssol_scene;
struct ssol_shape;
struct ssol_spectrum;
struct ssol_sun;
struct ssol_estimator;

enum ssol_side_flag {
  SSOL_FRONT = 1 << 0,
  SSOL_BACK = 1 << 1,
  SSOL_INVALID_SIDE = 1 << 2
};

enum ssol_path_type {
  SSOL_PATH_MISSING,
  SSOL_PATH_SHADOW,
  SSOL_PATH_SUCCESS,
  SSOL_PATH_ERROR
};

enum ssol_material_type {
  SSOL_MATERIAL_DIELECTRIC,
  SSOL_MATERIAL_MATTE,
  SSOL_MATERIAL_MIRROR,
  SSOL_MATERIAL_THIN_DIELECTRIC,
  SSOL_MATERIAL_VIRTUAL,
  SSOL_MATERIAL_TYPES_COUNT__
};

enum ssol_microfacet_distribution {
  SSOL_MICROFACET_BECKMANN,
  SSOL_MICROFACET_PILLBOX,
  SSOL_MICROFACET_DISTRIBUTIONS_COUNT__
};

enum ssol_clipping_op {
  SSOL_AND,
  SSOL_SUB,
  SSOL_CLIPPING_OPS_COUNT__
};

enum ssol_quadric_type {
  SSOL_QUADRIC_PLANE,
  SSOL_QUADRIC_PARABOL,
  SSOL_QUADRIC_PARABOL2F,
  SSOL_QUADRIC_HYPERBOL,
  SSOL_QUADRIC_PARABOLIC_CYLINDER,
  SSOL_QUADRIC_HEMISPHERE,
  SSOL_QUADRIC_TYPE_COUNT__
};

enum ssol_attrib_usage {
  SSOL_POSITION,
  SSOL_NORMAL,
  SSOL_TEXCOORD,
  SSOL_ATTRIBS_COUNT__
};

enum ssol_data_type {
  SSOL_DATA_REAL,
  SSOL_DATA_SPECTRUM
};

struct ssol_vertex_data {
  enum ssol_attrib_usage usage;
  void (*get)(const unsigned ivert, float value[], void* ctx);
};
#define SSOL_VERTEX_DATA_NULL__ { SSOL_ATTRIBS_COUNT__, NULL }
static const struct ssol_vertex_data SSOL_VERTEX_DATA_NULL =
  SSOL_VERTEX_DATA_NULL__;

/* Quadrics, all in their local frame */
struct ssol_quadric_plane { char dummy; };               /* z = 0 */
#define SSOL_QUADRIC_PLANE_DEFAULT__ { 0 }
static const struct ssol_quadric_plane SSOL_QUADRIC_PLANE_DEFAULT =
  SSOL_QUADRIC_PLANE_DEFAULT__;

struct ssol_quadric_parabol { double focal; };            /* x2+y2 = 4 f z */
#define SSOL_QUADRIC_PARABOL_NULL__ { -1.0 }
static const struct ssol_quadric_parabol SSOL_QUADRIC_PARABOL_NULL =
  SSOL_QUADRIC_PARABOL_NULL__;

struct ssol_quadric_parabol2f { double focal_x, focal_y; }; /* x2/4fx+y2/4fy = z */
#define SSOL_QUADRIC_PARABOL2F_NULL__ { -1.0, -1.0 }
static const struct ssol_quadric_parabol2f SSOL_QUADRIC_PARABOL2F_NULL =
  SSOL_QUADRIC_PARABOL2F_NULL__;

/* (x2+y2)/a2 - (z-1/2)2/b2 + 1 = 0, z > 0, with a2 = f - f2, b = f - 1/2 and
 * f = real_focal / (img_focal + real_focal), in units of the focal sum */
struct ssol_quadric_hyperbol { double img_focal, real_focal; };
#define SSOL_QUADRIC_HYPERBOL_NULL__ { -1.0, -1.0 }
static const struct ssol_quadric_hyperbol SSOL_QUADRIC_HYPERBOL_NULL =
  SSOL_QUADRIC_HYPERBOL_NULL__;

struct ssol_quadric_parabolic_cylinder { double focal; };  /* y2 = 4 f z */
#define SSOL_QUADRIC_PARABOLIC_CYLINDER_NULL__ { -1.0 }
static const struct ssol_quadric_parabolic_cylinder
SSOL_QUADRIC_PARABOLIC_CYLINDER_NULL = SSOL_QUADRIC_PARABOLIC_CYLINDER_NULL__;

struct ssol_quadric_hemisphere { double radius; };  /* x2+y2+(z-r)2 = r2, z<=r */
#define SSOL_QUADRIC_HEMISPHERE_NULL__ { -1.0 }
static const struct ssol_quadric_hemisphere SSOL_QUADRIC_HEMISPHERE_NULL =
  SSOL_QUADRIC_HEMISPHERE_NULL__;

struct ssol_quadric {
  enum ssol_quadric_type type;
  union {
    struct ssol_quadric_plane plane;
    struct ssol_quadric_parabol parabol;
    struct ssol_quadric_parabol2f parabol2f;
    struct ssol_quadric_hyperbol hyperbol;
    struct ssol_quadric_parabolic_cylinder parabolic_cylinder;
    struct ssol_quadric_hemisphere hemisphere;
  } data;
  double transform[12]; /* 3x4 column major, quadric -> object space */
  size_t slices_count_hint; /* SIZE_MAX <=> default discretisation */
};
#define SSOL_QUADRIC_DEFAULT__ {                                               \
  SSOL_QUADRIC_PLANE, {SSOL_QUADRIC_PLANE_DEFAULT__},                          \
  {1,0,0, 0,1,0, 0,0,1, 0,0,0}, SIZE_MAX }
static const struct ssol_quadric SSOL_QUADRIC_DEFAULT = SSOL_QUADRIC_DEFAULT__;

struct ssol_carving {
  void (*get)(const size_t ivert, double position[2], void* ctx);
  size_t nb_vertices;
  enum ssol_clipping_op operation;
  void* context;
};
#define SSOL_CARVING_NULL__ { NULL, 0, SSOL_CLIPPING_OPS_COUNT__, NULL }
static const struct ssol_carving SSOL_CARVING_NULL = SSOL_CARVING_NULL__;

struct ssol_punched_surface {
  struct ssol_quadric* quadric;
  struct ssol_carving* carvings;
  size_t nb_carvings;
};
#define SSOL_PUNCHED_SURFACE_NULL__ { NULL, NULL, 0 }
static const struct ssol_punched_surface SSOL_PUNCHED_SURFACE_NULL =
  SSOL_PUNCHED_SURFACE_NULL__;

struct ssol_data {
  enum ssol_data_type type;
  union {
    double real;
    struct ssol_spectrum* spectrum;
  } value;
};
#define SSOL_DATA_NULL__ {SSOL_DATA_REAL, {0.0}}
static const struct ssol_data SSOL_DATA_NULL = SSOL_DATA_NULL__;

struct ssol_medium {
  struct ssol_data extinction;
  struct ssol_data refractive_index;
};
#define SSOL_MEDIUM_VACUUM__ {{SSOL_DATA_REAL, {0}}, {SSOL_DATA_REAL, {1}}}
static const struct ssol_medium SSOL_MEDIUM_VACUUM = SSOL_MEDIUM_VACUUM__;

struct ssol_surface_fragment {
  double dir[3];  /* world incoming direction, pointing toward the surface */
  double P[3];    /* world position */
  double Ng[3];   /* normalised world geometry normal */
  double Ns[3];   /* normalised world shading normal */
  double uv[2];
  double dPdu[3];
  double dPdv[3];
};
#define SSOL_SURFACE_FRAGMENT_NULL__ \
  {{0,0,0}, {0,0,0}, {0,0,0}, {0,0,0}, {0,0}, {0,0,0}, {0,0,0}}
static const struct ssol_surface_fragment SSOL_SURFACE_FRAGMENT_NULL =
  SSOL_SURFACE_FRAGMENT_NULL__;

typedef void
(*ssol_shader_getter_T)
  (struct ssol_device* dev,
   struct ssol_param_buffer* buf,
   const double wavelength,
   const struct ssol_surface_fragment* fragment,
   double* val);

struct ssol_dielectric_shader { ssol_shader_getter_T normal; };
#define SSOL_DIELECTRIC_SHADER_NULL__ { NULL }
static const struct ssol_dielectric_shader SSOL_DIELECTRIC_SHADER_NULL =
  SSOL_DIELECTRIC_SHADER_NULL__;

struct ssol_mirror_shader {
  ssol_shader_getter_T normal;
  ssol_shader_getter_T reflectivity;
  ssol_shader_getter_T roughness;
};
#define SSOL_MIRROR_SHADER_NULL__ { NULL, NULL, NULL }
static const struct ssol_mirror_shader SSOL_MIRROR_SHADER_NULL =
  SSOL_MIRROR_SHADER_NULL__;

struct ssol_matte_shader {
  ssol_shader_getter_T normal;
  ssol_shader_getter_T reflectivity;
};
#define SSOL_MATTE_SHADER_NULL__ { NULL, NULL }
static const struct ssol_matte_shader SSOL_MATTE_SHADER_NULL =
  SSOL_MATTE_SHADER_NULL__;

struct ssol_thin_dielectric_shader { ssol_shader_getter_T normal; };
#define SSOL_THIN_DIELECTRIC_SHADER_NULL__ { NULL }
static const struct ssol_thin_dielectric_shader
SSOL_THIN_DIELECTRIC_SHADER_NULL = SSOL_THIN_DIELECTRIC_SHADER_NULL__;

struct ssol_instantiated_shaded_shape {
  struct ssol_shape* shape;
  struct ssol_material* mtl_front;
  struct ssol_material* mtl_back;
  double R__[9];
  double T__[3];
  double R_invtrans__[9];
};
#define SSOL_INSTANTIATED_SHADED_SHAPE_NULL__ { 0 }
static const struct ssol_instantiated_shaded_shape
SSOL_INSTANTIATED_SHADED_SHAPE_NULL = SSOL_INSTANTIATED_SHADED_SHAPE_NULL__;

struct ssol_path_tracker {
  double sun_ray_length;       /* < 0 <=> default */
  double infinite_ray_length;  /* < 0 <=> default */
};
#define SSOL_PATH_TRACKER_DEFAULT__ {-1, -1}
static const struct ssol_path_tracker SSOL_PATH_TRACKER_DEFAULT =
  SSOL_PATH_TRACKER_DEFAULT__;

struct ssol_path { const void* path__; };

struct ssol_path_vertex {
  double pos[3];
  double weight;
};

struct ssol_mc_result {
  double E;   /* expectation */
  double V;   /* variance */
  double SE;  /* standard error */
};
#define SSOL_MC_RESULT_NULL__ {0, 0, 0}
static const struct ssol_mc_result SSOL_MC_RESULT_NULL = SSOL_MC_RESULT_NULL__;

struct ssol_mc_global {
  struct ssol_mc_result cos_factor;
  struct ssol_mc_result absorbed_by_receivers;
  struct ssol_mc_result shadowed;
  struct ssol_mc_result missing;
  struct ssol_mc_result extinguished_by_atmosphere;
  struct ssol_mc_result other_absorbed;
};
#define SSOL_MC_GLOBAL_NULL__ { SSOL_MC_RESULT_NULL__, SSOL_MC_RESULT_NULL__, \
  SSOL_MC_RESULT_NULL__, SSOL_MC_RESULT_NULL__, SSOL_MC_RESULT_NULL__,        \
  SSOL_MC_RESULT_NULL__ }
static const struct ssol_mc_global SSOL_MC_GLOBAL_NULL = SSOL_MC_GLOBAL_NULL__;

struct ssol_mc_receiver {
  struct ssol_mc_result incoming_flux;
  struct ssol_mc_result incoming_if_no_atm_loss;
  struct ssol_mc_result incoming_if_no_field_loss;
  struct ssol_mc_result incoming_lost_in_field;
  struct ssol_mc_result incoming_lost_in_atmosphere;
  struct ssol_mc_result absorbed_flux;
  struct ssol_mc_result absorbed_if_no_atm_loss;
  struct ssol_mc_result absorbed_if_no_field_loss;
  struct ssol_mc_result absorbed_lost_in_field;
  struct ssol_mc_result absorbed_lost_in_atmosphere;
  /* Internal data */
  size_t N__;
  void* mc__;
  const struct ssol_instance* instance__;
};

struct ssol_mc_shape {
  size_t N__;
  void* mc__;
  const struct ssol_shape* shape__;
};
#define SSOL_MC_SHAPE_NULL__ { 0, NULL, NULL }
static const struct ssol_mc_shape SSOL_MC_SHAPE_NULL = SSOL_MC_SHAPE_NULL__;

struct ssol_mc_sampled {
  struct ssol_mc_result cos_factor;
  struct ssol_mc_result shadowed;
  size_t nb_samples;
};

struct ssol_mc_primitive {
  struct ssol_mc_result incoming_flux;
  struct ssol_mc_result incoming_if_no_atm_loss;
  struct ssol_mc_result incoming_if_no_field_loss;
  struct ssol_mc_result incoming_lost_in_field;
  struct ssol_mc_result incoming_lost_in_atmosphere;
  struct ssol_mc_result absorbed_flux;
  struct ssol_mc_result absorbed_if_no_atm_loss;
  struct ssol_mc_result absorbed_if_no_field_loss;
  struct ssol_mc_result absorbed_lost_in_field;
  struct ssol_mc_result absorbed_lost_in_atmosphere;
};

/* Device (reference src/ssol.h:548-559) */
SSOL_API res_T ssol_device_create
  (struct logger* logger, struct mem_allocator* allocator,
   const unsigned nthreads_hint, const int verbose, struct ssol_device** dev);
SSOL_API res_T ssol_device_ref_get(struct ssol_device* dev);
SSOL_API res_T ssol_device_ref_put(struct ssol_device* dev);

/* Scene (reference src/ssol.h:637-700) */
SSOL_API res_T ssol_scene_create(struct ssol_device* dev, struct ssol_scene** scn);
SSOL_API res_T ssol_scene_ref_get(struct ssol_scene* scn);
SSOL_API res_T ssol_scene_ref_put(struct ssol_scene* scn);
SSOL_API res_T ssol_scene_attach_instance
  (struct ssol_scene* scn, struct ssol_instance* instance);
SSOL_API res_T ssol_scene_detach_instance
  (struct ssol_scene* scn, struct ssol_instance* instance);
SSOL_API res_T ssol_scene_compute_aabb
  (const struct ssol_scene* scn, float lower[3], float upper[3]);
SSOL_API res_T ssol_scene_clear(struct ssol_scene* scn);
SSOL_API res_T ssol_scene_attach_sun(struct ssol_scene* scn, struct ssol_sun* sun);
SSOL_API res_T ssol_scene_detach_sun(struct ssol_scene* scn, struct ssol_sun* sun);
SSOL_API res_T ssol_scene_attach_atmosphere
  (struct ssol_scene* scn, struct ssol_atmosphere* atm);
SSOL_API res_T ssol_scene_detach_atmosphere
  (struct ssol_scene* scn, struct ssol_atmosphere* atm);
SSOL_API res_T ssol_scene_for_each_instance
  (struct ssol_scene* scn,
   res_T (*func)(struct ssol_instance* instance, void* ctx),
   void* ctx);

/* Shape (reference src/ssol.h:705-775) */
SSOL_API res_T ssol_shape_create_mesh(struct ssol_device* dev, struct ssol_shape** shape);
SSOL_API res_T ssol_shape_create_punched_surface
  (struct ssol_device* dev, struct ssol_shape** shape);
SSOL_API res_T ssol_shape_ref_get(struct ssol_shape* shape);
SSOL_API res_T ssol_shape_ref_put(struct ssol_shape* shape);
SSOL_API res_T ssol_shape_get_vertices_count
  (const struct ssol_shape* shape, unsigned* nverts);
SSOL_API res_T ssol_shape_get_vertex_attrib
  (const struct ssol_shape* shape, const unsigned ivert,
   const enum ssol_attrib_usage usage, double value[]);
SSOL_API res_T ssol_shape_get_triangles_count
  (const struct ssol_shape* shape, unsigned* ntris);
SSOL_API res_T ssol_shape_get_triangle_indices
  (const struct ssol_shape* shape, const unsigned itri, unsigned ids[3]);
SSOL_API res_T ssol_punched_surface_setup
  (struct ssol_shape* shape, const struct ssol_punched_surface* punched_surface);
SSOL_API res_T ssol_mesh_setup
  (struct ssol_shape* shape,
   const unsigned ntris,
   void (*get_indices)(const unsigned itri, unsigned ids[3], void* ctx),
   const unsigned nverts,
   const struct ssol_vertex_data attribs[],
   const unsigned nattribs,
   void* data);

/* Material (reference src/ssol.h:780-868) */
SSOL_API res_T ssol_material_create_dielectric
  (struct ssol_device* dev, struct ssol_material** mtl);
SSOL_API res_T ssol_material_create_mirror
  (struct ssol_device* dev, struct ssol_material** mtl);
SSOL_API res_T ssol_material_create_matte
  (struct ssol_device* dev, struct ssol_material** mtl);
SSOL_API res_T ssol_material_create_virtual
  (struct ssol_device* dev, struct ssol_material** mtl);
SSOL_API res_T ssol_material_create_thin_dielectric
  (struct ssol_device* dev, struct ssol_material** mtl);
SSOL_API res_T ssol_material_get_type
  (const struct ssol_material* mtl, enum ssol_material_type* type);
SSOL_API res_T ssol_material_ref_get(struct ssol_material* mtl);
SSOL_API res_T ssol_material_ref_put(struct ssol_material* mtl);
SSOL_API res_T ssol_material_set_param_buffer
  (struct ssol_material* mtl, struct ssol_param_buffer* buf);
SSOL_API res_T ssol_dielectric_setup
  (struct ssol_material* mtl, const struct ssol_dielectric_shader* shader,
   const struct ssol_medium* outside_medium,
   const struct ssol_medium* inside_medium);
SSOL_API res_T ssol_mirror_setup
  (struct ssol_material* mtl, const struct ssol_mirror_shader* shader,
   const enum ssol_microfacet_distribution distrib);
SSOL_API res_T ssol_matte_setup
  (struct ssol_material* mtl, const struct ssol_matte_shader* shader);
SSOL_API res_T ssol_thin_dielectric_setup
  (struct ssol_material* mtl, const struct ssol_thin_dielectric_shader* shader,
   const struct ssol_medium* outside_medium,
   const struct ssol_medium* slab_medium,
   const double thickness);

/* Object (reference src/ssol.h:873-902) */
SSOL_API res_T ssol_object_create(struct ssol_device* dev, struct ssol_object** obj);
SSOL_API res_T ssol_object_ref_get(struct ssol_object* obj);
SSOL_API res_T ssol_object_ref_put(struct ssol_object* obj);
SSOL_API res_T ssol_object_add_shaded_shape
  (struct ssol_object* object, struct ssol_shape* shape,
   struct ssol_material* mtl_front, struct ssol_material* mtl_back);
SSOL_API res_T ssol_object_clear(struct ssol_object* object);
SSOL_API res_T ssol_object_get_area(const struct ssol_object* object, double* area);

/* Instance (reference src/ssol.h:907-968) */
SSOL_API res_T ssol_object_instantiate
  (struct ssol_object* object, struct ssol_instance** instance);
SSOL_API res_T ssol_instance_ref_get(struct ssol_instance* instance);
SSOL_API res_T ssol_instance_ref_put(struct ssol_instance* intance);
SSOL_API res_T ssol_instance_set_transform
  (struct ssol_instance* instance, const double transform[12]);
SSOL_API res_T ssol_instance_set_receiver
  (struct ssol_instance* instance, const int mask, const int per_primitive);
SSOL_API res_T ssol_instance_is_receiver
  (struct ssol_instance* instance, int* mask, int* per_primitive);
SSOL_API res_T ssol_instance_sample(struct ssol_instance* instance, const int sample);
SSOL_API res_T ssol_instance_get_id(const struct ssol_instance* instance, uint32_t* id);
SSOL_API res_T ssol_instance_get_area(const struct ssol_instance* instance, double* area);
SSOL_API res_T ssol_instance_get_shaded_shapes_count
  (const struct ssol_instance* instance, size_t* nshaded_shapes);
SSOL_API res_T ssol_instance_get_shaded_shape
  (const struct ssol_instance* instance, const size_t ishaded_shape,
   struct ssol_instantiated_shaded_shape* shaded_shape_instance);
SSOL_API res_T ssol_instantiated_shaded_shape_get_vertex_attrib
  (const struct ssol_instantiated_shaded_shape* sshape, const unsigned ivert,
   const enum ssol_attrib_usage usage, double value[]);

/* Parameter buffer (reference src/ssol.h:973-1004) */
SSOL_API res_T ssol_param_buffer_create
  (struct ssol_device* dev, const size_t capacity, struct ssol_param_buffer** buf);
SSOL_API res_T ssol_param_buffer_ref_get(struct ssol_param_buffer* buf);
SSOL_API res_T ssol_param_buffer_ref_put(struct ssol_param_buffer* buf);
SSOL_API void* ssol_param_buffer_allocate
  (struct ssol_param_buffer* buf, const size_t size, const size_t alignment,
   void (*release)(void*));
SSOL_API void* ssol_param_buffer_get(struct ssol_param_buffer* buf);
SSOL_API res_T ssol_param_buffer_clear(struct ssol_param_buffer* buf);

/* Spectrum (reference src/ssol.h:1009-1026) */
SSOL_API res_T ssol_spectrum_create
  (struct ssol_device* dev, struct ssol_spectrum** spectrum);
SSOL_API res_T ssol_spectrum_ref_get(struct ssol_spectrum* spectrum);
SSOL_API res_T ssol_spectrum_ref_put(struct ssol_spectrum* spectrum);
SSOL_API res_T ssol_spectrum_setup
  (struct ssol_spectrum* spectrum,
   void (*get)(const size_t iwlen, double* wlen, double* data, void* ctx),
   const size_t nwlens,
   void* ctx);

/* Sun (reference src/ssol.h:1031-1086) */
SSOL_API res_T ssol_sun_create_directional(struct ssol_device* dev, struct ssol_sun** sun);
SSOL_API res_T ssol_sun_create_pillbox(struct ssol_device* dev, struct ssol_sun** sun);
SSOL_API res_T ssol_sun_create_gaussian(struct ssol_device* dev, struct ssol_sun** sun);
SSOL_API res_T ssol_sun_create_buie(struct ssol_device* dev, struct ssol_sun** sun);
SSOL_API res_T ssol_sun_ref_get(struct ssol_sun* sun);
SSOL_API res_T ssol_sun_ref_put(struct ssol_sun* sun);
SSOL_API res_T ssol_sun_set_direction(struct ssol_sun* sun, const double direction[3]);
SSOL_API res_T ssol_sun_get_direction(const struct ssol_sun* sun, double direction[3]);
SSOL_API res_T ssol_sun_set_dni(struct ssol_sun* sun, const double dni);
SSOL_API res_T ssol_sun_get_dni(const struct ssol_sun* sun, double* dni);
SSOL_API res_T ssol_sun_set_spectrum(struct ssol_sun* sun, struct ssol_spectrum* spectrum);
SSOL_API res_T ssol_sun_pillbox_set_half_angle(struct ssol_sun* sun, const double half_angle);
SSOL_API res_T ssol_sun_gaussian_set_std_dev(struct ssol_sun* sun, const double std_dev);
SSOL_API res_T ssol_sun_set_buie_param(struct ssol_sun* sun, const double param);

/* Atmosphere (reference src/ssol.h:1091-1102) */
SSOL_API res_T ssol_atmosphere_create
  (struct ssol_device* dev, struct ssol_atmosphere** atmosphere);
SSOL_API res_T ssol_atmosphere_ref_get(struct ssol_atmosphere* atmosphere);
SSOL_API res_T ssol_atmosphere_ref_put(struct ssol_atmosphere* atmosphere);
SSOL_API res_T ssol_atmosphere_set_extinction
  (struct ssol_atmosphere* atmosphere, struct ssol_data* extinction);

/* Estimator (reference src/ssol.h:1107-1211) */
SSOL_API res_T ssol_estimator_ref_get(struct ssol_estimator* estimator);
SSOL_API res_T ssol_estimator_ref_put(struct ssol_estimator* estimator);
SSOL_API res_T ssol_estimator_get_mc_global
  (struct ssol_estimator* estimator, struct ssol_mc_global* mc_global);
SSOL_API res_T ssol_estimator_get_mc_sampled_x_receiver
  (struct ssol_estimator* estimator,
   const struct ssol_instance* prim_instance,
   const struct ssol_instance* recv_instance,
   const enum ssol_side_flag side,
   struct ssol_mc_receiver* rcv);
SSOL_API res_T ssol_estimator_get_realisation_count
  (const struct ssol_estimator* estimator, size_t* count);
SSOL_API res_T ssol_estimator_get_failed_count
  (const struct ssol_estimator* estimator, size_t* count);
SSOL_API res_T ssol_estimator_get_sampled_area
  (const struct ssol_estimator* estimator, double* area);
SSOL_API res_T ssol_estimator_get_sampled_count
  (const struct ssol_estimator* estimator, size_t* count);
SSOL_API res_T ssol_estimator_get_mc_sampled
  (struct ssol_estimator* estimator, const struct ssol_instance* samp_instance,
   struct ssol_mc_sampled* sampled);
SSOL_API res_T ssol_estimator_get_rng_state
  (const struct ssol_estimator* estimator, const struct ssp_rng** rng_state);
SSOL_API res_T ssol_estimator_get_tracked_paths_count
  (const struct ssol_estimator* estimator, size_t* npaths);
SSOL_API res_T ssol_estimator_get_tracked_path
  (const struct ssol_estimator* estimator, const size_t ipath,
   struct ssol_path* path);
SSOL_API res_T ssol_path_get_vertices_count
  (const struct ssol_path* path, size_t* nvertices);
SSOL_API res_T ssol_path_get_vertex
  (const struct ssol_path* path, const size_t ivertex,
   struct ssol_path_vertex* vertex);
SSOL_API res_T ssol_path_get_type
  (const struct ssol_path* path, enum ssol_path_type* type);
SSOL_API res_T ssol_estimator_get_mc_receiver
  (struct ssol_estimator* estimator, const struct ssol_instance* instance,
   const enum ssol_side_flag side, struct ssol_mc_receiver* rcv);
SSOL_API res_T ssol_mc_receiver_get_mc_shape
  (struct ssol_mc_receiver* rcv, const struct ssol_shape* shape,
   struct ssol_mc_shape* mc);
SSOL_API res_T ssol_mc_shape_get_mc_primitive
  (struct ssol_mc_shape* shape, const unsigned i,
   struct ssol_mc_primitive* prim);

/* Solver (reference src/ssol.h:1216-1223, src/ssol_solver.c:613-842) */
SSOL_API res_T ssol_solve
  (struct ssol_scene* scn,
   const struct ssp_rng* rng,
   const size_t realisations_count,
   const size_t max_failed_count,
   const struct ssol_path_tracker* tracker, /* NULL <=> no tracked paths */
   struct ssol_estimator** estimator);

/* Data (reference src/ssol.h:1248-1271) */
SSOL_API struct ssol_data* ssol_data_set_real(struct ssol_data* data, const double real);
SSOL_API struct ssol_data* ssol_data_set_spectrum
  (struct ssol_data* data, struct ssol_spectrum* spectrum);
SSOL_API struct ssol_data* ssol_data_clear(struct ssol_data* data);
SSOL_API struct ssol_data* ssol_data_copy(struct ssol_data* dst, const struct ssol_data* src);
SSOL_API double ssol_data_get_value(const struct ssol_data* data, const double wavelength);

static inline struct ssol_medium*
ssol_medium_clear(struct ssol_medium* medium)
{
  ssol_data_clear(&medium->extinction);
  ssol_data_clear(&medium->refractive_index);
  return medium;
}

static inline struct ssol_medium*
ssol_medium_copy(struct ssol_medium* dst, const struct ssol_medium* src)
{
  ssol_data_copy(&dst->extinction, &src->extinction);
  ssol_data_copy(&dst->refractive_index, &src->refractive_index);
  return dst;
}

/*******************************************************************************
 * Extensions of this build (no counterpart in the reference header)
 ******************************************************************************/
/* Arithmetic of the path state on the GPU. FP64 is the default and follows
 * the reference's double-precision weights. FP32 keeps ray/BSDF/sun
 * arithmetic in single precision with per-instance origin rebasing; the
 * quadric root solve and the tallies stay FP64 in both forms. */
enum ssol_precision {
  SSOL_PRECISION_FP64 = 0,
  SSOL_PRECISION_FP32 = 1
};
SSOL_API res_T ssol_device_set_precision
  (struct ssol_device* dev, const enum ssol_precision precision);

/* Number of in-flight path slots of the wavefront (0 <=> automatic). */
SSOL_API res_T ssol_device_set_wavefront_size
  (struct ssol_device* dev, const size_t path_slots);

/* The device's GPU context (an ssolh_context*, include/ssolh.h), created by
 * its first solve and holding the scene that solve uploaded: for calling the
 * device C-ABI directly. RES_BAD_OP before the first solve. */
SSOL_API res_T ssol_device_get_gpu_context(const struct ssol_device* dev, void** ctx);

/* Counters of the last solve: traced ray segments (including shadow rays)
 * and GPU kernel time in milliseconds. */
struct ssol_solve_stats {
  uint64_t segments;
  uint64_t paths;
  uint64_t pending_spills; /* per-path pending tally entries flushed early */
  double kernel_ms;
  double trace_ms;
  double shade_ms;
  unsigned iterations;
  unsigned wavefronts; /* wavefronts started: chunks of realisations all in
                          flight (a narrow width), or 1 */
  unsigned refilled;   /* 1: slots took further realisations in place */
};
SSOL_API res_T ssol_estimator_get_solve_stats
  (const struct ssol_estimator* estimator, struct ssol_solve_stats* stats);

/* Wavefront iterations captured per HIP graph launch (0, the default: one
 * host round trip per iteration). */
SSOL_API res_T ssol_device_set_graph_iterations
  (struct ssol_device* dev, const unsigned iterations);

/* Several sun positions over one scene in one solve. Sun j runs
 * realisations[j] realisations; its global estimates (same definitions as
 * ssol_estimator_get_mc_global) land in globals[j] and its realisation
 * count in counts[j] (counts may be NULL). Every sun must use the spectrum
 * object of the sun attached to the scene. The RNG streams continue across
 * the suns from `rng`, so the batch equals consecutive single-sun solves
 * with the counter advanced by each count. */
SSOL_API res_T ssol_solve_sun_batch
  (struct ssol_scene* scn, const struct ssp_rng* rng, const size_t nsuns,
   struct ssol_sun* const* suns, const size_t* realisations,
   struct ssol_mc_global* globals, size_t* counts);

/* Adjoint of a solve: with the same scene, rng and realisations as a
 * forward ssol_solve, returns per parameterised instance (heliostat p)
 *   grad[p*3 + 0], grad[p*3 + 1]: derivatives with respect to rotations of
 *     the instance about the world axes axes[p*6 + 0..2], axes[p*6 + 3..5];
 *   grad[p*3 + 2]: derivative with respect to its glossy mirror's slope
 *     error (Beckmann roughness = sqrt(2) slope error)
 * of d_absorbed * absorbed_by_receivers + sum_t d_map[t] * F_t, F_t being
 * the front-side per-primitive absorbed flux (W/m2) of map_shape on
 * map_receiver (d_map may be NULL). Score-function estimator of the
 * microfacet reflection plus the pathwise derivative of the start weight's
 * cosine; blocking and shading changes of the moving instance are not
 * differentiated. FP64 form only. */
SSOL_API res_T ssol_solve_adjoint
  (struct ssol_scene* scn, const struct ssp_rng* rng, const size_t realisations,
   const size_t nparams, struct ssol_instance* const* heliostats, const double* axes,
   const double d_absorbed, const struct ssol_instance* map_receiver,
   const struct ssol_shape* map_shape, const double* d_map, double* grad);

/* Autotuner: times probe solves of `probe` realisations over candidate
 * wavefront widths (path slots) and graph lengths, keeps the fastest as the
 * device's setting and reports it. */
struct ssol_autotune_result {
  unsigned path_slots;
  unsigned graph_iterations;
  double paths_per_second;   /* of the kept candidate */
  unsigned candidates;       /* number of candidates timed */
};
SSOL_API res_T ssol_device_autotune
  (struct ssol_device* dev, struct ssol_scene* scn, const size_t probe,
   struct ssol_autotune_result* out);
/* The same for sun batches: candidate widths and graph lengths timed on a
 * probe run of ssol_solve_sun_batch over the given suns and realisation
 * counts (a sample of the sweep to come). */
SSOL_API res_T ssol_device_autotune_batch
  (struct ssol_device* dev, struct ssol_scene* scn, const size_t nsuns,
   struct ssol_sun* const* suns, const size_t* realisations,
   struct ssol_autotune_result* out);

/* Rendering (reference src/ssol.h: camera 545-575, image 577-631,
 * ssol_draw_draft / ssol_draw_pt 1225-1244) ---------------------------- */
enum ssol_pixel_format {
  SSOL_PIXEL_DOUBLE3,
  SSOL_PIXEL_FORMATS_COUNT__
};

enum ssol_filter_mode {
  SSOL_FILTER_LINEAR,
  SSOL_FILTER_NEAREST
};

enum ssol_address_mode {
  SSOL_ADDRESS_CLAMP,
  SSOL_ADDRESS_REPEAT
};

struct ssol_image_layout {
  size_t row_pitch;     /* bytes between two consecutive rows */
  size_t offset;        /* byte offset where the image begins */
  size_t size;          /* width x height, as the reference reports it */
  size_t width, height; /* pixels in X and Y */
  enum ssol_pixel_format pixel_format;
};
#define SSOL_IMAGE_LAYOUT_NULL__ { 0, 0, 0, 0, 0, SSOL_PIXEL_FORMATS_COUNT__ }

/* Receives `size` pixels (row ordered, X fastest) of the image at `origin` */
typedef res_T
(*ssol_write_pixels_T)
  (void* context, const size_t origin[2], const size_t size[2],
   const enum ssol_pixel_format fmt, const void* pixels);

static inline size_t
ssol_sizeof_pixel_format(const enum ssol_pixel_format format)
{
  return format == SSOL_PIXEL_DOUBLE3 ? 3 * sizeof(double) : 0;
}

/* Pinhole camera: at the origin looking along -Z with +Y up and a 90 degree
 * horizontal field of view until set */
SSOL_API res_T ssol_camera_create(struct ssol_device* dev, struct ssol_camera** cam);
SSOL_API res_T ssol_camera_ref_get(struct ssol_camera* cam);
SSOL_API res_T ssol_camera_ref_put(struct ssol_camera* cam);
SSOL_API res_T ssol_camera_set_proj_ratio(struct ssol_camera* cam, const double proj_ratio);
SSOL_API res_T ssol_camera_set_fov(struct ssol_camera* cam, const double fov); /* radians */
SSOL_API res_T ssol_camera_look_at
  (struct ssol_camera* cam, const double position[3], const double target[3],
   const double up[3]);

SSOL_API res_T ssol_image_create(struct ssol_device* dev, struct ssol_image** image);
SSOL_API res_T ssol_image_ref_get(struct ssol_image* image);
SSOL_API res_T ssol_image_ref_put(struct ssol_image* image);
SSOL_API res_T ssol_image_setup
  (struct ssol_image* image, const size_t width, const size_t height,
   const enum ssol_pixel_format format);
SSOL_API res_T ssol_image_get_layout
  (const struct ssol_image* image, struct ssol_image_layout* layout);
SSOL_API res_T ssol_image_map(const struct ssol_image* image, char** memory);
SSOL_API res_T ssol_image_unmap(const struct ssol_image* image);
SSOL_API res_T ssol_image_sample
  (const struct ssol_image* image, const enum ssol_filter_mode filter,
   const enum ssol_address_mode address_u, const enum ssol_address_mode address_v,
   const double uv[2], void* val);
/* Matches ssol_write_pixels_T: copies the pixels into the image */
SSOL_API res_T ssol_image_write
  (void* image, const size_t origin[2], const size_t size[2],
   const enum ssol_pixel_format fmt, const void* pixels);

/* Per pixel, the mean over `spp` samples of the cosine between the first
 * hit's normal and the view ray (0 on a miss) */
SSOL_API res_T ssol_draw_draft
  (struct ssol_scene* scn, struct ssol_camera* cam, const size_t width,
   const size_t height, const size_t spp, ssol_write_pixels_T writer, void* writer_data);

/* Path-traced radiance per pixel: unit sun sampled at every hit, sky 0.1
 * towards `up` */
SSOL_API res_T ssol_draw_pt
  (struct ssol_scene* scn, struct ssol_camera* cam, const size_t width,
   const size_t height, const size_t spp, const double up[3],
   ssol_write_pixels_T writer, void* writer_data);

#ifdef __cplusplus
}
#endif

#endif /* SSOL_H */
