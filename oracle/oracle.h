/* ORACLE -- test infrastructure only.
 *
 * A CPU, FP64, OpenMP restatement of the reference flux integrator
 * (solstice-solver 0.9: src/ssol_solver.c, src/ssol_scene.c:438-536,
 * src/ssol_shape.c, src/ssol_ranst_sun_dir.c, src/ssol_ranst_sun_wl.c,
 * src/ssol_material.c, src/ssol_spectrum.c, src/ssol_estimator_c.h,
 * src/ssol_estimator.c, src/ssol_mc_receiver.c). It is the checker of the
 * GPU path and the CPU baseline leg of the benchmark. Only tests, the smoke
 * check and the benchmark's CPU-baseline leg may load it; the product
 * library never links it.
 *
 * It takes a high-level scene description (the same information a client
 * passes through the ssol_* scene API) and builds everything else itself:
 * punched-surface tessellation and carving, areas, its own BVH over
 * world-space FP64 triangles, the sun and wavelength samplers. It shares no
 * code with the product.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *  - threefry4x64_20: Random123 known-answer vectors.
 *  - solve_second / quadric hits: closed forms (plane, paraboloid focus,
 *    hyperboloid foci, parabolic-cylinder focal line) and brute-force root
 *    bisection in long double.
 *  - sun samplers: inverse-CDF identities (pillbox sin2 theta = u sin2 max,
 *    Gaussian sin theta = sigma sqrt(-2 ln u)), Buie KS test against the
 *    reference pdf_theta.
 *  - spectrum interpolation: reference src/test_ssol_data.c:49-76 values.
 *  - solver: the reference's known-answer solver tests
 *    (src/test_ssol_solver{1,2,2b,3,4,5,6,7,9,10,11,12}.c,
 *    src/test_ssol_by_receiver_integration.c) restated with the same values
 *    and tolerances, plus energy conservation (src/ssol_solver.c:520-558).
 *  - parity unpinned: microfacet (Beckmann / pillbox) sampling, dielectric
 *    interface, thin dielectric slab (Star-SF 0.6 is absent and no
 *    reference test exercises them; checked against invariants instead),
 *    SUB carvings and non-convex carving polygons (not supported here).
 */
#ifndef ORACLE_H
#define ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Scene description. Enumerations use the numbering of the reference's
 * public header (ssol_material_type, ssol_quadric_type, ssol_side_flag,
 * ssol_clipping_op, ssol_microfacet_distribution). */
struct odata {
  double real;
  int spectrum;  /* index into oscene.spectra or -1 */
};

struct omedium {
  struct odata extinction;
  struct odata refractive_index;
};

struct ospectrum {
  size_t n;
  const double* wavelengths;
  const double* values;
};

struct omaterial {
  int type;            /* 0 dielectric, 1 matte, 2 mirror, 3 thin, 4 virtual */
  int distribution;    /* mirror: 0 Beckmann, 1 pillbox */
  int normal_from_ng;  /* shading normal = Ng (else Ns; both equal here) */
  double reflectivity;
  double roughness;
  struct omedium medium_out; /* dielectric: outside; thin: outside */
  struct omedium medium_in;  /* dielectric: inside; thin: = outside */
  struct omedium slab;       /* thin dielectric slab */
  double thickness;
};

struct ocarving {
  int operation;       /* 0 AND, 1 SUB */
  size_t nverts;
  const double* xy;    /* [nverts][2] */
};

struct oshape {
  int type;            /* 0 mesh, 1 punched surface */
  /* mesh */
  unsigned nverts, ntris;
  const float* positions;   /* [nverts][3] object space */
  const unsigned* indices;  /* [ntris][3] */
  /* punched surface */
  int quadric;              /* ssol_quadric_type */
  double qdata[2];          /* focal | focal_x,focal_y | img,real | radius */
  double qtransform[12];    /* quadric -> object, column major 3x4 */
  size_t slices_hint;       /* SIZE_MAX <=> default */
  size_t ncarvings;
  const struct ocarving* carvings;
};

struct osshape {
  int shape, mtl_front, mtl_back;
};

struct oobject {
  size_t nsshapes;
  const struct osshape* sshapes;
};

struct oinstance {
  int object;
  double transform[12];
  int receiver_mask;
  int per_primitive;
  int sample;
};

struct osun {
  int type;             /* 0 directional, 1 pillbox, 2 gaussian, 3 buie */
  double dir[3];
  double dni;
  double param;         /* half angle | std dev | CSR ratio */
  int spectrum;
};

struct oscene {
  size_t nspectra;
  const struct ospectrum* spectra;
  size_t nmaterials;
  const struct omaterial* materials;
  size_t nshapes;
  const struct oshape* shapes;
  size_t nobjects;
  const struct oobject* objects;
  size_t ninstances;
  const struct oinstance* instances;
  struct osun sun;
  int has_atmosphere;
  struct odata atmosphere_extinction;
  int track_sampled_x_receiver; /* 0 disables the per sampled x receiver
                                   tallies (large fields) */
};

struct ooracle;

/* Build tessellation, areas and BVH. Returns NULL on invalid scenes (the
 * cases where the reference's setup calls return RES_BAD_ARG). */
struct ooracle* oracle_create(const struct oscene* scn, int* err);
void oracle_destroy(struct ooracle* o);

/* res_T codes of the reference: 0 OK, 1 BAD_ARG, 5 BAD_OP */
int oracle_solve
  (struct ooracle* o, uint64_t rng_key, uint64_t rng_counter,
   uint64_t realisations, uint64_t max_failed, int nthreads);

/* Solve that records every path like the reference's path tracker
 * (src/ssol_solver.c:898-1127; src/ssol_estimator.c:270-322), in one thread,
 * paths in realisation order. A negative length takes the reference's
 * default, 0.75 x the largest extent of the scene's box
 * (src/ssol_solver.c:562-570, 1211-1220). Types are enum ssol_path_type
 * values (0 missing, 1 shadow, 2 success, 3 error). */
int oracle_solve_tracked
  (struct ooracle* o, uint64_t rng_key, uint64_t rng_counter,
   uint64_t realisations, uint64_t max_failed, double sun_ray_length,
   double infinite_ray_length);
size_t oracle_tracked_count(const struct ooracle* o);
/* vertices: [nvertices][4] = x, y, z, weight; valid until the next solve */
int oracle_tracked_path
  (const struct ooracle* o, size_t ipath, int* type, size_t* nvertices,
   const double** vertices);
double oracle_tracked_extend(const struct ooracle* o);

/* Results: E, V, SE triples like the reference getters */
void oracle_get_global(struct ooracle* o, double out[6][3]);
int oracle_get_receiver
  (struct ooracle* o, size_t inst, int side, double out[10][3]);
int oracle_get_sampled
  (struct ooracle* o, size_t inst, double out[2][3], uint64_t* nb_samples);
int oracle_get_sampled_x_receiver
  (struct ooracle* o, size_t samp_inst, size_t rcv_inst, int side,
   double out[10][3]);
int oracle_get_primitive
  (struct ooracle* o, size_t inst, size_t sshape, unsigned prim, int side,
   double out[10][3]);
uint64_t oracle_realisation_count(const struct ooracle* o);
uint64_t oracle_failed_count(const struct ooracle* o);
double oracle_sampled_area(const struct ooracle* o);
double oracle_sampled_area_proxy(const struct ooracle* o);
uint64_t oracle_segments(const struct ooracle* o);
unsigned oracle_shape_ntris(const struct ooracle* o, size_t shape);
double oracle_shape_rt_area(const struct ooracle* o, size_t shape);
double oracle_shape_samp_area(const struct ooracle* o, size_t shape);
/* Tessellated RT mesh of a shape in object space, for geometry checks */
unsigned oracle_shape_rt_vertex_count(const struct ooracle* o, size_t shape);
void oracle_shape_rt_vertex(const struct ooracle* o, size_t shape,
  unsigned ivert, double pos[3]);

/* Deterministic kernels exposed for known-answer tests */
void oracle_threefry4x64_20
  (const uint64_t ctr[4], const uint64_t key[4], uint64_t out[4]);
int oracle_solve_second
  (double a, double b, double c, double hint, double dist[2]);
/* qdata as in oshape. Returns 0 when no root; t = INFINITY then. */
int oracle_quadric_intersect_local
  (int quadric, const double qdata[2], const double org[3],
   const double dir[3], double hint, double* t, double grad[3]);
double oracle_quadric_z(int quadric, const double qdata[2], const double xy[2]);
void oracle_basis(const double dir[3], double basis[9]);
/* u: uniforms in [0,1) consumed in order; returns how many were used */
int oracle_sun_sample
  (int type, double param, const double dir[3], const double* u, int nu,
   double out[3]);
double oracle_buie_pdf_theta(double csr, double theta);
void oracle_buie_params(double csr, double out[8]);
double oracle_spectrum_interpolate
  (const double* wl, const double* val, size_t n, double wavelength);
double oracle_piecewise_linear_sample
  (const double* wl, const double* val, size_t n, double u);
/* BSDF sampling: wi points away from the surface, on N's side.
 * Returns the MC weight in [0,1]; *type: 1 reflection, 2 transmission. */
double oracle_bsdf_sample
  (const struct omaterial* m, double eta_out, double eta_in,
   double slab_eta, double slab_absorption, const double wi[3],
   const double N[3], const double* u, int nu, double wo[3], int* type);

#ifdef __cplusplus
}
#endif

#endif /* ORACLE_H */
