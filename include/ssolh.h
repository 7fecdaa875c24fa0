/* ssolh -- thin C-ABI between the host library (libssol) and the gfx950 HIP
 * kernels (libssol_hip).
 *
 * The reference has no device boundary: ssol_solve runs its OpenMP
 * realisation loop (src/ssol_solver.c:703-732) straight over Star-3D scene
 * views. In this build ssol_solve flattens the scene into the plain arrays
 * below (host memory, fixed-width types, no C++ or torch types), hands them
 * to ssolh_scene_upload, runs ssolh_solve, and reads the tallies back with
 * ssolh_tallies_download. Every entry point returns 0 on success, a negative
 * HIP error or one of the SSOLH_ERR_* codes otherwise; ssolh_last_error
 * gives the message.
 *
 * Conventions
 *  - 3x4 transforms are column major like the reference's: element (row r,
 *    column c) of the linear part is m[c*3+r], the translation is m[9..11].
 *  - Triangle front side: the geometric normal is (v2-v0) x (v1-v0), the
 *    left-handed convention of Star-3D hits (front = clockwise vertices).
 *  - Tally slots hold {sum of weights, sum of squared weights} pairs, each
 *    realisation's contributions summed before squaring
 *    (reference src/ssol_estimator_c.h:66-88).
 */
#ifndef SSOLH_H
#define SSOLH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SSOLH_API extern __attribute__((visibility("default")))

enum ssolh_error {
  SSOLH_OK = 0,
  SSOLH_ERR_ARG = 1,       /* invalid descriptor */
  SSOLH_ERR_NO_DEVICE = 2, /* no HIP device / wrong architecture */
  SSOLH_ERR_MEM = 3,       /* device allocation failed */
  SSOLH_ERR_LIMIT = 4      /* a fixed capacity of the kernels was exceeded */
};

enum ssolh_shape_type { SSOLH_SHAPE_MESH = 0, SSOLH_SHAPE_PUNCHED = 1 };

/* Same numbering as enum ssol_quadric_type */
enum ssolh_quadric {
  SSOLH_Q_PLANE = 0, SSOLH_Q_PARABOL, SSOLH_Q_PARABOL2F, SSOLH_Q_HYPERBOL,
  SSOLH_Q_PARABOLIC_CYLINDER, SSOLH_Q_HEMISPHERE
};

/* Same numbering as enum ssol_material_type */
enum ssolh_material_type {
  SSOLH_MTL_DIELECTRIC = 0, SSOLH_MTL_MATTE, SSOLH_MTL_MIRROR,
  SSOLH_MTL_THIN_DIELECTRIC, SSOLH_MTL_VIRTUAL
};

enum ssolh_sun_type {
  SSOLH_SUN_DIRECTIONAL = 0, SSOLH_SUN_PILLBOX, SSOLH_SUN_GAUSSIAN,
  SSOLH_SUN_BUIE
};

/* A value that is a constant or a spectrum of the scene's spectrum pool */
struct ssolh_data {
  double real;        /* used when spectrum < 0 */
  int32_t spectrum;   /* index into ssolh_scene.spectra, or -1 */
  int32_t pad__;
};

struct ssolh_spectrum {
  uint32_t offset;    /* into spectra_wl / spectra_val */
  uint32_t count;
};

/* Media are deduplicated by the host with the reference's media_ceq rule
 * (src/ssol_material.c:730-736): equal ids <=> media_ceq */
struct ssolh_medium {
  struct ssolh_data extinction;
  struct ssolh_data refractive_index;
};

/* Material, with its shaders already evaluated (baked) by the host */
struct ssolh_material {
  int32_t type;             /* enum ssolh_material_type */
  int32_t distribution;     /* mirror: 0 Beckmann, 1 pillbox */
  int32_t medium_in;        /* dielectric / thin dielectric media ids */
  int32_t medium_out;
  int32_t normal_from_ng;   /* 1: shader returns Ng, 0: returns Ns */
  int32_t pad__;
  double reflectivity;      /* mirror, matte */
  double roughness;         /* mirror; 0 <=> specular */
  double thickness;         /* thin dielectric */
  struct ssolh_data slab_eta;        /* thin dielectric slab index */
  struct ssolh_data slab_absorption; /* thin dielectric slab extinction */
};

/* Ray-tracing shape. Mesh shapes and punched surfaces share the tessellated
 * triangle representation (punched: the lifted quadric mesh, reference
 * src/ssol_shape.c:694-763); punched shapes add the analytic quadric used by
 * the hit refinement (src/ssol_shape.c:1449-1494). */
struct ssolh_shape {
  int32_t type;             /* enum ssolh_shape_type */
  int32_t quadric;          /* enum ssolh_quadric, punched only */
  uint32_t ntris;
  uint32_t pad__;
  double qparam[4];         /* private quadric data, see below */
  double qtransform[12];    /* quadric -> object space */
};
/* qparam by quadric:
 *  PARABOL:   {focal, 1/(4 focal)}
 *  PARABOL2F: {1/(4 fx), 1/(4 fy)}
 *  HYPERBOL:  {g/2, a2/b2, 1/a2, |b|} (reference priv_hyperbol_data)
 *  PARABOLIC_CYLINDER: {focal, 1/(4 focal)}
 *  HEMISPHERE: {radius, radius2} */

/* One shaded shape of an object */
struct ssolh_sshape {
  uint32_t shape;
  uint32_t mtl_front;
  uint32_t mtl_back;
  uint32_t pad__;
};

/* Object = BLAS over the triangles of its shaded shapes */
#define SSOLH_OBJECT_PLANAR 1u  /* mesh shapes only, all triangles in one
                                   plane: a ray leaving the object can meet it
                                   again only at t ~ 0, which the hit filter
                                   rejects (t <= 1e-6) */
struct ssolh_object {
  uint32_t sshape_first;    /* into sshapes */
  uint32_t sshape_count;
  uint32_t blas_root;       /* into nodes */
  uint32_t flags;           /* SSOLH_OBJECT_* */
};

/* BVH node (BVH2, both child boxes stored in the parent). A child with
 * count == 0 is the inner node `child`; count == SSOLH_BVH_EMPTY marks an
 * absent child; otherwise it is a leaf over [child, child+count) of the leaf
 * primitives (triangles for a BLAS, TLAS leaf slots for the TLAS). */
#define SSOLH_BVH_EMPTY 0xFFFFFFFFu
struct ssolh_bvh_node {
  float lo[2][3];
  float hi[2][3];
  uint32_t child[2];
  uint32_t count[2];
};

/* BLAS triangle, object space, reference winding */
struct ssolh_triangle {
  float v0[3];
  float v1[3];
  float v2[3];
  uint32_t sshape;          /* shaded-shape index local to its object */
  uint32_t prim;            /* triangle index within its shape */
  uint32_t pad__;
};

/* Instance of an object */
struct ssolh_instance {
  double transform[12];     /* object -> world */
  double inv_linear[9];     /* inverse of the linear part, column major */
  uint32_t object;
  uint32_t receiver_mask;   /* ssol_side_flag bits */
  int32_t receiver;         /* receiver tally slot or -1 */
  int32_t sampled;          /* sampled tally slot or -1 */
  uint32_t isshape_first;   /* into inst_sshapes: one per shaded shape */
  uint32_t user_id;
};

/* Per (instance, shaded shape): world <-> quadric frame for punched shapes
 * and the per-primitive tally base (reference src/ssol_shape.c:1461-1466,
 * composite of instance and quadric transforms) */
struct ssolh_inst_sshape {
  double R[9];              /* quadric -> world linear part */
  double T[3];              /* quadric -> world translation */
  double R_inv[9];          /* world -> quadric linear part */
  int64_t prim_tally;       /* first per-primitive slot or -1 */
};

/* Sampled triangle of the sampling view (flat proxy for punched shapes,
 * reference src/ssol_shape.c:767-811), object space */
struct ssolh_sample_tri {
  float v0[3];
  float v1[3];
  float v2[3];
  uint32_t instance;
  uint32_t sshape;          /* local to the instance's object */
  uint32_t prim;
};

/* Alias table entry over the sampled triangles, area weighted */
struct ssolh_alias {
  double prob;
  uint32_t alias;
  uint32_t pad__;
};

struct ssolh_sun {
  int32_t type;             /* enum ssolh_sun_type */
  int32_t wl_spectrum;      /* sun spectrum index in the pool */
  double dir[3];            /* normalised main direction */
  double basis[9];          /* column major frame whose z column is dir */
  double dni;
  double param[8];          /* pillbox: sin2 theta_max; gaussian: sigma;
                               buie: thetaSD, dThetaCSSD, gamma,
                               e^k 1000^gamma, alpha, h1, h2, p_rect1 */
  double wl_cdf_total;      /* piecewise-linear wavelength CDF normaliser */
};

struct ssolh_scene {
  uint32_t nspectra, spectra_len;
  const struct ssolh_spectrum* spectra;
  const double* spectra_wl;
  const double* spectra_val;
  const double* spectra_cdf; /* cumulative trapezoid areas, per spectrum */

  uint32_t nmedia, air_medium;
  const struct ssolh_medium* media;

  uint32_t nmaterials;
  const struct ssolh_material* materials;

  uint32_t nshapes;
  const struct ssolh_shape* shapes;

  uint32_t nsshapes;
  const struct ssolh_sshape* sshapes;

  uint32_t nobjects;
  const struct ssolh_object* objects;
  const float* object_boxes;  /* [nobjects][6]: object-space lo xyz, hi xyz
                                 (lo > hi: empty), the boxes the TLAS's
                                 instance boxes are made of */

  uint32_t nnodes;
  const struct ssolh_bvh_node* nodes;   /* BLASes then TLAS */
  uint32_t tlas_root;
  uint32_t stack_depth;       /* TLAS depth + deepest BLAS + 2 of the BVH2
                                 (informational: the upload derives the
                                 stack bound of its 4-wide nodes) */

  uint32_t ntriangles;
  const struct ssolh_triangle* triangles;

  uint32_t ninstances;
  const struct ssolh_instance* instances;
  const uint32_t* tlas_instances;       /* TLAS leaf order -> instance */

  uint32_t ninst_sshapes;
  const struct ssolh_inst_sshape* inst_sshapes;

  uint32_t nsample_tris;
  const struct ssolh_sample_tri* sample_tris;
  const struct ssolh_alias* sample_alias;

  struct ssolh_sun sun;
  double sampled_area;        /* S: sum of RT areas of sampled instances */
  double sampled_area_proxy;  /* S_proxy: sum of sampling-mesh areas */

  /* Tally layout */
  uint32_t nreceivers;        /* receiver slots (2 sides each) */
  uint32_t nsampled;          /* sampled-instance slots */
  uint64_t nprim_tallies;     /* per-primitive slots (2 sides each) */
};

enum ssolh_precision { SSOLH_FP64 = 0, SSOLH_FP32 = 1 };

struct ssolh_solve_args {
  uint64_t realisations;
  uint64_t max_failed;        /* reference max_failed_count */
  uint64_t rng_key;           /* Threefry key word 0 */
  uint64_t rng_counter;       /* first realisation counter */
  uint32_t precision;         /* enum ssolh_precision */
  uint32_t path_slots;        /* 0 <=> automatic */
  uint32_t pending_entries;   /* per-path pending tally entries, 0 => 4 */
  uint32_t graph_iterations;  /* 0: iterations driven from the host; n > 0:
                                 n (rounded up to even) iterations captured
                                 in a HIP graph and replayed until done */
};

/* Several suns over the uploaded scene in one solve: sun j runs
 * realisations[j] realisations (rng counters continue across suns). The
 * scene's receivers, sampled and per-primitive tallies accumulate over all
 * suns; the global fields and the realisation count are kept per sun. */
struct ssolh_batch_args {
  uint32_t nsuns;
  uint32_t pad__;
  const struct ssolh_sun* suns;     /* [nsuns], same spectrum pool as the scene */
  const uint64_t* realisations;     /* [nsuns] */
  uint64_t max_failed;
  uint64_t rng_key;
  uint64_t rng_counter;
  uint32_t precision;
  uint32_t path_slots;
  uint32_t pending_entries;
  uint32_t graph_iterations;
};

struct ssolh_batch_result {
  double global[6][2];        /* SSOLH_GLOBAL_FIELDS order, {sum w, sum w2} */
  uint64_t realisation_count;
};

/* Number of doubles of each tally array (pairs {sum w, sum w2}) */
#define SSOLH_GLOBAL_FIELDS 6   /* order: cos, absorbed_by_receivers,
                                   shadowed, missing, extinguished,
                                   other_absorbed */
#define SSOLH_RCV_FIELDS 10     /* ssol_mc_receiver field order */

struct ssolh_tallies {
  double global[SSOLH_GLOBAL_FIELDS][2];
  double* receivers;          /* [nreceivers][2 sides][10][2] */
  double* sampled;            /* [nsampled][2: cos, shadowed][2] */
  uint64_t* sampled_count;    /* [nsampled] nb_samples */
  double* sampled_x_receivers;/* [nsampled][nreceivers][2][10][2] */
  double* primitives;         /* [nprim_tallies][2 sides][10][2] */
  uint64_t realisation_count; /* successful realisations */
  uint64_t failed_count;
  uint64_t segments;          /* traced segments incl. shadow rays */
  uint64_t pending_spills;
  uint32_t aborted;           /* failures reached max_failed */
  uint32_t iterations;
  uint32_t wavefronts;        /* wavefronts started: chunks of realisations all
                                 in flight, or 1 */
  uint32_t refilled;          /* 1: slots took further realisations in place */
  double kernel_ms, trace_ms, shade_ms;
};

typedef struct ssolh_context ssolh_context;

SSOLH_API int ssolh_create(int device, ssolh_context** ctx);
SSOLH_API int ssolh_destroy(ssolh_context* ctx);
SSOLH_API const char* ssolh_last_error(void);
SSOLH_API int ssolh_device_count(void);

/* Copy the flat scene to the device (replaces the previous one) */
SSOLH_API int ssolh_scene_upload(ssolh_context* ctx, const struct ssolh_scene* scn);

/* Instances moved, nothing else changed (reference: the views rebuilt at
 * every solve, src/ssol_scene.c:346-433): new instance records and
 * (instance, shape) frames, same counts and objects as uploaded, replace the
 * device copies, and the TLAS is refit on the GPU to the instances' new
 * world boxes (same tree, new boxes; the LDS-staged top levels follow).
 * *cost_ratio (if not NULL) = the refit tree's summed child-box surface area
 * over the tree's as uploaded: the caller rebuilds and uploads the scene
 * when the motion has made the tree too loose. Ordered on the context's
 * stream; returns after the refit. */
SSOLH_API int ssolh_scene_move_instances
  (ssolh_context* ctx, uint32_t ninstances, const struct ssolh_instance* instances,
   uint32_t ninst_sshapes, const struct ssolh_inst_sshape* inst_sshapes,
   double* cost_ratio);

/* Builds the TLAS anew on the GPU over the instances' current world boxes
 * (linear BVH over Morton codes of the box centres, collapsed to 4-wide
 * nodes), in place of the uploaded one: for instances that moved too far
 * for a refit (reference: the views rebuilt at every solve,
 * src/ssol_scene.c:346-433). Uses the instance records the last upload or
 * ssolh_scene_move_instances left on the device; returns after the build. */
SSOLH_API int ssolh_scene_build_tlas(ssolh_context* ctx);

/* Image of the uploaded scene through a pinhole camera (reference
 * ssol_draw_draft, src/ssol_draw_draft.c, and ssol_draw_pt,
 * src/ssol_draw_pt.c:126-285), traced by the solver's kernels. Pixel
 * (x, y) averages `spp` camera rays through (x + jx, y + jy) / (width,
 * height) of the image plane, ray direction = axis_x (2 sx - 1) + axis_y
 * (2 sy - 1) + axis_z (the reference camera_ray). mode 0 (draft): the
 * cosine between the first hit's normal and the view ray, offsets
 * draft_samples[2k], [2k + 1] shared by every pixel; mode 1 (path
 * tracing): offsets and paths from Threefry streams keyed by (rng_key,
 * pixel * spp + k), sky 0.1 towards `up`, the unit sun sampled at every
 * hit. pixels: [height][width] host doubles, one radiance per pixel. */
struct ssolh_render_args {
  uint32_t width, height, spp, mode;
  float position[3], axis_x[3], axis_y[3], axis_z[3];
  double up[3];
  uint64_t rng_key;
  const float* draft_samples;   /* [spp][2], mode 0 */
};
SSOLH_API int ssolh_render(ssolh_context* ctx, const struct ssolh_render_args* args,
                           double* pixels);

/* Run the realisations of one solve over the uploaded scene */
SSOLH_API int ssolh_solve(ssolh_context* ctx, const struct ssolh_solve_args* args);

/* Adjoint solve: the same realisations as the forward solve of `solve`
 * (same RNG key and counter), with per-path score-function gradients of
 *   d_absorbed * absorbed_by_receivers + sum_k gmap[k] * absorbed flux on
 *   per-primitive tally key k (= (slot * 2 + side))
 * with respect to, for each parameterised instance (heliostat), two aim
 * angles (rotations about the world axes given in axes[p][0], axes[p][1])
 * and the slope error of its glossy Beckmann mirror. FP64 form only. */
struct ssolh_adjoint_args {
  struct ssolh_solve_args solve;
  uint32_t nparams;
  uint32_t pad__;
  const int32_t* inst_param;  /* [ninstances]: parameter slot or -1 */
  const double* axes;         /* [nparams][2][3] */
  const double* gmap;         /* [nprim_tallies][2] upstream weights, or NULL */
  double d_absorbed;
};

/* grad: [nparams][3] = d/d angle 1, d/d angle 2, d/d slope error, per
 * realisation (divided by the realisation count). */
SSOLH_API int ssolh_solve_adjoint
  (ssolh_context* ctx, const struct ssolh_adjoint_args* args, double* grad);

/* Path recording (reference path tracker, src/ssol_solver.c:898-1127, kept
 * per solve in src/ssol_estimator.c:270-322): every realisation records its
 * vertices (start point, each hit, and the sun and escape points at the given
 * lengths when they are > 0) with the flux carried toward each vertex, and
 * its path type. FP64 form, host-driven iterations. */
enum ssolh_path_type {
  SSOLH_PATH_MISSING = 0,  /* same values as enum ssol_path_type */
  SSOLH_PATH_SHADOW = 1,
  SSOLH_PATH_SUCCESS = 2,
  SSOLH_PATH_ERROR = 3,
  SSOLH_PATH_DROPPED = 255 /* cancelled by Russian roulette, or never run */
};

struct ssolh_track_args {
  double sun_ray_length;       /* > 0: vertex at start - length * sun dir */
  double infinite_ray_length;  /* > 0: vertex at last point + length * dir */
  uint64_t max_vertices;       /* capacity of the device vertex buffer */
};

struct ssolh_track_vertex {
  uint64_t path;               /* realisation id (0 .. realisations-1) */
  uint32_t index;              /* position of the vertex in its path */
  uint32_t pad__;
  double pos[3];
  double weight;
};

/* Solve of `args` that records paths. *nvertices receives the number of
 * vertices the paths produced; when it exceeds max_vertices the buffer kept
 * only the first max_vertices appended, and the same solve run again with a
 * larger buffer records them all (the realisations are the same). Tallies
 * are those of the solve, as after ssolh_solve. */
SSOLH_API int ssolh_solve_tracked
  (ssolh_context* ctx, const struct ssolh_solve_args* args,
   const struct ssolh_track_args* track, uint64_t* nvertices);

/* Copy the recorded vertices (unordered, min(nvertices, max_vertices) of
 * them) and the type of each realisation (npaths = realisations). */
SSOLH_API int ssolh_track_download
  (ssolh_context* ctx, struct ssolh_track_vertex* vertices, uint64_t nvertices,
   uint8_t* types, uint64_t npaths);

/* Batched solve (see ssolh_batch_args); out: [nsuns] */
SSOLH_API int ssolh_solve_batch
  (ssolh_context* ctx, const struct ssolh_batch_args* args,
   struct ssolh_batch_result* out);

/* Copy the tallies of the last solve into host arrays the caller sized
 * from the uploaded scene's tally layout */
SSOLH_API int ssolh_tallies_download(ssolh_context* ctx, struct ssolh_tallies* out);

/* ssolh_solve followed by ssolh_tallies_download, allocating nothing: no
 * counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). Every buffer of the solve is carved from a 64 MiB
 * device array that the library holds on each GPU it runs on until the
 * process ends; the scene uploaded by ssolh_scene_upload is the input and
 * `out` (host arrays, as for ssolh_tallies_download) the output. The
 * wavefront is as wide as fits beside the tallies (args->path_slots, when
 * set, must fit too); a scene whose tallies need more than half of the array
 * is rejected with SSOLH_ERR_LIMIT and a message. Plain single-sun solves
 * driven from the host only: args->graph_iterations must be 0. The array is
 * shared by every call on one GPU and keeps the last call's contents; calls
 * that overlap on one GPU (two contexts, two threads) race. */
SSOLH_API int ssolh_solve_noalloc
  (ssolh_context* ctx, const struct ssolh_solve_args* args, struct ssolh_tallies* out);

/* Deterministic kernels, one launch over n inputs each, for known-answer
 * tests against the oracle. Inputs and outputs are host arrays. */
SSOLH_API int ssolh_test_quadric_hits
  (ssolh_context* ctx, int quadric, const double qparam[4], uint32_t n,
   const double* org, const double* dir, const double* hint,
   double* t, double* normal);
/* Sun directions from fixed uniforms: sample i draws from u[i*nu ...]
 * (0.5 once its nu values are used up); used[i] = uniforms consumed */
SSOLH_API int ssolh_test_sun_dirs
  (ssolh_context* ctx, const struct ssolh_sun* sun, uint32_t n,
   const double* u, uint32_t nu, double* dir, uint32_t* used);
/* BSDF samples from fixed uniforms (same convention as above): wi points
 * away from the surface on the normal's side; lobe 1 reflect, 2 transmit */
SSOLH_API int ssolh_test_bsdf
  (ssolh_context* ctx, const struct ssolh_material* m, double eta_out,
   double eta_in, double slab_eta, double slab_absorption, uint32_t n,
   const double* wi, const double* normal, const double* u, uint32_t nu,
   double* wo, double* weight, int* lobe);
SSOLH_API int ssolh_test_threefry
  (ssolh_context* ctx, uint32_t n, const uint64_t* ctr, const uint64_t* key,
   uint64_t* out);

/* Wave-wide sum used by the tallies: out[w*64 + i] = sum of in[w*64 ..
 * w*64+63] for every lane i, nwaves waves of 64 lanes */
SSOLH_API int ssolh_test_wave_sum
  (ssolh_context* ctx, uint32_t nwaves, const double* in, double* out);

/* Traversal stack of the trace kernels: caps the entries each lane keeps in
 * LDS (16, 24 or 32; 0 restores the default, the smallest that holds the
 * scene's bound) for the solves that follow, so that the rest of the bound
 * goes to the device-memory overflow; *bound (if not NULL) = the uploaded
 * scene's bound (0 without a scene). The results do not depend on the cap. */
SSOLH_API int ssolh_test_trace_stack(ssolh_context* ctx, uint32_t lds_cap, uint32_t* bound);

/* Slack of the trace kernels' box tests, per unit of the largest |origin /
 * direction| component of the ray (default 2^-21, which bounds the
 * rounding of the FMA plane distances with margin): the scale for this
 * context's current and later scenes, in [0, 1e-3]. Results do not change
 * with any scale at or above the default; 0 lets rounding cull boxes that
 * rays cross near their faces. */
SSOLH_API int ssolh_test_slab_slack(ssolh_context* ctx, float scale);

/* Log of per-primitive entries (finished paths append their entries; the
 * log is sorted by key and summed into the per-primitive tally after each
 * wavefront and whenever a host-driven iteration leaves it half full; a
 * full log sends further entries to per-lane atomics): caps it at `records`
 * entries for the solves that follow (0 restores the default, one per
 * realisation up to 2^27). *flushes and *overflowed (if not NULL) = the last
 * solve's sorted reductions and the entries it sent to the atomics. The
 * results do not depend on the cap beyond the order of the additions. */
SSOLH_API int ssolh_test_prim_log(ssolh_context* ctx, uint32_t records, uint32_t* flushes,
                                  uint64_t* overflowed);

/* Scene uploads, TLAS refits and GPU TLAS builds this context has done */
SSOLH_API int ssolh_test_scene_counters(ssolh_context* ctx, uint32_t* uploads, uint32_t* refits,
                                        uint32_t* builds);

/* Bandwidth of a plain device copy of `bytes` bytes repeated `reps` times
 * (read + write bytes over the HIP-event time), in GB/s. Calibrates the HBM
 * byte counters of the profiler; no counterpart in the reference. */
SSOLH_API int ssolh_copy_bandwidth(ssolh_context* ctx, size_t bytes, int reps, double* gbs);

#ifdef __cplusplus
}
#endif

#endif /* SSOLH_H */
