// Internal types of the host library (libssol). Objects mirror the
// reference's ref-counted handles (src/ssol_*_c.h); geometry is kept on the
// host in the flat form the GPU upload needs.
#pragma once

#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "ssol.h"
#include "ssolh.h"
#include "ssp.h"

namespace ssol {

void log_error(const ssol_device* dev, const char* fmt, ...);
void log_warning(const ssol_device* dev, const char* fmt, ...);

}  // namespace ssol

// Reference counting shared by every handle
struct ssol_refcounted {
  int ref = 1;
  uint64_t gen = 0;  // bumped by every setter; part of the solve cache key
};

struct solve_cache;  // flattened scene of the last solve (ssol_solve.cpp)

struct ssol_device : ssol_refcounted {
  int verbose = 0;
  unsigned nthreads = 1;
  enum ssol_precision precision = SSOL_PRECISION_FP64;
  size_t path_slots = 0;
  unsigned graph_iterations = 0;  // 0: host-driven wavefront iterations
  ssolh_context* gpu = nullptr;  // created at the first solve
  solve_cache* cache = nullptr;
};

struct ssol_spectrum : ssol_refcounted {
  ssol_device* dev = nullptr;
  std::vector<double> wavelengths, intensities;
  uint64_t checksum[2] = {0, 0};
};

struct ssol_param_buffer : ssol_refcounted {
  ssol_device* dev = nullptr;
  std::vector<unsigned char> mem;
  size_t used = 0;
  struct Release { void* p; void (*fn)(void*); };
  std::vector<Release> releases;
};

enum class shape_kind { mesh, punched };

struct ssol_shape : ssol_refcounted {
  ssol_device* dev = nullptr;
  shape_kind kind = shape_kind::mesh;
  uint32_t uid = 0;
  bool ready = false;
  // ray-tracing mesh, object space (punched: lifted quadric)
  std::vector<float> pos;       // [nverts][3]
  std::vector<unsigned> ids;    // [ntris][3]
  std::vector<float> normals;   // optional [nverts][3]
  std::vector<float> uvs;       // optional [nverts][2]
  // sampling mesh (punched: flat proxy, same ids); meshes sample pos
  std::vector<float> samp_pos;
  double rt_area = 0, samp_area = 0;
  // punched surface
  int quadric = 0;
  double qparam[4] = {0, 0, 0, 0};
  double qtransform[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};
};

struct ssol_material : ssol_refcounted {
  ssol_device* dev = nullptr;
  enum ssol_material_type type = SSOL_MATERIAL_VIRTUAL;
  ssol_shader_getter_T normal = nullptr;
  ssol_shader_getter_T reflectivity = nullptr;
  ssol_shader_getter_T roughness = nullptr;
  enum ssol_microfacet_distribution distrib = SSOL_MICROFACET_BECKMANN;
  ssol_medium in_medium = SSOL_MEDIUM_VACUUM__;
  ssol_medium out_medium = SSOL_MEDIUM_VACUUM__;
  ssol_medium slab_medium = SSOL_MEDIUM_VACUUM__;
  double thickness = 0;
  ssol_param_buffer* buf = nullptr;
};

struct shaded_shape {
  ssol_shape* shape;
  ssol_material* front;
  ssol_material* back;
};

struct ssol_object : ssol_refcounted {
  ssol_device* dev = nullptr;
  std::vector<shaded_shape> sshapes;
  double rt_area = 0, samp_area = 0;
};

struct ssol_instance : ssol_refcounted {
  ssol_device* dev = nullptr;
  ssol_object* object = nullptr;
  double transform[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};
  int receiver_mask = 0;
  int per_primitive = 0;
  int sample = 1;
  uint32_t id = 0;
};

enum class sun_kind { directional, pillbox, gaussian, buie };

struct ssol_sun : ssol_refcounted {
  ssol_device* dev = nullptr;
  sun_kind kind = sun_kind::directional;
  double direction[3] = {0, 0, 0};
  double dni = 0;
  ssol_spectrum* spectrum = nullptr;
  double param = 0;
  ssol_scene* scene = nullptr;
};

struct ssol_atmosphere : ssol_refcounted {
  ssol_device* dev = nullptr;
  ssol_data extinction = SSOL_DATA_NULL__;
  ssol_scene* scene = nullptr;
};

struct ssol_scene : ssol_refcounted {
  ssol_device* dev = nullptr;
  std::vector<ssol_instance*> instances;
  ssol_sun* sun = nullptr;
  ssol_atmosphere* atmosphere = nullptr;
};

struct ssp_rng {
  int ref = 1;
  enum ssp_rng_type type = SSP_RNG_THREEFRY;
  uint64_t key = 0;
  uint64_t counter = 0;
  uint64_t buf[4] = {0, 0, 0, 0};
  int avail = 0;
};

// Estimator -----------------------------------------------------------------
struct estimator_shape_block {
  const ssol_shape* shape;
  size_t first;  // per-primitive slot of triangle 0
};

struct ssol_estimator : ssol_refcounted {
  ssol_device* dev = nullptr;
  size_t realisation_count = 0;
  size_t failed_count = 0;
  double sampled_area = 0;
  double global[6][2] = {};
  std::vector<double> rcv, samp, sxr, prim;
  std::vector<uint64_t> samp_count;
  std::map<uint32_t, int> rcv_slot;   // instance id -> receiver slot
  std::map<uint32_t, int> samp_slot;  // instance id -> sampled slot
  std::map<uint32_t, int> rcv_mask;   // instance id -> receiver mask
  std::map<std::pair<uint32_t, uint32_t>, size_t> prim_first;  // (inst id, shape uid)
  size_t nrcv = 0, nsamp = 0;
  ssp_rng* rng = nullptr;
  ssol_solve_stats stats = {};
  // receiver handles hand out pointers to these
  struct rcv_ref { ssol_estimator* est; int slot; int side; const ssol_instance* inst; };
  std::vector<std::unique_ptr<rcv_ref>> refs;
  struct shape_ref { ssol_estimator* est; size_t first; const ssol_shape* shape; };
  std::vector<std::unique_ptr<shape_ref>> shape_refs;
  // paths recorded by a solve with a path tracker, in realisation order;
  // ssol_path::path__ points to one of `paths`
  struct tracked_path { int type; const ssol_path_vertex* v; size_t n; };
  std::vector<tracked_path> paths;
  std::vector<ssol_path_vertex> path_vertices;
};

namespace ssol {

// Refcount helpers
template <typename T> inline void ref_get(T* o) { ++o->ref; }

void device_ref_put_(ssol_device* d);
void solve_cache_free(ssol_device* d);
// The scene flattened and on the GPU as ssol_solve would have it (sun
// checked, cache reused or refreshed, context created); `who` names the
// caller in log lines (ssol_solve.cpp)
res_T scene_for_gpu(ssol_scene* scn, const char* who);
uint64_t next_gen();  // monotonic stamp for ssol_refcounted::gen

// Punched-surface tessellation (ssol_shape.cpp)
res_T punched_setup(ssol_shape* shape, const ssol_punched_surface* ps);

// BVH (ssol_bvh.cpp): builds a BVH2 over boxes, appends nodes, returns root
// and the leaf order of the primitives; max depth reported.
struct aabb { float lo[3], hi[3]; };
uint32_t bvh_build(const std::vector<aabb>& boxes, unsigned max_leaf,
                   std::vector<ssolh_bvh_node>& nodes, std::vector<uint32_t>& order,
                   unsigned* depth);

// FNV-1a 64 over raw bytes
uint64_t fnv64(const void* data, size_t n);

}  // namespace ssol
