// ssolh -- wavefront flux integrator on one MI355X (gfx950).
//
// Replaces the reference's OpenMP realisation loop (src/ssol_solver.c:
// 1223-1248 running trace_radiative_path, :873-1127) by a wavefront of path
// slots in HBM. Every per-path record (path state, next ray, last hit) is
// indexed by its slot; the work list of an iteration is the list of slots
// that have a ray to trace, kept in increasing slot order so that a wave
// reads and writes contiguous runs of every record.
//
//   k_generate  fills the slots with new realisations (start point, sun
//               direction, wavelength, weights) and writes their shadow rays;
//   k_trace     one thread per listed slot: two-level BVH traversal in FP32
//               boxes with the reference's any-hit filter inline
//               (src/ssol_scene.c:438-536) and the FP64 quadric refinement
//               (src/ssol_shape.c:1449-1494);
//   k_shade     one thread per listed slot: the tail of the reference loop
//               iteration that traced the ray (miss, segment extinction
//               tallies, Russian roulette, point update) and the head of the
//               next one (extinction, BSDF or virtual pass-through, receiver
//               tallies); finished realisations flush their pending tallies
//               with wave-aggregated FP64 atomics and, when the solve has
//               more realisations than slots, the slot is refilled with its
//               next realisation in place;
//   k_scan_chunks / k_scan_top / k_scatter  order-preserving compaction of
//               the slots that wrote a ray into the next work list (per-block
//               counts, two-level scan of the counts, scatter).
//
// The host loop runs trace / shade / scan / scatter until the list is empty.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <float.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "ssolh.h"
#include "ssolh_device.hpp"
#include "ssolh_device32.hpp"

using namespace ssolh;

namespace {

constexpr int BLOCK = 128;           // threads per block (2 waves)
constexpr int WAVES = BLOCK / 64;
constexpr int MAX_STACK = 32;        // traversal stack entries per thread in LDS, largest kernel
constexpr int SCAN_THREADS = 1024;   // threads of a scan block
constexpr uint32_t SENTINEL = 0xFFFFFFFFu;
constexpr uint32_t LEAF = 0x80000000u;
constexpr uint32_t NONE = 0xFFFFFFFFu;
constexpr uint32_t TOP_REF = 0x40000000u;  // inner-node reference into the LDS copy
constexpr uint32_t EMPTY4 = 0xFFFFFFFEu;   // absent child of a 4-wide node
constexpr int TOP_NODES = 31;              // TLAS nodes staged in LDS (3.9 KB): 3 levels
constexpr float SLAB_SLACK = 0x1p-21f;     // box-test slack (FRay), per unit of max |o*|
constexpr int MAX_PEND = 8;          // compile-time bound of pending entries

// Per-thread message of the last failing call
thread_local char g_err[512] = "";

void set_err(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

#define HIP_TRY(expr)                                                        \
  do {                                                                       \
    hipError_t e__ = (expr);                                                 \
    if (e__ != hipSuccess) {                                                 \
      set_err("%s: %s (%s:%d)", #expr, hipGetErrorString(e__), __FILE__,     \
              __LINE__);                                                     \
      return -(int)e__;                                                      \
    }                                                                        \
  } while (0)

// -------------------------------------------------------------------------
// Device scene
// -------------------------------------------------------------------------
// IssInfo::obj_flags bit (beside the SSOLH_OBJECT_* ones): the object has
// one shaded shape
constexpr uint32_t OBJ_ONE_SHAPE = 1u << 31;

// What a hit on (instance, shaded shape) needs, gathered at upload so that a
// hit costs one lookup: global index = instance.isshape_first + sshape.
struct alignas(8) IssInfo {
  uint32_t inst;
  uint32_t shape;
  uint32_t mtl_front, mtl_back;
  int32_t receiver;        // receiver slot of the instance or -1
  uint32_t receiver_mask;  // ssol_side_flag bits
  uint32_t shape_type;     // enum ssolh_shape_type
  uint32_t obj_flags;      // SSOLH_OBJECT_* of the instance's object, OBJ_ONE_SHAPE
  int64_t prim_tally;      // first per-primitive slot or -1
};

// BVH node of the traversal: four child boxes as one 16-byte row per axis
// and side (box[axis][0] the four lower planes, box[axis][1] the upper
// ones; an absent child has lower planes +inf and upper planes -inf, which
// no ray crosses) and four child references in the form the traversal
// follows (inner node index, TOP_REF | index into the LDS copy,
// LEAF | (count - 1) << 28 | first item, or EMPTY4). Built at upload by
// collapsing the host's BVH2 (each node takes the children and grandchildren
// of a BVH2 node, opening the largest child boxes first), which halves the
// depth a ray descends.
struct alignas(16) Node4 {
  float box[3][2][4];
  uint32_t child[4];
  uint32_t pad[4];
};
static_assert(sizeof(Node4) == 128, "Node4 layout");

struct StartAlias {
  double prob;
  uint32_t inst, inst_alias;
};

struct DScene {
  Spectra spectra;
  const ssolh_medium* media;
  uint32_t air;
  const ssolh_material* mtls;
  const ssolh_shape* shapes;
  const ssolh_object* objects;
  const Node4* nodes;
  uint32_t tlas_root;
  // copies of the TLAS's top levels (breadth first, children in the set
  // referenced as TOP_REF | index), staged in LDS by the trace kernels
  const Node4* top;
  uint32_t ntop;
  uint32_t root_ref;         // first node reference of a traversal
  float slack_scale;         // box-test slack per unit of max |o*| (SLAB_SLACK)
  const ssolh_triangle* tris;
  const ssolh_instance* insts;
  const uint32_t* tlas_inst;
  const ssolh_inst_sshape* isshapes;
  const IssInfo* iss;
  const ssolh_sample_tri* stris;
  const ssolh_alias* alias;
  // per alias cell: its probability and the instances of its two sampled
  // triangles (the start instance in one load, for k_start_keys)
  const struct StartAlias* salias;
  const float* obj_box;  // [object][6]: object-space lo xyz, hi xyz
  uint32_t nstris;
  ssolh_sun sun;
  double S, S_proxy;
  uint32_t nrcv, nsamp;
};

// Ray and hit records, indexed by slot ------------------------------------
enum : uint32_t {
  RF_SIDE_SHIFT = 16,     // 2 bits: side of the origin point
  RF_DISCARD_VIRTUAL = 1u << 18,
  RF_REVERSED = 1u << 19,
  RF_ANY_HIT = 1u << 20
};

struct alignas(16) RayRec {
  double ox, oy, oz;
  double dx, dy, dz;
  float tmin;
  uint32_t iss_from;   // (instance, shaded shape) of the emitting point
  uint32_t prim_from;  // its triangle
  uint32_t flags;      // side << RF_SIDE_SHIFT | RF_*
};
static_assert(sizeof(RayRec) == 64, "RayRec layout");

struct alignas(16) HitRec {
  double t;             // distance of the nearest accepted triangle
  double qdst;          // nearest refined quadric distance (punched shapes)
  double nx, ny, nz;    // unnormalised world normal: triangle Ng, or quadric gradient
  uint32_t iss, prim;   // iss == NONE <=> miss
};
static_assert(sizeof(HitRec) == 48, "HitRec layout");

// Path state, indexed by slot ---------------------------------------------
enum : uint32_t {
  PF_STAGE_BOUNCE = 1u << 0,   // else: waiting for the shadow ray
  PF_HIT_VIRTUAL = 1u << 1,    // the point that emitted the ray is virtual
  PF_IN_ATM = 1u << 2,         // the incoming segment of that point was in air
  PF_ACTIVE = 1u << 3,
  PF_HIT_RECEIVER = 1u << 4    // the path reached a receiver (path tracking)
};

// The FP64 form keeps per slot only what a bounce changes plus a few
// records written once per path; the position and direction of the point a
// ray leaves are those of the ray record (its origin is the last real point,
// also past a virtual one).
struct PathState {
  double2* cst;        // wavelength, cos factor (written once per path)
  double4* pw;         // outgoing weights of the last real point (3); w: the
                       // initial flux until the first shading, then the
                       // incoming flux of the last virtual point crossed
  double4* nrm;        // normal of the start point (written once per path)
  double4* pend;       // pending absorbed_by_receivers, extinguished, other (first
                       // written by the path's first shading, which starts from zeros)
  uint4* st;           // path id lo/hi, rng block, flags
  uint4* sm;           // (instance, shaded shape) index, prim, depth | side << 30,
                       // in medium | out medium << 16
  uint4* aux;          // sampled slot, job (written once per path), survivor
                       // score, roulette interval (rewritten past depth 64)
  uint32_t* pkey;      // [npend][slots]: receiver key (slot*2+side) or NONE
  double* pval;        // [npend][10][slots]
  uint32_t* qkey;      // [npend][slots]: per-primitive key or NONE
  double* qval;        // [npend][10][slots]
};

struct Counters {
  unsigned long long failed;
  unsigned long long spills;
  unsigned long long segments;      // traced rays (added by k_scan_top, one thread)
  unsigned int aborted;
  unsigned int pad;
};

// Tallies that every wave updates (global fields, counts) or that few keys
// share (receivers, sampled instances) are striped: block b adds into
// stripe b & mask, so concurrent atomics spread over many cache lines; the
// stripes are folded into stripe 0 after the solve. The stripe count of a
// tally shrinks with its size (2^21 doubles at most over all its stripes).
struct Tallies {
  double* global;   // [mask_g+1][njobs][16]: 6 fields x {w, w2}, then the
                    // realisation count (exact below 2^53); one 128-byte line
  double* rcv;      // [mask_r+1][nrcv][2][10][2]
  double* samp;     // [mask_s+1][nsamp][2][2]
  unsigned long long* samp_count;    // [mask_s+1][nsamp]
  double* sxr;      // [mask_x+1][nsamp][nrcv][2][10][2]
  double* prim;     // [mask_p+1][nprim][2][10][2]
  uint32_t mask_g, mask_r, mask_s, mask_x, mask_p;
  uint32_t nrcv, nsamp, njobs;
  uint64_t nprim;
  // Log of finished paths' per-primitive entries (prim_add), reduced into
  // `prim` by key after sorting (k_prim_reduce): record i is the key
  // log_key[i] and the 10 fields' weights log_val[10 i ..]. log_n counts the
  // records asked for, also past log_cap (those went to atomics).
  uint32_t* log_key;
  double* log_val;
  unsigned long long* log_n;
  uint64_t log_cap;
};

struct SolveParams {
  uint64_t N;
  uint64_t max_failed;
  uint64_t key;
  uint64_t counter;
  uint32_t slots;
  uint32_t npend;
  // several suns in one solve: realisations [job_first[j], job_first[j+1])
  // run under job_suns[j]; without job_suns the one job is the scene's sun
  const uint64_t* job_first;
  const ssolh_sun* job_suns;
  uint32_t njobs;
};

// Job of realisation id: the last j with job_first[j] <= id (empty jobs are
// skipped because their successor starts at the same id)
__device__ __forceinline__ uint32_t job_of(const SolveParams& prm, uint64_t id) {
  if (prm.njobs <= 1) return 0;
  uint32_t lo = 0, hi = prm.njobs;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (prm.job_first[mid] <= id) lo = mid; else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ const ssolh_sun& job_sun(const DScene& S, const SolveParams& prm,
                                                     uint32_t job) {
  return prm.job_suns ? prm.job_suns[job] : S.sun;
}

// Adjoint solve (score-function gradients, see k_shade): null grad <=> off.
struct AdjParams {
  double* grad;              // [nparams][3]: d/d aim angle 1, 2, d/d slope error
  const int32_t* inst_param; // [ninstances]: parameter slot of an instance or -1
  const double* axes;        // [nparams][2][3]: world rotation axes of the aim angles
  const double* gmap;        // [nprim_tallies][2]: upstream weight per unit absorbed flux
  double d_total;            // upstream weight of absorbed_by_receivers
  double4* s;                // [slots]: scores of the path (aim 1, aim 2, slope), q
  uint32_t* p;               // [slots]: parameter slot of the path or NONE
};

// Path recording (src/ssol_solver.c:898-1127): null vtx <=> off. Vertices
// are appended in arrival order, tagged with their realisation and index.
struct TrackParams {
  ssolh_track_vertex* vtx;
  unsigned long long* nvtx;  // vertices produced (may exceed cap)
  unsigned long long cap;
  uint8_t* type;             // [N] SSOLH_PATH_* of each realisation
  uint32_t* seq;             // [slots] vertices recorded by the slot's path
  double4* pos;              // [slots] the point the slot's ray leaves
  double sun_len, inf_len;
};

__device__ void track_vertex(const TrackParams& K, uint32_t slot, uint64_t path, d3 p, double w) {
  const uint32_t s = K.seq[slot];
  K.seq[slot] = s + 1;
  const unsigned long long i = atomicAdd(K.nvtx, 1ull);
  if (i < K.cap) {
    ssolh_track_vertex v;
    v.path = path;
    v.index = s;
    v.pad__ = 0;
    v.pos[0] = p.x; v.pos[1] = p.y; v.pos[2] = p.z;
    v.weight = w;
    K.vtx[i] = v;
  }
}

// Work list: slot of entry q (a null list is the identity)
__device__ __forceinline__ uint32_t slot_of(const uint32_t* list, uint32_t q) {
  return list ? list[q] : q;
}

// -------------------------------------------------------------------------
// Wave and block helpers (64-wide waves)
// -------------------------------------------------------------------------
// Value of lane (i ^ pattern) within DPP rows, both halves of a double
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

__device__ __forceinline__ double read_lane(double v, int lane) {
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, lane);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), lane);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Sum over the 64 lanes of a wave; every lane must be active. DPP moves
// sum within quads (quad_perm [1,0,3,2], [2,3,0,1]), half rows
// (row_half_mirror) and rows of 16 (row_mirror); the four row sums are
// then read from lanes 0, 16, 32, 48.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_move<0xB1>(v);
  v += dpp_move<0x4E>(v);
  v += dpp_move<0x141>(v);
  v += dpp_move<0x140>(v);
  return (read_lane(v, 0) + read_lane(v, 16)) + (read_lane(v, 32) + read_lane(v, 48));
}

__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }

__device__ __forceinline__ void atomic_add(double* p, double v) {
  if (v != 0) unsafeAtomicAdd(p, v);
}

// Adds the NV values of each participating lane at base[key*NV + i]. Lanes
// sharing the first participant's key are reduced across the wave and
// flushed by one lane; the others use per-lane atomics. Must be called by
// the whole wave.
template <int NV, int STRIDE = NV>
__device__ void keyed_add(bool has, uint64_t key, const double* v, double* base) {
  const uint64_t mask = __ballot(has);
  if (!mask) return;
  const int leader = __ffsll((unsigned long long)mask) - 1;
  const uint64_t lkey = __shfl(key, leader, 64);
  const bool match = has && key == lkey;
  if (__popcll(__ballot(match)) > 1) {
    // several lanes share the key: one reduction, one set of atomics
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const double s = wave_sum(match ? v[i] : 0.0);
      if ((int)lane_id() == leader) atomic_add(base + lkey * STRIDE + i, s);
    }
    if (has && !match) {
#pragma unroll
      for (int i = 0; i < NV; ++i) atomic_add(base + key * STRIDE + i, v[i]);
    }
  } else if (has) {
    // keys all distinct so far: per-lane atomics, no reduction
#pragma unroll
    for (int i = 0; i < NV; ++i) atomic_add(base + key * STRIDE + i, v[i]);
  }
}

__device__ __forceinline__ unsigned long long wave_count_add(bool pred, unsigned long long* ctr) {
  const uint64_t mask = __ballot(pred);
  if (!mask) return 0;
  const int leader = __ffsll((unsigned long long)mask) - 1;
  unsigned long long base = 0;
  if ((int)lane_id() == leader) base = atomicAdd(ctr, (unsigned long long)__popcll(mask));
  base = __shfl(base, leader, 64);
  const uint64_t below = mask & ((1ull << lane_id()) - 1ull);
  return base + (unsigned long long)__popcll(below);
}

// A finished path's per-primitive entry: the key and the 10 fields' weights
// a[f] * factor. FP64 atomics issue at one device-wide rate whatever their
// addresses, and the entries of a wave fall on as many primitives as it has
// lanes, so the 20 atomics of an entry (x and x^2 per field) would cost far
// more than the 84 bytes of a record in the log, which k_prim_reduce sums
// by key after sorting. A full log takes the atomics. Must be called by the
// whole wave.
template <typename R>
__device__ void prim_add(const Tallies& T, bool has, uint32_t key, const R* a, double factor) {
  const unsigned long long pos = T.log_cap ? wave_count_add(has, T.log_n) : ~0ull;
  if (!has) return;
  if (pos < T.log_cap) {
    T.log_key[pos] = key;
    double2* v = (double2*)(T.log_val + pos * 10u);
#pragma unroll
    for (int f = 0; f < 5; ++f)
      v[f] = make_double2((double)a[2 * f] * factor, (double)a[2 * f + 1] * factor);
    return;
  }
  double* pb = T.prim + (size_t)(blockIdx.x & T.mask_p) * T.nprim * 40u + (size_t)key * 20u;
#pragma unroll
  for (int f = 0; f < 10; ++f) {
    const double x = (double)a[f] * factor;
    atomic_add(pb + 2 * f, x);
    atomic_add(pb + 2 * f + 1, x * x);
  }
}

// The sorted log (keys skey, record indices sidx, m records): each wave
// takes PRIM_ROUNDS rounds of 64 consecutive records. While a round's
// records all carry the key of the sums the wave holds, each lane adds its
// record's x and x^2 of the 10 fields to its own sums; at a change of key
// the sums are reduced across the wave and added with one set of 20
// atomics, and a round of several keys reduces each of its runs of equal
// keys so.
constexpr uint32_t PRIM_ROUNDS = 8;

__device__ __forceinline__ void prim_flush_run(double* pb, uint32_t key, const double* acc,
                                               int leader) {
#pragma unroll
  for (int f = 0; f < 20; ++f) {
    const double s = wave_sum(acc[f]);
    if ((int)lane_id() == leader) atomic_add(pb + (size_t)key * 20u + f, s);
  }
}

__global__ void __launch_bounds__(256) k_prim_reduce(const uint32_t* __restrict__ skey,
                                                     const uint32_t* __restrict__ sidx,
                                                     const double* __restrict__ val, uint32_t m,
                                                     double* prim, uint32_t mask_p, uint64_t nprim) {
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / 64u;
  const uint32_t base = wave * 64u * PRIM_ROUNDS;
  double* pb = prim + (size_t)(blockIdx.x & mask_p) * nprim * 40u;
  double acc[20];
#pragma unroll
  for (int f = 0; f < 20; ++f) acc[f] = 0.0;
  uint32_t acc_key = NONE;  // uniform over the wave: the key of acc
  for (uint32_t r = 0; r < PRIM_ROUNDS; ++r) {
    const uint32_t i = base + r * 64u + lane_id();
    const bool in = i < m;
    const uint64_t inmask = __ballot(in);
    if (!inmask) break;  // lane 0 holds the round's first record from here on
    const uint32_t key = in ? skey[i] : 0u;
    double x[10];
    if (in) {
      const double2* v = (const double2*)(val + (size_t)sidx[i] * 10u);
#pragma unroll
      for (int f = 0; f < 5; ++f) {
        const double2 t = v[f];
        x[2 * f] = t.x;
        x[2 * f + 1] = t.y;
      }
    } else {
#pragma unroll
      for (int f = 0; f < 10; ++f) x[f] = 0.0;
    }
    const uint32_t k0 = __shfl(key, 0, 64);
    const bool one_key = __ballot(in && key == k0) == inmask;
    if (!(one_key && k0 == acc_key) && acc_key != NONE) {
      prim_flush_run(pb, acc_key, acc, 0);
#pragma unroll
      for (int f = 0; f < 20; ++f) acc[f] = 0.0;
      acc_key = NONE;
    }
    if (one_key) {
      acc_key = k0;
#pragma unroll
      for (int f = 0; f < 10; ++f) {
        acc[2 * f] += x[f];
        acc[2 * f + 1] += x[f] * x[f];
      }
      continue;
    }
    // several keys in the round: each run of equal keys on its own
    bool todo = in;
    for (uint64_t mask = __ballot(todo); mask; mask = __ballot(todo)) {
      const int leader = __ffsll((unsigned long long)mask) - 1;
      const uint32_t lkey = __shfl(key, leader, 64);
      const bool match = todo && key == lkey;
      double run[20];
#pragma unroll
      for (int f = 0; f < 10; ++f) {
        run[2 * f] = match ? x[f] : 0.0;
        run[2 * f + 1] = match ? x[f] * x[f] : 0.0;
      }
      prim_flush_run(pb, lkey, run, leader);
      todo = todo && !match;
    }
  }
  if (acc_key != NONE) prim_flush_run(pb, acc_key, acc, 0);
}

__global__ void k_iota32(uint32_t* a, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = i;
}

// The logical blocks (of BLOCK list entries) of a looping, graph-captured
// launch: the first 7/8 in static shares (block b to workgroup b mod grid),
// the rest handed out in chunks from *fetch, which k_scan_top zeroes before
// the next iteration. With static shares alone the kernel ends with its
// slowest share; the handed-out tail lets the workgroups finish within about
// one chunk of each other. A chunk is sized for about four tickets per
// workgroup: one block per ticket made the atomics on *fetch (one address)
// the bottleneck when the list stays full. Calls f(b) once for every block;
// every thread of the workgroup must call it.
template <typename F>
__device__ __forceinline__ void for_each_block(uint32_t nb, unsigned* fetch, F&& f) {
  __shared__ uint32_t s_b;
  const uint32_t stat = nb - nb / 8;
  const uint32_t chunk = max(1u, (nb - stat) / (gridDim.x * 4u));
  uint32_t b = blockIdx.x, end = 0;  // uniform over the workgroup; [b, end): current chunk
  for (;;) {
    if (b >= stat && b >= end) {  // static share or chunk done: a new ticket
      __syncthreads();  // every thread has read the previous ticket
      if (threadIdx.x == 0) s_b = stat + atomicAdd(fetch, 1u) * chunk;
      __syncthreads();
      b = s_b;
      if (b >= nb) break;
      end = min(b + chunk, nb);
    }
    f(b);  // one call site: the block body is inlined once
    b = b < stat ? b + gridDim.x : b + 1;
  }
}

// Number of set predicates in the block, written by thread 0 to out[b]
// (b: the logical block). Every thread of the block must call it.
__device__ void block_count(bool pred, uint32_t* out, uint32_t b) {
  __shared__ uint32_t wcount[WAVES];
  const uint64_t m = __ballot(pred);
  if (lane_id() == 0) wcount[threadIdx.x / 64] = (uint32_t)__popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t s = 0;
    for (int w = 0; w < WAVES; ++w) s += wcount[w];
    out[b] = s;
  }
  __syncthreads();  // wcount is reused by the block's next logical block
}

// -------------------------------------------------------------------------
// Traversal with the reference hit filter
// -------------------------------------------------------------------------
// Ray of the box tests: the distance to the plane x = p is p * ix + ox (one
// FMA), and the rows of a node that hold the near planes of the ray's
// direction signs are read by offset, so a slab needs no min/max. The
// rounding of ox = -o * ix puts up to |ox| 2^-24 of error on every plane
// distance, which with km-scale origins is far more than the (p - o) * ix
// form's error near the origin (about 6e-5 m against 1e-6 m at 1.5 km):
// a box is crossed when tnear - slack <= tfar, slack = 2^-21 max |o*|
// (SLAB_SLACK), so the test culls no box the subtracting form keeps. ox, oy, oz are kept
// finite; a ray with a non-finite origin or direction gets
// ox = oy = oz = +inf, so tnear - slack is NaN and it crosses no box.
struct FRay {
  float ix, iy, iz;      // 1 / direction (components at least 1e-30 in magnitude)
  float ox, oy, oz;      // -origin / direction
  uint32_t nx, ny, nz;   // byte offset in a Node4 of each axis's near-plane row
};

__device__ __forceinline__ FRay make_fray32(float ox, float oy, float oz, float dx, float dy,
                                            float dz) {
  auto inv = [](float f) { return 1.0f / (fabsf(f) > 1e-30f ? f : copysignf(1e-30f, f)); };
  auto off = [](float o, float i) { return fminf(fmaxf(-o * i, -FLT_MAX), FLT_MAX); };
  FRay r;
  r.ix = inv(dx);
  r.iy = inv(dy);
  r.iz = inv(dz);
  r.ox = off(ox, r.ix);
  r.oy = off(oy, r.iy);
  r.oz = off(oz, r.iz);
  r.nx = r.ix < 0.f ? 16u : 0u;
  r.ny = r.iy < 0.f ? 48u : 32u;
  r.nz = r.iz < 0.f ? 80u : 64u;
  const bool finite = fabsf(ox) <= FLT_MAX && fabsf(oy) <= FLT_MAX && fabsf(oz) <= FLT_MAX &&
                      fabsf(dx) <= FLT_MAX && fabsf(dy) <= FLT_MAX && fabsf(dz) <= FLT_MAX;
  if (!finite) {
    r.ix = r.iy = r.iz = 1.f;
    r.ox = r.oy = r.oz = INFINITY;
  }
  return r;
}

__device__ __forceinline__ FRay make_fray(d3 o, d3 d) {
  return make_fray32((float)o.x, (float)o.y, (float)o.z, (float)d.x, (float)d.y, (float)d.z);
}

// Ray in the shear frame of the watertight ray/triangle test (Woop,
// Benthin, Wald, JCGT 2013): computed once per ray (and per instance).
struct Shear {
  int kx, ky, kz;
  double Sx, Sy, Sz;
};

__device__ __forceinline__ Shear make_shear(d3 d) {
  const double a[3] = {fabs(d.x), fabs(d.y), fabs(d.z)};
  const double v[3] = {d.x, d.y, d.z};
  Shear s;
  s.kz = a[0] > a[1] ? (a[0] > a[2] ? 0 : 2) : (a[1] > a[2] ? 1 : 2);
  s.kx = (s.kz + 1) % 3;
  s.ky = (s.kx + 1) % 3;
  if (v[s.kz] < 0) { const int t = s.kx; s.kx = s.ky; s.ky = t; }
  s.Sx = v[s.kx] / v[s.kz];
  s.Sy = v[s.ky] / v[s.kz];
  s.Sz = 1.0 / v[s.kz];
  return s;
}

__device__ __forceinline__ double comp(d3 p, int k) { return k == 0 ? p.x : (k == 1 ? p.y : p.z); }

// Watertight: a ray through a shared edge or vertex hits at least one of
// the triangles. The edge functions of a shared edge must come out as exact
// negations of each other in the two triangles, so no product here may be
// fused into an FMA. Returns the distance or -1.
__device__ __forceinline__ double ray_triangle(d3 o, const Shear& s, d3 v0, d3 v1, d3 v2) {
#pragma clang fp contract(off)
  const d3 A = v0 - o, B = v1 - o, C = v2 - o;
  const double Az = comp(A, s.kz), Bz = comp(B, s.kz), Cz = comp(C, s.kz);
  const double Ax = comp(A, s.kx) - s.Sx * Az, Ay = comp(A, s.ky) - s.Sy * Az;
  const double Bx = comp(B, s.kx) - s.Sx * Bz, By = comp(B, s.ky) - s.Sy * Bz;
  const double Cx = comp(C, s.kx) - s.Sx * Cz, Cy = comp(C, s.ky) - s.Sy * Cz;
  const double U = Cx * By - Cy * Bx;
  const double V = Ax * Cy - Ay * Cx;
  const double W = Bx * Ay - By * Ax;
  if ((U < 0 || V < 0 || W < 0) && (U > 0 || V > 0 || W > 0)) return -1;
  const double det = U + V + W;
  if (det == 0) return -1;
  const double T = U * s.Sz * Az + V * s.Sz * Bz + W * s.Sz * Cz;
  return T / det;
}

struct TraceState {
  double tbest;
  double qdst;     // nearest refined quadric distance
  uint32_t qiss;   // its (instance, shaded shape): the normal is rebuilt after the traversal
  uint32_t tri;  // the best triangle (its world normal is rebuilt after the traversal)
  uint32_t iss, prim;
  bool punched;  // the best triangle belongs to a punched surface
};

// World geometry normal (unnormalised) of mesh triangle `tri` of instance `inst`
__device__ __forceinline__ d3 mesh_ng(const DScene& S, uint32_t inst, const ssolh_triangle& tri) {
  const d3 v0 = ld3(tri.v0), v1 = ld3(tri.v1), v2 = ld3(tri.v2);
  return tmul(S.insts[inst].inv_linear, cross(v2 - v0, v1 - v0));
}

// Returns true if the candidate is accepted (src/ssol_scene.c:438-536).
__device__ bool filter(const DScene& S, const RayRec& r, d3 o, d3 d, uint32_t iss_id,
                       const ssolh_triangle& tri, double t, TraceState& ts, d3& ng_out,
                       bool& punched) {
  const uint32_t side_from = (r.flags >> RF_SIDE_SHIFT) & 3u;
  const IssInfo is = S.iss[iss_id];
  uint32_t side;
  double dst = 0;
  d3 qn = mk(0, 0, 0);
  if (t < r.tmin) return false;
  punched = is.shape_type != SSOLH_SHAPE_MESH;
  if (!punched) {
    if (t <= 1.e-6 || t <= r.tmin || (tri.prim == r.prim_from && iss_id == r.iss_from))
      return false;
    ng_out = mesh_ng(S, is.inst, tri);
    side = dot(ng_out, d) < 0 ? SSOL_FRONT_ : SSOL_BACK_;
  } else {
    const ssolh_inst_sshape& q = S.isshapes[iss_id];
    const ssolh_shape& sh = S.shapes[is.shape];
    const d3 lo = mul(q.R_inv, o - ld3(q.T));
    const d3 ldir = mul(q.R_inv, d);
    d3 grad;
    if (!quadric_hit(sh.quadric, sh.qparam, lo, ldir, t, dst, grad)) return false;
    if (!(dst < (double)FLT_MAX)) return false;
    if ((double)(float)dst <= r.tmin) return false;
    qn = normalize(tmul(q.R_inv, grad));
    side = dot(d, qn) < 0 ? SSOL_FRONT_ : SSOL_BACK_;
    if (iss_id == r.iss_from && side != side_from) return false;
  }
  if (r.flags & RF_REVERSED) side = side == SSOL_FRONT_ ? SSOL_BACK_ : SSOL_FRONT_;
  const ssolh_material& m = S.mtls[side == SSOL_FRONT_ ? is.mtl_front : is.mtl_back];
  if (m.type == SSOLH_MTL_VIRTUAL) {
    if (r.flags & RF_DISCARD_VIRTUAL) return false;
    if (!(is.receiver_mask & side)) return false;
  }
  if (punched && ts.qdst >= dst) {
    ts.qdst = dst;
    ts.qiss = iss_id;
  }
  return true;
}

// Distance at which the ray (o, d) leaves box b (lo xyz, hi xyz); an axis
// the ray runs parallel to bounds nothing.
__device__ double box_exit(const float* b, d3 o, d3 d) {
  const double oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
  double t = INFINITY;
  for (int a = 0; a < 3; ++a) {
    if (dd[a] == 0) continue;
    const double t0 = ((double)b[a] - oo[a]) / dd[a], t1 = ((double)b[3 + a] - oo[a]) / dd[a];
    t = fmin(t, fmax(t0, t1));
  }
  return t;
}

// Distance at which the world ray (o, d) leaves the object box of instance
// inst_id (in the instance's object space, where the trace's triangle
// distances are taken)
__device__ double object_exit(const DScene& S, uint32_t inst_id, d3 o, d3 d) {
  const ssolh_instance& inst = S.insts[inst_id];
  const d3 T = mk(inst.transform[9], inst.transform[10], inst.transform[11]);
  return box_exit(S.obj_box + 6 * (size_t)inst.object, mul(inst.inv_linear, o - T),
                  mul(inst.inv_linear, d));
}

// Whether root k of the n roots t[] of a ray / quadric equation can be the
// one the filter's refinement takes for a candidate at a distance in
// [0, t_hi]: the root nearest the candidate's distance
__device__ bool root_choosable(const double t[2], int n, int k, double t_hi) {
  if (n == 1) return true;
  const double other = t[1 - k];
  if (!(fabs(other) < (double)FLT_MAX) || t[k] <= other) return true;
  return !(t_hi * 1.001 + 1e-3 < 0.5 * (t[k] + other));
}

// Whether the filter can accept a candidate on (instance, shaded shape)
// iss_id, a punched shape, for a ray that leaves it at tmin 0 (world ray o,
// d) and whose candidates on it lie within distance t_hi. The filter
// refines every candidate of the shape to the root of the same ray /
// quadric equation nearest the candidate's distance, and accepts it only
// past tmin and on the side the ray left. A root that passes these checks
// (computed here exactly as the filter computes them) can be chosen when it
// is the only root, when it is the nearer one, or when the candidates reach
// the midpoint between the roots; otherwise no candidate of the shape can be
// accepted.
__device__ bool punched_rehit_possible(const DScene& S, const RayRec& r, d3 o, d3 d,
                                       uint32_t iss_id, double t_hi) {
  const uint32_t side_from = (r.flags >> RF_SIDE_SHIFT) & 3u;
  const IssInfo is = S.iss[iss_id];
  const ssolh_inst_sshape& q = S.isshapes[iss_id];
  const ssolh_shape& sh = S.shapes[is.shape];
  const d3 lo = mul(q.R_inv, o - ld3(q.T));
  const d3 ldir = mul(q.R_inv, d);
  double a, b, c, t[2];
  if (!quadric_coeffs(sh.quadric, sh.qparam, lo, ldir, a, b, c)) return false;
  const int n = quadric_roots(a, b, c, t);
  for (int k = 0; k < n; ++k) {
    if (!(t[k] < (double)FLT_MAX)) continue;
    if ((double)(float)t[k] <= r.tmin) continue;
    const d3 grad = quadric_grad(sh.quadric, sh.qparam, lo + ldir * t[k]);
    const d3 qn = normalize(tmul(q.R_inv, grad));
    const uint32_t side = dot(d, qn) < 0 ? SSOL_FRONT_ : SSOL_BACK_;
    if (side != side_from) continue;
    if (root_choosable(t, n, k, t_hi)) return true;
  }
  return false;
}

// Unit world normal of the quadric of (instance, shaded shape) `iss` where the
// world ray (o, d) meets it at distance dst: the gradient of quadric_hit's
// refinement at the same point, as the filter took it
__device__ d3 quadric_normal(const DScene& S, uint32_t iss, d3 o, d3 d, double dst) {
  const ssolh_inst_sshape& q = S.isshapes[iss];
  const ssolh_shape& sh = S.shapes[S.iss[iss].shape];
  const d3 lo = mul(q.R_inv, o - ld3(q.T));
  const d3 ldir = mul(q.R_inv, d);
  return normalize(tmul(q.R_inv, quadric_grad(sh.quadric, sh.qparam, lo + ldir * dst)));
}

typedef float v4f __attribute__((ext_vector_type(4)));
typedef uint32_t v4u __attribute__((ext_vector_type(4)));

// What a visit reads of a node: the near- and far-plane rows of the ray's
// direction signs and the child references
struct NodeRows {
  v4f n[3], f[3];
  v4u child;
};

// Rows of node `ref` of the device array (32-bit byte offsets from the
// array's base: the loads take the base in scalar registers)
__device__ __forceinline__ NodeRows node_rows(const Node4* __restrict__ nodes, uint32_t ref,
                                              const FRay& r) {
  const char* b = (const char*)nodes;
  const uint32_t base = ref * (uint32_t)sizeof(Node4);
  const uint32_t o[3] = {base + r.nx, base + r.ny, base + r.nz};
  NodeRows w;
  for (int a = 0; a < 3; ++a) {
    w.n[a] = *(const v4f*)(b + o[a]);
    w.f[a] = *(const v4f*)(b + (o[a] ^ 16u));
  }
  w.child = *(const v4u*)(b + base + 96u);
  return w;
}

// The same from the LDS copy of the top levels, through an LDS (address
// space 3) pointer: the loads are ds_reads, where a pointer that may address
// either LDS or global memory would make them flat accesses
__device__ __forceinline__ NodeRows lds_rows(const Node4* lds, uint32_t i, const FRay& r) {
  typedef const __attribute__((address_space(3))) char* lds_b;
  const lds_b b = (lds_b)lds;
  const uint32_t base = i * (uint32_t)sizeof(Node4);
  const uint32_t o[3] = {base + r.nx, base + r.ny, base + r.nz};
  typedef const __attribute__((address_space(3))) v4f* lf;
  typedef const __attribute__((address_space(3))) v4u* lu;
  NodeRows w;
  for (int a = 0; a < 3; ++a) {
    w.n[a] = *(lf)(b + o[a]);
    w.f[a] = *(lf)(b + (o[a] ^ 16u));
  }
  w.child = *(lu)(b + base + 96u);
  return w;
}

// Traversal stack of one lane: its first STACK entries in LDS (stride
// BLOCK), the rest, when the scene's bound exceeds STACK, in a global array
// of `extra` entries per work-list position (stride: the list's capacity)
struct TStack {
  uint32_t* lds;
  uint32_t* ovf;   // this lane's first overflow entry, or null
  uint32_t stride, extra;
  int sp;
};

template <int STACK>
__device__ __forceinline__ void tpush(TStack& s, uint32_t v) {
  if (s.sp < STACK) s.lds[s.sp * BLOCK] = v;
  else if ((uint32_t)(s.sp - STACK) < s.extra) s.ovf[(size_t)(s.sp - STACK) * s.stride] = v;
  ++s.sp;
}

// false: the stack is empty (or, never with the host's bound, past its end)
template <int STACK>
__device__ __forceinline__ bool tpop(TStack& s, uint32_t& ref) {
  if (s.sp == 0) return false;
  --s.sp;
  if (s.sp < STACK) {
    ref = s.lds[s.sp * BLOCK];
  } else {
    if ((uint32_t)(s.sp - STACK) >= s.extra) return false;
    ref = s.ovf[(size_t)(s.sp - STACK) * s.stride];
  }
  return true;
}

// One visit of a 4-wide inner node: the children whose boxes the ray
// crosses in [tmin, tmax], nearest first; the nearest becomes `ref`, the
// others are pushed farthest first (so the next nearest is on top). With no
// child crossed, `ref` is popped; false when the stack is empty.
template <int STACK>
__device__ __forceinline__ bool visit4(const NodeRows& w, const FRay& fr, float tmin, float tmax,
                                       float slack_scale, uint32_t& ref, TStack& st) {
  float k[4];
  uint32_t r[4];
  const float slack = slack_scale * fmaxf(fabsf(fr.ox), fmaxf(fabsf(fr.oy), fabsf(fr.oz)));
  for (int c = 0; c < 4; ++c) {
    const float tn = fmaxf(fmaxf(fmaf(w.n[0][c], fr.ix, fr.ox), fmaf(w.n[1][c], fr.iy, fr.oy)),
                           fmaxf(fmaf(w.n[2][c], fr.iz, fr.oz), tmin));
    const float tf = fminf(fminf(fmaf(w.f[0][c], fr.ix, fr.ox), fmaf(w.f[1][c], fr.iy, fr.oy)),
                           fminf(fmaf(w.f[2][c], fr.iz, fr.oz), tmax));
    k[c] = tn - slack <= tf ? tn : INFINITY;
    r[c] = w.child[c];
  }
  // sorting network (0,1)(2,3)(0,2)(1,3)(1,2)
  auto cs = [&](int a, int b) {
    const bool sw = k[b] < k[a];
    const float ka = k[a], kb = k[b];
    const uint32_t ra = r[a], rb = r[b];
    k[a] = sw ? kb : ka; k[b] = sw ? ka : kb;
    r[a] = sw ? rb : ra; r[b] = sw ? ra : rb;
  };
  cs(0, 1); cs(2, 3); cs(0, 2); cs(1, 3); cs(1, 2);
  const int cnt = (k[0] < INFINITY) + (k[1] < INFINITY) + (k[2] < INFINITY) + (k[3] < INFINITY);
  if (cnt == 0) return tpop<STACK>(st, ref);
  if (cnt >= 4) tpush<STACK>(st, r[3]);
  if (cnt >= 3) tpush<STACK>(st, r[2]);
  if (cnt >= 2) tpush<STACK>(st, r[1]);
  ref = r[0];
  return true;
}

// Overflow of the traversal stacks past their LDS part: `extra` entries per
// work-list position, `stride` = the list's capacity (null when the scene's
// stack bound fits in LDS)
struct TOvf {
  uint32_t* base;
  uint32_t stride, extra;
};

// STACK: LDS stack entries per lane (the smaller the stack, the more blocks
// share a CU); the scene's bound beyond them goes to the overflow array. Four
// waves per SIMD caps the kernel at 128 VGPRs; the few values that spill are
// cold.
// Entry q of the work list: traversal, filter, hit record
template <int STACK>
__device__ __forceinline__ void trace_ray(const DScene& S, const uint32_t* __restrict__ list,
                                          uint32_t q, const RayRec* __restrict__ rays,
                                          HitRec* __restrict__ hits, uint32_t* lds_stack,
                                          const Node4* lds_top, const TOvf& ovf) {

  const uint32_t slot = slot_of(list, q);
  const RayRec r = rays[slot];
  const d3 ow = mk(r.ox, r.oy, r.oz), dw = mk(r.dx, r.dy, r.dz);
  const bool any_hit = (r.flags & RF_ANY_HIT) != 0;
  TraceState ts;
  ts.tbest = INFINITY;
  ts.qdst = INFINITY;
  ts.qiss = NONE;
  ts.tri = 0;
  ts.iss = NONE;
  ts.prim = 0;
  ts.punched = false;

  TStack st{lds_stack + threadIdx.x, ovf.base ? ovf.base + q : nullptr, ovf.stride, ovf.extra, 0};
  FRay fr = make_fray(ow, dw);
  d3 o = ow, d = dw;        // current-level ray (double)
  Shear sh = make_shear(dw);
  bool in_blas = false;
  uint32_t iss_first = 0;   // (instance, sshape) base of the BLAS being traversed
  const float ftmin = r.tmin;
  // A ray that leaves a point of a planar object can meet that object again
  // only at t ~ 0, which the filter rejects: its instance is not entered.
  // Nor is the instance of a one-shape object whose punched shape the ray
  // leaves when the filter could accept no candidate on it.
  uint32_t skip_inst = NONE;
  if (r.tmin == 0.f) {
    const IssInfo e = S.iss[r.iss_from];
    if ((e.obj_flags & SSOLH_OBJECT_PLANAR) ||
        ((e.obj_flags & OBJ_ONE_SHAPE) && e.shape_type != SSOLH_SHAPE_MESH &&
         !punched_rehit_possible(S, r, ow, dw, r.iss_from, object_exit(S, e.inst, ow, dw))))
      skip_inst = e.inst;
  }

  // "while-while" traversal: the lanes of a wave first descend through inner
  // nodes together, then process their leaves together, which keeps the
  // node, instance and triangle code paths from interleaving per lane.
  // SENTINEL (all bits set) reads as a leaf reference and ends the descent.
  uint32_t ref = S.root_ref;
  bool alive = true;
  // box culling distance: the best hit so far, widened by 2 ulp (FP32)
  float tmax = FLT_MAX;
  while (alive) {
    while (!(ref & LEAF)) {
      // the whole node in one round of loads (LDS or global memory, each
      // through its own address space), then its four boxes
      const NodeRows n =
          (ref & TOP_REF) ? lds_rows(lds_top, ref & 0x0FFFFFFFu, fr) : node_rows(S.nodes, ref, fr);
      if (!visit4<STACK>(n, fr, ftmin, tmax, S.slack_scale, ref, st)) {
        alive = false;
        break;
      }
    }
    if (!alive) break;
    if (ref == SENTINEL) {
      // back to the TLAS
      in_blas = false;
      fr = make_fray(ow, dw);
      o = ow;
      d = dw;
      alive = tpop<STACK>(st, ref);
      continue;
    }
    const uint32_t first = ref & 0x0FFFFFFFu;
    if (!in_blas) {
      // TLAS leaf: one instance, enter its BLAS
      const uint32_t ii = S.tlas_inst[first];
      if (ii == skip_inst) {
        alive = tpop<STACK>(st, ref);
        continue;
      }
      const ssolh_instance& inst = S.insts[ii];
      const d3 T = mk(inst.transform[9], inst.transform[10], inst.transform[11]);
      o = mul(inst.inv_linear, ow - T);
      d = mul(inst.inv_linear, dw);
      fr = make_fray(o, d);
      sh = make_shear(d);
      iss_first = inst.isshape_first;
      in_blas = true;
      tpush<STACK>(st, SENTINEL);
      ref = S.objects[inst.object].blas_root;
      continue;
    }
    const uint32_t cnt = ((ref >> 28) & 7u) + 1u;
    bool done = false;
    for (uint32_t k = 0; k < cnt; ++k) {
      const ssolh_triangle tri = S.tris[first + k];
      const double t = ray_triangle(o, sh, ld3(tri.v0), ld3(tri.v1), ld3(tri.v2));
      if (!(t >= r.tmin && t < ts.tbest)) continue;
      d3 ng;
      bool punched;
      const uint32_t iss = iss_first + tri.sshape;
      if (!filter(S, r, ow, dw, iss, tri, t, ts, ng, punched)) continue;
      ts.tbest = t;
      tmax = t < (double)FLT_MAX ? (float)t * 1.0000002f + 1e-30f : FLT_MAX;
      ts.iss = iss;
      ts.prim = tri.prim;
      ts.tri = first + k;
      ts.punched = punched;
      if (any_hit) { done = true; break; }
    }
    if (done) break;
    alive = tpop<STACK>(st, ref);
  }

  HitRec h;
  h.t = ts.tbest;
  h.qdst = ts.qdst;
  d3 n = mk(0, 0, 0);
  if (ts.punched)
    n = quadric_normal(S, ts.qiss, ow, dw, ts.qdst);
  else if (ts.iss != NONE)
    n = mesh_ng(S, S.iss[ts.iss].inst, S.tris[ts.tri]);
  h.nx = n.x; h.ny = n.y; h.nz = n.z;
  h.iss = ts.iss;
  h.prim = ts.prim;
  hits[slot] = h;
}

// LOOP: graph-captured launch of a fixed grid, each block taking every
// gridDim.x-th logical block of BLOCK list entries; else one logical block
// per block
template <int STACK, bool LOOP>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4)))
k_trace(DScene S, const uint32_t* __restrict__ list, const unsigned* __restrict__ count_ptr,
        const RayRec* __restrict__ rays, HitRec* __restrict__ hits, unsigned* fetch, TOvf ovf) {
  __shared__ uint32_t lds_stack[STACK * BLOCK];
  __shared__ Node4 lds_top[TOP_NODES];
  const uint32_t count = *count_ptr;
  if (blockIdx.x * BLOCK >= count) return;  // whole blocks only: the staging syncs
  for (uint32_t i = threadIdx.x; i < S.ntop; i += BLOCK) lds_top[i] = S.top[i];
  __syncthreads();
  if (LOOP) {
    for_each_block((count + BLOCK - 1) / BLOCK, fetch, [&](uint32_t b) {
      const uint32_t q = b * BLOCK + threadIdx.x;
      if (q < count) trace_ray<STACK>(S, list, q, rays, hits, lds_stack, lds_top, ovf);
    });
  } else {
    const uint32_t q = blockIdx.x * BLOCK + threadIdx.x;
    if (q < count) trace_ray<STACK>(S, list, q, rays, hits, lds_stack, lds_top, ovf);
  }
}

// -------------------------------------------------------------------------
// TLAS refit for moved instances (the reference rebuilds its scene views on
// every solve, src/ssol_scene.c:346-433): the 4-wide TLAS keeps its
// topology and takes the world boxes of the instances' new frames, level by
// level from the deepest, in one workgroup; its LDS-staged copy of the top
// levels follows. *cost = sum over the TLAS's child boxes of their surface
// areas (x*y + y*z + z*x), the quality measure the host compares with the
// tree's at build time.
// -------------------------------------------------------------------------
constexpr int REFIT_THREADS = 1024;

// World box of instance ii: the 8 corners of its object's box under the
// instance transform, in FP64 and in the order of the host's TLAS build,
// rounded to FP32, widened by 2e-6 of its largest coordinate + 1e-6 (an
// empty object gets the degenerate box at the origin, widened alike)
__device__ void instance_world_box(const ssolh_instance& in, const float* ob, float lo[3],
                                   float hi[3]) {
  for (int k = 0; k < 3; ++k) { lo[k] = FLT_MAX; hi[k] = -FLT_MAX; }
  if (ob[0] > ob[3]) {
    for (int k = 0; k < 3; ++k) lo[k] = hi[k] = 0.f;
  } else {
    for (int c = 0; c < 8; ++c) {
      const double p[3] = {(double)((c & 1) ? ob[3] : ob[0]), (double)((c & 2) ? ob[4] : ob[1]),
                           (double)((c & 4) ? ob[5] : ob[2])};
      for (int k = 0; k < 3; ++k) {
        const double w = __dadd_rn(__dadd_rn(__dadd_rn(__dmul_rn(in.transform[k], p[0]),
                                                       __dmul_rn(in.transform[3 + k], p[1])),
                                             __dmul_rn(in.transform[6 + k], p[2])),
                                   in.transform[9 + k]);
        lo[k] = fminf(lo[k], (float)w);
        hi[k] = fmaxf(hi[k], (float)w);
      }
    }
  }
  float m = 0.f;
  for (int k = 0; k < 3; ++k) m = fmaxf(m, fmaxf(fabsf(lo[k]), fabsf(hi[k])));
  const float e = __fadd_rn(__fmul_rn(m, 2e-6f), 1e-6f);
  for (int k = 0; k < 3; ++k) { lo[k] = __fsub_rn(lo[k], e); hi[k] = __fadd_rn(hi[k], e); }
}

__global__ void __launch_bounds__(REFIT_THREADS)
k_tlas_refit(const ssolh_instance* __restrict__ insts, const uint32_t* __restrict__ tlas_inst,
             const float* __restrict__ obj_box, Node4* nodes, const uint32_t* __restrict__ order,
             const uint32_t* __restrict__ level_end, uint32_t nlevels, Node4* top,
             const uint32_t* __restrict__ top_src, uint32_t ntop, double* cost) {
  __shared__ double part[REFIT_THREADS];
  double acc = 0;
  uint32_t begin = 0;
  for (uint32_t l = 0; l < nlevels; ++l) {
    const uint32_t end = level_end[l];
    for (uint32_t i = begin + threadIdx.x; i < end; i += REFIT_THREADS) {
      Node4& n = nodes[order[i]];
      for (int c = 0; c < 4; ++c) {
        const uint32_t ref = n.child[c];
        if (ref == EMPTY4) continue;
        float lo[3], hi[3];
        if (ref & LEAF) {
          const ssolh_instance& in = insts[tlas_inst[ref & 0x0FFFFFFFu]];
          instance_world_box(in, obj_box + 6 * (size_t)in.object, lo, hi);
        } else {
          const Node4& m = nodes[ref];
          for (int a = 0; a < 3; ++a) {
            lo[a] = fminf(fminf(m.box[a][0][0], m.box[a][0][1]), fminf(m.box[a][0][2], m.box[a][0][3]));
            hi[a] = fmaxf(fmaxf(m.box[a][1][0], m.box[a][1][1]), fmaxf(m.box[a][1][2], m.box[a][1][3]));
          }
        }
        for (int a = 0; a < 3; ++a) {
          n.box[a][0][c] = lo[a];
          n.box[a][1][c] = hi[a];
        }
        const double x = (double)hi[0] - lo[0], y = (double)hi[1] - lo[1],
                     z = (double)hi[2] - lo[2];
        acc += x * y + y * z + z * x;
      }
    }
    begin = end;
    __syncthreads();  // the next level reads this one's boxes
  }
  for (uint32_t i = threadIdx.x; i < ntop; i += REFIT_THREADS)
    for (int a = 0; a < 3; ++a)
      for (int sd = 0; sd < 2; ++sd)
        for (int c = 0; c < 4; ++c) top[i].box[a][sd][c] = nodes[top_src[i]].box[a][sd][c];
  part[threadIdx.x] = acc;
  __syncthreads();
  for (uint32_t w = REFIT_THREADS / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *cost = part[0];
}

// -------------------------------------------------------------------------
// TLAS build on the GPU (linear BVH, T. Karras, "Maximizing parallelism in
// the construction of BVHs, octrees, and k-d trees", HPG 2012): when moved
// instances have made a refit tree too loose, the TLAS is built anew over
// the instances' new world boxes without the host. Keys: 30-bit Morton codes
// of the box centres in the centres' bounds (ties broken by sorted position);
// a radix sort of (key, instance); the binary radix tree; boxes bottom-up;
// every BVH2 node at even depth becomes a 4-wide node over its children and
// grandchildren. Two single-workgroup kernels around the sort: the TLAS of a
// field holds tens of thousands of instances at most, which one workgroup
// handles in microseconds.
// -------------------------------------------------------------------------
constexpr int LBVH_THREADS = 1024;
constexpr int LBVH_MAX_LEVELS = 66;  // 4-wide levels of a tree of 30-bit keys + 32-bit ties

struct LbvhResult {
  uint32_t nodes4;  // 4-wide nodes written
  uint32_t depth4;  // deepest 4-wide level (root 0)
  uint32_t ntop;    // nodes of the LDS copy
  uint32_t pad;
  double cost;      // summed child-box surface area
};

__device__ __forceinline__ uint32_t spread10(uint32_t x) {
  x &= 0x3FFu;
  x = (x | (x << 16)) & 0x030000FFu;
  x = (x | (x << 8)) & 0x0300F00Fu;
  x = (x | (x << 4)) & 0x030C30C3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}

__global__ void __launch_bounds__(LBVH_THREADS)
k_lbvh_keys(const ssolh_instance* __restrict__ insts, const float* __restrict__ obj_box,
            uint32_t n, float* __restrict__ lbox, uint32_t* __restrict__ key,
            uint32_t* __restrict__ val) {
  __shared__ float red[6][LBVH_THREADS];
  float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  for (uint32_t i = threadIdx.x; i < n; i += LBVH_THREADS) {
    float lo[3], hi[3];
    instance_world_box(insts[i], obj_box + 6 * (size_t)insts[i].object, lo, hi);
    for (int a = 0; a < 3; ++a) {
      lbox[6 * (size_t)i + a] = lo[a];
      lbox[6 * (size_t)i + 3 + a] = hi[a];
      const float cen = 0.5f * (lo[a] + hi[a]);
      mn[a] = fminf(mn[a], cen);
      mx[a] = fmaxf(mx[a], cen);
    }
  }
  for (int a = 0; a < 3; ++a) { red[a][threadIdx.x] = mn[a]; red[3 + a][threadIdx.x] = mx[a]; }
  __syncthreads();
  for (uint32_t w = LBVH_THREADS / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int a = 0; a < 3; ++a) {
        red[a][threadIdx.x] = fminf(red[a][threadIdx.x], red[a][threadIdx.x + w]);
        red[3 + a][threadIdx.x] = fmaxf(red[3 + a][threadIdx.x], red[3 + a][threadIdx.x + w]);
      }
    __syncthreads();
  }
  float lo[3], sc[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = red[a][0];
    const float ext = red[3 + a][0] - red[a][0];
    sc[a] = ext > 0.f ? 1023.f / ext : 0.f;
  }
  for (uint32_t i = threadIdx.x; i < n; i += LBVH_THREADS) {
    uint32_t q[3];
    for (int a = 0; a < 3; ++a) {
      const float cen = 0.5f * (lbox[6 * (size_t)i + a] + lbox[6 * (size_t)i + 3 + a]);
      q[a] = (uint32_t)fminf(fmaxf((cen - lo[a]) * sc[a], 0.f), 1023.f);
    }
    key[i] = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
    val[i] = i;
  }
}

// Common-prefix length of sorted keys i and j (-1 out of range; equal keys
// compare their positions)
__device__ __forceinline__ int lbvh_delta(const uint32_t* k, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  const uint32_t a = k[i], b = k[j];
  return a == b ? 32 + __clz(i ^ j) : __clz(a ^ b);
}

// Work arrays of the build: child pairs and parents of the binary tree
// (internal nodes 0..n-2, leaf i as n-1+i), arrival counters, internal
// boxes, depths, 4-wide indices and levels
struct LbvhWork {
  int2* child;
  int* parent;
  unsigned* flag;
  float* ibox;
  uint32_t* depth;
  uint32_t* idx4;
  uint32_t* dep4;
};

__device__ __forceinline__ void lbvh_box(const LbvhWork& w, const float* lbox,
                                         const uint32_t* sval, int n, int x, float b[6]) {
  const float* src = x < n - 1 ? w.ibox + 6 * (size_t)x : lbox + 6 * (size_t)sval[x - (n - 1)];
  for (int k = 0; k < 6; ++k) b[k] = __hip_atomic_load(src + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void __launch_bounds__(LBVH_THREADS)
k_lbvh_finish(uint32_t n, const uint32_t* __restrict__ skey, const uint32_t* __restrict__ sval,
              const float* __restrict__ lbox, LbvhWork w, Node4* tlas, uint32_t tlas_base,
              uint32_t* tlas_inst, uint32_t* order, uint32_t* level_end, Node4* top,
              uint32_t* top_src, LbvhResult* res) {
  __shared__ uint32_t s_scan[LBVH_THREADS];
  __shared__ uint32_t s_lvl[LBVH_MAX_LEVELS];
  __shared__ uint32_t s_cur[LBVH_MAX_LEVELS];
  __shared__ double s_cost[LBVH_THREADS];
  __shared__ uint32_t s_max;
  const int N = (int)n;
  const uint32_t tid = threadIdx.x;
  for (uint32_t i = tid; i < n; i += LBVH_THREADS) tlas_inst[i] = sval[i];
  if (N == 1) {
    // one instance: a root with one leaf
    if (tid == 0) {
      float b[6];
      for (int k = 0; k < 6; ++k) b[k] = lbox[6 * (size_t)sval[0] + k];
      Node4 o;
      for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 4; ++c) {
          o.box[a][0][c] = c == 0 ? b[a] : INFINITY;
          o.box[a][1][c] = c == 0 ? b[3 + a] : -INFINITY;
        }
      for (int c = 0; c < 4; ++c) { o.child[c] = c == 0 ? LEAF : EMPTY4; o.pad[c] = 0; }
      tlas[0] = o;
      top[0] = o;
      top_src[0] = tlas_base;
      order[0] = tlas_base;
      level_end[0] = 1;
      const double x = (double)b[3] - b[0], y = (double)b[4] - b[1], z = (double)b[5] - b[2];
      *res = LbvhResult{1u, 0u, 1u, 0u, x * y + y * z + z * x};
    }
    return;
  }
  // binary radix tree
  for (uint32_t t = tid; t < n - 1; t += LBVH_THREADS) {
    const int i = (int)t;
    w.flag[i] = 0;
    const int d = lbvh_delta(skey, N, i, i + 1) - lbvh_delta(skey, N, i, i - 1) >= 0 ? 1 : -1;
    const int dmin = lbvh_delta(skey, N, i, i - d);
    int lmax = 2;
    while (lbvh_delta(skey, N, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int s = lmax / 2; s >= 1; s /= 2)
      if (lbvh_delta(skey, N, i, i + (l + s) * d) > dmin) l += s;
    const int j = i + l * d;
    const int dnode = lbvh_delta(skey, N, i, j);
    int sp = 0;
    for (int s = (l + 1) / 2;; s = (s + 1) / 2) {
      if (lbvh_delta(skey, N, i, i + (sp + s) * d) > dnode) sp += s;
      if (s == 1) break;
    }
    const int g = i + sp * d + min(d, 0);
    const int left = min(i, j) == g ? N - 1 + g : g;
    const int right = max(i, j) == g + 1 ? N - 1 + g + 1 : g + 1;
    w.child[i] = make_int2(left, right);
    w.parent[left] = i;
    w.parent[right] = i;
  }
  if (tid == 0) w.parent[0] = -1;
  __threadfence();
  __syncthreads();
  // boxes bottom-up: the second of two children to arrive at a node forms it
  for (uint32_t t = tid; t < n; t += LBVH_THREADS) {
    int node = w.parent[N - 1 + (int)t];
    while (node >= 0) {
      const unsigned prev =
          __hip_atomic_fetch_add(w.flag + node, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
      if (prev == 0) break;
      float a[6], b[6];
      lbvh_box(w, lbox, sval, N, w.child[node].x, a);
      lbvh_box(w, lbox, sval, N, w.child[node].y, b);
      float* dst = w.ibox + 6 * (size_t)node;
      for (int k = 0; k < 3; ++k) {
        __hip_atomic_store(dst + k, fminf(a[k], b[k]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(dst + 3 + k, fmaxf(a[3 + k], b[3 + k]), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      }
      node = w.parent[node];
    }
  }
  __threadfence();
  __syncthreads();
  // depths, then the 4-wide index of every even-depth node (block scan)
  for (uint32_t t = tid; t < n - 1; t += LBVH_THREADS) {
    uint32_t dd = 0;
    for (int p = w.parent[t]; p >= 0; p = w.parent[p]) ++dd;
    w.depth[t] = dd;
  }
  if (tid == 0) s_max = 0;
  __threadfence();
  __syncthreads();
  uint32_t run = 0;
  for (uint32_t base = 0; base < n - 1; base += LBVH_THREADS) {
    const uint32_t t = base + tid;
    const uint32_t v = (t < n - 1 && !(w.depth[t] & 1u)) ? 1u : 0u;
    s_scan[tid] = v;
    __syncthreads();
    for (uint32_t off = 1; off < LBVH_THREADS; off <<= 1) {
      const uint32_t x = tid >= off ? s_scan[tid - off] : 0u;
      __syncthreads();
      s_scan[tid] += x;
      __syncthreads();
    }
    if (t < n - 1) w.idx4[t] = run + s_scan[tid] - v;
    run += s_scan[LBVH_THREADS - 1];
    __syncthreads();
  }
  __threadfence();
  __syncthreads();
  // the 4-wide nodes
  double acc = 0;
  for (uint32_t t = tid; t < n - 1; t += LBVH_THREADS) {
    if (w.depth[t] & 1u) continue;
    int ent[4], ne = 0;
    const int2 c2 = w.child[t];
    for (int side = 0; side < 2; ++side) {
      const int c = side ? c2.y : c2.x;
      if (c >= N - 1) {
        ent[ne++] = c;
      } else {
        ent[ne++] = w.child[c].x;
        ent[ne++] = w.child[c].y;
      }
    }
    Node4 o;
    for (int e = 0; e < 4; ++e) {
      if (e >= ne) {
        for (int a = 0; a < 3; ++a) { o.box[a][0][e] = INFINITY; o.box[a][1][e] = -INFINITY; }
        o.child[e] = EMPTY4;
        o.pad[e] = 0;
        continue;
      }
      float b[6];
      lbvh_box(w, lbox, sval, N, ent[e], b);
      for (int a = 0; a < 3; ++a) { o.box[a][0][e] = b[a]; o.box[a][1][e] = b[3 + a]; }
      o.child[e] = ent[e] >= N - 1 ? (LEAF | (uint32_t)(ent[e] - (N - 1)))
                                   : tlas_base + w.idx4[ent[e]];
      o.pad[e] = 0;
      const double x = (double)b[3] - b[0], y = (double)b[4] - b[1], z = (double)b[5] - b[2];
      acc += x * y + y * z + z * x;
    }
    const uint32_t k = w.idx4[t];
    tlas[k] = o;
    const uint32_t d4 = w.depth[t] / 2;
    w.dep4[k] = d4;
    atomicMax(&s_max, d4);
  }
  s_cost[tid] = acc;
  __threadfence();
  __syncthreads();
  for (uint32_t s = LBVH_THREADS / 2; s > 0; s >>= 1) {
    if (tid < s) s_cost[tid] += s_cost[tid + s];
    __syncthreads();
  }
  // levels for the refit, deepest first
  const uint32_t n4 = run, dmax = s_max;
  for (uint32_t l = tid; l < LBVH_MAX_LEVELS; l += LBVH_THREADS) s_lvl[l] = 0;
  __syncthreads();
  for (uint32_t k = tid; k < n4; k += LBVH_THREADS) atomicAdd(&s_lvl[w.dep4[k]], 1u);
  __syncthreads();
  if (tid == 0) {
    uint32_t acc_l = 0;
    for (int l = (int)dmax; l >= 0; --l) {
      s_cur[l] = acc_l;
      acc_l += s_lvl[l];
      level_end[dmax - l] = acc_l;
    }
  }
  __syncthreads();
  for (uint32_t k = tid; k < n4; k += LBVH_THREADS)
    order[atomicAdd(&s_cur[w.dep4[k]], 1u)] = tlas_base + k;
  __threadfence();
  __syncthreads();
  // the LDS copy of the top levels, breadth first
  if (tid == 0) {
    uint32_t q[TOP_NODES];
    uint32_t nq = 1;
    q[0] = tlas_base;
    for (uint32_t h = 0; h < nq && nq < (uint32_t)TOP_NODES; ++h)
      for (int c = 0; c < 4 && nq < (uint32_t)TOP_NODES; ++c) {
        const uint32_t ch = tlas[q[h] - tlas_base].child[c];
        if (!(ch & LEAF)) q[nq++] = ch;
      }
    for (uint32_t h = 0; h < nq; ++h) {
      Node4 o = tlas[q[h] - tlas_base];
      for (int c = 0; c < 4; ++c) {
        if (o.child[c] & LEAF) continue;
        for (uint32_t g = 0; g < nq; ++g)
          if (q[g] == o.child[c]) { o.child[c] = TOP_REF | g; break; }
      }
      top[h] = o;
      top_src[h] = q[h];
    }
    *res = LbvhResult{n4, dmax, nq, 0u, s_cost[0]};
  }
}

// -------------------------------------------------------------------------
// Path generation (reference point_init, src/ssol_solver.c:207-358)
// -------------------------------------------------------------------------
struct Emit {
  bool has;
  RayRec r;
};

__device__ uint32_t start_medium(const DScene& S, const ssolh_material& m, uint32_t side) {
  if (m.type == SSOLH_MTL_DIELECTRIC || m.type == SSOLH_MTL_THIN_DIELECTRIC)
    return (uint32_t)(side == SSOL_FRONT_ ? m.medium_in : m.medium_out);
  return S.air;
}

// Sample triangle of a start point from the first uniform of its stream
// (alias table over the sampling proxy, by area)
__device__ __forceinline__ uint32_t alias_pick(const DScene& S, double u0) {
  const double x = u0 * (double)S.nstris;
  uint32_t i = (uint32_t)x;
  if (i >= S.nstris) i = S.nstris - 1;
  const ssolh_alias al = S.alias[i];
  return (x - (double)i) < al.prob ? i : al.alias;
}

__device__ void generate(const DScene& S, const PathState& P, const SolveParams& prm,
                         uint32_t slot, uint64_t path, Emit& e, int& samp_slot,
                         const AdjParams& A, const TrackParams& K) {
  PathRng rng;
  rng.init(prm.key, prm.counter + path, 0);
  // area-weighted primitive through the alias table, uniform point
  const double u0 = rng.next(), u1 = rng.next(), u2 = rng.next();
  const uint32_t idx = alias_pick(S, u0);
  const ssolh_sample_tri st = S.stris[idx];
  const ssolh_instance& inst = S.insts[st.instance];
  const double sq = sqrt(u1);
  const double b0 = 1 - sq, b1 = sq * (1 - u2), b2 = sq * u2;
  const d3 v0 = ld3(st.v0), v1 = ld3(st.v1), v2 = ld3(st.v2);
  const d3 T = mk(inst.transform[9], inst.transform[10], inst.transform[11]);
  d3 pos = mul(inst.transform, v0 * b0 + v1 * b1 + v2 * b2) + T;
  const d3 ns = normalize(tmul(inst.inv_linear, cross(v2 - v0, v1 - v0)));

  const uint32_t job = job_of(prm, path);
  const ssolh_sun& sun = job_sun(S, prm, job);
  const d3 dir = sample_sun(sun, rng);
  const double wl = sample_wavelength(S.spectra, sun.wl_spectrum, rng.next());

  const uint32_t iss = inst.isshape_first + st.sshape;
  const IssInfo is = S.iss[iss];
  d3 N = ns;
  const bool punched = is.shape_type == SSOLH_SHAPE_PUNCHED;
  if (punched) {
    // lift onto the quadric (src/ssol_shape.c:1408-1447)
    const ssolh_shape& sh = S.shapes[is.shape];
    const ssolh_inst_sshape& q = S.isshapes[iss];
    d3 loc = mul(q.R_inv, pos - ld3(q.T));
    loc.z = quadric_z(sh.quadric, sh.qparam, loc.x, loc.y);
    N = normalize(tmul(q.R_inv, quadric_grad(sh.quadric, sh.qparam, loc)));
    pos = mul(q.R, loc) + ld3(q.T);
  }
  uint32_t side = SSOL_FRONT_;
  if (dot(N, dir) >= 0) { side = SSOL_BACK_; N = -N; }

  const d3 sun0 = ld3(sun.dir);
  const double surface_sun_cos = dot(N, dir);
  const double surface_sun0_cos = fabs(dot(sun0, N));
  const double sun0_sun_cos = dot(sun0, dir);
  const double proxy_cos = punched ? fabs(dot(ns, N)) : 1.0;
  const double w0 = sun.dni * S.S_proxy * fabs(surface_sun_cos / (proxy_cos * sun0_sun_cos));
  const double cosf = S.S_proxy / S.S * surface_sun0_cos / proxy_cos;

  const ssolh_material& m = S.mtls[side == SSOL_FRONT_ ? is.mtl_front : is.mtl_back];
  const uint32_t med = start_medium(S, m, side);

  if (K.vtx) {
    // sun point and start point (src/ssol_solver.c:909-921)
    K.seq[slot] = 0;
    if (K.sun_len > 0) track_vertex(K, slot, path, pos - dir * K.sun_len, sun.dni);
    track_vertex(K, slot, path, pos, w0);
  }

  if (A.grad) {
    // w0 is proportional to |N.w|: d log w0 / d angle = ((a x N).w) / (N.w)
    const int32_t pp = A.inst_param[st.instance];
    double4 sc = make_double4(0, 0, 0, 0);
    if (pp >= 0) {
      const d3 a1 = ld3(A.axes + 6 * pp), a2 = ld3(A.axes + 6 * pp + 3);
      sc.x = dot(cross(a1, N), dir) / surface_sun_cos;
      sc.y = dot(cross(a2, N), dir) / surface_sun_cos;
    }
    A.s[slot] = sc;
    A.p[slot] = pp >= 0 ? (uint32_t)pp : NONE;
  }

  if (K.vtx) K.pos[slot] = make_double4(pos.x, pos.y, pos.z, 0);
  P.cst[slot] = make_double2(wl, cosf);
  P.pw[slot] = make_double4(w0, w0, w0, w0);
  P.nrm[slot] = make_double4(N.x, N.y, N.z, 0);
  P.st[slot] = make_uint4((uint32_t)path, (uint32_t)(path >> 32), rng.block, PF_ACTIVE);
  P.sm[slot] = make_uint4(iss, st.prim, 0u | (side << 30), med | (med << 16));
  P.aux[slot] = make_uint4((uint32_t)inst.sampled, job, 0, 64u);
  // the slot's pending keys are all NONE already (cleared with the state,
  // and by the flush of the slot's previous realisation)
  samp_slot = inst.sampled;

  // shadow ray toward the sun (src/ssol_solver.c:325-352)
  e.has = true;
  e.r.ox = pos.x; e.r.oy = pos.y; e.r.oz = pos.z;
  e.r.dx = -dir.x; e.r.dy = -dir.y; e.r.dz = -dir.z;
  e.r.tmin = 0;
  e.r.iss_from = iss;
  e.r.prim_from = st.prim;
  e.r.flags = (side << RF_SIDE_SHIFT) | RF_DISCARD_VIRTUAL | RF_REVERSED | RF_ANY_HIT;
}

__device__ void flush_sampled_count(bool has, int samp_slot, const Tallies& T) {
  const uint64_t key = has ? (uint64_t)samp_slot : 0;
  const uint64_t mask = __ballot(has);
  if (!mask) return;
  unsigned long long* base = T.samp_count + (size_t)(blockIdx.x & T.mask_s) * T.nsamp;
  const int leader = __ffsll((unsigned long long)mask) - 1;
  const uint64_t lkey = __shfl(key, leader, 64);
  const bool match = has && key == lkey;
  const uint64_t mm = __ballot(match);
  if ((int)lane_id() == leader) atomicAdd(base + lkey, (unsigned long long)__popcll(mm));
  if (has && !match) atomicAdd(base + key, 1ull);
}

// First realisations of a wavefront: slot s holds realisation first +
// order[s] (first + s when order is null), s < count, order a permutation
// of [0, count). With refill, slot s then runs that realisation + k * slots,
// k = 1, 2, ... while below N. Four waves per SIMD (at most 128 VGPRs, 44
// bytes per lane spilled) run it faster than three without spills (10k
// field: 8.20 -> 7.59 ms per 1e8 realisations; five: 9.65 ms).
template <bool ADJ>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4)))
k_generate(DScene S, PathState P, SolveParams prm, Tallies T, RayRec* __restrict__ rays,
           uint64_t first, uint32_t count, AdjParams A_, TrackParams K,
           const uint32_t* __restrict__ order) {
  const AdjParams A = ADJ ? A_ : AdjParams{};
  const uint32_t slot = blockIdx.x * BLOCK + threadIdx.x;
  const bool live = slot < count;
  Emit e;
  e.has = false;
  int samp_slot = 0;
  if (live) generate(S, P, prm, slot, first + (order ? order[slot] : slot), e, samp_slot, A, K);
  flush_sampled_count(live, samp_slot, T);
  if (e.has) rays[slot] = e.r;
}

// -------------------------------------------------------------------------
// Shading (reference trace_radiative_path loop body, src/ssol_solver.c:939-1086)
// -------------------------------------------------------------------------
enum { FIN_NONE = 0, FIN_DONE = 1, FIN_KILLED = 2, FIN_FAILED = 3 };

struct Finish {
  int kind;
  double g[6];      // cos, absorbed_by_receivers, shadowed, missing, extinct, other
  double factor;    // survivor factor applied to pending entries
  int samp_slot;
  uint32_t job;     // sun of the realisation
};

// Add the 10 receiver weights of a hit to the per-path pending entries;
// false when every entry holds another key.
__device__ bool pend_add(uint32_t* keys, double* vals, size_t stride, uint32_t npend,
                         uint32_t key, const double* w) {
  int slot_k = -1, free_k = -1;
  uint32_t kk[MAX_PEND];
#pragma unroll
  for (int k = 0; k < MAX_PEND; ++k) kk[k] = (uint32_t)k < npend ? keys[k * stride] : 0u;
#pragma unroll
  for (int k = MAX_PEND - 1; k >= 0; --k) {
    if ((uint32_t)k >= npend) continue;
    if (kk[k] == key) slot_k = k;
    if (kk[k] == NONE) free_k = k;
  }
  if (slot_k < 0 && free_k >= 0) {
    slot_k = free_k;
    keys[slot_k * stride] = key;
    for (int f = 0; f < 10; ++f) vals[((size_t)slot_k * 10 + f) * stride] = 0;
  }
  if (slot_k < 0) return false;
  for (int f = 0; f < 10; ++f) vals[((size_t)slot_k * 10 + f) * stride] += w[f];
  return true;
}

// Weights that found no pending entry are flushed as a term of their own
__device__ void spill(double* base, uint64_t key, const double* w, unsigned long long* spills) {
  atomicAdd(spills, 1ull);
  for (int f = 0; f < 10; ++f) {
    atomic_add(base + key * 20 + 2 * f, w[f]);
    atomic_add(base + key * 20 + 2 * f + 1, w[f] * w[f]);
  }
}

// Logical block b (list entries b*BLOCK ...) of k_shade; ADJ: the adjoint
// solve's variant, the plain one carries no score code; REFILL: slots take
// further realisations (fewer slots than realisations), without it no
// generation code.
template <bool ADJ, bool REFILL>
__device__ __forceinline__ void shade_block(
    uint32_t b, const DScene& S, const PathState& P, const SolveParams& prm, const Tallies& T,
    const uint32_t* __restrict__ list, uint32_t count, RayRec* __restrict__ rays,
    const HitRec* __restrict__ hits, uint8_t* __restrict__ emitted,
    uint32_t* __restrict__ block_emits, Counters* C, const AdjParams& A_, const TrackParams& K) {
  const AdjParams A = ADJ ? A_ : AdjParams{};
  const uint32_t q = b * BLOCK + threadIdx.x;
  const bool live = q < count;
  const size_t stride = prm.slots;

  Emit e;
  e.has = false;
  Finish fin;
  fin.kind = FIN_NONE;
  fin.factor = 1;
  fin.samp_slot = 0;
  fin.job = 0;
  for (int i = 0; i < 6; ++i) fin.g[i] = 0;
  const uint32_t slot = live ? slot_of(list, q) : 0;
  // receiver weights of a hit that ends the path: flushed from registers
  // below rather than stored as pending entries and read back (dkey, dpkey:
  // receiver and per-primitive keys, NONE: no such hit)
  uint32_t dkey = NONE, dpkey = NONE;
  double din[4] = {0, 0, 0, 0};  // in0, in1, in2, kabs of that hit

  uint64_t path = 0;
  if (live) {
    const HitRec h = hits[slot];
    const uint4 st = P.st[slot];
    const uint4 sm = P.sm[slot];
    const double2 cw = P.cst[slot];
    // outgoing weights of the last real point (3); w: the initial flux
    // before the first shading, then the incoming flux of the last virtual
    // point the path crossed
    double4 pv = P.pw[slot];
    double4 pd = (st.w & PF_STAGE_BOUNCE) ? P.pend[slot] : make_double4(0, 0, 0, 0);
    // the ray of the emitting point: its origin (the last real point, also
    // past a virtual point, whose ray keeps that origin) and direction
    const RayRec& rr = rays[slot];
    const d3 rorg0 = mk(rr.ox, rr.oy, rr.oz), rdir0 = mk(rr.dx, rr.dy, rr.dz);
    path = (uint64_t)st.x | ((uint64_t)st.y << 32);
    PathRng rng;
    rng.init(prm.key, prm.counter + path, st.z);
    const uint32_t flags0 = st.w;
    uint32_t flags = flags0;
    uint32_t iss = sm.x, prim = sm.y;
    uint32_t depth = sm.z & 0x3FFFFFFFu, side = sm.z >> 30;
    uint32_t in_med = sm.w & 0xFFFFu, out_med = sm.w >> 16;
    // sampled slot, job, survivor score and roulette interval: read at the
    // end of the path and past the first roulette depth only
    uint4 aux = make_uint4(0, 0, 0, 64);
    bool have_aux = false;
    auto load_aux = [&]() {
      if (!have_aux) { aux = P.aux[slot]; have_aux = true; }
    };
    const double wl = cw.x;
    d3 pos = rorg0;
    d3 dir = (flags & PF_STAGE_BOUNCE) ? rdir0 : -rdir0;  // the shadow ray points at the sun
    d3 N = mk(0, 0, 0);
    if (!(flags & PF_STAGE_BOUNCE)) {
      // normal of the start point (later points take theirs from the hit)
      const double4 nt = P.nrm[slot];
      N = mk(nt.x, nt.y, nt.z);
    }
    double hit_t = 0;
    d3 rorg = pos, rdir = dir;  // origin/direction of the ray of the point
    bool head = false;

    if (!(flags & PF_STAGE_BOUNCE)) {
      // result of the shadow ray
      if (h.iss != NONE) {
        fin.kind = FIN_DONE;
        fin.g[2] = pv.w;      // shadowed = initial flux
      } else {
        head = true;
        hit_t = 0;
      }
    } else {
      // tail of the iteration that traced this ray
      const bool found = h.iss != NONE;
      const bool hit_virtual = (flags & PF_HIT_VIRTUAL) != 0;
      const bool in_atm = (flags & PF_IN_ATM) != 0;
      // flux that left the emitting point
      const double left = hit_virtual ? pv.w : pv.x;
      if (!found) {
        // escape point (src/ssol_solver.c:1012-1019)
        if (K.vtx && K.inf_len > 0) {
          const double4 kp = K.pos[slot];
          track_vertex(K, slot, path, mk(kp.x, kp.y, kp.z) + dir * K.inf_len, left);
        }
        if (out_med != S.air) {
          fin.kind = FIN_FAILED;  // ray lost outside air
        } else {
          // past a virtual point the segment from the last real point
          // ends at the virtual one (a real point's segment was counted
          // at that point)
          if (hit_virtual) {
            const double absorbed = pv.x - pv.w;
            if (in_atm) pd.y += absorbed; else pd.z += absorbed;
          }
          fin.kind = FIN_DONE;
          fin.g[3] = left;  // missing
        }
      }
      if (fin.kind == FIN_NONE) {
        depth += hit_virtual ? 0u : 1u;
        if (depth > 64) {
          load_aux();
          const uint32_t rinterval = aux.w;
          if (depth > rinterval && depth % rinterval == 1) {
            if (rng.next() > 0.5) {
              aux.z += 1;
              aux.w = 16;
              P.aux[slot] = aux;
            } else {
              fin.kind = FIN_KILLED;
            }
          }
        }
      }
      if (fin.kind == FIN_NONE) {
        // point_update_from_hit (src/ssol_solver.c:360-411)
        rorg = rorg0;
        rdir = dir;
        iss = h.iss;
        prim = h.prim;
        N = normalize(mk(h.nx, h.ny, h.nz));
        if (S.iss[iss].shape_type == SSOLH_SHAPE_MESH) pos = rorg + rdir * h.t;
        else pos = rorg + dir * h.qdst;
        if (dot(dir, N) < 0) side = SSOL_FRONT_;
        else { side = SSOL_BACK_; N = -N; }
        in_med = out_med;
        hit_t = h.t;
        head = true;
        if (K.vtx) track_vertex(K, slot, path, pos, left);  // src/ssol_solver.c:1080-1083
      }
    }

    if (head) {
      if (!(flags & PF_STAGE_BOUNCE)) { rorg = pos; rdir = dir; }
      const IssInfo is = S.iss[iss];
      const ssolh_material& m = S.mtls[side == SSOL_FRONT_ ? is.mtl_front : is.mtl_back];
      const bool in_atm = in_med == S.air;
      const bool receiver = (is.receiver_mask & side) != 0;
      const bool virt = m.type == SSOLH_MTL_VIRTUAL;
      if (receiver) flags |= PF_HIT_RECEIVER;
      double trans = 1;
      if (hit_t > 0) {
        const double k = data_value(S.spectra, S.media[in_med].extinction, wl);
        if (k > 0) trans = exp(-k * hit_t);
      }
      const double in0 = pv.x * trans;
      const double in1 = in_atm ? pv.y : pv.y * trans;
      const double in2 = !in_atm ? pv.z : pv.z * trans;
      double out0, out1, out2, kabs;
      uint32_t next_med = in_med;
      if (virt) {
        out0 = in0; out1 = in1; out2 = in2; kabs = 0;
      } else {
        double eta_out = 1, eta_in = 1, slab_eta = 1, slab_abs = 0;
        if (m.type == SSOLH_MTL_DIELECTRIC) {
          if ((uint32_t)m.medium_out != in_med) fin.kind = FIN_FAILED;
          eta_out = data_value(S.spectra, S.media[m.medium_out].refractive_index, wl);
          eta_in = data_value(S.spectra, S.media[m.medium_in].refractive_index, wl);
        } else if (m.type == SSOLH_MTL_THIN_DIELECTRIC) {
          eta_out = data_value(S.spectra, S.media[m.medium_out].refractive_index, wl);
          slab_eta = data_value(S.spectra, m.slab_eta, wl);
          slab_abs = data_value(S.spectra, m.slab_absorption, wl);
        }
        double w = 0;
        int lobe = LOBE_NONE;
        const d3 wi = -dir;
        if (fin.kind == FIN_NONE && dot(wi, N) > 0) {
          d3 wo, hm = N;
          w = bsdf_sample(m, eta_out, eta_in, slab_eta, slab_abs, wi, N, rng, wo, lobe, &hm);
          const double c = dot(N, wo);
          if ((c > 0 && lobe == LOBE_TRANSMIT) || (c < 0 && lobe == LOBE_REFLECT)) w = 0;
          dir = wo;
          if (A.grad && !(flags & PF_STAGE_BOUNCE) && w > 0 && m.type == SSOLH_MTL_MIRROR &&
              m.roughness > 0 && m.distribution == 0) {
            // first reflection of a path that starts on a parameterised
            // heliostat: score of the Beckmann microfacet density
            // p(h) = exp(-tan^2/m^2) / (pi m^2 cos^3) under a rotation of
            // the macro normal N about axis a (dN = a x N) and under m:
            //   dlog p/d angle = (2 sec^2/m^2 - 3) sec (h . (a x N))
            //   dlog p/dm      = 2 tan^2/m^3 - 2/m,  m = sqrt(2) slope error
            const uint32_t pp = A.p[slot];
            if (pp != NONE) {
              const double ct = fmin(dot(hm, N), 1.0);
              const double sec = 1.0 / ct, t2 = sec * sec - 1.0;
              const double mr = m.roughness;
              const double k = (2.0 * sec * sec / (mr * mr) - 3.0) * sec;
              const d3 a1 = ld3(A.axes + 6 * pp), a2 = ld3(A.axes + 6 * pp + 3);
              double4 sc = A.s[slot];
              sc.x += k * dot(hm, cross(a1, N));
              sc.y += k * dot(hm, cross(a2, N));
              sc.z += 1.4142135623730951 * (2.0 * t2 / (mr * mr * mr) - 2.0 / mr);
              A.s[slot] = sc;
            }
          }
        }
        kabs = 1 - w;
        out0 = in0 * w;
        out1 = in1 * w;
        out2 = receiver ? in2 * w : in2;
        if (lobe == LOBE_TRANSMIT && m.type == SSOLH_MTL_DIELECTRIC)
          next_med = (uint32_t)(m.medium_out == (int32_t)in_med ? m.medium_in : m.medium_out);
      }
      if (fin.kind == FIN_NONE) {
        if (receiver) {
          pd.x += in0 - out0;
          double w10[10];
          w10[0] = in0; w10[1] = in1; w10[2] = in2;
          w10[3] = in2 - in0; w10[4] = in1 - in0;
          w10[5] = in0 * kabs; w10[6] = in1 * kabs; w10[7] = in2 * kabs;
          w10[8] = (in2 - in0) * kabs; w10[9] = (in1 - in0) * kabs;
          const uint32_t rkey = (uint32_t)is.receiver * 2u + (side == SSOL_FRONT_ ? 0u : 1u);
          // the path ends at this point (weight zero below): its weights
          // join the end-of-path flush directly
          const bool direct = out0 == 0;
          if (direct) {
            dkey = rkey;
            din[0] = in0; din[1] = in1; din[2] = in2; din[3] = kabs;
          } else if (!pend_add(P.pkey + slot, P.pval + slot, stride, prm.npend, rkey, w10)) {
            load_aux();
            spill(T.rcv + (size_t)(blockIdx.x & T.mask_r) * T.nrcv * 40u, rkey, w10, &C->spills);
            spill(T.sxr + (size_t)(blockIdx.x & T.mask_x) * T.nsamp * T.nrcv * 40u, (uint64_t)aux.x * S.nrcv * 2u + rkey, w10, &C->spills);
          }
          if (is.prim_tally >= 0) {
            const uint32_t pkey = (uint32_t)(((uint64_t)is.prim_tally + prim) * 2u +
                                             (side == SSOL_FRONT_ ? 0u : 1u));
            if (direct)
              dpkey = pkey;
            else if (!pend_add(P.qkey + slot, P.qval + slot, stride, prm.npend, pkey, w10))
              spill(T.prim + (size_t)(blockIdx.x & T.mask_p) * T.nprim * 40u, pkey, w10, &C->spills);
            if (A.grad && A.p[slot] != NONE) A.s[slot].w += A.gmap[pkey] * w10[5];
          }
          if (A.grad && A.p[slot] != NONE) A.s[slot].w += A.d_total * (in0 - out0);
        } else {
          pd.z += in0 * kabs;
        }
        // extinction of the segment that reached this point, counted here at
        // a real point (past virtual points the segment runs from the last
        // real point); a virtual point counts it only if the path ends
        if (!virt || out0 == 0) {
          const double absorbed = pv.x - in0;
          if (in_atm) pd.y += absorbed; else pd.z += absorbed;
        }
        if (out0 == 0) {
          // weight is zero: the segment ends here
          fin.kind = FIN_DONE;
          fin.g[3] = 0;
        } else {
          e.has = true;
          if (virt) {
            e.r.ox = rorg.x; e.r.oy = rorg.y; e.r.oz = rorg.z;
            e.r.dx = rdir.x; e.r.dy = rdir.y; e.r.dz = rdir.z;
            e.r.tmin = nextafterf((float)hit_t, FLT_MAX);
            pv.w = in0;
          } else {
            e.r.ox = pos.x; e.r.oy = pos.y; e.r.oz = pos.z;
            e.r.dx = dir.x; e.r.dy = dir.y; e.r.dz = dir.z;
            e.r.tmin = 0;
            pv = make_double4(out0, out1, out2, pv.w);
          }
          e.r.iss_from = iss;
          e.r.prim_from = prim;
          e.r.flags = side << RF_SIDE_SHIFT;
          flags = (flags & PF_HIT_RECEIVER) | PF_ACTIVE | PF_STAGE_BOUNCE |
                  (virt ? PF_HIT_VIRTUAL : 0u) | (in_atm ? PF_IN_ATM : 0u);
          out_med = next_med;
        }
      }
    }

    if (fin.kind != FIN_NONE) {
      load_aux();
      fin.samp_slot = (int)aux.x;
      fin.job = aux.y;
    }
    if (K.vtx && fin.kind != FIN_NONE) {
      // path type (src/ssol_solver.c:931, 1092-1094, 1111-1125)
      uint8_t ty;
      if (fin.kind == FIN_KILLED) ty = SSOLH_PATH_DROPPED;
      else if (fin.kind == FIN_FAILED) ty = SSOLH_PATH_ERROR;
      else if (!(flags0 & PF_STAGE_BOUNCE) && h.iss != NONE) ty = SSOLH_PATH_SHADOW;
      else ty = (flags & PF_HIT_RECEIVER) ? SSOLH_PATH_SUCCESS : SSOLH_PATH_MISSING;
      K.type[path] = ty;
    }
    if (fin.kind == FIN_DONE) {
      fin.g[0] = cw.y;  // cos factor
      fin.g[1] = pd.x;
      fin.g[4] = pd.y;
      fin.g[5] = pd.z;
      fin.factor = (double)(1u << aux.z);
      for (int i = 0; i < 6; ++i) fin.g[i] *= fin.factor;
    }
    if (fin.kind == FIN_NONE) {
      P.pw[slot] = pv;
      P.pend[slot] = pd;
      P.st[slot] = make_uint4(st.x, st.y, rng.block, flags);
      P.sm[slot] = make_uint4(iss, prim, depth | (side << 30), in_med | (out_med << 16));
      if (K.vtx) K.pos[slot] = make_double4(pos.x, pos.y, pos.z, 0);
    }
  }

  // --- convergent flush of finished realisations -------------------------
  const bool done = fin.kind == FIN_DONE;
  double* rcv_base = T.rcv + (size_t)(blockIdx.x & T.mask_r) * T.nrcv * 40u;
  {
    // global fields and the realisation count of the path's job (sun)
    const bool counted = done || fin.kind == FIN_KILLED;
    double g2[13];
    for (int i = 0; i < 6; ++i) {
      g2[2 * i] = done ? fin.g[i] : 0.0;
      g2[2 * i + 1] = done ? fin.g[i] * fin.g[i] : 0.0;
    }
    g2[12] = 1.0;
    keyed_add<13, 16>(counted, counted ? fin.job : 0, g2,
                      T.global + (size_t)(blockIdx.x & T.mask_g) * T.njobs * 16u);
    double s4[4] = {fin.g[0], fin.g[0] * fin.g[0], fin.g[2], fin.g[2] * fin.g[2]};
    keyed_add<4>(done, (uint64_t)fin.samp_slot, s4,
                 T.samp + (size_t)(blockIdx.x & T.mask_s) * T.nsamp * 4u);
    const unsigned long long nf = wave_count_add(fin.kind == FIN_FAILED, &C->failed);
    if (fin.kind == FIN_FAILED && nf + 1 >= prm.max_failed) atomicExch(&C->aborted, 1u);
  }
  // pending entries fill from entry 0 up (pend_add takes the lowest free
  // one), so a path's keys end at its first empty entry and the wave stops
  // at the first entry none of its finished paths uses
  // The weights of the hit that ended the path (a direct entry) join the
  // pending entry of the same key if the path holds one, before squaring;
  // else they take the path's first empty entry position, one past the
  // pending entries at most.
  bool more = live && fin.kind != FIN_NONE;
  bool dpend = done && dkey != NONE;
  double dw[10];
  dw[0] = din[0]; dw[1] = din[1]; dw[2] = din[2];
  dw[3] = din[2] - din[0]; dw[4] = din[1] - din[0];
  dw[5] = din[0] * din[3]; dw[6] = din[1] * din[3]; dw[7] = din[2] * din[3];
  dw[8] = (din[2] - din[0]) * din[3]; dw[9] = (din[1] - din[0]) * din[3];
  for (uint32_t k = 0; k <= prm.npend; ++k) {
    uint32_t key = NONE;
    double a10[10];
    if (more && k < prm.npend) {
      key = P.pkey[k * stride + slot];
      if (key != NONE) {
        const bool merge = dpend && key == dkey;
        for (int f = 0; f < 10; ++f)
          a10[f] = P.pval[((size_t)k * 10 + f) * stride + slot] + (merge ? dw[f] : 0.0);
        if (merge) dpend = false;
        P.pkey[k * stride + slot] = NONE;
      }
    }
    if (key == NONE) {
      more = false;
      if (dpend) {
        key = dkey;
        for (int f = 0; f < 10; ++f) a10[f] = dw[f];
        dpend = false;
      }
    }
    if (!__ballot(key != NONE)) break;
    const bool has = done && key != NONE;
    double v[20];
    for (int f = 0; f < 10; ++f) {
      const double x = has ? a10[f] * fin.factor : 0.0;
      v[2 * f] = x;
      v[2 * f + 1] = x * x;
    }
    keyed_add<20>(has, has ? key : 0, v, rcv_base);
    const uint64_t xkey = has ? (uint64_t)fin.samp_slot * (uint64_t)S.nrcv * 2u + key : 0;
    keyed_add<20>(has, xkey, v, T.sxr + (size_t)(blockIdx.x & T.mask_x) * T.nsamp * T.nrcv * 40u);
  }
  more = live && fin.kind != FIN_NONE;  // per-primitive entries, the same way
  dpend = done && dpkey != NONE;
  for (uint32_t k = 0; k <= prm.npend; ++k) {
    uint32_t key = NONE;
    double a10[10];
    if (more && k < prm.npend) {
      key = P.qkey[k * stride + slot];
      if (key != NONE) {
        const bool merge = dpend && key == dpkey;
        for (int f = 0; f < 10; ++f)
          a10[f] = P.qval[((size_t)k * 10 + f) * stride + slot] + (merge ? dw[f] : 0.0);
        if (merge) dpend = false;
        P.qkey[k * stride + slot] = NONE;
      }
    }
    if (key == NONE) {
      more = false;
      if (dpend) {
        key = dpkey;
        for (int f = 0; f < 10; ++f) a10[f] = dw[f];
        dpend = false;
      }
    }
    if (!__ballot(key != NONE)) break;
    prim_add(T, done && key != NONE, key, a10, fin.factor);
  }

  // --- adjoint: the finished path's weighted contribution times its scores,
  // wave-reduced (in start-instance order a wave's paths share a heliostat)
  if (A.grad) {
    const uint32_t pp = done ? A.p[slot] : NONE;
    const bool has = pp != NONE;
    double v[3] = {0.0, 0.0, 0.0};
    if (has) {
      const double4 sc = A.s[slot];
      const double q = sc.w * fin.factor;
      v[0] = q * sc.x;
      v[1] = q * sc.y;
      v[2] = q * sc.z;
    }
    keyed_add<3>(has, has ? pp : 0u, v, A.grad);
  }

  // --- refill finished slots with their next realisation -------------------
  const bool want = live && fin.kind != FIN_NONE && !C->aborted;
  const uint64_t id = want ? path + prm.slots : 0;
  int samp_slot = 0;
  bool gen = false;
  if (REFILL && want && id < prm.N) {
    generate(S, P, prm, slot, id, e, samp_slot, A, K);
    gen = true;
  }
  flush_sampled_count(gen, samp_slot, T);

  // --- the slot's next ray and its place in the next work list -----------
  if (e.has) rays[slot] = e.r;
  if (live) emitted[slot] = e.has ? 1 : 0;
  block_count(e.has, block_emits, b);
}

// The arguments of a shade kernel in one struct: the struct is the kernel's
// argument segment (offset 0), which a looping launch reads again for each
// logical block (kernel_args)
struct ShadeArgs {
  DScene S;
  PathState P;
  SolveParams prm;
  Tallies T;
  const uint32_t* list;
  const unsigned* count;
  RayRec* rays;
  const HitRec* hits;
  uint8_t* emitted;
  uint32_t* block_emits;
  Counters* C;
  AdjParams A;
  TrackParams K;
  unsigned* fetch;  // tail-block counter of a looping launch (for_each_block)
};

// The kernel's argument segment as an Args, behind a pointer the compiler
// cannot see through. A looping kernel takes it anew for each logical
// block, so the scalar loads of the arguments stay next to their uses: hoisted
// out of the loop, the dozens of argument words would be held in scalar
// registers across it and spill (through vector registers into scratch).
template <typename Args>
__device__ __forceinline__ const Args& kernel_args() {
  auto p = (const __attribute__((address_space(4))) Args*)__builtin_amdgcn_kernarg_segment_ptr();
  asm volatile("" : "+s"(p));
  return *(const Args*)p;
}

template <bool ADJ, bool REFILL, bool LOOP>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(3)))
k_shade(ShadeArgs a) {
  const uint32_t count = *a.count;
  // whole logical blocks only: block_count syncs (LOOP: as in k_trace)
  if (LOOP) {
    for_each_block((count + BLOCK - 1) / BLOCK, a.fetch, [&](uint32_t b) {
      const ShadeArgs& f = kernel_args<ShadeArgs>();
      shade_block<ADJ, REFILL>(b, f.S, f.P, f.prm, f.T, f.list, count, f.rays, f.hits, f.emitted,
                               f.block_emits, f.C, f.A, f.K);
    });
  } else if (blockIdx.x * BLOCK < count) {
    shade_block<ADJ, REFILL>(blockIdx.x, a.S, a.P, a.prm, a.T, a.list, count, a.rays, a.hits,
                             a.emitted, a.block_emits, a.C, a.A, a.K);
  }
}

// -------------------------------------------------------------------------
// Order-preserving compaction of the work list
// -------------------------------------------------------------------------
// Exclusive scan of the per-block emit counts of k_shade in two levels:
// k_scan_chunks scans each chunk of SCAN_CHUNK counts (offsets within the
// chunk) and records the chunk's total; k_scan_top scans the chunk totals in
// place into chunk offsets and writes the grand total, the next list's
// length. Entry b of the scan is offsets[b] + chunk[b / SCAN_CHUNK].
constexpr int SCAN_PER = 4;
constexpr int SCAN_CHUNK = SCAN_THREADS * SCAN_PER;
static_assert((1u << 28) / BLOCK / SCAN_CHUNK <= SCAN_THREADS, "one k_scan_top block scans every chunk");

__device__ void block_inclusive_scan(uint32_t* part) {
  const uint32_t tid = threadIdx.x;
  for (uint32_t off = 1; off < SCAN_THREADS; off <<= 1) {
    const uint32_t v = tid >= off ? part[tid - off] : 0u;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
}

__global__ void __launch_bounds__(SCAN_THREADS)
k_scan_chunks(const uint32_t* __restrict__ counts, const unsigned* __restrict__ count_ptr,
              uint32_t* __restrict__ offsets, uint32_t* __restrict__ chunk) {
  __shared__ uint32_t part[SCAN_THREADS];
  const uint32_t nblocks = (*count_ptr + BLOCK - 1) / BLOCK;
  const uint32_t base = blockIdx.x * SCAN_CHUNK;
  if (base >= nblocks) return;  // whole blocks only
  const uint32_t tid = threadIdx.x;
  const uint32_t b0 = base + tid * SCAN_PER;
  uint32_t v[SCAN_PER];
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < SCAN_PER; ++k) {
    v[k] = b0 + k < nblocks ? counts[b0 + k] : 0u;
    sum += v[k];
  }
  part[tid] = sum;
  __syncthreads();
  block_inclusive_scan(part);
  uint32_t run = tid ? part[tid - 1] : 0u;
#pragma unroll
  for (int k = 0; k < SCAN_PER; ++k) {
    if (b0 + k < nblocks) offsets[b0 + k] = run;
    run += v[k];
  }
  if (tid == SCAN_THREADS - 1) chunk[blockIdx.x] = part[SCAN_THREADS - 1];
}

__global__ void __launch_bounds__(SCAN_THREADS)
k_scan_top(const unsigned* __restrict__ count_ptr, uint32_t* __restrict__ chunk,
           unsigned int* total, unsigned long long* segments, unsigned* fetch) {
  __shared__ uint32_t part[SCAN_THREADS];
  const uint32_t tid = threadIdx.x;
  const uint32_t nblocks = (*count_ptr + BLOCK - 1) / BLOCK;
  const uint32_t nchunks = (nblocks + SCAN_CHUNK - 1) / SCAN_CHUNK;
  if (tid == 0) {
    *segments += *count_ptr;
    fetch[0] = fetch[1] = 0;  // tail-block counters of the next trace and shade
  }
  const uint32_t v = tid < nchunks ? chunk[tid] : 0u;
  part[tid] = v;
  __syncthreads();
  block_inclusive_scan(part);
  if (tid < nchunks) chunk[tid] = part[tid] - v;
  if (tid == SCAN_THREADS - 1) *total = part[SCAN_THREADS - 1];
}

// Entry q of the next list: the q-th slot (in list order) that wrote a ray.
// Same block partition as k_shade.
template <bool LOOP>
__global__ void __launch_bounds__(BLOCK)
k_scatter(const uint32_t* __restrict__ list, const unsigned* __restrict__ count_ptr,
          const uint8_t* __restrict__ emitted, const uint32_t* __restrict__ offsets,
          const uint32_t* __restrict__ chunk, uint32_t* __restrict__ next) {
  __shared__ uint32_t wcount[WAVES];
  const uint32_t count = *count_ptr;
  const uint32_t b_end = LOOP ? count : min(count, (blockIdx.x + 1) * BLOCK);
  for (uint32_t b = blockIdx.x; b * BLOCK < b_end; b += gridDim.x) {  // whole logical blocks
    const uint32_t q = b * BLOCK + threadIdx.x;
    const bool live = q < count;
    const uint32_t slot = live ? slot_of(list, q) : 0;
    const bool pred = live && emitted[slot];
    const uint64_t m = __ballot(pred);
    const uint32_t w = threadIdx.x / 64;
    if (lane_id() == 0) wcount[w] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t base = offsets[b] + chunk[b / SCAN_CHUNK];
    for (uint32_t i = 0; i < w; ++i) base += wcount[i];
    const uint64_t below = m & ((1ull << lane_id()) - 1ull);
    if (pred) next[base + (uint32_t)__popcll(below)] = slot;
    __syncthreads();  // wcount is reused by the next logical block
  }
}

#include "ssolh_fp32.hpp"
#include "ssolh_render.hpp"

// First work list: slots [0, n) in order, and its length
__global__ void k_iota(uint32_t* __restrict__ list, uint32_t n, unsigned* __restrict__ count) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) list[i] = i;
  if (i == 0) *count = n;
}

// Slot order of the first realisations by (sun, start instance). The
// realisations are drawn at random over the field, so in realisation order a
// wave's rays start on unrelated heliostats; placed by start instance, a
// wave's rays leave neighbouring points in like directions (coherent
// traversal) and its finished paths share tally keys (wave-reduced atomics).
// The sun (job) comes first so a batch keeps each wave on one sun, as the
// realisation order did. k_start_keys draws each realisation's start
// instance exactly as generate does; a stable radix sort of (key, index)
// pairs then gives the order, realisations of one key in index order.
template <bool FP32>
__device__ __forceinline__ uint32_t start_key(const DScene& S, const SolveParams& prm,
                                              uint64_t path, uint32_t ninst) {
  // S.stris[alias_pick(...)].instance (alias_pick32 for FP32) from one
  // StartAlias load
  uint32_t inst;
  if (FP32) {
    PathRng32 rng;
    rng.init(prm.key, prm.counter + path, 0);
    const float u0 = rng.next(), uc = rng.next();
    uint32_t i = (uint32_t)(u0 * (float)S.nstris);
    if (i >= S.nstris) i = S.nstris - 1;
    const StartAlias e = S.salias[i];
    inst = (double)uc < e.prob ? e.inst : e.inst_alias;
  } else {
    PathRng rng;
    rng.init(prm.key, prm.counter + path, 0);
    const double x = rng.next() * (double)S.nstris;
    uint32_t i = (uint32_t)x;
    if (i >= S.nstris) i = S.nstris - 1;
    const StartAlias e = S.salias[i];
    inst = (x - (double)i) < e.prob ? e.inst : e.inst_alias;
  }
  return job_of(prm, path) * ninst + inst;
}

// The start key of each of the wavefront's realisations and its index
template <bool FP32>
__global__ void __launch_bounds__(256)
k_start_keys(DScene S, SolveParams prm, uint64_t first, uint32_t count, uint32_t ninst,
             uint32_t* __restrict__ key, uint32_t* __restrict__ idx) {
  const uint32_t q = blockIdx.x * 256 + threadIdx.x;
  if (q >= count) return;
  key[q] = start_key<FP32>(S, prm, first + q, ninst);
  idx[q] = q;
}

// Bits of the start keys (below nkeys) and the sort's work bytes for count
// realisations
uint32_t start_key_bits(uint64_t nkeys) {
  uint32_t bits = 1;
  while (bits < 32 && (uint64_t(1) << bits) < nkeys) ++bits;
  return bits;
}

hipError_t start_sort_bytes(uint32_t count, uint32_t bits, hipStream_t stream, size_t& bytes) {
  bytes = 0;
  return hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint32_t*)nullptr,
                                            (uint32_t*)nullptr, (const uint32_t*)nullptr,
                                            (uint32_t*)nullptr, (int)count, 0, (int)bits, stream);
}

// Sum of the stripes of a striped tally into stripe 0
template <typename T>
__global__ void k_fold(T* base, size_t n, uint32_t stripes) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  T s = base[i];
  for (uint32_t k = 1; k < stripes; ++k) s += base[(size_t)k * n + i];
  base[i] = s;
}

// -------------------------------------------------------------------------
// Deterministic test kernels
// -------------------------------------------------------------------------
__global__ void k_test_quadric(int quadric, double4 qp, uint32_t n, const double* org,
                               const double* dir, const double* hint, double* t, double* nrm) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double p[4] = {qp.x, qp.y, qp.z, qp.w};
  double tt;
  d3 g;
  if (quadric_hit(quadric, p, ld3(org + 3 * i), ld3(dir + 3 * i), hint[i], tt, g)) {
    t[i] = tt;
    nrm[3 * i] = g.x; nrm[3 * i + 1] = g.y; nrm[3 * i + 2] = g.z;
  } else {
    t[i] = INFINITY;
    nrm[3 * i] = nrm[3 * i + 1] = nrm[3 * i + 2] = 0;
  }
}

// One wave_sum per wave of 64 inputs (blocks of 64 threads: every lane active)
__global__ void __launch_bounds__(64) k_test_wave_sum(const double* in, double* out) {
  const double s = wave_sum(in[blockIdx.x * 64 + threadIdx.x]);
  out[blockIdx.x * 64 + threadIdx.x] = s;
}

// Uniforms from an array (0.5 once it is exhausted), for fixed-input tests
struct ArrayRng {
  const double* u;
  uint32_t n, i;
  __device__ double next() { return i < n ? u[i++] : 0.5; }
};

__global__ void k_test_sun(ssolh_sun sun, uint32_t n, const double* u, uint32_t nu,
                           double* out, uint32_t* used) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  ArrayRng rng{u + (size_t)i * nu, nu, 0};
  const d3 d = sample_sun(sun, rng);
  out[3 * i] = d.x; out[3 * i + 1] = d.y; out[3 * i + 2] = d.z;
  used[i] = rng.i;
}

__global__ void k_test_bsdf(ssolh_material m, double4 eta, uint32_t n, const double* wi,
                            const double* nrm, const double* u, uint32_t nu, double* wo,
                            double* w, int* lobe) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  ArrayRng rng{u + (size_t)i * nu, nu, 0};
  d3 o;
  int l;
  w[i] = bsdf_sample(m, eta.x, eta.y, eta.z, eta.w, ld3(wi + 3 * i), ld3(nrm + 3 * i), rng, o, l);
  wo[3 * i] = o.x; wo[3 * i + 1] = o.y; wo[3 * i + 2] = o.z;
  lobe[i] = l;
}

// Plain device-to-device copy: reads and writes exactly 16 B per element.
// Used to calibrate the HBM byte counters and as the achievable-bandwidth
// yardstick for the path-tracing kernels.
__global__ void __launch_bounds__(256) k_copy(const uint4* __restrict__ a, uint4* __restrict__ b,
                                             size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

__global__ void k_test_threefry(uint32_t n, const uint64_t* ctr, const uint64_t* key, uint64_t* out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  threefry4x64_20(ctr + 4 * i, key + 4 * i, out + 4 * i);
}

}  // namespace

// -------------------------------------------------------------------------
// Host side
// -------------------------------------------------------------------------
struct DevBuf {
  void* p = nullptr;
  size_t n = 0;
};

// Device memory of ssolh_solve_noalloc: every buffer of its solve (path
// state, rays, hits, work lists, scan arrays, counters, tallies) is carved
// from this array, which each GPU holds from the library's load to the end
// of the process. Host code never names it as an address: each device's own
// copy is reached through the pointer hipGetSymbolAddress returns with that
// device current (kept per device by ssolh_create).
constexpr size_t SCRATCH_BYTES = size_t(64) << 20;
// records of the per-primitive log (84 B each plus 12 B of sort arrays)
constexpr uint64_t LOG_MAX_RECORDS = uint64_t(1) << 27;
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch[MAX_DEVICES] = {};

struct ssolh_context {
  int device = 0;
  hipStream_t stream = nullptr;
  void* scratch = nullptr;  // this device's copy of the scratch array
  hipEvent_t ev[4] = {};
  // graph drain: the list length after each of two replays in flight, in
  // pinned host memory, and the events that mark it copied
  hipEvent_t gev[2] = {};
  unsigned* pinned = nullptr;
  // scene
  std::vector<DevBuf> scene_bufs;
  DScene scn{};
  bool has_scene = false;
  uint32_t nrcv = 0, nsamp = 0;
  uint64_t nprim = 0;
  uint32_t stack_depth = MAX_STACK;  // traversal stack bound of the scene's 4-wide nodes
  uint32_t lds_stack_cap = 0;        // 16, 24 or 32: LDS stack entries at most (0: MAX_STACK)
  // The pending keys of every slot are NONE when a solve starts: a solve
  // that drains leaves them so (each finished path clears its own), so they
  // are cleared only when the path state's layout changed since the last
  // solve that drained (state buffer, slots, entries, form) or it came from
  // the scratch array.
  struct KeysLayout {
    void* state;
    uint64_t slots;
    uint32_t npend;
    bool fp32;
    bool operator==(const KeysLayout& o) const {
      return state == o.state && slots == o.slots && npend == o.npend && fp32 == o.fp32;
    }
  };
  KeysLayout keys_layout{};
  bool keys_clean = false;
  float slack_scale = SLAB_SLACK;     // box-test slack of the scenes uploaded next
  // TLAS refit of moved instances: per-object boxes, the TLAS nodes by level
  // (deepest first) and the top copy's sources, all in scene_bufs
  const float* obj_box = nullptr;
  const uint32_t* tlas_order = nullptr;
  const uint32_t* tlas_level_end = nullptr;
  uint32_t tlas_levels = 0;
  const uint32_t* top_src = nullptr;
  double* refit_cost = nullptr;
  double tlas_cost0 = 0;              // the TLAS's cost as built
  std::vector<uint32_t> inst_object;  // object of each instance as uploaded
  uint32_t ninst_sshapes = 0;
  uint32_t n_uploads = 0, n_refits = 0, n_builds = 0;
  // GPU TLAS builds: the TLAS's region of `nodes` (first index, capacity),
  // the BLASes' part of the stack bound, the build's work buffer
  uint32_t tlas_base = 0, tlas_cap = 0, blas_pb = 0;
  DevBuf lbvh;
  DevBuf render;  // renderer work buffer (ssolh_render)
  // solve buffers
  DevBuf start_sort;  // work bytes of the start-key sort of the first realisations
  DevBuf state, rays, hits, lists[2], emitted, block_emits, block_offsets, scan_chunks, counters,
      tallies, jobs, trace_ovf;
  // per-primitive log: keys, sorted keys, indices, sorted indices, values,
  // count, sort work (Tallies::log_*); log_sort, log_sidx, log_skey and
  // log_idx within it, log_sort_bytes of sort work
  DevBuf plog;
  void* log_sort = nullptr;
  size_t log_sort_bytes = 0;
  uint32_t *log_skey = nullptr, *log_idx = nullptr, *log_sidx = nullptr;
  uint32_t log_key_bits = 1;
  uint32_t log_cap_test = 0;          // ssolh_test_prim_log's cap (0: the default)
  uint32_t log_flushes = 0;           // the last solve's sorted reductions
  uint64_t log_overflowed = 0;        // and the entries it sent to the atomics
  DevBuf adj_state, adj_in, adj_grad;
  DevBuf trk_vtx, trk_type, trk_seq, trk_count, trk_pos;
  uint64_t trk_paths = 0;           // realisations of the last tracked solve
  unsigned long long trk_nvtx = 0;  // vertices it produced
  uint64_t trk_cap = 0;
  uint32_t ninstances = 0;
  uint32_t njobs = 1;
  PathState P{};
  PathState32 P32{};
  const ShapeQ* shapeq = nullptr;  // per shape, in scene_bufs
  Tallies T{};
  size_t tally_bytes = 0;
  // last solve
  Counters last{};
  uint64_t segments = 0;
  double kernel_ms = 0, trace_ms = 0, shade_ms = 0;
  uint32_t iterations = 0;
  uint32_t wavefronts = 0;    // of the last solve
  uint32_t refilled = 0;
  // grid of each looping kernel: its workgroups resident at once
  std::vector<std::pair<const void*, uint32_t>> resident;
};

namespace {

// Workgroups of each looping (graph-captured) kernel resident on the device
// at once: occupancy per CU x CUs. A looping launch uses that grid, so every
// workgroup starts at once and takes an equal share of the logical blocks;
// with a larger grid the workgroups past the resident ones wait for whole
// shares of others to end, and the tail runs on part of the chip.
int init_resident_grids(ssolh_context* c, uint32_t cus) {
  const void* fns[] = {
      (const void*)k_trace<16, true>, (const void*)k_trace<24, true>,
      (const void*)k_trace<MAX_STACK, true>,
      (const void*)k_trace32<16, true>, (const void*)k_trace32<24, true>,
      (const void*)k_trace32<MAX_STACK, true>,
      (const void*)k_shade<false, false, true>, (const void*)k_shade<false, true, true>,
      (const void*)k_shade<true, false, true>, (const void*)k_shade<true, true, true>,
      (const void*)k_shade32<false, true>, (const void*)k_shade32<true, true>,
      (const void*)k_scatter<true>};
  for (const void* f : fns) {
    int per_cu = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, BLOCK, 0));
    c->resident.push_back({f, (uint32_t)std::max(per_cu, 1) * cus});
  }
  return 0;
}

uint32_t resident_grid(const ssolh_context* c, const void* f) {
  for (const auto& e : c->resident)
    if (e.first == f) return e.second;
  return 1024;  // not reached: init_resident_grids lists every looping kernel
}

int dev_alloc(DevBuf& b, size_t bytes) {
  if (b.n >= bytes && b.p) return 0;
  if (b.p) { (void)hipFree(b.p); b.p = nullptr; b.n = 0; }
  if (!bytes) bytes = 16;
  HIP_TRY(hipMalloc(&b.p, bytes));
  b.n = bytes;
  return 0;
}

// A solve's buffers carved from caller-independent device memory that is
// never freed (the library's static scratch array): each take advances a
// cursor, 256-byte aligned; nothing is allocated or released.
struct Arena {
  char* base;
  size_t size;
  size_t used;
};

int arena_take(Arena& a, DevBuf& b, size_t bytes) {
  if (!bytes) bytes = 16;
  // taken earlier in this solve and large enough (the start-instance
  // histogram of each chunk)
  if (b.p && b.n >= bytes && (char*)b.p >= a.base && (char*)b.p < a.base + a.size) return 0;
  const size_t off = (a.used + 255) & ~size_t(255);
  if (off > a.size || bytes > a.size - off) {
    set_err("ssolh_solve_noalloc: the solve needs more than the %zu bytes of the scratch array "
            "(%zu in use, %zu more asked)", a.size, off, bytes);
    return SSOLH_ERR_LIMIT;
  }
  b.p = a.base + off;
  b.n = bytes;
  a.used = off + bytes;
  return 0;
}

template <typename T>
int upload(ssolh_context* c, const T* src, size_t count, const T** dst) {
  DevBuf b;
  const size_t bytes = count * sizeof(T);
  int r = dev_alloc(b, bytes ? bytes : 16);
  if (r) return r;
  if (bytes) HIP_TRY(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, c->stream));
  c->scene_bufs.push_back(b);
  *dst = (const T*)b.p;
  return 0;
}

void free_scene(ssolh_context* c) {
  for (auto& b : c->scene_bufs) if (b.p) (void)hipFree(b.p);
  c->scene_bufs.clear();
  c->has_scene = false;
}

// Inverse of a column-major 3x3 matrix; false when singular
bool inv3(const double* m, double* o) {
  const double a = m[0], b = m[3], cc = m[6], d = m[1], e = m[4], f = m[7], g = m[2], h = m[5],
               i = m[8];
  const double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
  const double det = a * A + b * B + cc * C;
  if (det == 0 || !std::isfinite(det)) return false;
  const double r = 1.0 / det;
  // rows of the inverse (row-major r0..r2) stored column-major
  const double inv[3][3] = {{A * r, -(b * i - cc * h) * r, (b * f - cc * e) * r},
                            {B * r, (a * i - cc * g) * r, -(a * f - cc * d) * r},
                            {C * r, -(a * h - b * g) * r, (a * e - b * d) * r}};
  for (int row = 0; row < 3; ++row)
    for (int col = 0; col < 3; ++col) o[col * 3 + row] = inv[row][col];
  return true;
}

// Per (instance, shaded shape) records of the hit path, from the flat scene
int build_iss(const ssolh_scene* s, std::vector<IssInfo>& out) {
  out.assign(s->ninst_sshapes, IssInfo{});
  for (uint32_t i = 0; i < s->ninstances; ++i) {
    const ssolh_instance& in = s->instances[i];
    if (in.object >= s->nobjects) return -1;
    const ssolh_object& o = s->objects[in.object];
    for (uint32_t k = 0; k < o.sshape_count; ++k) {
      const uint64_t g = (uint64_t)in.isshape_first + k;
      const uint32_t ss = o.sshape_first + k;
      if (g >= s->ninst_sshapes || ss >= s->nsshapes) return -1;
      const ssolh_sshape& sh = s->sshapes[ss];
      if (sh.shape >= s->nshapes || sh.mtl_front >= s->nmaterials || sh.mtl_back >= s->nmaterials)
        return -1;
      IssInfo& d = out[g];
      d.inst = i;
      d.shape = sh.shape;
      d.mtl_front = sh.mtl_front;
      d.mtl_back = sh.mtl_back;
      d.receiver = in.receiver;
      d.receiver_mask = in.receiver_mask;
      d.shape_type = (uint32_t)s->shapes[sh.shape].type;
      d.obj_flags = o.flags | (o.sshape_count == 1 ? OBJ_ONE_SHAPE : 0u);
      d.prim_tally = s->inst_sshapes[g].prim_tally;
    }
  }
  return 0;
}

// LDS stack entries per lane of the trace kernel: the smallest of 16, 24, 32
// that holds the scene's bound (the more blocks then share a CU), within the
// context's cap; the rest of the bound goes to the overflow array
uint32_t trace_lds_stack(const ssolh_context* c) {
  uint32_t e = c->stack_depth <= 16 ? 16 : c->stack_depth <= 24 ? 24 : (uint32_t)MAX_STACK;
  if (c->lds_stack_cap && c->lds_stack_cap < e) e = c->lds_stack_cap;
  return e;
}

// BVH2 -> 4-wide nodes. A 4-wide node takes the children of a BVH2 node
// and, while it holds fewer than four and one of them is an inner node,
// replaces the inner child of largest box area by that child's children;
// each inner child left becomes a 4-wide node of its own. pb[i]: the most
// stack entries a descent from node i pushes (each visited node pushes all
// but one of the children a ray crosses).
struct Collapse4 {
  struct Ent {
    float lo[3], hi[3];
    uint32_t child, count;  // count 0: inner BVH2 node `child`
  };
  const ssolh_bvh_node* in;
  uint32_t nin;
  std::vector<Node4> out;
  std::vector<uint32_t> pb;
  int err = 0;
  char msg[160] = "";

  Collapse4(const ssolh_bvh_node* nodes, uint32_t n) : in(nodes), nin(n) {}

  static float area(const Ent& e) {
    const float x = e.hi[0] - e.lo[0], y = e.hi[1] - e.lo[1], z = e.hi[2] - e.lo[2];
    return x * y + y * z + z * x;
  }
  void fail(int code, const char* fmt, uint32_t a, uint32_t b) {
    if (!err) snprintf(msg, sizeof(msg), fmt, a, b);
    err = code;
  }
  // children of BVH2 node b appended to e[n..]
  void take(uint32_t b, Ent* e, int& n) {
    const ssolh_bvh_node& m = in[b];
    for (int k = 0; k < 2; ++k) {
      if (m.count[k] == SSOLH_BVH_EMPTY) continue;
      Ent& x = e[n++];
      for (int a = 0; a < 3; ++a) { x.lo[a] = m.lo[k][a]; x.hi[a] = m.hi[k][a]; }
      x.child = m.child[k];
      x.count = m.count[k];
    }
  }
  uint32_t root(uint32_t b) { return node(b, 0); }
  uint32_t node(uint32_t b, int depth) {
    if (b >= nin) { fail(SSOLH_ERR_ARG, "BVH node %u out of %u nodes", b, nin); return 0; }
    if (depth > 256) { fail(SSOLH_ERR_LIMIT, "BVH deeper than %u levels at node %u", 256, b); return 0; }
    Ent e[5];
    int n = 0;
    take(b, e, n);
    while (n < 4) {
      int best = -1;
      float ba = -1.f;
      for (int i = 0; i < n; ++i)
        if (e[i].count == 0 && area(e[i]) > ba) { ba = area(e[i]); best = i; }
      if (best < 0) break;
      const uint32_t b2 = e[best].child;
      if (b2 >= nin) { fail(SSOLH_ERR_ARG, "BVH child %u out of %u nodes", b2, nin); return 0; }
      e[best] = e[--n];
      take(b2, e, n);
    }
    const uint32_t me = (uint32_t)out.size();
    out.emplace_back();
    pb.push_back(0);
    uint32_t ref[4], deeper = 0;
    for (int i = 0; i < 4; ++i) {
      if (i >= n) { ref[i] = EMPTY4; continue; }
      if (e[i].count == 0) {
        ref[i] = node(e[i].child, depth + 1);
        if (err) return 0;
        deeper = std::max(deeper, pb[ref[i]]);
      } else {
        if (e[i].count > 8 || e[i].child >= (1u << 28)) {
          fail(SSOLH_ERR_LIMIT, "BVH leaf of %u items at %u (at most 8 below 2^28)", e[i].count,
               e[i].child);
          return 0;
        }
        ref[i] = LEAF | ((e[i].count - 1u) << 28) | e[i].child;
      }
    }
    Node4& o = out[me];
    for (int i = 0; i < 4; ++i) {
      for (int a = 0; a < 3; ++a) {
        o.box[a][0][i] = i < n ? e[i].lo[a] : INFINITY;
        o.box[a][1][i] = i < n ? e[i].hi[a] : -INFINITY;
      }
      o.child[i] = ref[i];
      o.pad[i] = 0;
    }
    pb[me] = (n > 1 ? (uint32_t)n - 1u : 0u) + deeper;
    return me;
  }
};

}  // namespace

extern "C" {

SSOLH_API const char* ssolh_last_error(void) { return g_err; }

SSOLH_API int ssolh_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

SSOLH_API int ssolh_create(int device, ssolh_context** out) {
  if (!out) { set_err("ssolh_create: NULL output"); return SSOLH_ERR_ARG; }
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= device) {
    set_err("ssolh_create: no HIP device %d (found %d)", device, n);
    return SSOLH_ERR_NO_DEVICE;
  }
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device));
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    set_err("ssolh_create: device %d is %s, the kernels are built for gfx950", device,
            prop.gcnArchName);
    return SSOLH_ERR_NO_DEVICE;
  }
  if (device < 0 || device >= MAX_DEVICES) {
    set_err("ssolh_create: device %d beyond the %d this library serves", device, MAX_DEVICES);
    return SSOLH_ERR_NO_DEVICE;
  }
  HIP_TRY(hipSetDevice(device));
  void* scratch_ptr = nullptr;
  {
    // once per device, with the device current: the address of its copy of
    // the scratch array
    // (without it only ssolh_solve_noalloc is unavailable, and says so)
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) {
      void* p = nullptr;
      const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e == hipSuccess && p) {
        g_scratch[device] = p;
      } else {
        (void)hipGetLastError();
        fprintf(stderr, "ssolh_create: hipGetSymbolAddress(scratch) on device %d: %s\n", device,
                hipGetErrorString(e));
      }
    }
    scratch_ptr = g_scratch[device];
  }
  ssolh_context* c = new ssolh_context();
  c->device = device;
  c->scratch = scratch_ptr;
  HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  for (auto& e : c->ev) HIP_TRY(hipEventCreate(&e));
  for (auto& e : c->gev) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  HIP_TRY(hipHostMalloc((void**)&c->pinned, 2 * sizeof(unsigned), hipHostMallocDefault));
  if (int r = init_resident_grids(c, (uint32_t)prop.multiProcessorCount)) {
    ssolh_destroy(c);
    return r;
  }
  *out = c;
  return 0;
}

SSOLH_API int ssolh_destroy(ssolh_context* c) {
  if (!c) return SSOLH_ERR_ARG;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_scene(c);
  for (DevBuf* b : {&c->state, &c->rays, &c->hits, &c->lists[0], &c->lists[1], &c->emitted,
                    &c->block_emits, &c->block_offsets, &c->scan_chunks, &c->counters, &c->tallies,
                    &c->jobs, &c->start_sort, &c->trace_ovf, &c->lbvh, &c->render, &c->plog,
                    &c->adj_state, &c->adj_in, &c->adj_grad, &c->trk_vtx, &c->trk_type,
                    &c->trk_seq, &c->trk_count, &c->trk_pos})
    if (b->p) (void)hipFree(b->p);
  for (auto& e : c->ev) if (e) (void)hipEventDestroy(e);
  for (auto& e : c->gev) if (e) (void)hipEventDestroy(e);
  if (c->pinned) (void)hipHostFree(c->pinned);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

SSOLH_API int ssolh_scene_upload(ssolh_context* c, const struct ssolh_scene* s) {
  if (!c || !s) { set_err("ssolh_scene_upload: NULL argument"); return SSOLH_ERR_ARG; }
  if (!s->nsample_tris || !s->ninstances || !s->nnodes) {
    set_err("ssolh_scene_upload: empty scene (sampled triangles %u, instances %u)",
            s->nsample_tris, s->ninstances);
    return SSOLH_ERR_ARG;
  }
  std::vector<ShapeQ> shapeq(s->nshapes);
  for (uint32_t i = 0; i < s->nshapes; ++i) {
    const double* m = s->shapes[i].qtransform;
    ShapeQ& q = shapeq[i];
    if (!inv3(m, q.Rinv)) {
      set_err("ssolh_scene_upload: singular quadric transform of shape %u", i);
      return SSOLH_ERR_ARG;
    }
    for (int k = 0; k < 3; ++k) q.t[k] = m[9 + k];
  }
  std::vector<IssInfo> iss;
  if (build_iss(s, iss)) {
    set_err("ssolh_scene_upload: inconsistent instance / shaded-shape indices");
    return SSOLH_ERR_ARG;
  }
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipStreamSynchronize(c->stream));
  free_scene(c);
  DScene& d = c->scn;
  int r = 0;
  const double* cdf = nullptr;
  r |= upload(c, s->spectra, s->nspectra, &d.spectra.sp);
  r |= upload(c, s->spectra_wl, s->spectra_len, &d.spectra.wl);
  r |= upload(c, s->spectra_val, s->spectra_len, &d.spectra.val);
  r |= upload(c, s->spectra_cdf, s->spectra_len, &cdf);
  d.spectra.cdf = cdf;
  r |= upload(c, s->media, s->nmedia, &d.media);
  r |= upload(c, s->materials, s->nmaterials, &d.mtls);
  r |= upload(c, s->shapes, s->nshapes, &d.shapes);
  // 4-wide traversal nodes from the host's BVH2: each object's BLAS (one
  // copy per distinct root) and the TLAS, with the objects' roots remapped
  Collapse4 col(s->nodes, s->nnodes);
  std::vector<ssolh_object> objects(s->objects, s->objects + s->nobjects);
  std::map<uint32_t, uint32_t> blas_of;  // BVH2 root -> 4-wide root
  uint32_t blas_pb = 0;
  for (ssolh_object& o : objects) {
    auto it = blas_of.find(o.blas_root);
    if (it == blas_of.end()) it = blas_of.emplace(o.blas_root, col.root(o.blas_root)).first;
    o.blas_root = it->second;
    blas_pb = std::max(blas_pb, col.pb[it->second]);
  }
  const uint32_t troot = col.root(s->tlas_root);
  if (col.err) { set_err("ssolh_scene_upload: %s", col.msg); return col.err; }
  if (col.out.size() >= (1u << 25)) {  // byte offsets of the nodes fit 32 bits
    set_err("ssolh_scene_upload: %zu BVH nodes", col.out.size());
    return SSOLH_ERR_LIMIT;
  }
  r |= upload(c, objects.data(), objects.size(), &d.objects);
  // the TLAS region holds any tree over the instances (a GPU build writes
  // at most ninstances 4-wide nodes)
  c->tlas_base = troot;
  c->tlas_cap = std::max<uint32_t>((uint32_t)col.out.size() - troot, s->ninstances);
  col.out.resize((size_t)troot + c->tlas_cap);
  r |= upload(c, col.out.data(), col.out.size(), &d.nodes);
  // host arrays of the uploads below, alive until the closing synchronize
  std::vector<Node4> top;
  std::vector<uint32_t> top_order{troot}, tlas_order, level_end;
  static const double zero = 0;
  {
    // the TLAS's first TOP_NODES nodes, breadth first, with the references
    // between them rewritten to TOP_REF | index
    std::vector<uint32_t>& order = top_order;
    for (size_t h = 0; h < order.size() && order.size() < (size_t)TOP_NODES; ++h)
      for (int k = 0; k < 4 && order.size() < (size_t)TOP_NODES; ++k) {
        const uint32_t ch = col.out[order[h]].child[k];
        if (!(ch & LEAF)) order.push_back(ch);
      }
    std::map<uint32_t, uint32_t> where;
    for (size_t i = 0; i < order.size(); ++i) where[order[i]] = (uint32_t)i;
    for (uint32_t g : order) {
      Node4 n = col.out[g];
      for (int k = 0; k < 4; ++k) {
        auto w = (n.child[k] & LEAF) ? where.end() : where.find(n.child[k]);
        if (w != where.end()) n.child[k] = TOP_REF | w->second;
      }
      top.push_back(n);
    }
    d.ntop = (uint32_t)top.size();
    d.root_ref = TOP_REF;
    d.slack_scale = c->slack_scale;
    top.resize(TOP_NODES);
    r |= upload(c, top.data(), top.size(), &d.top);
    order.resize(TOP_NODES);
    r |= upload(c, order.data(), order.size(), &c->top_src);
  }
  {
    // TLAS nodes by depth, deepest level first (the refit's order), and the
    // tree's cost as built
    std::vector<std::vector<uint32_t>> levels;
    std::vector<std::pair<uint32_t, uint32_t>> todo{{troot, 0u}};
    double cost = 0;
    while (!todo.empty()) {
      const auto [nd, depth] = todo.back();
      todo.pop_back();
      if (levels.size() <= depth) levels.resize(depth + 1);
      levels[depth].push_back(nd);
      const Node4& n = col.out[nd];
      for (int k = 0; k < 4; ++k) {
        if (n.child[k] == EMPTY4) continue;
        const double x = (double)n.box[0][1][k] - n.box[0][0][k];
        const double y = (double)n.box[1][1][k] - n.box[1][0][k];
        const double z = (double)n.box[2][1][k] - n.box[2][0][k];
        cost += x * y + y * z + z * x;
        if (!(n.child[k] & LEAF)) todo.push_back({n.child[k], depth + 1});
      }
    }
    for (size_t l = levels.size(); l-- > 0;) {
      tlas_order.insert(tlas_order.end(), levels[l].begin(), levels[l].end());
      level_end.push_back((uint32_t)tlas_order.size());
    }
    c->tlas_levels = (uint32_t)level_end.size();
    tlas_order.resize(std::max<size_t>(tlas_order.size(), s->ninstances));
    level_end.resize(std::max<size_t>(level_end.size(), LBVH_MAX_LEVELS));
    r |= upload(c, tlas_order.data(), tlas_order.size(), &c->tlas_order);
    r |= upload(c, level_end.data(), level_end.size(), &c->tlas_level_end);
    c->tlas_cost0 = cost;
    r |= upload(c, &zero, 1, (const double**)&c->refit_cost);
  }
  r |= upload(c, s->object_boxes, (size_t)s->nobjects * 6, &c->obj_box);
  d.obj_box = c->obj_box;
  // a lane's stack: the TLAS descent's pushes, the instance sentinel, the
  // deepest BLAS descent's pushes
  c->stack_depth = col.pb[troot] + 1 + blas_pb;
  c->blas_pb = blas_pb;
  r |= upload(c, s->triangles, s->ntriangles, &d.tris);
  r |= upload(c, s->instances, s->ninstances, &d.insts);
  r |= upload(c, s->tlas_instances, s->ninstances, &d.tlas_inst);
  r |= upload(c, s->inst_sshapes, s->ninst_sshapes, &d.isshapes);
  r |= upload(c, iss.data(), iss.size(), &d.iss);
  r |= upload(c, shapeq.data(), shapeq.size(), &c->shapeq);
  r |= upload(c, s->sample_tris, s->nsample_tris, &d.stris);
  r |= upload(c, s->sample_alias, s->nsample_tris, &d.alias);
  std::vector<StartAlias> salias(s->nsample_tris);
  for (uint32_t i = 0; i < s->nsample_tris; ++i)
    salias[i] = {s->sample_alias[i].prob, s->sample_tris[i].instance,
                 s->sample_tris[s->sample_alias[i].alias].instance};
  r |= upload(c, salias.data(), salias.size(), &d.salias);
  if (r) return r;
  d.air = s->air_medium;
  d.tlas_root = troot;
  d.nstris = s->nsample_tris;
  d.sun = s->sun;
  d.S = s->sampled_area;
  d.S_proxy = s->sampled_area_proxy;
  d.nrcv = s->nreceivers;
  d.nsamp = s->nsampled;
  c->nrcv = s->nreceivers;
  c->nsamp = s->nsampled;
  c->nprim = s->nprim_tallies;
  c->ninstances = s->ninstances;
  c->ninst_sshapes = s->ninst_sshapes;
  c->inst_object.resize(s->ninstances);
  for (uint32_t i = 0; i < s->ninstances; ++i) c->inst_object[i] = s->instances[i].object;
  HIP_TRY(hipStreamSynchronize(c->stream));  // the host arrays may go after return
  c->has_scene = true;
  ++c->n_uploads;
  return 0;
}

SSOLH_API int ssolh_scene_move_instances(ssolh_context* c, uint32_t ninstances,
                                         const struct ssolh_instance* instances,
                                         uint32_t ninst_sshapes,
                                         const struct ssolh_inst_sshape* inst_sshapes,
                                         double* cost_ratio) {
  if (!c || !instances || (ninst_sshapes && !inst_sshapes)) {
    set_err("ssolh_scene_move_instances: NULL argument");
    return SSOLH_ERR_ARG;
  }
  if (!c->has_scene || ninstances != c->ninstances || ninst_sshapes != c->ninst_sshapes) {
    set_err("ssolh_scene_move_instances: %u instances, %u (instance, shape) frames for a scene "
            "of %u and %u", ninstances, ninst_sshapes, c->ninstances, c->ninst_sshapes);
    return SSOLH_ERR_ARG;
  }
  for (uint32_t i = 0; i < ninstances; ++i)
    if (instances[i].object != c->inst_object[i]) {
      set_err("ssolh_scene_move_instances: instance %u changed object", i);
      return SSOLH_ERR_ARG;
    }
  HIP_TRY(hipSetDevice(c->device));
  DScene& d = c->scn;
  HIP_TRY(hipMemcpyAsync((void*)d.insts, instances, ninstances * sizeof(ssolh_instance),
                         hipMemcpyHostToDevice, c->stream));
  if (ninst_sshapes)
    HIP_TRY(hipMemcpyAsync((void*)d.isshapes, inst_sshapes,
                           ninst_sshapes * sizeof(ssolh_inst_sshape), hipMemcpyHostToDevice,
                           c->stream));
  k_tlas_refit<<<1, REFIT_THREADS, 0, c->stream>>>(
      d.insts, d.tlas_inst, c->obj_box, (Node4*)d.nodes, c->tlas_order, c->tlas_level_end,
      c->tlas_levels, (Node4*)d.top, c->top_src, d.ntop, c->refit_cost);
  HIP_TRY(hipGetLastError());
  double cost = 0;
  HIP_TRY(hipMemcpyAsync(&cost, c->refit_cost, sizeof(cost), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (cost_ratio) *cost_ratio = c->tlas_cost0 > 0 ? cost / c->tlas_cost0 : 1.0;
  ++c->n_refits;
  return 0;
}

SSOLH_API int ssolh_scene_build_tlas(ssolh_context* c) {
  if (!c || !c->has_scene) {
    set_err("ssolh_scene_build_tlas: no scene");
    return SSOLH_ERR_ARG;
  }
  const uint32_t n = c->ninstances;
  HIP_TRY(hipSetDevice(c->device));
  // work buffer: sort keys and values (twice), the radix sort's temporary,
  // instance boxes, the binary tree, results
  size_t sort_bytes = 0;
  HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (const uint32_t*)nullptr,
                                             (uint32_t*)nullptr, (const uint32_t*)nullptr,
                                             (uint32_t*)nullptr, (int)n, 0, 30, c->stream));
  auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t nn = n, ni = n > 1 ? n - 1 : 1;
  const size_t sizes[] = {nn * 4, nn * 4, nn * 4, nn * 4, sort_bytes, nn * 24, ni * 8,
                          (2 * nn) * 4, ni * 4, ni * 24, ni * 4, ni * 4, nn * 4,
                          sizeof(LbvhResult)};
  size_t total = 0;
  for (size_t b : sizes) total += al(b);
  if (int e = dev_alloc(c->lbvh, total)) return e;
  char* p = (char*)c->lbvh.p;
  void* part[14];
  for (int i = 0; i < 14; ++i) { part[i] = p; p += al(sizes[i]); }
  uint32_t *key = (uint32_t*)part[0], *skey = (uint32_t*)part[1], *val = (uint32_t*)part[2],
           *sval = (uint32_t*)part[3];
  LbvhWork w{(int2*)part[6], (int*)part[7], (unsigned*)part[8], (float*)part[9],
             (uint32_t*)part[10], (uint32_t*)part[11], (uint32_t*)part[12]};
  LbvhResult* res = (LbvhResult*)part[13];
  DScene& d = c->scn;
  k_lbvh_keys<<<1, LBVH_THREADS, 0, c->stream>>>(d.insts, c->obj_box, n, (float*)part[5], key,
                                                 val);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipcub::DeviceRadixSort::SortPairs(part[4], sort_bytes, key, skey, val, sval, (int)n, 0,
                                             30, c->stream));
  k_lbvh_finish<<<1, LBVH_THREADS, 0, c->stream>>>(
      n, skey, sval, (const float*)part[5], w, (Node4*)d.nodes + c->tlas_base, c->tlas_base,
      (uint32_t*)d.tlas_inst, (uint32_t*)c->tlas_order, (uint32_t*)c->tlas_level_end,
      (Node4*)d.top, (uint32_t*)c->top_src, res);
  HIP_TRY(hipGetLastError());
  LbvhResult h{};
  HIP_TRY(hipMemcpyAsync(&h, res, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (h.nodes4 == 0 || h.nodes4 > c->tlas_cap || h.ntop == 0 || h.ntop > (uint32_t)TOP_NODES) {
    set_err("ssolh_scene_build_tlas: inconsistent build (%u nodes, %u top)", h.nodes4, h.ntop);
    return SSOLH_ERR_LIMIT;
  }
  d.ntop = h.ntop;
  d.root_ref = TOP_REF;
  c->tlas_levels = h.depth4 + 1;
  c->tlas_cost0 = h.cost;
  // a descent pushes at most 3 per 4-wide level above the leaves
  c->stack_depth = 3u * (h.depth4 + 1u) + 1u + c->blas_pb;
  ++c->n_builds;
  return 0;
}

SSOLH_API int ssolh_render(ssolh_context* c, const struct ssolh_render_args* a, double* pixels) {
  if (!c || !a || !pixels || !a->width || !a->height || !a->spp || a->mode > 1 ||
      (a->mode == 0 && !a->draft_samples)) {
    set_err("ssolh_render: invalid arguments");
    return SSOLH_ERR_ARG;
  }
  if (!c->has_scene) { set_err("ssolh_render: no scene"); return SSOLH_ERR_ARG; }
  const uint64_t npix = (uint64_t)a->width * a->height, total = npix * a->spp;
  if (npix > (1u << 28) || a->spp > (1u << 20)) {
    set_err("ssolh_render: %llu pixels x %u samples", (unsigned long long)npix, a->spp);
    return SSOLH_ERR_LIMIT;
  }
  HIP_TRY(hipSetDevice(c->device));
  const uint32_t batch = (uint32_t)std::min<uint64_t>(total, 1u << 21);
  const uint32_t nslot = 2 * batch;
  const uint32_t lds_stack = trace_lds_stack(c);
  const uint32_t extra = c->stack_depth > lds_stack ? c->stack_depth - lds_stack : 0;
  auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t sizes[] = {(size_t)batch * sizeof(RenderState), (size_t)nslot * sizeof(RayRec),
                          (size_t)nslot * sizeof(HitRec), (size_t)nslot * 4, (size_t)nslot * 4,
                          npix * 8, (size_t)a->spp * 8, 16, (size_t)nslot * extra * 4 + 4};
  size_t bytes = 0;
  for (size_t b : sizes) bytes += al(b);
  if (int e = dev_alloc(c->render, bytes)) return e;
  char* p = (char*)c->render.p;
  void* part[9];
  for (int i = 0; i < 9; ++i) { part[i] = p; p += al(sizes[i]); }
  RenderState* st = (RenderState*)part[0];
  RayRec* rays = (RayRec*)part[1];
  HitRec* hits = (HitRec*)part[2];
  uint32_t* lists[2] = {(uint32_t*)part[3], (uint32_t*)part[4]};
  double* pix = (double*)part[5];
  float* draft = (float*)part[6];
  unsigned* cnt = (unsigned*)part[7];  // [0], [1]: list lengths, [2]: unused ticket
  const TOvf ovf{extra ? (uint32_t*)part[8] : nullptr, nslot, extra};
  HIP_TRY(hipMemsetAsync(pix, 0, npix * 8, c->stream));
  if (a->mode == 0)
    HIP_TRY(hipMemcpyAsync(draft, a->draft_samples, (size_t)a->spp * 8, hipMemcpyHostToDevice,
                           c->stream));
  RenderParams rp{};
  rp.width = a->width; rp.height = a->height; rp.spp = a->spp; rp.mode = a->mode;
  for (int k = 0; k < 3; ++k) {
    rp.pos[k] = a->position[k]; rp.ax[k] = a->axis_x[k]; rp.ay[k] = a->axis_y[k];
    rp.az[k] = a->axis_z[k]; rp.up[k] = a->up[k];
  }
  rp.key = a->rng_key;
  rp.draft = draft;
  for (uint64_t first = 0; first < total; first += batch) {
    rp.first = first;
    rp.count = (uint32_t)std::min<uint64_t>(total - first, batch);
    const uint32_t blocks = (rp.count + BLOCK - 1) / BLOCK;
    HIP_TRY(hipMemsetAsync(cnt, 0, 16, c->stream));
    k_render_start<<<blocks, BLOCK, 0, c->stream>>>(c->scn, rp, st, rays, lists[0], cnt);
    HIP_TRY(hipGetLastError());
    int cur = 0;
    for (uint32_t iter = 0;; ++iter) {
      unsigned n = 0;
      HIP_TRY(hipMemcpyAsync(&n, cnt + cur, sizeof(n), hipMemcpyDeviceToHost, c->stream));
      HIP_TRY(hipStreamSynchronize(c->stream));
      if (!n) break;
      if (iter >= 100000) { set_err("ssolh_render: paths did not end"); return SSOLH_ERR_LIMIT; }
      const uint32_t tb = (n + BLOCK - 1) / BLOCK;
      if (lds_stack == 16)
        k_trace<16, false><<<tb, BLOCK, 0, c->stream>>>(c->scn, lists[cur], cnt + cur, rays, hits,
                                                       cnt + 2, ovf);
      else if (lds_stack == 24)
        k_trace<24, false><<<tb, BLOCK, 0, c->stream>>>(c->scn, lists[cur], cnt + cur, rays, hits,
                                                       cnt + 2, ovf);
      else
        k_trace<MAX_STACK, false><<<tb, BLOCK, 0, c->stream>>>(c->scn, lists[cur], cnt + cur,
                                                              rays, hits, cnt + 2, ovf);
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipMemsetAsync(cnt + (cur ^ 1), 0, sizeof(unsigned), c->stream));
      k_render_shade<<<blocks, BLOCK, 0, c->stream>>>(c->scn, rp, st, rays, hits, pix,
                                                      lists[cur ^ 1], cnt + (cur ^ 1));
      HIP_TRY(hipGetLastError());
      cur ^= 1;
    }
  }
  HIP_TRY(hipMemcpyAsync(pixels, pix, npix * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

SSOLH_API int ssolh_test_slab_slack(ssolh_context* c, float scale) {
  if (!c || !(scale >= 0.f && scale <= 1e-3f)) {
    set_err("ssolh_test_slab_slack: scale outside [0, 1e-3]");
    return SSOLH_ERR_ARG;
  }
  c->slack_scale = scale;
  c->scn.slack_scale = c->slack_scale;
  return 0;
}

SSOLH_API int ssolh_test_scene_counters(ssolh_context* c, uint32_t* uploads, uint32_t* refits,
                                        uint32_t* builds) {
  if (!c) return SSOLH_ERR_ARG;
  if (uploads) *uploads = c->n_uploads;
  if (refits) *refits = c->n_refits;
  if (builds) *builds = c->n_builds;
  return 0;
}

}  // extern "C"

namespace {

// One solve over `njobs` suns (njobs == 1: the scene's sun). Realisation
// ids [0, N) map to jobs through job_first. graph_iters > 0 captures that
// many iterations (rounded up to even) into a HIP graph replayed until the
// work list is empty; 0 runs the iterations from the host one by one.
// Host description of an adjoint solve (host pointers)
struct TrackHost {
  double sun_len, inf_len;
  uint64_t cap;
};

struct AdjHost {
  uint32_t nparams;
  const int32_t* inst_param;  // [ninstances]
  const double* axes;         // [nparams][2][3]
  const double* gmap;         // [nprim_tallies][2] or null
  double d_total;
};

int run_solve(ssolh_context* c, uint64_t N, uint64_t max_failed, uint64_t key, uint64_t counter,
              bool fp32, uint32_t path_slots, uint32_t npend, uint32_t njobs,
              const uint64_t* job_first_h, const ssolh_sun* job_suns_h, uint32_t graph_iters,
              const AdjHost* adj = nullptr, const TrackHost* trk = nullptr,
              Arena* arena = nullptr) {
  // buffers come from the device allocator, or all from the arena
  auto alloc = [&](DevBuf& b, size_t bytes) -> int {
    return arena ? arena_take(*arena, b, bytes) : dev_alloc(b, bytes);
  };
  if (arena && (adj || trk || job_suns_h || graph_iters)) {
    set_err("ssolh_solve_noalloc: plain single-sun solves driven from the host only");
    return SSOLH_ERR_ARG;
  }
  if (adj && fp32) { set_err("ssolh_solve_adjoint: the adjoint runs in the FP64 form"); return SSOLH_ERR_ARG; }
  if (trk && (fp32 || adj || job_suns_h)) {
    set_err("ssolh_solve_tracked: paths are recorded by plain FP64 solves only");
    return SSOLH_ERR_ARG;
  }
  if (trk) graph_iters = 0;
  else c->trk_paths = 0;  // the recorded paths belong to the last solve only
  if (npend > MAX_PEND) { set_err("ssolh_solve: at most %d pending entries", MAX_PEND); return SSOLH_ERR_LIMIT; }
  // Tally layout: each tally is striped over up to 256 copies while its
  // stripes hold at most 2^21 doubles; from the scratch array every stripe
  // count is halved until all tallies take at most half of it.
  uint32_t sg, sr, ss, sx, sp;
  size_t n_global, n_rcv, n_samp, n_sxr, n_prim, n_u64, tally_bytes;
  for (size_t cap = 256;; cap >>= 1) {
    auto pow2_stripes = [cap](size_t per_stripe_doubles) {
      size_t k = cap;
      while (k > 1 && k * per_stripe_doubles > (size_t(1) << 21)) k >>= 1;
      return (uint32_t)k;
    };
    sg = pow2_stripes((size_t)njobs * 16);
    sr = pow2_stripes((size_t)c->nrcv * 40);
    ss = pow2_stripes((size_t)c->nsamp * 5);
    sx = pow2_stripes((size_t)c->nsamp * c->nrcv * 40);
    sp = pow2_stripes((size_t)c->nprim * 40);
    n_global = (size_t)sg * njobs * 16;
    n_rcv = (size_t)sr * c->nrcv * 40;
    n_samp = (size_t)ss * c->nsamp * 4;
    n_sxr = (size_t)sx * c->nsamp * c->nrcv * 40;
    n_prim = (size_t)sp * c->nprim * 40;
    n_u64 = (size_t)ss * c->nsamp;
    tally_bytes = (n_global + n_rcv + n_samp + n_sxr + n_prim + n_u64) * 8 + 64;
    if (!arena || tally_bytes <= arena->size / 2 || cap == 1) break;
  }
  if (arena && tally_bytes > arena->size / 2) {
    set_err("ssolh_solve_noalloc: the scene's tallies take %zu bytes; at most %zu fit in the "
            "scratch array", tally_bytes, arena->size / 2);
    return SSOLH_ERR_LIMIT;
  }
  // Wavefront width. By default every realisation is in flight at once when
  // 2/5 of the device memory holds their state: the realisations then start
  // together and each iteration traces one generation of rays (shadow rays,
  // first reflections, ...), which keeps the waves coherent; larger solves
  // run as many as fit, each slot taking its next realisation in place.
  uint64_t slots64 = path_slots;
  const size_t per_slot = (fp32 ? 5 * sizeof(float4) + 3 * sizeof(uint4)
                               : sizeof(double2) + 3 * sizeof(double4) + 3 * sizeof(uint4)) +
                          (size_t)npend * 2 * (sizeof(uint32_t) + 10 * (fp32 ? sizeof(float) : sizeof(double))) +
                          (fp32 ? sizeof(RayRec32) + sizeof(HitRec32) : sizeof(RayRec) + sizeof(HitRec)) +
                          2 * sizeof(uint32_t) + 1 + (adj ? sizeof(double4) + sizeof(uint32_t) : 0) +
                          (trk ? sizeof(uint32_t) + sizeof(double4) : 0);
  if (!slots64 && arena) {
    // from the scratch array: what is left beside the tallies, the
    // per-block scan arrays, counters and start-instance histogram
    size_t sort_bytes = 0;
    const uint64_t most = std::min<uint64_t>(std::min<uint64_t>(N, arena->size / per_slot), 1u << 28);
    HIP_TRY(start_sort_bytes((uint32_t)most, start_key_bits((uint64_t)njobs * c->ninstances),
                             c->stream, sort_bytes));
    const size_t fixed = tally_bytes + (sort_bytes + 255) / 256 * 256 + (64u << 10);
    const size_t rest = arena->size > fixed ? arena->size - fixed : 0;
    slots64 = rest / (per_slot + 1) / BLOCK * BLOCK;  // + 8 B of per-block counts per block
    if (slots64 < BLOCK) {
      set_err("ssolh_solve_noalloc: no path slot fits beside the tallies in the scratch array");
      return SSOLH_ERR_LIMIT;
    }
    if (slots64 > N) slots64 = N;
  }
  if (!slots64) {
    size_t free_b = 0, total_b = 0;
    HIP_TRY(hipMemGetInfo(&free_b, &total_b));
    const uint64_t fit = (uint64_t)(total_b / 5 * 2) / per_slot;
    slots64 = std::min<uint64_t>(std::max<uint64_t>(fit, 1u << 20), 1ull << 27);
    // and within what is free now (this context's own solve buffers count
    // as free: they are reallocated), when other work holds device memory
    size_t held = 0;
    for (const DevBuf* b : {&c->state, &c->rays, &c->hits, &c->lists[0], &c->lists[1], &c->emitted,
                            &c->adj_state, &c->trk_seq})
      held += b->n;
    const uint64_t fit_free = (uint64_t)((free_b + held) / 10 * 9) / per_slot;
    slots64 = std::min<uint64_t>(slots64, std::max<uint64_t>(fit_free, 1u << 16));
  }
  if (slots64 > N) slots64 = N;
  slots64 = (slots64 + BLOCK - 1) / BLOCK * BLOCK;
  if (slots64 > (1ull << 28)) { set_err("ssolh_solve: at most 2^28 path slots"); return SSOLH_ERR_LIMIT; }
  const uint32_t slots = (uint32_t)slots64;
  const uint32_t max_blocks = slots / BLOCK;

  // path state: arrays of 16- or 32-byte records and the pending entries,
  // each array starting on a 256-byte boundary
  const size_t n = slots;
  const size_t real = fp32 ? sizeof(float) : sizeof(double);
  const size_t pend_keys = npend * n * sizeof(uint32_t), pend_vals = npend * n * 10 * real;
  std::vector<size_t> sizes;
  if (fp32)
    sizes = {n * sizeof(float4), n * sizeof(float4), n * sizeof(float4), n * sizeof(float4),
             n * sizeof(float4), n * sizeof(uint4), n * sizeof(uint4), n * sizeof(uint4),
             pend_keys, pend_keys, pend_vals, pend_vals};
  else
    sizes = {n * sizeof(double2), n * sizeof(double4), n * sizeof(double4), n * sizeof(double4),
             n * sizeof(uint4), n * sizeof(uint4), n * sizeof(uint4), pend_keys, pend_keys,
             pend_vals, pend_vals};
  size_t bytes_state = 0;
  for (size_t sz : sizes) bytes_state += (sz + 255) / 256 * 256;
  int r = alloc(c->state, bytes_state);
  if (r) return r;
  {
    char* p = (char*)c->state.p;
    std::vector<void*> arr(sizes.size());
    for (size_t i = 0; i < sizes.size(); ++i) { arr[i] = p; p += (sizes[i] + 255) / 256 * 256; }
    if (fp32) {
      c->P32.pos = (float4*)arr[0];
      c->P32.prev_init = (float4*)arr[1];
      c->P32.out_in0 = (float4*)arr[2];
      c->P32.nrm_t = (float4*)arr[3];
      c->P32.pend_cos = (float4*)arr[4];
      c->P32.meta = (uint4*)arr[5];
      c->P32.cur = (uint4*)arr[6];
      c->P32.misc = (uint4*)arr[7];
      c->P32.pkey = (uint32_t*)arr[8];
      c->P32.qkey = (uint32_t*)arr[9];
      c->P32.pval = (float*)arr[10];
      c->P32.qval = (float*)arr[11];
    } else {
      c->P.cst = (double2*)arr[0];
      c->P.pw = (double4*)arr[1];
      c->P.nrm = (double4*)arr[2];
      c->P.pend = (double4*)arr[3];
      c->P.st = (uint4*)arr[4];
      c->P.sm = (uint4*)arr[5];
      c->P.aux = (uint4*)arr[6];
      c->P.pkey = (uint32_t*)arr[7];
      c->P.qkey = (uint32_t*)arr[8];
      c->P.pval = (double*)arr[9];
      c->P.qval = (double*)arr[10];
    }
  }
  {
    const ssolh_context::KeysLayout lay{c->state.p, n, npend, fp32};
    const bool clean = !arena && c->keys_clean && c->keys_layout == lay;
    c->keys_clean = false;  // until this solve drains
    c->keys_layout = lay;
    if (!clean && npend) {
      HIP_TRY(hipMemsetAsync(fp32 ? c->P32.pkey : c->P.pkey, 0xFF, pend_keys, c->stream));
      HIP_TRY(hipMemsetAsync(fp32 ? c->P32.qkey : c->P.qkey, 0xFF, pend_keys, c->stream));
    }
  }
  if ((r = alloc(c->rays, n * (fp32 ? sizeof(RayRec32) : sizeof(RayRec))))) return r;
  if ((r = alloc(c->hits, n * (fp32 ? sizeof(HitRec32) : sizeof(HitRec))))) return r;
  for (int i = 0; i < 2; ++i) if ((r = alloc(c->lists[i], n * sizeof(uint32_t)))) return r;
  if ((r = alloc(c->emitted, n))) return r;
  if ((r = alloc(c->block_emits, max_blocks * sizeof(uint32_t)))) return r;
  if ((r = alloc(c->block_offsets, max_blocks * sizeof(uint32_t)))) return r;
  if ((r = alloc(c->scan_chunks, ((max_blocks + SCAN_CHUNK - 1) / SCAN_CHUNK) * sizeof(uint32_t))))
    return r;
  if ((r = alloc(c->counters, sizeof(Counters) + 4 * sizeof(unsigned)))) return r;
  // traversal stack entries past the LDS ones (scenes whose bound exceeds
  // MAX_STACK), per work-list position
  const uint32_t lds_stack = trace_lds_stack(c);
  TOvf ovf{nullptr, slots, 0};
  if (c->stack_depth > lds_stack) {
    ovf.extra = c->stack_depth - lds_stack;
    if ((r = alloc(c->trace_ovf, n * ovf.extra * sizeof(uint32_t)))) return r;
    ovf.base = (uint32_t*)c->trace_ovf.p;
  }
  // adjoint buffers
  AdjParams A{};
  if (adj) {
    const size_t ninst = c->ninstances, ngm = (size_t)c->nprim * 2;
    const size_t in_bytes = ninst * sizeof(int32_t) + 64 + (size_t)adj->nparams * 6 * sizeof(double) +
                            ngm * sizeof(double) + 64;
    if ((r = alloc(c->adj_state, n * (sizeof(double4) + sizeof(uint32_t)) + 256))) return r;
    if ((r = alloc(c->adj_in, in_bytes))) return r;
    if ((r = alloc(c->adj_grad, (size_t)adj->nparams * 3 * sizeof(double) + 8))) return r;
    char* ai = (char*)c->adj_in.p;
    int32_t* d_ip = (int32_t*)ai;
    double* d_axes = (double*)(ai + (ninst * sizeof(int32_t) + 63) / 64 * 64);
    double* d_gm = d_axes + (size_t)adj->nparams * 6;
    HIP_TRY(hipMemcpyAsync(d_ip, adj->inst_param, ninst * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(d_axes, adj->axes, (size_t)adj->nparams * 6 * sizeof(double),
                           hipMemcpyHostToDevice, c->stream));
    if (adj->gmap && ngm)
      HIP_TRY(hipMemcpyAsync(d_gm, adj->gmap, ngm * sizeof(double), hipMemcpyHostToDevice, c->stream));
    else if (ngm)
      HIP_TRY(hipMemsetAsync(d_gm, 0, ngm * sizeof(double), c->stream));
    HIP_TRY(hipMemsetAsync(c->adj_grad.p, 0, (size_t)adj->nparams * 3 * sizeof(double), c->stream));
    A.grad = (double*)c->adj_grad.p;
    A.inst_param = d_ip;
    A.axes = d_axes;
    A.gmap = d_gm;
    A.d_total = adj->d_total;
    A.s = (double4*)c->adj_state.p;
    A.p = (uint32_t*)((char*)c->adj_state.p + (n * sizeof(double4) + 255) / 256 * 256);
  }
  // path recording buffers
  TrackParams K{};
  if (trk) {
    const uint64_t cap = std::max<uint64_t>(trk->cap, 1);
    if ((r = alloc(c->trk_vtx, cap * sizeof(ssolh_track_vertex)))) return r;
    if ((r = alloc(c->trk_type, N))) return r;
    if ((r = alloc(c->trk_seq, n * sizeof(uint32_t)))) return r;
    if ((r = alloc(c->trk_pos, n * sizeof(double4)))) return r;
    if ((r = alloc(c->trk_count, sizeof(unsigned long long)))) return r;
    HIP_TRY(hipMemsetAsync(c->trk_type.p, SSOLH_PATH_DROPPED, N, c->stream));
    HIP_TRY(hipMemsetAsync(c->trk_count.p, 0, sizeof(unsigned long long), c->stream));
    K.vtx = (ssolh_track_vertex*)c->trk_vtx.p;
    K.nvtx = (unsigned long long*)c->trk_count.p;
    K.cap = cap;
    K.type = (uint8_t*)c->trk_type.p;
    K.seq = (uint32_t*)c->trk_seq.p;
    K.pos = (double4*)c->trk_pos.p;
    K.sun_len = trk->sun_len;
    K.inf_len = trk->inf_len;
    c->trk_cap = cap;
    c->trk_paths = N;
  }
  // jobs
  if (job_suns_h) {
    if ((r = alloc(c->jobs, (njobs + 1) * sizeof(uint64_t) + njobs * sizeof(ssolh_sun)))) return r;
    HIP_TRY(hipMemcpyAsync(c->jobs.p, job_first_h, (njobs + 1) * sizeof(uint64_t),
                           hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync((char*)c->jobs.p + (njobs + 1) * sizeof(uint64_t), job_suns_h,
                           njobs * sizeof(ssolh_sun), hipMemcpyHostToDevice, c->stream));
  }

  // tallies (striped where many waves meet, see Tallies)
  if ((r = alloc(c->tallies, tally_bytes))) return r;
  c->tally_bytes = tally_bytes;
  double* tb = (double*)c->tallies.p;
  c->T.global = tb;
  c->T.rcv = c->T.global + n_global;
  c->T.samp = c->T.rcv + n_rcv;
  c->T.sxr = c->T.samp + n_samp;
  c->T.prim = c->T.sxr + n_sxr;
  c->T.samp_count = (unsigned long long*)(c->T.prim + n_prim);
  c->T.mask_g = sg - 1;
  c->T.mask_r = sr - 1;
  c->T.mask_s = ss - 1;
  c->T.mask_x = sx - 1;
  c->T.mask_p = sp - 1;
  c->T.nprim = c->nprim;
  c->T.nrcv = c->nrcv;
  c->T.nsamp = c->nsamp;
  c->T.njobs = njobs;
  c->njobs = njobs;
  HIP_TRY(hipMemsetAsync(c->tallies.p, 0, tally_bytes, c->stream));
  HIP_TRY(hipMemsetAsync(c->counters.p, 0, sizeof(Counters) + 4 * sizeof(unsigned), c->stream));
  // per-primitive log: room for one record per realisation, at most
  // LOG_MAX_RECORDS (12.9 GB); none from the scratch array
  c->T.log_key = nullptr;
  c->T.log_val = nullptr;
  c->T.log_n = nullptr;
  c->T.log_cap = 0;
  c->log_flushes = 0;
  c->log_overflowed = 0;
  if (!arena && c->nprim) {
    uint32_t cap = (uint32_t)std::min<uint64_t>(N, LOG_MAX_RECORDS);
    if (c->log_cap_test) cap = std::min(cap, c->log_cap_test);
    uint32_t bits = 1;
    while (bits < 32 && (uint64_t(1) << bits) < (uint64_t)c->nprim * 2) ++bits;
    size_t sort_bytes = 0;
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (const uint32_t*)nullptr,
                                               (uint32_t*)nullptr, (const uint32_t*)nullptr,
                                               (uint32_t*)nullptr, (int)cap, 0, (int)bits,
                                               c->stream));
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t k4 = up((size_t)cap * 4);
    const size_t bytes = 4 * k4 + up((size_t)cap * 80) + 256 + up(sort_bytes);
    if (alloc(c->plog, bytes)) {
      // no room left beside the path state: every entry takes the atomics
      (void)hipGetLastError();
      c->plog = DevBuf{};
      cap = 0;
    }
    if (cap) {
      char* q = (char*)c->plog.p;
      c->T.log_key = (uint32_t*)q;
      c->log_skey = (uint32_t*)(q + k4);
      c->log_idx = (uint32_t*)(q + 2 * k4);
      c->log_sidx = (uint32_t*)(q + 3 * k4);
      c->T.log_val = (double*)(q + 4 * k4);
      c->T.log_n = (unsigned long long*)(q + 4 * k4 + up((size_t)cap * 80));
      c->log_sort = q + 4 * k4 + up((size_t)cap * 80) + 256;
      c->log_sort_bytes = sort_bytes;
      c->log_key_bits = bits;
      c->T.log_cap = cap;
      HIP_TRY(hipMemsetAsync(c->T.log_n, 0, sizeof(unsigned long long), c->stream));
    }
  }
  // Sums the logged records into the per-primitive tally and empties the
  // log (outside graph capture: the stream is the solve's)
  auto log_flush = [&]() -> int {
    if (!c->T.log_cap) return 0;
    unsigned long long n = 0;
    HIP_TRY(hipMemcpyAsync(&n, c->T.log_n, sizeof(n), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    const uint32_t m = (uint32_t)std::min<unsigned long long>(n, c->T.log_cap);
    c->log_overflowed += n - m;
    if (m) {
      ++c->log_flushes;
      k_iota32<<<(m + 255) / 256, 256, 0, c->stream>>>(c->log_idx, m);
      HIP_TRY(hipGetLastError());
      size_t sb = c->log_sort_bytes;
      HIP_TRY(hipcub::DeviceRadixSort::SortPairs(c->log_sort, sb, c->T.log_key, c->log_skey,
                                                 c->log_idx, c->log_sidx, (int)m, 0,
                                                 (int)c->log_key_bits, c->stream));
      const uint32_t per_block = 256 * PRIM_ROUNDS;
      k_prim_reduce<<<(m + per_block - 1) / per_block, 256, 0, c->stream>>>(
          c->log_skey, c->log_sidx, c->T.log_val, m, c->T.prim, c->T.mask_p, c->T.nprim);
      HIP_TRY(hipGetLastError());
    }
    if (n) HIP_TRY(hipMemsetAsync(c->T.log_n, 0, sizeof(unsigned long long), c->stream));
    return 0;
  };

  SolveParams prm;
  prm.N = N;
  prm.max_failed = max_failed;
  prm.key = key;
  prm.counter = counter;
  prm.slots = slots;
  prm.npend = npend;
  prm.njobs = njobs;
  prm.job_first = job_suns_h ? (const uint64_t*)c->jobs.p : nullptr;
  prm.job_suns = job_suns_h ? (const ssolh_sun*)((char*)c->jobs.p + (njobs + 1) * sizeof(uint64_t))
                            : nullptr;

  Counters* ctr = (Counters*)c->counters.p;
  unsigned* cnt = (unsigned*)((char*)c->counters.p + sizeof(Counters));  // [2] list lengths
  // tail-block counters of the looping trace and shade kernels (zeroed with
  // the counters, then by every k_scan_top)
  unsigned* fetch = cnt + 2;
  uint32_t* lists[2] = {(uint32_t*)c->lists[0].p, (uint32_t*)c->lists[1].p};
  uint8_t* emitted = (uint8_t*)c->emitted.p;
  uint32_t* block_emits = (uint32_t*)c->block_emits.p;
  uint32_t* block_offsets = (uint32_t*)c->block_offsets.p;
  uint32_t* scan_chunks = (uint32_t*)c->scan_chunks.p;

  // With fewer slots than realisations and something to order by (more than
  // one sun or sampled instance), the solve runs in chunks of `slots`
  // realisations, each all in flight and in start-instance order, instead of
  // refilling slots in place (100-paraboloid field at 2^20 slots: 449 -> 629
  // Mpaths/s; its 512-sun sweep, 51,712 keys: 1,680 -> 2,290).
  const uint64_t nkeys = (uint64_t)njobs * c->ninstances;
  const bool orderable = (uint64_t)njobs * c->nsamp > 1 && nkeys <= (1u << 24);
  const bool chunked = slots < N && orderable;

  // one iteration: trace, shade, compaction of list i into list i^1
  auto iteration = [&](int i, uint32_t blocks) -> int {
    // blocks: logical blocks of the list (of its largest length under graph
    // capture). The host loop launches one workgroup per logical block; a
    // captured iteration launches each kernel's resident grid, whose
    // workgroups loop over the logical blocks, however long or short the
    // list turns out to be
    const bool loop = graph_iters != 0;
    auto grid_of = [&](const void* f) {
      return loop ? std::min<uint32_t>(blocks, resident_grid(c, f)) : blocks;
    };
    const uint32_t* list = lists[i];
    const unsigned* count = cnt + i;
    const bool refill = slots < N && !chunked;
    if (fp32) {
      RayRec32* r32 = (RayRec32*)c->rays.p;
      HitRec32* h32 = (HitRec32*)c->hits.p;
      auto launch = [&](auto kernel) {
        kernel<<<grid_of((const void*)kernel), BLOCK, 0, c->stream>>>(c->scn, c->shapeq, list,
                                                                     count, r32, h32, fetch, ovf);
      };
      if (lds_stack == 16)
        loop ? launch(k_trace32<16, true>) : launch(k_trace32<16, false>);
      else if (lds_stack == 24)
        loop ? launch(k_trace32<24, true>) : launch(k_trace32<24, false>);
      else
        loop ? launch(k_trace32<MAX_STACK, true>) : launch(k_trace32<MAX_STACK, false>);
    } else {
      RayRec* r64 = (RayRec*)c->rays.p;
      HitRec* h64 = (HitRec*)c->hits.p;
      auto launch = [&](auto kernel) {
        kernel<<<grid_of((const void*)kernel), BLOCK, 0, c->stream>>>(c->scn, list, count, r64,
                                                                     h64, fetch, ovf);
      };
      if (lds_stack == 16)
        loop ? launch(k_trace<16, true>) : launch(k_trace<16, false>);
      else if (lds_stack == 24)
        loop ? launch(k_trace<24, true>) : launch(k_trace<24, false>);
      else
        loop ? launch(k_trace<MAX_STACK, true>) : launch(k_trace<MAX_STACK, false>);
    }
    HIP_TRY(hipGetLastError());
    if (!graph_iters) HIP_TRY(hipEventRecord(c->ev[2], c->stream));
    if (fp32) {
      const Shade32Args args{c->scn,     c->shapeq, c->P32, prm, c->T, list, count,
                             (RayRec32*)c->rays.p, (const HitRec32*)c->hits.p, emitted,
                             block_emits, ctr, fetch + 1};
      auto launch = [&](auto kernel) {
        kernel<<<grid_of((const void*)kernel), BLOCK, 0, c->stream>>>(args);
      };
      if (refill) loop ? launch(k_shade32<true, true>) : launch(k_shade32<true, false>);
      else loop ? launch(k_shade32<false, true>) : launch(k_shade32<false, false>);
    } else {
      const ShadeArgs args{c->scn,  c->P,     prm,         c->T, list, count, (RayRec*)c->rays.p,
                           (const HitRec*)c->hits.p, emitted, block_emits, ctr,  A,    K,
                           fetch + 1};
      auto launch = [&](auto kernel) {
        kernel<<<grid_of((const void*)kernel), BLOCK, 0, c->stream>>>(args);
      };
      if (A.grad && refill)
        loop ? launch(k_shade<true, true, true>) : launch(k_shade<true, true, false>);
      else if (A.grad)
        loop ? launch(k_shade<true, false, true>) : launch(k_shade<true, false, false>);
      else if (refill)
        loop ? launch(k_shade<false, true, true>) : launch(k_shade<false, true, false>);
      else
        loop ? launch(k_shade<false, false, true>) : launch(k_shade<false, false, false>);
    }
    HIP_TRY(hipGetLastError());
    if (!graph_iters) HIP_TRY(hipEventRecord(c->ev[3], c->stream));
    k_scan_chunks<<<(blocks + SCAN_CHUNK - 1) / SCAN_CHUNK, SCAN_THREADS, 0, c->stream>>>(
        block_emits, count, block_offsets, scan_chunks);
    HIP_TRY(hipGetLastError());
    k_scan_top<<<1, SCAN_THREADS, 0, c->stream>>>(count, scan_chunks, cnt + (i ^ 1), &ctr->segments,
                                                  fetch);
    HIP_TRY(hipGetLastError());
    if (loop)
      k_scatter<true><<<grid_of((const void*)k_scatter<true>), BLOCK, 0, c->stream>>>(
          list, count, emitted, block_offsets, scan_chunks, lists[i ^ 1]);
    else
      k_scatter<false><<<blocks, BLOCK, 0, c->stream>>>(list, count, emitted, block_offsets,
                                                        scan_chunks, lists[i ^ 1]);
    HIP_TRY(hipGetLastError());
    return 0;
  };

  float ms_trace = 0, ms_shade = 0;
  HIP_TRY(hipEventRecord(c->ev[0], c->stream));

  // The first realisations of a wavefront, first .. first + count - 1, fill
  // slots [0, count) (list 0 is 0..count-1): in start-instance order when
  // no slot is refilled (a chunk, or every realisation in flight; refilled
  // slots take realisations r + slots, which in id order stay on one sun per
  // wave and in this order would not). The keys and indices before and
  // after the sort take the hit records, the order the second work list,
  // neither of which is used before the first trace.
  auto start = [&](uint64_t first, uint32_t count) -> int {
    k_iota<<<(count + 255) / 256, 256, 0, c->stream>>>(lists[0], count, cnt);
    HIP_TRY(hipGetLastError());
    const uint32_t* order = nullptr;
    if (count && (chunked || first + count == N) && orderable) {
      const uint32_t bits = start_key_bits(nkeys), gg = (count + 255) / 256;
      size_t sort_bytes = 0;
      HIP_TRY(start_sort_bytes(count, bits, c->stream, sort_bytes));
      if (int e = alloc(c->start_sort, sort_bytes)) return e;
      const size_t part = (size_t)(count + 63) / 64 * 64;  // 256-byte aligned parts
      uint32_t* key = (uint32_t*)c->hits.p;
      uint32_t* skey = key + part;
      uint32_t* idx = skey + part;
      if (fp32)
        k_start_keys<true><<<gg, 256, 0, c->stream>>>(c->scn, prm, first, count, c->ninstances,
                                                       key, idx);
      else
        k_start_keys<false><<<gg, 256, 0, c->stream>>>(c->scn, prm, first, count, c->ninstances,
                                                        key, idx);
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipcub::DeviceRadixSort::SortPairs(c->start_sort.p, sort_bytes, key, skey, idx,
                                                 lists[1], (int)count, 0, (int)bits, c->stream));
      order = lists[1];
    }
    const uint32_t gb = (count + BLOCK - 1) / BLOCK;
    if (fp32)
      k_generate32<<<gb, BLOCK, 0, c->stream>>>(c->scn, c->shapeq, c->P32, prm, c->T,
                                                (RayRec32*)c->rays.p, first, count, order);
    else if (A.grad)
      k_generate<true><<<gb, BLOCK, 0, c->stream>>>(c->scn, c->P, prm, c->T, (RayRec*)c->rays.p,
                                                    first, count, A, K, order);
    else
      k_generate<false><<<gb, BLOCK, 0, c->stream>>>(c->scn, c->P, prm, c->T, (RayRec*)c->rays.p,
                                                     first, count, A, K, order);
    HIP_TRY(hipGetLastError());
    return 0;
  };

  // iterations of the whole solve (reported) and of the current wavefront
  // (bounded: a wavefront drains within its deepest path, which roulette
  // ends with probability 1 - 2^-k after 16k bounces past depth 65)
  uint32_t iter = 0, wave_iter = 0;
  constexpr uint32_t MAX_WAVE_ITER = 1000000u;
  // host-driven iterations until the list is empty
  auto drain_host = [&](uint32_t count) -> int {
    int cur = 0;
    while (count) {
      if (wave_iter >= MAX_WAVE_ITER) {
        set_err("ssolh_solve: the wavefront did not drain after %u iterations", wave_iter);
        return SSOLH_ERR_LIMIT;
      }
      HIP_TRY(hipEventRecord(c->ev[1], c->stream));
      if (int e = iteration(cur, (count + BLOCK - 1) / BLOCK)) return e;
      unsigned total = 0;
      unsigned long long logged = 0;
      HIP_TRY(hipMemcpyAsync(&total, cnt + (cur ^ 1), sizeof(total), hipMemcpyDeviceToHost, c->stream));
      if (c->T.log_cap)
        HIP_TRY(hipMemcpyAsync(&logged, c->T.log_n, sizeof(logged), hipMemcpyDeviceToHost, c->stream));
      HIP_TRY(hipStreamSynchronize(c->stream));
      float t1 = 0, t2 = 0;
      (void)hipEventElapsedTime(&t1, c->ev[1], c->ev[2]);
      (void)hipEventElapsedTime(&t2, c->ev[2], c->ev[3]);
      ms_trace += t1;
      ms_shade += t2;
      if (total > count) {
        set_err("ssolh_solve: work list grew from %u to %u entries", count, total);
        return SSOLH_ERR_LIMIT;
      }
      cur ^= 1;
      count = total;
      ++iter;
      ++wave_iter;
      // past half full, the log is summed before it overflows into atomics
      if (count && c->T.log_cap && logged >= c->T.log_cap / 2)
        if (int e = log_flush()) return e;
    }
    return 0;
  };

  // graph of an even number of iterations: list 0 -> 1 -> 0 ...; every
  // kernel reads its work count from device memory, so iterations past the
  // end of the work are empty launches. Captured once, replayed for every
  // chunk until its list is empty.
  const uint32_t gi = (graph_iters + 1) & ~1u;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // graph replays until the list is empty, two in flight: the list length
  // after replay k is copied to pinned memory and read while replay k+1
  // runs, so the GPU does not wait for the host between replays; a replay
  // past the end runs empty launches
  auto drain_graph = [&](uint32_t count) -> int {
    uint32_t launched = 0, seen = 0;
    while (count) {
      while (launched - seen < 2) {
        if (wave_iter >= MAX_WAVE_ITER) {
          set_err("ssolh_solve: the wavefront did not drain after %u iterations", wave_iter);
          (void)hipStreamSynchronize(c->stream);
          return SSOLH_ERR_LIMIT;
        }
        const uint32_t k = launched & 1u;
        hipError_t e = hipGraphLaunch(exec, c->stream);
        if (e == hipSuccess)
          e = hipMemcpyAsync(c->pinned + k, cnt, sizeof(unsigned), hipMemcpyDeviceToHost,
                             c->stream);
        if (e == hipSuccess) e = hipEventRecord(c->gev[k], c->stream);
        if (e != hipSuccess) {
          set_err("graph solve: %s", hipGetErrorString(e));
          (void)hipStreamSynchronize(c->stream);
          return -(int)e;
        }
        ++launched;
        iter += gi;
        wave_iter += gi;
      }
      const uint32_t k = seen & 1u;
      const hipError_t se = hipEventSynchronize(c->gev[k]);
      if (se != hipSuccess) { set_err("graph solve: %s", hipGetErrorString(se)); return -(int)se; }
      count = c->pinned[k];
      ++seen;
    }
    const hipError_t se = hipStreamSynchronize(c->stream);
    if (se != hipSuccess) { set_err("graph solve: %s", hipGetErrorString(se)); return -(int)se; }
    return 0;
  };

  const uint64_t chunk = chunked ? slots : N;  // realisations started per wavefront
  int gerr = 0;
  uint32_t waves = 0;
  for (uint64_t first = 0; first < N && !gerr; first += chunk) {
    const uint32_t count = (uint32_t)std::min<uint64_t>(N - first, slots);
    wave_iter = 0;
    ++waves;
    if ((gerr = start(first, count))) break;
    if (!graph_iters) {
      gerr = drain_host(count);
    } else {
      if (!exec) {
        HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int cr = 0;
        for (uint32_t k = 0; k < gi && !cr; ++k) cr = iteration((int)(k & 1), max_blocks);
        const hipError_t ce = hipStreamEndCapture(c->stream, &graph);
        if (cr) { if (graph) (void)hipGraphDestroy(graph); return cr; }
        HIP_TRY(ce);
        const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (ie != hipSuccess) {
          (void)hipGraphDestroy(graph);
          HIP_TRY(ie);
        }
      }
      gerr = drain_graph(count);
    }
    if (!gerr) gerr = log_flush();
    if (!chunked || gerr) break;  // refill: one wavefront runs every realisation
    // Too many failed paths: like the reference's loop, which skips every
    // realisation not yet started once the limit is hit
    // (src/ssol_solver.c:1227, 1241-1245), no further chunk starts. The
    // stream is idle here (the drain synchronised it).
    unsigned aborted = 0;
    const hipError_t ae = hipMemcpy(&aborted, &ctr->aborted, sizeof(aborted), hipMemcpyDeviceToHost);
    if (ae != hipSuccess) { set_err("ssolh_solve: %s", hipGetErrorString(ae)); gerr = -(int)ae; break; }
    if (aborted) break;
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (gerr) return gerr;
  // fold the stripes
  auto fold = [&](auto* base, size_t nn, uint32_t stripes) -> int {
    if (stripes > 1 && nn) {
      k_fold<<<(unsigned)((nn + 255) / 256), 256, 0, c->stream>>>(base, nn, stripes);
      HIP_TRY(hipGetLastError());
    }
    return 0;
  };
  if ((r = fold(c->T.global, (size_t)njobs * 16, sg))) return r;
  if ((r = fold(c->T.rcv, (size_t)c->nrcv * 40, sr))) return r;
  if ((r = fold(c->T.samp, (size_t)c->nsamp * 4, ss))) return r;
  if ((r = fold(c->T.samp_count, (size_t)c->nsamp, ss))) return r;
  if ((r = fold(c->T.sxr, (size_t)c->nsamp * c->nrcv * 40, sx))) return r;
  if ((r = fold(c->T.prim, (size_t)c->nprim * 40, sp))) return r;
  float total_ms = 0;
  HIP_TRY(hipEventRecord(c->ev[3], c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  (void)hipEventElapsedTime(&total_ms, c->ev[0], c->ev[3]);
  HIP_TRY(hipMemcpy(&c->last, c->counters.p, sizeof(Counters), hipMemcpyDeviceToHost));
  c->segments = c->last.segments;
  if (trk) HIP_TRY(hipMemcpy(&c->trk_nvtx, c->trk_count.p, sizeof(c->trk_nvtx), hipMemcpyDeviceToHost));
  c->kernel_ms = total_ms;
  c->trace_ms = ms_trace;
  c->shade_ms = ms_shade;
  c->iterations = iter;
  c->wavefronts = waves;
  c->refilled = slots < N && !chunked;
  c->keys_clean = !arena;  // drained: every finished path cleared its keys
  return 0;
}

}  // namespace

extern "C" {

SSOLH_API int ssolh_solve(ssolh_context* c, const struct ssolh_solve_args* a) {
  if (!c || !a || !a->realisations) { set_err("ssolh_solve: invalid arguments"); return SSOLH_ERR_ARG; }
  if (!c->has_scene) { set_err("ssolh_solve: no scene uploaded"); return SSOLH_ERR_ARG; }
  if (a->precision != SSOLH_FP64 && a->precision != SSOLH_FP32) {
    set_err("ssolh_solve: unknown precision %u", a->precision);
    return SSOLH_ERR_ARG;
  }
  HIP_TRY(hipSetDevice(c->device));
  const uint32_t npend = a->pending_entries ? a->pending_entries : 4;
  return run_solve(c, a->realisations, a->max_failed, a->rng_key, a->rng_counter,
                   a->precision == SSOLH_FP32, a->path_slots, npend, 1, nullptr, nullptr,
                   a->graph_iterations);
}

SSOLH_API int ssolh_solve_noalloc(ssolh_context* c, const struct ssolh_solve_args* a,
                                  struct ssolh_tallies* out) {
  if (!c || !a || !out || !a->realisations) {
    set_err("ssolh_solve_noalloc: invalid arguments");
    return SSOLH_ERR_ARG;
  }
  if (!c->has_scene) { set_err("ssolh_solve_noalloc: no scene uploaded"); return SSOLH_ERR_ARG; }
  if (a->precision != SSOLH_FP64 && a->precision != SSOLH_FP32) {
    set_err("ssolh_solve_noalloc: unknown precision %u", a->precision);
    return SSOLH_ERR_ARG;
  }
  if (a->graph_iterations) {
    set_err("ssolh_solve_noalloc: iterations are driven from the host (graph_iterations 0)");
    return SSOLH_ERR_ARG;
  }
  if (!c->scratch) { set_err("ssolh_solve_noalloc: no scratch array on this device"); return SSOLH_ERR_MEM; }
  HIP_TRY(hipSetDevice(c->device));
  // The solve's buffers come from the scratch array: the context's own
  // buffers are set aside for the call (nothing is freed or allocated) and
  // put back after it.
  DevBuf* bufs[] = {&c->state,        &c->rays,         &c->hits,        &c->lists[0],
                    &c->lists[1],     &c->emitted,      &c->block_emits, &c->block_offsets,
                    &c->scan_chunks,  &c->counters,     &c->tallies,     &c->start_sort,
                    &c->trace_ovf};
  constexpr size_t NB = sizeof(bufs) / sizeof(bufs[0]);
  DevBuf saved[NB];
  for (size_t i = 0; i < NB; ++i) {
    saved[i] = *bufs[i];
    *bufs[i] = DevBuf{};
  }
  Arena arena{(char*)c->scratch, SCRATCH_BYTES, 0};
  const uint32_t npend = a->pending_entries ? a->pending_entries : 4;
  int r = run_solve(c, a->realisations, a->max_failed, a->rng_key, a->rng_counter,
                    a->precision == SSOLH_FP32, a->path_slots, npend, 1, nullptr, nullptr, 0,
                    nullptr, nullptr, &arena);
  if (!r) r = ssolh_tallies_download(c, out);
  for (size_t i = 0; i < NB; ++i) *bufs[i] = saved[i];
  return r;
}

SSOLH_API int ssolh_solve_batch(ssolh_context* c, const struct ssolh_batch_args* a,
                                struct ssolh_batch_result* out) {
  if (!c || !a || !out || !a->nsuns || !a->suns || !a->realisations) {
    set_err("ssolh_solve_batch: invalid arguments");
    return SSOLH_ERR_ARG;
  }
  if (!c->has_scene) { set_err("ssolh_solve_batch: no scene uploaded"); return SSOLH_ERR_ARG; }
  if (a->precision != SSOLH_FP64 && a->precision != SSOLH_FP32) {
    set_err("ssolh_solve_batch: unknown precision %u", a->precision);
    return SSOLH_ERR_ARG;
  }
  if (a->nsuns > (1u << 24)) { set_err("ssolh_solve_batch: at most 2^24 suns"); return SSOLH_ERR_LIMIT; }
  std::vector<uint64_t> first(a->nsuns + 1, 0);
  for (uint32_t j = 0; j < a->nsuns; ++j) first[j + 1] = first[j] + a->realisations[j];
  if (!first[a->nsuns]) { set_err("ssolh_solve_batch: no realisation"); return SSOLH_ERR_ARG; }
  HIP_TRY(hipSetDevice(c->device));
  const uint32_t npend = a->pending_entries ? a->pending_entries : 4;
  const uint32_t njobs = a->nsuns;
  std::vector<ssolh_sun> suns(a->suns, a->suns + a->nsuns);
  int r = run_solve(c, first[a->nsuns], a->max_failed, a->rng_key, a->rng_counter,
                    a->precision == SSOLH_FP32, a->path_slots, npend, njobs, first.data(),
                    suns.data(), a->graph_iterations);
  if (r) return r;
  std::vector<double> g((size_t)njobs * 16);
  HIP_TRY(hipMemcpy(g.data(), c->T.global, g.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (uint32_t j = 0; j < njobs; ++j) {
    memcpy(out[j].global, &g[(size_t)j * 16], 12 * sizeof(double));
    out[j].realisation_count = (uint64_t)g[(size_t)j * 16 + 12];
  }
  return 0;
}

SSOLH_API int ssolh_solve_adjoint(ssolh_context* c, const struct ssolh_adjoint_args* a,
                                  double* grad) {
  if (!c || !a || !grad || !a->solve.realisations || !a->nparams || !a->inst_param || !a->axes) {
    set_err("ssolh_solve_adjoint: invalid arguments");
    return SSOLH_ERR_ARG;
  }
  if (!c->has_scene) { set_err("ssolh_solve_adjoint: no scene uploaded"); return SSOLH_ERR_ARG; }
  if (a->solve.precision != SSOLH_FP64) {
    set_err("ssolh_solve_adjoint: the adjoint runs in the FP64 form");
    return SSOLH_ERR_ARG;
  }
  for (uint32_t i = 0; i < c->ninstances; ++i) {
    if (a->inst_param[i] >= (int32_t)a->nparams) {
      set_err("ssolh_solve_adjoint: instance %u has parameter slot %d >= %u", i, a->inst_param[i],
              a->nparams);
      return SSOLH_ERR_ARG;
    }
  }
  HIP_TRY(hipSetDevice(c->device));
  AdjHost h{a->nparams, a->inst_param, a->axes, a->gmap, a->d_absorbed};
  const uint32_t npend = a->solve.pending_entries ? a->solve.pending_entries : 4;
  int r = run_solve(c, a->solve.realisations, a->solve.max_failed, a->solve.rng_key,
                    a->solve.rng_counter, false, a->solve.path_slots, npend, 1, nullptr, nullptr,
                    a->solve.graph_iterations, &h);
  if (r) return r;
  HIP_TRY(hipMemcpy(grad, c->adj_grad.p, (size_t)a->nparams * 3 * sizeof(double),
                    hipMemcpyDeviceToHost));
  double n = 0;
  HIP_TRY(hipMemcpy(&n, c->T.global + 12, sizeof(double), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < (size_t)a->nparams * 3; ++i) grad[i] = n > 0 ? grad[i] / n : 0.0;
  return 0;
}

SSOLH_API int ssolh_solve_tracked(ssolh_context* c, const struct ssolh_solve_args* a,
                                  const struct ssolh_track_args* t, uint64_t* nvertices) {
  if (!c || !a || !t || !nvertices || !a->realisations) {
    set_err("ssolh_solve_tracked: invalid arguments");
    return SSOLH_ERR_ARG;
  }
  if (!c->has_scene) { set_err("ssolh_solve_tracked: no scene uploaded"); return SSOLH_ERR_ARG; }
  if (a->precision != SSOLH_FP64) {
    set_err("ssolh_solve_tracked: paths are recorded by the FP64 form");
    return SSOLH_ERR_ARG;
  }
  if (!(t->sun_ray_length == t->sun_ray_length) || !(t->infinite_ray_length == t->infinite_ray_length)) {
    set_err("ssolh_solve_tracked: NaN ray length");
    return SSOLH_ERR_ARG;
  }
  HIP_TRY(hipSetDevice(c->device));
  TrackHost h{t->sun_ray_length, t->infinite_ray_length, t->max_vertices};
  const uint32_t npend = a->pending_entries ? a->pending_entries : 4;
  int r = run_solve(c, a->realisations, a->max_failed, a->rng_key, a->rng_counter, false,
                    a->path_slots, npend, 1, nullptr, nullptr, 0, nullptr, &h);
  if (r) return r;
  *nvertices = c->trk_nvtx;
  return 0;
}

SSOLH_API int ssolh_track_download(ssolh_context* c, struct ssolh_track_vertex* v, uint64_t nv,
                                   uint8_t* types, uint64_t npaths) {
  if (!c || (nv && !v) || (npaths && !types)) {
    set_err("ssolh_track_download: invalid arguments");
    return SSOLH_ERR_ARG;
  }
  if (!c->trk_paths) { set_err("ssolh_track_download: no tracked solve"); return SSOLH_ERR_ARG; }
  if (nv > std::min<uint64_t>(c->trk_nvtx, c->trk_cap) || npaths > c->trk_paths) {
    set_err("ssolh_track_download: asked for %llu vertices and %llu paths, the solve holds %llu and %llu",
            (unsigned long long)nv, (unsigned long long)npaths,
            (unsigned long long)std::min<uint64_t>(c->trk_nvtx, c->trk_cap),
            (unsigned long long)c->trk_paths);
    return SSOLH_ERR_ARG;
  }
  HIP_TRY(hipSetDevice(c->device));
  if (nv) HIP_TRY(hipMemcpy(v, c->trk_vtx.p, nv * sizeof(ssolh_track_vertex), hipMemcpyDeviceToHost));
  if (npaths) HIP_TRY(hipMemcpy(types, c->trk_type.p, npaths, hipMemcpyDeviceToHost));
  return 0;
}

SSOLH_API int ssolh_tallies_download(ssolh_context* c, struct ssolh_tallies* o) {
  if (!c || !o) { set_err("ssolh_tallies_download: NULL argument"); return SSOLH_ERR_ARG; }
  HIP_TRY(hipSetDevice(c->device));
  const size_t n_rcv = (size_t)c->nrcv * 40, n_samp = (size_t)c->nsamp * 4;
  const size_t n_sxr = (size_t)c->nsamp * c->nrcv * 40, n_prim = (size_t)c->nprim * 40;
  HIP_TRY(hipMemcpy(o->global, c->T.global, 12 * sizeof(double), hipMemcpyDeviceToHost));
  if (o->receivers && n_rcv) HIP_TRY(hipMemcpy(o->receivers, c->T.rcv, n_rcv * 8, hipMemcpyDeviceToHost));
  if (o->sampled && n_samp) HIP_TRY(hipMemcpy(o->sampled, c->T.samp, n_samp * 8, hipMemcpyDeviceToHost));
  if (o->sampled_count && c->nsamp)
    HIP_TRY(hipMemcpy(o->sampled_count, c->T.samp_count, c->nsamp * 8, hipMemcpyDeviceToHost));
  if (o->sampled_x_receivers && n_sxr)
    HIP_TRY(hipMemcpy(o->sampled_x_receivers, c->T.sxr, n_sxr * 8, hipMemcpyDeviceToHost));
  if (o->primitives && n_prim) HIP_TRY(hipMemcpy(o->primitives, c->T.prim, n_prim * 8, hipMemcpyDeviceToHost));
  double cnt0 = 0;
  HIP_TRY(hipMemcpy(&cnt0, c->T.global + 12, sizeof(double), hipMemcpyDeviceToHost));
  o->realisation_count = (uint64_t)cnt0;
  o->failed_count = c->last.failed;
  o->segments = c->segments;
  o->pending_spills = c->last.spills;
  o->aborted = c->last.aborted;
  o->iterations = c->iterations;
  o->wavefronts = c->wavefronts;
  o->refilled = c->refilled;
  o->kernel_ms = c->kernel_ms;
  o->trace_ms = c->trace_ms;
  o->shade_ms = c->shade_ms;
  return 0;
}

SSOLH_API int ssolh_test_quadric_hits(ssolh_context* c, int quadric, const double qparam[4],
                                      uint32_t n, const double* org, const double* dir,
                                      const double* hint, double* t, double* normal) {
  if (!c || !n) return SSOLH_ERR_ARG;
  HIP_TRY(hipSetDevice(c->device));
  double *d_o, *d_d, *d_h, *d_t, *d_n;
  HIP_TRY(hipMalloc(&d_o, n * 24)); HIP_TRY(hipMalloc(&d_d, n * 24));
  HIP_TRY(hipMalloc(&d_h, n * 8)); HIP_TRY(hipMalloc(&d_t, n * 8)); HIP_TRY(hipMalloc(&d_n, n * 24));
  HIP_TRY(hipMemcpy(d_o, org, n * 24, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(d_d, dir, n * 24, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(d_h, hint, n * 8, hipMemcpyHostToDevice));
  k_test_quadric<<<(n + 127) / 128, 128, 0, c->stream>>>(
      quadric, make_double4(qparam[0], qparam[1], qparam[2], qparam[3]), n, d_o, d_d, d_h, d_t, d_n);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(t, d_t, n * 8, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(normal, d_n, n * 24, hipMemcpyDeviceToHost));
  (void)hipFree(d_o); (void)hipFree(d_d); (void)hipFree(d_h); (void)hipFree(d_t); (void)hipFree(d_n);
  return 0;
}

SSOLH_API int ssolh_test_sun_dirs(ssolh_context* c, const struct ssolh_sun* sun, uint32_t n,
                                  const double* u, uint32_t nu, double* dir, uint32_t* used) {
  if (!c || !sun || !n || !nu || !u || !dir || !used) return SSOLH_ERR_ARG;
  HIP_TRY(hipSetDevice(c->device));
  double *d_u, *d_o;
  uint32_t* d_used;
  HIP_TRY(hipMalloc(&d_u, (size_t)n * nu * 8)); HIP_TRY(hipMalloc(&d_o, n * 24));
  HIP_TRY(hipMalloc(&d_used, n * 4));
  HIP_TRY(hipMemcpy(d_u, u, (size_t)n * nu * 8, hipMemcpyHostToDevice));
  k_test_sun<<<(n + 127) / 128, 128, 0, c->stream>>>(*sun, n, d_u, nu, d_o, d_used);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(dir, d_o, n * 24, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(used, d_used, n * 4, hipMemcpyDeviceToHost));
  (void)hipFree(d_u); (void)hipFree(d_o); (void)hipFree(d_used);
  return 0;
}

SSOLH_API int ssolh_test_bsdf(ssolh_context* c, const struct ssolh_material* m, double eta_out,
                              double eta_in, double slab_eta, double slab_absorption, uint32_t n,
                              const double* wi, const double* normal, const double* u, uint32_t nu,
                              double* wo, double* weight, int* lobe) {
  if (!c || !m || !n || !nu || !wi || !normal || !u || !wo || !weight || !lobe) return SSOLH_ERR_ARG;
  HIP_TRY(hipSetDevice(c->device));
  double *d_wi, *d_n, *d_u, *d_wo, *d_w;
  int* d_l;
  HIP_TRY(hipMalloc(&d_wi, n * 24)); HIP_TRY(hipMalloc(&d_n, n * 24));
  HIP_TRY(hipMalloc(&d_u, (size_t)n * nu * 8)); HIP_TRY(hipMalloc(&d_wo, n * 24));
  HIP_TRY(hipMalloc(&d_w, n * 8)); HIP_TRY(hipMalloc(&d_l, n * 4));
  HIP_TRY(hipMemcpy(d_wi, wi, n * 24, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(d_n, normal, n * 24, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(d_u, u, (size_t)n * nu * 8, hipMemcpyHostToDevice));
  k_test_bsdf<<<(n + 127) / 128, 128, 0, c->stream>>>(
      *m, make_double4(eta_out, eta_in, slab_eta, slab_absorption), n, d_wi, d_n, d_u, nu, d_wo,
      d_w, d_l);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(wo, d_wo, n * 24, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(weight, d_w, n * 8, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(lobe, d_l, n * 4, hipMemcpyDeviceToHost));
  (void)hipFree(d_wi); (void)hipFree(d_n); (void)hipFree(d_u); (void)hipFree(d_wo);
  (void)hipFree(d_w); (void)hipFree(d_l);
  return 0;
}

SSOLH_API int ssolh_test_threefry(ssolh_context* c, uint32_t n, const uint64_t* ctr,
                                  const uint64_t* key, uint64_t* out) {
  if (!c || !n) return SSOLH_ERR_ARG;
  HIP_TRY(hipSetDevice(c->device));
  uint64_t *d_c, *d_k, *d_o;
  HIP_TRY(hipMalloc(&d_c, n * 32)); HIP_TRY(hipMalloc(&d_k, n * 32)); HIP_TRY(hipMalloc(&d_o, n * 32));
  HIP_TRY(hipMemcpy(d_c, ctr, n * 32, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(d_k, key, n * 32, hipMemcpyHostToDevice));
  k_test_threefry<<<(n + 127) / 128, 128, 0, c->stream>>>(n, d_c, d_k, d_o);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(out, d_o, n * 32, hipMemcpyDeviceToHost));
  (void)hipFree(d_c); (void)hipFree(d_k); (void)hipFree(d_o);
  return 0;
}

SSOLH_API int ssolh_test_wave_sum(ssolh_context* c, uint32_t nwaves, const double* in,
                                  double* out) {
  if (!c || !nwaves || !in || !out) return SSOLH_ERR_ARG;
  HIP_TRY(hipSetDevice(c->device));
  double *d_in, *d_out;
  const size_t bytes = (size_t)nwaves * 64 * sizeof(double);
  HIP_TRY(hipMalloc(&d_in, bytes)); HIP_TRY(hipMalloc(&d_out, bytes));
  HIP_TRY(hipMemcpy(d_in, in, bytes, hipMemcpyHostToDevice));
  k_test_wave_sum<<<nwaves, 64, 0, c->stream>>>(d_in, d_out);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(out, d_out, bytes, hipMemcpyDeviceToHost));
  (void)hipFree(d_in); (void)hipFree(d_out);
  return 0;
}

SSOLH_API int ssolh_test_trace_stack(ssolh_context* c, uint32_t lds_cap, uint32_t* bound) {
  if (!c || !(lds_cap == 0 || lds_cap == 16 || lds_cap == 24 || lds_cap == 32)) {
    set_err("ssolh_test_trace_stack: cap %u not 0, 16, 24 or 32", lds_cap);
    return SSOLH_ERR_ARG;
  }
  c->lds_stack_cap = lds_cap;
  if (bound) *bound = c->has_scene ? c->stack_depth : 0;
  return 0;
}

SSOLH_API int ssolh_test_prim_log(ssolh_context* c, uint32_t records, uint32_t* flushes,
                                  uint64_t* overflowed) {
  if (!c) return SSOLH_ERR_ARG;
  c->log_cap_test = records;
  if (flushes) *flushes = c->log_flushes;
  if (overflowed) *overflowed = c->log_overflowed;
  return 0;
}

SSOLH_API int ssolh_copy_bandwidth(ssolh_context* c, size_t bytes, int reps, double* gbs) {
  if (!c || bytes < 16 || reps < 1 || !gbs) return SSOLH_ERR_ARG;
  HIP_TRY(hipSetDevice(c->device));
  const size_t n = bytes / 16;
  uint4 *a, *b;
  HIP_TRY(hipMalloc(&a, n * 16));
  HIP_TRY(hipMalloc(&b, n * 16));
  HIP_TRY(hipMemsetAsync(a, 0x3c, n * 16, c->stream));
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0)); HIP_TRY(hipEventCreate(&e1));
  const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
  k_copy<<<grid, 256, 0, c->stream>>>(a, b, n);  // warm
  HIP_TRY(hipEventRecord(e0, c->stream));
  for (int r = 0; r < reps; ++r) k_copy<<<grid, 256, 0, c->stream>>>(a, b, n);
  HIP_TRY(hipEventRecord(e1, c->stream));
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventSynchronize(e1));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
  *gbs = 2.0 * (double)(n * 16) * reps / (ms * 1e-3) / 1e9;
  (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
  (void)hipFree(a); (void)hipFree(b);
  return 0;
}

}  // extern "C"
