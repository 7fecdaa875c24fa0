// Single-precision form of the wavefront (included by ssolh.hip inside its
// anonymous namespace, after the shared records and helpers).
//
// Positions are rebased per instance: a point is (instance b, float offset
// from b's translation), so a float carries it with the precision of the
// instance's own size however far the instance sits from the origin. Rays
// are traversed in FP32 (TLAS in world space around the float origin, BLAS
// and watertight triangle test in the instance's frame), BSDFs and weights
// are FP32; the quadric root refinement (src/ssol_shape.c:864-1135) and the
// instance-entry transform run in FP64, and tallies accumulate in FP64.

struct alignas(16) RayRec32 {
  float ox, oy, oz;     // origin, offset from the translation of instance `base`
  uint32_t base;
  float dx, dy, dz;
  float tmin;
  uint32_t iss_from, prim_from, flags;
  float wl;             // the path's wavelength (the shade's, not the trace's)
};
static_assert(sizeof(RayRec32) == 48, "RayRec32 layout");

struct alignas(16) HitRec32 {
  float px, py, pz;     // hit point, offset from the translation of the hit instance
  float t;              // distance of the nearest accepted triangle
  float nx, ny, nz;     // unit world normal (triangle Ng or quadric gradient)
  uint32_t iss, prim;   // iss == NONE <=> miss
  uint32_t pad[3];
};
static_assert(sizeof(HitRec32) == 48, "HitRec32 layout");

struct PathState32 {
  float4* pos;        // start point: offset from the base instance's translation, base
                      // (bits); read by the path's first shading only (later points
                      // come from the hit records)
  float4* prev_init;  // previous outgoing weights (3), initial flux
  float4* out_in0;    // outgoing weights of the last shaded point (3), its incoming flux
  float4* nrm_t;      // normal of the current point, last hit distance
  float4* pend_cos;   // pending absorbed_by_receivers, extinguished, other; cos factor
  uint4* meta;        // path id lo/hi, rng block, flags
  uint4* cur;         // (instance, shaded shape) index, prim, side, sampled slot
  uint4* misc;        // depth, in medium | out medium << 16, survivor score, roulette interval
  uint32_t* pkey;     // [npend][slots]
  float* pval;        // [npend][10][slots]
  uint32_t* qkey;
  float* qval;
};

// Object -> quadric frame of each punched shape: p_q = Rinv (p_obj - t)
struct ShapeQ {
  double Rinv[9];
  double t[3];
};

__device__ __forceinline__ d3 inst_T(const ssolh_instance& in) {
  return mk(in.transform[9], in.transform[10], in.transform[11]);
}

// Ray origin and direction in the frame of instance i, FP64 (the
// translation difference is exact before the rotation)
__device__ __forceinline__ void to_instance(const DScene& S, uint32_t base, f3 off, f3 dir,
                                            uint32_t i, d3& o, d3& d) {
  const ssolh_instance& in = S.insts[i];
  const d3 rel = (inst_T(S.insts[base]) - inst_T(in)) + tod(off);
  o = mul(in.inv_linear, rel);
  d = mul(in.inv_linear, tod(dir));
}

struct FShear {
  int kx, ky, kz;
  float Sx, Sy, Sz;
};

__device__ __forceinline__ FShear make_shear32(f3 d) {
  const float a[3] = {fabsf(d.x), fabsf(d.y), fabsf(d.z)};
  const float v[3] = {d.x, d.y, d.z};
  FShear s;
  s.kz = a[0] > a[1] ? (a[0] > a[2] ? 0 : 2) : (a[1] > a[2] ? 1 : 2);
  s.kx = (s.kz + 1) % 3;
  s.ky = (s.kx + 1) % 3;
  if (v[s.kz] < 0) { const int t = s.kx; s.kx = s.ky; s.ky = t; }
  s.Sx = v[s.kx] / v[s.kz];
  s.Sy = v[s.ky] / v[s.kz];
  s.Sz = 1.0f / v[s.kz];
  return s;
}

__device__ __forceinline__ float compf(f3 p, int k) { return k == 0 ? p.x : (k == 1 ? p.y : p.z); }

// Watertight ray/triangle in FP32 (Woop, Benthin, Wald, JCGT 2013): an edge
// function that comes out exactly zero is recomputed in FP64, as the paper
// prescribes, so shared edges stay closed. No FMA contraction (exact
// negation of shared edges). Returns the distance or -1.
__device__ __forceinline__ float ray_triangle32(f3 o, const FShear& s, f3 v0, f3 v1, f3 v2) {
#pragma clang fp contract(off)
  const f3 A = v0 - o, B = v1 - o, C = v2 - o;
  const float Az = compf(A, s.kz), Bz = compf(B, s.kz), Cz = compf(C, s.kz);
  const float Ax = compf(A, s.kx) - s.Sx * Az, Ay = compf(A, s.ky) - s.Sy * Az;
  const float Bx = compf(B, s.kx) - s.Sx * Bz, By = compf(B, s.ky) - s.Sy * Bz;
  const float Cx = compf(C, s.kx) - s.Sx * Cz, Cy = compf(C, s.ky) - s.Sy * Cz;
  float U = Cx * By - Cy * Bx;
  float V = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  if (U == 0.0f || V == 0.0f || W == 0.0f) {
    U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
    V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
    W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
  }
  if ((U < 0 || V < 0 || W < 0) && (U > 0 || V > 0 || W > 0)) return -1;
  const float det = U + V + W;
  if (det == 0) return -1;
  const float T = U * s.Sz * Az + V * s.Sz * Bz + W * s.Sz * Cz;
  return T / det;
}

struct TraceState32 {
  float tbest;
  uint32_t iss, prim;
  bool punched;       // the best triangle belongs to a punched surface
  f3 nobj;            // object-space normal of the best (mesh) triangle
  double qdst;        // nearest refined quadric distance
  uint32_t qiss;      // its (instance, shaded shape)
};

// Reference any-hit filter (src/ssol_scene.c:438-536) in the frame of the
// instance being traversed (d: its FP32 ray direction); the quadric
// refinement rebuilds that ray in FP64 from the rebased origin.
__device__ bool filter32(const DScene& S, const ShapeQ* SQ, const RayRec32& r, f3 o, f3 d,
                         uint32_t iss_id, const ssolh_triangle& tri, float t,
                         TraceState32& ts, f3& nobj, bool& punched) {
  const uint32_t side_from = (r.flags >> RF_SIDE_SHIFT) & 3u;
  const IssInfo is = S.iss[iss_id];
  uint32_t side;
  double dst = 0;
  if (t < r.tmin) return false;
  punched = is.shape_type != SSOLH_SHAPE_MESH;
  if (!punched) {
    if (t <= 1.e-6f || t <= r.tmin || (tri.prim == r.prim_from && iss_id == r.iss_from))
      return false;
    const f3 v0 = ldf(tri.v0), v1 = ldf(tri.v1), v2 = ldf(tri.v2);
    nobj = cross(v2 - v0, v1 - v0);
    side = dot(nobj, d) < 0 ? SSOL_FRONT_ : SSOL_BACK_;
  } else {
    const ssolh_shape& sh = S.shapes[is.shape];
    const ShapeQ& q = SQ[is.shape];
    d3 od, dd;
    to_instance(S, r.base, mkf(r.ox, r.oy, r.oz), mkf(r.dx, r.dy, r.dz), is.inst, od, dd);
    const d3 lo = mul(q.Rinv, od - ld3(q.t));
    const d3 ld = mul(q.Rinv, dd);
    d3 grad;
    if (!quadric_hit(sh.quadric, sh.qparam, lo, ld, (double)t, dst, grad)) return false;
    if (!(dst < (double)FLT_MAX)) return false;
    if ((float)dst <= r.tmin) return false;
    side = dot(ld, grad) < 0 ? SSOL_FRONT_ : SSOL_BACK_;
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

// punched_rehit_possible for the FP32 form, with the FP32 filter's own
// quadric frame and side test (ray rebased at base, candidates within
// distance t_hi)
__device__ bool punched_rehit_possible32(const DScene& S, const ShapeQ* __restrict__ SQ,
                                         const RayRec32& r, uint32_t iss_id, double t_hi) {
  const uint32_t side_from = (r.flags >> RF_SIDE_SHIFT) & 3u;
  const IssInfo is = S.iss[iss_id];
  const ssolh_shape& sh = S.shapes[is.shape];
  const ShapeQ& q = SQ[is.shape];
  d3 od, dd;
  to_instance(S, r.base, mkf(r.ox, r.oy, r.oz), mkf(r.dx, r.dy, r.dz), is.inst, od, dd);
  const d3 lo = mul(q.Rinv, od - ld3(q.t));
  const d3 ld = mul(q.Rinv, dd);
  double a, b, c, t[2];
  if (!quadric_coeffs(sh.quadric, sh.qparam, lo, ld, a, b, c)) return false;
  const int n = quadric_roots(a, b, c, t);
  for (int k = 0; k < n; ++k) {
    if (!(t[k] < (double)FLT_MAX)) continue;
    if ((float)t[k] <= r.tmin) continue;
    const d3 grad = quadric_grad(sh.quadric, sh.qparam, lo + ld * t[k]);
    const uint32_t side = dot(ld, grad) < 0 ? SSOL_FRONT_ : SSOL_BACK_;
    if (side != side_from) continue;
    if (root_choosable(t, n, k, t_hi)) return true;
  }
  return false;
}

// Entry q of the work list, FP32 form
template <int STACK>
__device__ __forceinline__ void trace_ray32(const DScene& S, const ShapeQ* __restrict__ SQ,
                                            const uint32_t* __restrict__ list, uint32_t q,
                                            const RayRec32* __restrict__ rays,
                                            HitRec32* __restrict__ hits, uint32_t* lds_stack,
                                            const Node4* lds_top, const TOvf& ovf) {

  const uint32_t slot = slot_of(list, q);
  const RayRec32 r = rays[slot];
  const f3 off = mkf(r.ox, r.oy, r.oz), dw = mkf(r.dx, r.dy, r.dz);
  const f3 ow = tof(inst_T(S.insts[r.base]) + tod(off));  // world origin, float
  const bool any_hit = (r.flags & RF_ANY_HIT) != 0;
  TraceState32 ts;
  ts.tbest = INFINITY;
  ts.iss = NONE;
  ts.prim = 0;
  ts.punched = false;
  ts.nobj = mkf(0, 0, 0);
  ts.qdst = INFINITY;
  ts.qiss = NONE;

  TStack st{lds_stack + threadIdx.x, ovf.base ? ovf.base + q : nullptr, ovf.stride, ovf.extra, 0};
  f3 o = ow, d = dw;           // current-level ray
  FRay fr = make_fray32(ow.x, ow.y, ow.z, dw.x, dw.y, dw.z);
  const FRay fw = fr;
  FShear sh = make_shear32(dw);
  // a ray leaving a planar object does not enter that object again, nor a
  // one-shape object's punched shape that no candidate could hit again
  // (as trace_ray)
  uint32_t skip_inst = NONE;
  if (r.tmin == 0.f) {
    const IssInfo e = S.iss[r.iss_from];
    bool skip = (e.obj_flags & SSOLH_OBJECT_PLANAR) != 0;
    if (!skip && (e.obj_flags & OBJ_ONE_SHAPE) && e.shape_type != SSOLH_SHAPE_MESH) {
      d3 od, dd;
      to_instance(S, r.base, off, dw, e.inst, od, dd);
      const double t_hi = box_exit(S.obj_box + 6 * (size_t)S.insts[e.inst].object, od, dd);
      skip = !punched_rehit_possible32(S, SQ, r, r.iss_from, t_hi);
    }
    if (skip) skip_inst = e.inst;
  }
  bool in_blas = false;
  uint32_t iss_first = 0;
  const float ftmin = r.tmin;

  uint32_t ref = S.root_ref;  // the TLAS's top levels come from the LDS copy (as trace_ray)
  bool alive = true;
  float tmax = FLT_MAX;  // box culling distance (as trace_ray)
  while (alive) {
    while (!(ref & LEAF)) {
      const NodeRows n =
          (ref & TOP_REF) ? lds_rows(lds_top, ref & 0x0FFFFFFFu, fr) : node_rows(S.nodes, ref, fr);
      if (!visit4<STACK>(n, fr, ftmin, tmax, S.slack_scale, ref, st)) {
        alive = false;
        break;
      }
    }
    if (!alive) break;
    if (ref == SENTINEL) {
      in_blas = false;
      fr = fw;
      alive = tpop<STACK>(st, ref);
      continue;
    }
    const uint32_t first = ref & 0x0FFFFFFFu;
    if (!in_blas) {
      const uint32_t ii = S.tlas_inst[first];
      if (ii == skip_inst) {
        alive = tpop<STACK>(st, ref);
        continue;
      }
      d3 od, dd;
      to_instance(S, r.base, off, dw, ii, od, dd);
      o = tof(od);
      d = tof(dd);
      fr = make_fray(od, dd);
      sh = make_shear32(d);
      iss_first = S.insts[ii].isshape_first;
      in_blas = true;
      tpush<STACK>(st, SENTINEL);
      ref = S.objects[S.insts[ii].object].blas_root;
      continue;
    }
    const uint32_t cnt = ((ref >> 28) & 7u) + 1u;
    bool done = false;
    for (uint32_t k = 0; k < cnt; ++k) {
      const ssolh_triangle tri = S.tris[first + k];
      const float t = ray_triangle32(o, sh, ldf(tri.v0), ldf(tri.v1), ldf(tri.v2));
      if (!(t >= r.tmin && t < ts.tbest)) continue;
      f3 nobj;
      bool punched;
      const uint32_t iss = iss_first + tri.sshape;
      if (!filter32(S, SQ, r, o, d, iss, tri, t, ts, nobj, punched)) continue;
      ts.tbest = t;
      tmax = t < FLT_MAX ? t * 1.0000002f + 1e-30f : FLT_MAX;
      ts.iss = iss;
      ts.prim = tri.prim;
      ts.nobj = nobj;
      ts.punched = punched;
      if (any_hit) { done = true; break; }
    }
    if (done) break;
    alive = tpop<STACK>(st, ref);
  }

  HitRec32 h;
  h.t = ts.tbest;
  h.iss = ts.iss;
  h.prim = ts.prim;
  h.pad[0] = h.pad[1] = h.pad[2] = 0;
  if (ts.iss != NONE) {
    // hit point relative to the hit instance, and the world normal, in FP64
    const uint32_t hiss = ts.punched ? ts.qiss : ts.iss;
    const IssInfo is = S.iss[hiss];
    const ssolh_instance& in = S.insts[is.inst];
    d3 od, dd;
    to_instance(S, r.base, off, dw, is.inst, od, dd);
    d3 p_obj, n_w;
    if (ts.punched) {
      const ssolh_shape& shp = S.shapes[is.shape];
      const ShapeQ& qq = SQ[is.shape];
      p_obj = od + dd * ts.qdst;
      const d3 pq = mul(qq.Rinv, p_obj - ld3(qq.t));
      n_w = tmul(S.isshapes[hiss].R_inv, quadric_grad(shp.quadric, shp.qparam, pq));
    } else {
      p_obj = od + dd * (double)ts.tbest;
      n_w = tmul(in.inv_linear, tod(ts.nobj));
    }
    const d3 p = mul(in.transform, p_obj);
    n_w = normalize(n_w);
    h.px = (float)p.x; h.py = (float)p.y; h.pz = (float)p.z;
    h.nx = (float)n_w.x; h.ny = (float)n_w.y; h.nz = (float)n_w.z;
  } else {
    h.px = h.py = h.pz = 0;
    h.nx = h.ny = h.nz = 0;
  }
  hits[slot] = h;
}

template <int STACK, bool LOOP>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4)))
k_trace32(DScene S, const ShapeQ* __restrict__ SQ, const uint32_t* __restrict__ list,
          const unsigned* __restrict__ count_ptr, const RayRec32* __restrict__ rays,
          HitRec32* __restrict__ hits, unsigned* fetch, TOvf ovf) {
  __shared__ uint32_t lds_stack[STACK * BLOCK];
  __shared__ Node4 lds_top[TOP_NODES];
  const uint32_t count = *count_ptr;
  if (blockIdx.x * BLOCK >= count) return;  // whole blocks only: the staging syncs
  for (uint32_t i = threadIdx.x; i < S.ntop; i += BLOCK) lds_top[i] = S.top[i];
  __syncthreads();
  if (LOOP) {  // graph-captured launch of a fixed grid (as k_trace)
    for_each_block((count + BLOCK - 1) / BLOCK, fetch, [&](uint32_t b) {
      const uint32_t q = b * BLOCK + threadIdx.x;
      if (q < count) trace_ray32<STACK>(S, SQ, list, q, rays, hits, lds_stack, lds_top, ovf);
    });
  } else {
    const uint32_t q = blockIdx.x * BLOCK + threadIdx.x;
    if (q < count) trace_ray32<STACK>(S, SQ, list, q, rays, hits, lds_stack, lds_top, ovf);
  }
}

// Path generation in FP32 (reference point_init, src/ssol_solver.c:207-358)
struct Emit32 {
  bool has;
  RayRec32 r;
};

// Sample triangle of a start point from the first two uniforms of its
// stream (FP32 form of alias_pick)
__device__ __forceinline__ uint32_t alias_pick32(const DScene& S, float u0, float uc) {
  uint32_t i = (uint32_t)(u0 * (float)S.nstris);
  if (i >= S.nstris) i = S.nstris - 1;
  const ssolh_alias al = S.alias[i];
  return (double)uc < al.prob ? i : al.alias;
}

__device__ void generate32(const DScene& S, const ShapeQ* SQ, const PathState32& P,
                           const SolveParams& prm, uint32_t slot, uint64_t path, Emit32& e,
                           int& samp_slot) {
  PathRng32 rng;
  rng.init(prm.key, prm.counter + path, 0);
  const float u0 = rng.next(), uc = rng.next(), u1 = rng.next(), u2 = rng.next();
  const uint32_t idx = alias_pick32(S, u0, uc);
  const ssolh_sample_tri st = S.stris[idx];
  const ssolh_instance& inst = S.insts[st.instance];
  const float sq = sqrtf(u1);
  const float b0 = 1 - sq, b1 = sq * (1 - u2), b2 = sq * u2;
  const f3 v0 = ldf(st.v0), v1 = ldf(st.v1), v2 = ldf(st.v2);
  const f3 p_obj = v0 * b0 + v1 * b1 + v2 * b2;
  f3 off = mulf(inst.transform, p_obj);
  const f3 ns = normalize(tmulf(inst.inv_linear, cross(v2 - v0, v1 - v0)));

  const uint32_t job = job_of(prm, path);
  const ssolh_sun& sun = job_sun(S, prm, job);
  const f3 dir = sample_sun32(sun, rng);
  const float wl = (float)sample_wavelength(S.spectra, sun.wl_spectrum, (double)rng.next());

  const uint32_t iss = inst.isshape_first + st.sshape;
  const IssInfo is = S.iss[iss];
  f3 N = ns;
  const bool punched = is.shape_type == SSOLH_SHAPE_PUNCHED;
  if (punched) {
    // lift onto the quadric in FP64 (src/ssol_shape.c:1408-1447)
    const ssolh_shape& sh = S.shapes[is.shape];
    const ShapeQ& q = SQ[is.shape];
    d3 loc = mul(q.Rinv, tod(p_obj) - ld3(q.t));
    loc.z = quadric_z(sh.quadric, sh.qparam, loc.x, loc.y);
    N = tof(normalize(tmul(S.isshapes[iss].R_inv, quadric_grad(sh.quadric, sh.qparam, loc))));
    const ssolh_inst_sshape& qs = S.isshapes[iss];
    const d3 pw = mul(qs.R, loc) + ld3(qs.T) - inst_T(inst);
    off = tof(pw);
  }
  uint32_t side = SSOL_FRONT_;
  if (dot(N, dir) >= 0) { side = SSOL_BACK_; N = -N; }

  const f3 sun0 = mkf((float)sun.dir[0], (float)sun.dir[1], (float)sun.dir[2]);
  const float surface_sun_cos = dot(N, dir);
  const float surface_sun0_cos = fabsf(dot(sun0, N));
  const float sun0_sun_cos = dot(sun0, dir);
  const float proxy_cos = punched ? fabsf(dot(ns, N)) : 1.0f;
  const float w0 = (float)(sun.dni * S.S_proxy) * fabsf(surface_sun_cos / (proxy_cos * sun0_sun_cos));
  const float cosf_ = (float)(S.S_proxy / S.S) * surface_sun0_cos / proxy_cos;

  const ssolh_material& m = S.mtls[side == SSOL_FRONT_ ? is.mtl_front : is.mtl_back];
  const uint32_t med = start_medium(S, m, side);

  P.pos[slot] = make_float4(off.x, off.y, off.z, __uint_as_float(st.instance));
  P.prev_init[slot] = make_float4(w0, w0, w0, w0);
  P.nrm_t[slot] = make_float4(N.x, N.y, N.z, 0);
  P.pend_cos[slot] = make_float4(0, 0, 0, cosf_);
  P.meta[slot] = make_uint4((uint32_t)path, (uint32_t)(path >> 32), rng.block, PF_ACTIVE);
  P.cur[slot] = make_uint4(iss, st.prim, side, (uint32_t)inst.sampled);
  P.misc[slot] = make_uint4(0, med | (med << 16), 0, 64u | (job << 8));
  // pending keys: NONE already (as generate)
  samp_slot = inst.sampled;

  e.has = true;
  e.r.ox = off.x; e.r.oy = off.y; e.r.oz = off.z;
  e.r.base = st.instance;
  e.r.dx = -dir.x; e.r.dy = -dir.y; e.r.dz = -dir.z;
  e.r.tmin = 0;
  e.r.iss_from = iss;
  e.r.prim_from = st.prim;
  e.r.flags = (side << RF_SIDE_SHIFT) | RF_DISCARD_VIRTUAL | RF_REVERSED | RF_ANY_HIT;
  e.r.wl = wl;
}

__global__ void __launch_bounds__(BLOCK)
k_generate32(DScene S, const ShapeQ* __restrict__ SQ, PathState32 P, SolveParams prm, Tallies T,
             RayRec32* __restrict__ rays, uint64_t first, uint32_t count,
             const uint32_t* __restrict__ order) {
  const uint32_t slot = blockIdx.x * BLOCK + threadIdx.x;
  const bool live = slot < count;
  Emit32 e;
  e.has = false;
  int samp_slot = 0;
  if (live) generate32(S, SQ, P, prm, slot, first + (order ? order[slot] : slot), e, samp_slot);  // as k_generate
  flush_sampled_count(live, samp_slot, T);
  if (e.has) rays[slot] = e.r;
}

__device__ bool pend_add32(uint32_t* keys, float* vals, size_t stride, uint32_t npend,
                           uint32_t key, const float* w) {
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

__device__ void spill32(double* base, uint64_t key, const float* w, unsigned long long* spills) {
  atomicAdd(spills, 1ull);
  for (int f = 0; f < 10; ++f) {
    const double x = w[f];
    atomic_add(base + key * 20 + 2 * f, x);
    atomic_add(base + key * 20 + 2 * f + 1, x * x);
  }
}

// Shading in FP32 (reference trace_radiative_path loop body,
// src/ssol_solver.c:939-1086), logical block b; same structure as k_shade
// (REFILL: slots take further realisations).
template <bool REFILL>
__device__ __forceinline__ void shade32_block(
    uint32_t b, const DScene& S, const ShapeQ* __restrict__ SQ, const PathState32& P,
    const SolveParams& prm, const Tallies& T, const uint32_t* __restrict__ list, uint32_t count,
    RayRec32* __restrict__ rays, const HitRec32* __restrict__ hits,
    uint8_t* __restrict__ emitted, uint32_t* __restrict__ block_emits, Counters* C) {
  const uint32_t q = b * BLOCK + threadIdx.x;
  const bool live = q < count;
  const size_t stride = prm.slots;

  Emit32 e;
  e.has = false;
  Finish fin;
  fin.kind = FIN_NONE;
  fin.factor = 1;
  fin.samp_slot = 0;
  fin.job = 0;
  for (int i = 0; i < 6; ++i) fin.g[i] = 0;
  const uint32_t slot = live ? slot_of(list, q) : 0;
  // receiver weights of a hit that ends the path, flushed from registers (as
  // shade_block)
  uint32_t dkey = NONE, dpkey = NONE;
  float din[4] = {0, 0, 0, 0};  // in0, in1, in2, kabs of that hit

  if (live) {
    const RayRec32 r = rays[slot];
    const HitRec32 h = hits[slot];
    uint4 meta = P.meta[slot];
    uint4 cur = P.cur[slot];
    uint4 misc = P.misc[slot];
    // the start point, shaded after its shadow ray; later points come from
    // the hit record
    float4 ps = (meta.w & PF_STAGE_BOUNCE) ? make_float4(0.f, 0.f, 0.f, 0.f) : P.pos[slot];
    float4 pv = P.prev_init[slot];
    float4 oi = (meta.w & PF_STAGE_BOUNCE) ? P.out_in0[slot] : make_float4(0, 0, 0, 0);
    float4 pc = P.pend_cos[slot];
    const uint64_t path = (uint64_t)meta.x | ((uint64_t)meta.y << 32);
    PathRng32 rng;
    rng.init(prm.key, prm.counter + path, meta.z);
    uint32_t flags = meta.w;
    fin.samp_slot = (int)cur.w;

    uint32_t in_med = misc.y & 0xFFFFu, out_med = misc.y >> 16;
    uint32_t depth = misc.x, score = misc.z, rinterval = misc.w & 0xFFu;
    fin.job = misc.w >> 8;
    f3 pos = mkf(ps.x, ps.y, ps.z);
    uint32_t base = __float_as_uint(ps.w);
    // the traced ray carries the path's direction (reversed for the shadow
    // ray toward the sun) and its wavelength
    const float wl = r.wl;
    f3 dir = (meta.w & PF_STAGE_BOUNCE) ? mkf(r.dx, r.dy, r.dz) : mkf(-r.dx, -r.dy, -r.dz);
    f3 N = mkf(0, 0, 0);
    if (!(flags & PF_STAGE_BOUNCE)) {
      const float4 nt = P.nrm_t[slot];
      N = mkf(nt.x, nt.y, nt.z);
    }
    float hit_t = 0;
    bool head = false;
    uint32_t iss = cur.x, prim = cur.y, side = cur.z;

    if (!(flags & PF_STAGE_BOUNCE)) {
      if (h.iss != NONE) {
        fin.kind = FIN_DONE;
        fin.g[0] = pc.w;
        fin.g[2] = pv.w;
      } else {
        head = true;
      }
    } else {
      const bool found = h.iss != NONE;
      const bool hit_virtual = (flags & PF_HIT_VIRTUAL) != 0;
      const bool in_atm = (flags & PF_IN_ATM) != 0;
      bool last = false;
      if (!found) {
        last = true;
        if (out_med != S.air) fin.kind = FIN_FAILED;
      }
      if (fin.kind == FIN_NONE && (last || !hit_virtual)) {
        const float absorbed = pv.x - oi.w;
        if (in_atm) pc.y += absorbed; else pc.z += absorbed;
        if (last) {
          fin.kind = FIN_DONE;
          fin.g[3] = oi.x;
        } else {
          pv.x = oi.x; pv.y = oi.y; pv.z = oi.z;
        }
      }
      if (fin.kind == FIN_NONE) {
        depth += hit_virtual ? 0u : 1u;
        if (depth > rinterval && depth % rinterval == 1) {
          if (rng.next() > 0.5f) {
            score += 1;
            rinterval = 16;
          } else {
            fin.kind = FIN_KILLED;
          }
        }
      }
      if (fin.kind == FIN_NONE) {
        iss = h.iss;
        prim = h.prim;
        base = S.iss[iss].inst;
        pos = mkf(h.px, h.py, h.pz);
        N = mkf(h.nx, h.ny, h.nz);
        if (dot(dir, N) < 0) side = SSOL_FRONT_;
        else { side = SSOL_BACK_; N = -N; }
        in_med = out_med;
        hit_t = h.t;
        head = true;
      }
    }

    if (head) {
      const IssInfo is = S.iss[iss];
      const ssolh_material& m = S.mtls[side == SSOL_FRONT_ ? is.mtl_front : is.mtl_back];
      const bool in_atm = in_med == S.air;
      const bool receiver = (is.receiver_mask & side) != 0;
      const bool virt = m.type == SSOLH_MTL_VIRTUAL;
      float trans = 1;
      if (hit_t > 0) {
        const float k = (float)data_value(S.spectra, S.media[in_med].extinction, (double)wl);
        if (k > 0) trans = expf(-k * hit_t);
      }
      const float in0 = pv.x * trans;
      const float in1 = in_atm ? pv.y : pv.y * trans;
      const float in2 = !in_atm ? pv.z : pv.z * trans;
      float out0, out1, out2, kabs;
      uint32_t next_med = in_med;
      if (virt) {
        out0 = in0; out1 = in1; out2 = in2; kabs = 0;
      } else {
        float eta_out = 1, eta_in = 1, slab_eta = 1, slab_abs = 0;
        if (m.type == SSOLH_MTL_DIELECTRIC) {
          if ((uint32_t)m.medium_out != in_med) fin.kind = FIN_FAILED;
          eta_out = (float)data_value(S.spectra, S.media[m.medium_out].refractive_index, wl);
          eta_in = (float)data_value(S.spectra, S.media[m.medium_in].refractive_index, wl);
        } else if (m.type == SSOLH_MTL_THIN_DIELECTRIC) {
          eta_out = (float)data_value(S.spectra, S.media[m.medium_out].refractive_index, wl);
          slab_eta = (float)data_value(S.spectra, m.slab_eta, wl);
          slab_abs = (float)data_value(S.spectra, m.slab_absorption, wl);
        }
        float w = 0;
        int lobe = LOBE_NONE;
        const f3 wi = -dir;
        if (fin.kind == FIN_NONE && dot(wi, N) > 0) {
          f3 wo;
          w = bsdf_sample32(m, eta_out, eta_in, slab_eta, slab_abs, wi, N, rng, wo, lobe);
          const float c = dot(N, wo);
          if ((c > 0 && lobe == LOBE_TRANSMIT) || (c < 0 && lobe == LOBE_REFLECT)) w = 0;
          dir = wo;
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
          pc.x += in0 - out0;
          float w10[10];
          w10[0] = in0; w10[1] = in1; w10[2] = in2;
          w10[3] = in2 - in0; w10[4] = in1 - in0;
          w10[5] = in0 * kabs; w10[6] = in1 * kabs; w10[7] = in2 * kabs;
          w10[8] = (in2 - in0) * kabs; w10[9] = (in1 - in0) * kabs;
          const uint32_t rkey = (uint32_t)is.receiver * 2u + (side == SSOL_FRONT_ ? 0u : 1u);
          const bool direct = out0 == 0;  // the path ends here
          if (direct) {
            dkey = rkey;
            din[0] = in0; din[1] = in1; din[2] = in2; din[3] = kabs;
          } else if (!pend_add32(P.pkey + slot, P.pval + slot, stride, prm.npend, rkey, w10)) {
            spill32(T.rcv + (size_t)(blockIdx.x & T.mask_r) * T.nrcv * 40u, rkey, w10, &C->spills);
            spill32(T.sxr + (size_t)(blockIdx.x & T.mask_x) * T.nsamp * T.nrcv * 40u, (uint64_t)cur.w * S.nrcv * 2u + rkey, w10, &C->spills);
          }
          if (is.prim_tally >= 0) {
            const uint32_t pkey = (uint32_t)(((uint64_t)is.prim_tally + prim) * 2u +
                                             (side == SSOL_FRONT_ ? 0u : 1u));
            if (direct)
              dpkey = pkey;
            else if (!pend_add32(P.qkey + slot, P.qval + slot, stride, prm.npend, pkey, w10))
              spill32(T.prim + (size_t)(blockIdx.x & T.mask_p) * T.nprim * 40u, pkey, w10, &C->spills);
          }
        } else {
          pc.z += in0 * kabs;
        }
        if (out0 == 0) {
          const float absorbed = pv.x - in0;
          if (in_atm) pc.y += absorbed; else pc.z += absorbed;
          fin.kind = FIN_DONE;
          fin.g[3] = 0;
        } else {
          e.has = true;
          if (virt) {
            // pass through: the same ray beyond this hit (from the start
            // point itself when the path starts on a virtual surface)
            if (flags & PF_STAGE_BOUNCE) {
              e.r = r;
            } else {
              e.r.ox = pos.x; e.r.oy = pos.y; e.r.oz = pos.z;
              e.r.base = base;
              e.r.dx = dir.x; e.r.dy = dir.y; e.r.dz = dir.z;
              e.r.wl = wl;
            }
            e.r.tmin = nextafterf(hit_t, FLT_MAX);
          } else {
            e.r.ox = pos.x; e.r.oy = pos.y; e.r.oz = pos.z;
            e.r.base = base;
            e.r.dx = dir.x; e.r.dy = dir.y; e.r.dz = dir.z;
            e.r.tmin = 0;
            e.r.wl = wl;
          }
          e.r.iss_from = iss;
          e.r.prim_from = prim;
          e.r.flags = side << RF_SIDE_SHIFT;
          flags = PF_ACTIVE | PF_STAGE_BOUNCE | (virt ? PF_HIT_VIRTUAL : 0u) | (in_atm ? PF_IN_ATM : 0u);
          oi = make_float4(out0, out1, out2, in0);
          out_med = next_med;
        }
      }
    }

    if (fin.kind == FIN_DONE) {
      fin.g[0] = pc.w;
      fin.g[1] = pc.x;
      fin.g[4] = pc.y;
      fin.g[5] = pc.z;
      fin.factor = (double)(1u << score);
      for (int i = 0; i < 6; ++i) fin.g[i] *= fin.factor;
    }
    if (fin.kind == FIN_NONE) {
      P.prev_init[slot] = pv;
      P.out_in0[slot] = oi;
      P.pend_cos[slot] = pc;
      P.meta[slot] = make_uint4(meta.x, meta.y, rng.block, flags);
      P.cur[slot] = make_uint4(iss, prim, side, cur.w);
      P.misc[slot] = make_uint4(depth, in_med | (out_med << 16), score, rinterval | (fin.job << 8));
    }
  }

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
  // pending entries end at a path's first empty one; the direct entry joins
  // the pending entry of its key (summed in FP32 as pend_add32 would) or
  // takes the first empty position (as shade_block)
  bool more = live && fin.kind != FIN_NONE;
  bool dpend = done && dkey != NONE;
  float dw[10];
  dw[0] = din[0]; dw[1] = din[1]; dw[2] = din[2];
  dw[3] = din[2] - din[0]; dw[4] = din[1] - din[0];
  dw[5] = din[0] * din[3]; dw[6] = din[1] * din[3]; dw[7] = din[2] * din[3];
  dw[8] = (din[2] - din[0]) * din[3]; dw[9] = (din[1] - din[0]) * din[3];
  for (uint32_t k = 0; k <= prm.npend; ++k) {
    uint32_t key = NONE;
    float a10[10];
    if (more && k < prm.npend) {
      key = P.pkey[k * stride + slot];
      if (key != NONE) {
        const bool merge = dpend && key == dkey;
        for (int f = 0; f < 10; ++f) {
          const float x = P.pval[((size_t)k * 10 + f) * stride + slot];
          a10[f] = merge ? x + dw[f] : x;
        }
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
      const double x = has ? (double)a10[f] * fin.factor : 0.0;
      v[2 * f] = x;
      v[2 * f + 1] = x * x;
    }
    keyed_add<20>(has, has ? key : 0, v, rcv_base);
    const uint64_t xkey = has ? (uint64_t)fin.samp_slot * (uint64_t)S.nrcv * 2u + key : 0;
    keyed_add<20>(has, xkey, v, T.sxr + (size_t)(blockIdx.x & T.mask_x) * T.nsamp * T.nrcv * 40u);
  }
  more = live && fin.kind != FIN_NONE;
  dpend = done && dpkey != NONE;
  for (uint32_t k = 0; k <= prm.npend; ++k) {
    uint32_t key = NONE;
    float a10[10];
    if (more && k < prm.npend) {
      key = P.qkey[k * stride + slot];
      if (key != NONE) {
        const bool merge = dpend && key == dpkey;
        for (int f = 0; f < 10; ++f) {
          const float x = P.qval[((size_t)k * 10 + f) * stride + slot];
          a10[f] = merge ? x + dw[f] : x;
        }
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

  const bool want = live && fin.kind != FIN_NONE && !C->aborted;
  const uint64_t id = want ? (((uint64_t)P.meta[slot].y << 32) | P.meta[slot].x) + prm.slots : 0;
  int samp_slot = 0;
  bool gen = false;
  if (REFILL && want && id < prm.N) {
    generate32(S, SQ, P, prm, slot, id, e, samp_slot);
    gen = true;
  } else if (live && fin.kind != FIN_NONE) {
    P.meta[slot] = make_uint4(0, 0, 0, 0);
  }
  flush_sampled_count(gen, samp_slot, T);

  if (e.has) rays[slot] = e.r;
  if (live) emitted[slot] = e.has ? 1 : 0;
  block_count(e.has, block_emits, b);
}

// The arguments of k_shade32 in one struct (as ShadeArgs)
struct Shade32Args {
  DScene S;
  const ShapeQ* SQ;
  PathState32 P;
  SolveParams prm;
  Tallies T;
  const uint32_t* list;
  const unsigned* count;
  RayRec32* rays;
  const HitRec32* hits;
  uint8_t* emitted;
  uint32_t* block_emits;
  Counters* C;
  unsigned* fetch;  // tail-block counter of a looping launch (for_each_block)
};

// Four waves per SIMD (at most 128 VGPRs, 28 bytes per lane spilled): on
// the 10k-heliostat field 25.7-26.3 ms per step where three waves took
// 26.3-29.6 ms.
template <bool REFILL, bool LOOP>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4)))
k_shade32(Shade32Args a) {
  const uint32_t count = *a.count;
  if (LOOP) {  // whole logical blocks, arguments read per block (as k_shade)
    for_each_block((count + BLOCK - 1) / BLOCK, a.fetch, [&](uint32_t b) {
      const Shade32Args& f = kernel_args<Shade32Args>();
      shade32_block<REFILL>(b, f.S, f.SQ, f.P, f.prm, f.T, f.list, count, f.rays, f.hits,
                            f.emitted, f.block_emits, f.C);
    });
  } else if (blockIdx.x * BLOCK < count) {
    shade32_block<REFILL>(blockIdx.x, a.S, a.SQ, a.P, a.prm, a.T, a.list, count, a.rays, a.hits,
                          a.emitted, a.block_emits, a.C);
  }
}
