// Renderer on the solver's kernels (reference ssol_draw_pt
// src/ssol_draw_pt.c:126-285, ssol_draw_draft src/ssol_draw_draft.c): the
// camera, bounce and sun rays of every pixel sample are traced by the
// solver's own trace kernel (`k_trace`, closest hit, or first hit for sun
// rays, virtual materials discarded), and one shading kernel per bounce
// follows each sample's radiance estimate. Samples are independent
// Threefry streams keyed by (seed, pixel * spp + sample), so an image does
// not depend on scheduling. Included by ssolh.hip.
//
// Radiance conventions of the reference path tracer: a ray that leaves the
// scene towards the sky (direction . up > 0) gathers 0.1; at every hit the
// sun is sampled with a shadow ray (BSDF value x cosine, unit sun); the
// bounce multiplies the throughput by |wi . N| times the sampling weight
// (the reference's double cosine, kept), and once the throughput falls
// below 0.1 Russian roulette on the sampling weight takes over. The draft
// mode returns the cosine between the first hit's normal and the view ray.

enum : uint32_t {
  RS_ALIVE = 1u << 0,    // the sample's bounce ray is in flight
  RS_SUN = 1u << 1,      // its sun ray is in flight, pend holds what it adds
  RS_ROULETTE = 1u << 2, // Russian roulette has taken over
  RS_DONE = 1u << 3      // the sample's radiance is in its pixel
};

struct RenderState {
  double throughput, L, pend, wl;
  uint32_t pixel, medium, flags, block;
};

struct RenderParams {
  uint32_t width, height, spp, mode;   // mode 0: draft, 1: path tracing
  float pos[3], ax[3], ay[3], az[3];   // camera frame (camera_ray)
  double up[3];
  uint64_t key;
  uint64_t first;                      // first sample of the batch
  uint32_t count;                      // samples in the batch
  const float* draft;                  // [spp][2] pixel offsets of the draft mode
};

// BSDF value for the sun sample of a hit (rendering variant of the
// materials, reference material_create_bsdf with `rendering` set,
// src/ssol_material.c:209-211): Lambertian rho/pi; a glossy mirror's
// microfacet reflection F D G / (4 |wi.n| |wo.n|) with the constant
// Fresnel term rho, the sampler's Beckmann or pillbox D and Smith's G for
// Beckmann (Walter et al. 2007, rational form; 1 for pillbox); specular
// and dielectric interfaces are Dirac: 0.
__device__ double smith_g1_beckmann(double m, double cos_v) {
  const double c = fmin(fabs(cos_v), 1.0);
  const double tan_v = sqrt(fmax(1 - c * c, 0.0)) / fmax(c, 1e-300);
  if (tan_v == 0) return 1;
  const double a = 1 / (m * tan_v);
  return a >= 1.6 ? 1 : (3.535 * a + 2.181 * a * a) / (1 + 2.276 * a + 2.577 * a * a);
}

__device__ double microfacet_d(int distrib, double m, double cos_h) {
  if (cos_h <= 0) return 0;
  const double c2 = cos_h * cos_h, t2 = (1 - c2) / c2;
  if (distrib == 0) return exp(-t2 / (m * m)) / (PI * m * m * c2 * c2);
  return t2 <= m * m ? 1 / (PI * m * m * c2 * c2) : 0;
}

__device__ double render_bsdf_eval(const ssolh_material& m, d3 wo, d3 n, d3 wi) {
  const double co = dot(wo, n), ci = dot(wi, n);
  if (co <= 0 || ci <= 0) return 0;
  if (m.type == SSOLH_MTL_MATTE) return m.reflectivity / PI;
  if (m.type == SSOLH_MTL_MIRROR && m.roughness > 0) {
    const d3 h = normalize(wo + wi);
    const double D = microfacet_d(m.distribution, m.roughness, dot(h, n));
    const double G = m.distribution == 0 ? smith_g1_beckmann(m.roughness, co) *
                                               smith_g1_beckmann(m.roughness, ci)
                                         : 1.0;
    return m.reflectivity * D * G / (4 * co * ci);
  }
  return 0;
}

// Bounce sample of the rendering variant: the glossy mirror reflects about
// one sampled microfacet normal (no resampling below the surface, weight
// F G |wi.h| / (|wi.n| |h.n|)); every other material samples as in the
// solver
template <class Rng>
__device__ double render_bsdf_sample(const ssolh_material& m, double eta_out, double eta_in,
                                     double slab_eta, double slab_abs, d3 wi, d3 n, Rng& rng,
                                     d3& wo, int& lobe) {
  if (m.type == SSOLH_MTL_MIRROR && m.roughness > 0) {
    lobe = LOBE_REFLECT;
    const double u1 = rng.next(), u2 = rng.next();
    const d3 h = sample_microfacet(m.distribution, m.roughness, u1, u2, n);
    const double c = dot(wi, h);
    if (c <= 0) { wo = n; return 0; }
    wo = reflect(wi, h, c);
    const double ci = dot(wi, n), co = dot(wo, n), ch = dot(h, n);
    if (co <= 0 || ci <= 0 || ch <= 0) return 0;
    const double G = m.distribution == 0 ? smith_g1_beckmann(m.roughness, ci) *
                                               smith_g1_beckmann(m.roughness, co)
                                         : 1.0;
    return fmin(m.reflectivity * G * c / (ci * ch), 1.0);
  }
  return bsdf_sample(m, eta_out, eta_in, slab_eta, slab_abs, wi, n, rng, wo, lobe);
}

// Appends `slot` to the work list (one atomic per wave)
__device__ __forceinline__ void render_push(bool has, uint32_t slot, uint32_t* list,
                                            unsigned* count) {
  const uint64_t mask = __ballot(has);
  if (!mask) return;
  const int leader = __ffsll((unsigned long long)mask) - 1;
  unsigned base = 0;
  if ((int)lane_id() == leader) base = atomicAdd(count, (unsigned)__popcll(mask));
  base = __shfl(base, leader, 64);
  if (has) list[base + __popcll(mask & ((1ull << lane_id()) - 1ull))] = slot;
}

// Camera rays of the batch's samples (slot 2s; slot 2s + 1 is the sun ray)
__global__ void __launch_bounds__(BLOCK)
k_render_start(DScene S, RenderParams rp, RenderState* __restrict__ st, RayRec* __restrict__ rays,
               uint32_t* list, unsigned* count) {
  const uint32_t s = blockIdx.x * BLOCK + threadIdx.x;
  const bool live = s < rp.count;
  if (live) {
    const uint64_t g = rp.first + s;
    const uint32_t pixel = (uint32_t)(g / rp.spp), k = (uint32_t)(g % rp.spp);
    const uint32_t px = pixel % rp.width, py = pixel / rp.width;
    PathRng rng;
    rng.init(rp.key, g, 0);
    double jx, jy;
    if (rp.mode == 0) {
      jx = rp.draft[2 * k];
      jy = rp.draft[2 * k + 1];
    } else {
      jx = rng.next();
      jy = rng.next();
    }
    // camera_ray (reference src/ssol_camera.h): FP32 frame and sample
    const float sx = ((float)px + (float)jx) / (float)rp.width;
    const float sy = ((float)py + (float)jy) / (float)rp.height;
    float d[3];
    for (int a = 0; a < 3; ++a)
      d[a] = rp.ax[a] * (sx * 2.f - 1.f) + rp.ay[a] * (sy * 2.f - 1.f) + rp.az[a];
    const float len = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    RayRec r;
    r.ox = rp.pos[0]; r.oy = rp.pos[1]; r.oz = rp.pos[2];
    r.dx = d[0] / len; r.dy = d[1] / len; r.dz = d[2] / len;
    r.tmin = 1e-30f;  // no emitting point: t > 0 only
    r.iss_from = NONE;
    r.prim_from = NONE;
    r.flags = RF_DISCARD_VIRTUAL;
    rays[2 * (size_t)s] = r;
    RenderState x;
    x.throughput = 1;
    x.L = 0;
    x.pend = 0;
    x.wl = rp.mode == 0 ? 0.0 : sample_wavelength(S.spectra, S.sun.wl_spectrum, rng.next());
    x.pixel = pixel;
    x.medium = S.air;
    x.flags = RS_ALIVE;
    x.block = rng.block;
    st[s] = x;
  }
  render_push(live, 2 * s, list, count);
}

// One bounce of every live sample: the sun ray's result, then the hit of
// the bounce ray (reference Li, src/ssol_draw_pt.c:126-262)
__global__ void __launch_bounds__(BLOCK)
k_render_shade(DScene S, RenderParams rp, RenderState* __restrict__ st,
               RayRec* __restrict__ rays, const HitRec* __restrict__ hits, double* pixels,
               uint32_t* list, unsigned* count) {
  const uint32_t s = blockIdx.x * BLOCK + threadIdx.x;
  bool emit_bounce = false, emit_sun = false;
  if (s < rp.count && !(st[s].flags & RS_DONE)) {
    RenderState x = st[s];
    if (x.flags & RS_SUN) {
      if (hits[2 * (size_t)s + 1].iss == NONE) x.L += x.pend;
      x.flags &= ~RS_SUN;
    }
    if (x.flags & RS_ALIVE) {
      const RayRec r = rays[2 * (size_t)s];
      const HitRec h = hits[2 * (size_t)s];
      const d3 dir = mk(r.dx, r.dy, r.dz);
      bool done = false;
      if (h.iss == NONE) {
        if (rp.mode == 1 && dot(dir, ld3(rp.up)) > 0) x.L += x.throughput * 1e-1;
        done = true;
      } else {
        const IssInfo is = S.iss[h.iss];
        d3 N = normalize(mk(h.nx, h.ny, h.nz));
        uint32_t side = SSOL_FRONT_;
        if (dot(N, dir) >= 0) { side = SSOL_BACK_; N = -N; }
        if (rp.mode == 0) {
          x.L += fmax(dot(N, -dir), 0.0);
          done = true;
        } else {
          const double t = is.shape_type == SSOLH_SHAPE_MESH ? h.t : h.qdst;
          const double k = data_value(S.spectra, S.media[x.medium].extinction, x.wl);
          if (k > 0) x.throughput *= exp(-k * t);
          if (x.throughput <= 0) {
            done = true;
          } else {
            const ssolh_material& m = S.mtls[side == SSOL_FRONT_ ? is.mtl_front : is.mtl_back];
            const d3 pos = mk(r.ox, r.oy, r.oz) + dir * t;
            const d3 wo = -dir;
            // sun: a shadow ray towards it
            const d3 ws = -ld3(S.sun.dir);
            const double cs = dot(ws, N);
            if (cs > 1e-6) {
              const double R = render_bsdf_eval(m, wo, N, ws);
              if (R > 0) {
                x.pend = x.throughput * R * cs;
                x.flags |= RS_SUN;
                RayRec sr;
                sr.ox = pos.x; sr.oy = pos.y; sr.oz = pos.z;
                sr.dx = ws.x; sr.dy = ws.y; sr.dz = ws.z;
                sr.tmin = 0;
                sr.iss_from = h.iss;
                sr.prim_from = h.prim;
                sr.flags = (side << RF_SIDE_SHIFT) | RF_DISCARD_VIRTUAL | RF_ANY_HIT;
                rays[2 * (size_t)s + 1] = sr;
                emit_sun = true;
              }
            }
            // bounce
            double eta_out = 1, eta_in = 1, slab_eta = 1, slab_abs = 0;
            if (m.type == SSOLH_MTL_DIELECTRIC) {
              eta_out = data_value(S.spectra, S.media[m.medium_out].refractive_index, x.wl);
              eta_in = data_value(S.spectra, S.media[m.medium_in].refractive_index, x.wl);
            } else if (m.type == SSOLH_MTL_THIN_DIELECTRIC) {
              eta_out = data_value(S.spectra, S.media[m.medium_out].refractive_index, x.wl);
              slab_eta = data_value(S.spectra, m.slab_eta, x.wl);
              slab_abs = data_value(S.spectra, m.slab_absorption, x.wl);
            }
            PathRng rng;
            rng.init(rp.key, rp.first + s, x.block);
            d3 wi;
            int lobe;
            double R = render_bsdf_sample(m, eta_out, eta_in, slab_eta, slab_abs, wo, N, rng, wi,
                                          lobe);
            const double cg = dot(N, wi);
            if ((cg > 0 && lobe == LOBE_TRANSMIT) || (cg < 0 && lobe == LOBE_REFLECT)) R = 0;
            if (lobe == LOBE_TRANSMIT && m.type == SSOLH_MTL_DIELECTRIC)
              x.medium = (uint32_t)(m.medium_out == (int32_t)x.medium ? m.medium_in : m.medium_out);
            if (!(x.flags & RS_ROULETTE)) {
              x.throughput *= fabs(cg) * R;
            } else if (rng.next() >= R) {
              done = true;
            } else {
              x.throughput *= fabs(cg);
            }
            x.block = rng.block;
            if (!done && x.throughput <= 0) done = true;
            if (!done) {
              if (x.throughput < 0.1) x.flags |= RS_ROULETTE;
              RayRec br;
              br.ox = pos.x; br.oy = pos.y; br.oz = pos.z;
              br.dx = wi.x; br.dy = wi.y; br.dz = wi.z;
              br.tmin = 0;
              br.iss_from = h.iss;
              br.prim_from = h.prim;
              br.flags = (side << RF_SIDE_SHIFT) | RF_DISCARD_VIRTUAL;
              rays[2 * (size_t)s] = br;
              emit_bounce = true;
            }
          }
        }
      }
      if (done) x.flags &= ~RS_ALIVE;
    }
    if (!(x.flags & (RS_ALIVE | RS_SUN))) {
      if (x.L != 0) atomicAdd(pixels + x.pixel, x.L / (double)rp.spp);
      x.flags |= RS_DONE;
    }
    st[s] = x;
  }
  render_push(emit_bounce, 2 * s, list, count);
  render_push(emit_sun, 2 * s + 1, list, count);
}
