/* ORACLE -- test infrastructure only (see oracle.h).
 * The realisation loop (src/ssol_solver.c:207-842), deferred MC tallies
 * (src/ssol_estimator_c.h:33-126) and estimator getters
 * (src/ssol_estimator.c:112-268, src/ssol_mc_receiver.c:32-173). */
#include "oracle_internal.h"

#include <float.h>
#include <omp.h>
#include <stdio.h>

enum { G_COS, G_ABS_RCV, G_SHADOWED, G_MISSING, G_EXTINCT, G_OTHER };
enum { R_IN, R_IN_NO_ATM, R_IN_NO_FIELD, R_IN_LOST_FIELD, R_IN_LOST_ATM,
       R_ABS, R_ABS_NO_ATM, R_ABS_NO_FIELD, R_ABS_LOST_FIELD, R_ABS_LOST_ATM };

/*******************************************************************************
 * Per-thread tallies
 ******************************************************************************/
struct otctx {
  struct omc global[6];
  struct omc* rcv;
  struct omc* samp;
  uint64_t* nb_samples;
  struct omc* sxr;
  struct omc* prim;
  struct omc** touched;
  size_t ntouched, cap;
  uint64_t realisations, segments;
  /* path recording: type of the last path, index of its first vertex */
  int path_type;
  size_t path_first;
};

/* Path types (enum ssol_path_type, src/ssol.h:70-75); -1: not registered */
enum { P_MISSING, P_SHADOW, P_SUCCESS, P_ERROR, P_DROPPED = -1 };

/* path_add_vertex (src/ssol_solver.c:588-600) */
static void track_vertex(struct ooracle* o, const double pos[3], double w)
{
  if(o->tv_n == o->tv_cap) {
    o->tv_cap = o->tv_cap ? 2 * o->tv_cap : 1024;
    o->tv = realloc(o->tv, o->tv_cap * 4 * sizeof(double));
  }
  o->tv[4*o->tv_n + 0] = pos[0];
  o->tv[4*o->tv_n + 1] = pos[1];
  o->tv[4*o->tv_n + 2] = pos[2];
  o->tv[4*o->tv_n + 3] = w;
  o->tv_n++;
}

/* path_register_and_clear (src/ssol_solver.c:572-588); dropped paths
 * (roulette) give their vertices back */
static void track_finish(struct ooracle* o, struct otctx* c, int r)
{
  const int type = r ? P_ERROR : c->path_type;
  if(type == P_DROPPED) {
    o->tv_n = c->path_first;
    return;
  }
  if(o->tp_n == o->tp_cap) {
    o->tp_cap = o->tp_cap ? 2 * o->tp_cap : 1024;
    o->tp_first = realloc(o->tp_first, o->tp_cap * sizeof(size_t));
    o->tp_count = realloc(o->tp_count, o->tp_cap * sizeof(size_t));
    o->tp_type = realloc(o->tp_type, o->tp_cap * sizeof(int));
  }
  o->tp_first[o->tp_n] = c->path_first;
  o->tp_count[o->tp_n] = o->tv_n - c->path_first;
  o->tp_type[o->tp_n] = type;
  o->tp_n++;
}

static void omc_flush(struct omc* d)
{
  d->w += d->tmp;
  d->w2 += d->tmp * d->tmp;
  d->tmp = 0;
}

static void omc_add(struct otctx* c, struct omc* d, uint64_t ir, double w)
{
  if(d->irealisation != ir) {
    omc_flush(d);
    d->irealisation = ir;
    if(c->ntouched == c->cap) {
      c->cap = c->cap ? c->cap * 2 : 64;
      c->touched = realloc(c->touched, c->cap * sizeof(*c->touched));
    }
    c->touched[c->ntouched++] = d;
  }
  d->tmp += w;
}

static void omc_apply_factor(struct otctx* c, uint64_t ir, double f)
{
  size_t i;
  for(i = 0; i < c->ntouched; ++i)
    if(c->touched[i]->irealisation == ir) c->touched[i]->tmp *= f;
}

static void omc_init(struct omc* d, size_t n)
{
  size_t i;
  for(i = 0; i < n; ++i) { d[i].irealisation = UINT64_MAX; d[i].tmp = d[i].w = d[i].w2 = 0; }
}

static void omc_merge(struct omc* dst, struct omc* src, size_t n)
{
  size_t i;
  for(i = 0; i < n; ++i) {
    omc_flush(dst + i);
    omc_flush(src + i);
    dst[i].w += src[i].w;
    dst[i].w2 += src[i].w2;
  }
}

/*******************************************************************************
 * Media (src/ssol_material.c:36-52, 730-736)
 ******************************************************************************/
static int odata_ceq(const struct oscene* s, const struct odata* a,
  const struct odata* b)
{
  const struct ospectrum *sa, *sb;
  if((a->spectrum < 0) != (b->spectrum < 0)) return 0;
  if(a->spectrum < 0) return a->real == b->real;
  if(a->spectrum == b->spectrum) return 1;
  sa = s->spectra + a->spectrum;
  sb = s->spectra + b->spectrum;
  return sa->n == sb->n
    && !memcmp(sa->wavelengths, sb->wavelengths, sa->n * sizeof(double))
    && !memcmp(sa->values, sb->values, sa->n * sizeof(double));
}

static int omedia_ceq(const struct oscene* s, const struct omedium* a,
  const struct omedium* b)
{
  return odata_ceq(s, &a->refractive_index, &b->refractive_index)
    && odata_ceq(s, &a->extinction, &b->extinction);
}

/*******************************************************************************
 * Random walk (src/ssol_solver.c:873-1127)
 ******************************************************************************/
struct opoint {
  unsigned inst, sshape, prim;
  int side;
  double N[3], pos[3], dir[3];
  double wl;
  const struct omaterial* mtl;
  double kabs;
  int survivor_score;
  double initial_flux, cos_factor;
  double prev_out[3];  /* flux, if_no_atm_loss, if_no_field_loss */
  double in[3];
  double out[3];
  int samp_slot;
};

static double rng_next(void* ctx) { return orng_canonical(ctx); }

static const struct omaterial*
point_material(const struct ooracle* o, const struct opoint* p)
{
  const struct osshape* ss = oget_sshape(o, p->inst, p->sshape);
  return o->scn.materials + (p->side == OSIDE_FRONT ? ss->mtl_front : ss->mtl_back);
}

static int point_is_receiver(const struct ooracle* o, const struct opoint* p)
{
  return (o->scn.instances[p->inst].receiver_mask & p->side) != 0;
}

static void
update_mc(struct ooracle* o, struct otctx* c, struct opoint* p, uint64_t ir)
{
  const int side = p->side == OSIDE_FRONT ? 0 : 1;
  const int rslot = o->rcv_slot[p->inst];
  const double k = p->kabs;
  double w[10];
  int f;
  w[R_IN] = p->in[0];
  w[R_IN_NO_ATM] = p->in[1];
  w[R_IN_NO_FIELD] = p->in[2];
  w[R_IN_LOST_FIELD] = p->in[2] - p->in[0];
  w[R_IN_LOST_ATM] = p->in[1] - p->in[0];
  w[R_ABS] = p->in[0] * k;
  w[R_ABS_NO_ATM] = p->in[1] * k;
  w[R_ABS_NO_FIELD] = p->in[2] * k;
  w[R_ABS_LOST_FIELD] = (p->in[2] - p->in[0]) * k;
  w[R_ABS_LOST_ATM] = (p->in[1] - p->in[0]) * k;
  omc_add(c, &c->global[G_ABS_RCV], ir, p->in[0] - p->out[0]);
  for(f = 0; f < 10; ++f)
    omc_add(c, c->rcv + ((size_t)rslot * 2 + side) * 10 + f, ir, w[f]);
  if(c->sxr) {
    const size_t base = (((size_t)p->samp_slot * o->nrcv + rslot) * 2 + side) * 10;
    for(f = 0; f < 10; ++f) omc_add(c, c->sxr + base + f, ir, w[f]);
  }
  if(o->scn.instances[p->inst].per_primitive) {
    const size_t b = o->prim_base[p->inst * o->max_sshapes + p->sshape];
    const size_t base = ((b + p->prim) * 2 + side) * 10;
    for(f = 0; f < 10; ++f) omc_add(c, c->prim + base + f, ir, w[f]);
  }
}

/* src/ssol_solver.c:419-494. Returns 0, or 5 (BAD_OP) on inconsistent media */
static int
point_shade(struct ooracle* o, struct opoint* p, const struct omedium* in_med,
  const struct omedium** out_med, struct orng* rng)
{
  const struct omaterial* m = p->mtl;
  const struct oscene* s = &o->scn;
  double wi[3], wo[3], propagated = 0;
  double eta_out = 1, eta_in = 1, slab_eta = 1, slab_abs = 0;
  int type = 0, i;

  if(m->type == OM_DIELECTRIC) {
    if(!omedia_ceq(s, in_med, &m->medium_out)) return 5;
    eta_out = odata_value(&m->medium_out.refractive_index, s, p->wl);
    eta_in = odata_value(&m->medium_in.refractive_index, s, p->wl);
  } else if(m->type == OM_THIN) {
    eta_out = odata_value(&m->medium_out.refractive_index, s, p->wl);
    slab_eta = odata_value(&m->slab.refractive_index, s, p->wl);
    slab_abs = odata_value(&m->slab.extinction, s, p->wl);
  }
  for(i = 0; i < 3; ++i) wi[i] = -p->dir[i];
  /* Baked shaders: the shading normal is the geometry normal */
  if(ov_dot(wi, p->N) > 0) {
    double c;
    propagated = obsdf_sample(m, eta_out, eta_in, slab_eta, slab_abs, wi,
      p->N, rng_next, rng, wo, &type);
    c = ov_dot(p->N, wo);
    if((c > 0 && type == 2) || (c < 0 && type == 1)) propagated = 0;
    ov_set(p->dir, wo);
  }
  p->kabs = 1 - propagated;
  p->out[0] = p->in[0] * propagated;
  p->out[1] = p->in[1] * propagated;
  p->out[2] = point_is_receiver(o, p) ? p->in[2] * propagated : p->in[2];
  if(type == 2 && m->type == OM_DIELECTRIC) {
    *out_med = omedia_ceq(s, &m->medium_out, in_med) ? &m->medium_in : &m->medium_out;
  } else {
    *out_med = in_med;
  }
  return 0;
}

static size_t cdf_search(const double* cdf, size_t n, double x)
{
  size_t lo = 0, hi = n;
  while(lo < hi) {
    const size_t mid = (lo + hi) / 2;
    if(cdf[mid] <= x) lo = mid + 1; else hi = mid;
  }
  return lo < n ? lo : n - 1;
}

static int
trace_path(struct ooracle* o, struct otctx* c, uint64_t ir, uint64_t key,
  uint64_t ctr0)
{
  const struct oscene* s = &o->scn;
  struct orng rng;
  struct opoint p;
  struct oray ray;
  const struct omedium* in_med;
  const struct omedium* out_med;
  const struct oshape_geo* g;
  double Ns[3], N[3], u[3];
  double surface_sun_cos, surface_sun0_cos, sun0_sun_cos, proxy_cos, w0;
  double org[3], rdir[3], tmin = 0, hit_t = 0;
  size_t idx;
  int i, depth = 0, roulette_interval = 64, killed = 0, is_lit;
  const int typical_max_depth = 16;

  memset(&p, 0, sizeof(p));
  c->ntouched = 0;
  c->path_first = o->tv_n;
  c->path_type = P_MISSING;
  orng_init(&rng, key, ctr0 + ir);

  /* Area-weighted start point on the sampling view (src/ssol_solver.c:237-252) */
  u[0] = orng_canonical(&rng);
  u[1] = orng_canonical(&rng);
  u[2] = orng_canonical(&rng);
  idx = cdf_search(o->samp_cdf, o->nsamp_tris, u[0] * o->samp_cdf[o->nsamp_tris-1]);
  p.inst = o->samp_inst[idx];
  p.sshape = o->samp_sshape[idx];
  p.prim = o->samp_prim[idx];
  g = o->shapes + oget_sshape(o, p.inst, p.sshape)->shape;
  {
    const double* a = g->samp_verts + 3 * g->rt_ids[3*p.prim];
    const double* b = g->samp_verts + 3 * g->rt_ids[3*p.prim+1];
    const double* cc = g->samp_verts + 3 * g->rt_ids[3*p.prim+2];
    const double sq = sqrt(u[1]);
    const double b0 = 1 - sq, b1 = sq * (1 - u[2]), b2 = sq * u[2];
    double loc[3], e1[3], e2[3], n[3];
    for(i = 0; i < 3; ++i) {
      loc[i] = b0 * a[i] + b1 * b[i] + b2 * cc[i];
      e1[i] = b[i] - a[i];
      e2[i] = cc[i] - a[i];
    }
    ov_mat_mul(p.pos, o->inst_R + 9*p.inst, loc);
    for(i = 0; i < 3; ++i) p.pos[i] += o->inst_T[3*p.inst + i];
    ov_cross(n, e2, e1);
    ov_mat_mul(Ns, o->inst_Rinvt + 9*p.inst, n);
    ov_normalize(Ns);
  }

  /* Sun direction and wavelength (src/ssol_solver.c:260-264) */
  osun_sample(&o->sun, rng_next, &rng, p.dir);
  {
    const struct ospectrum* sp = s->spectra + s->sun.spectrum;
    p.wl = sp->n <= 1 ? sp->wavelengths[0]
      : oracle_piecewise_linear_sample(sp->wavelengths, sp->values, sp->n,
          orng_canonical(&rng));
  }

  if(g->type == 0) {
    ov_set(N, Ns);
  } else {
    opunched_project(o, p.inst, p.sshape, p.pos, p.pos, N);
  }
  if(ov_dot(N, p.dir) < 0) {
    p.side = OSIDE_FRONT;
  } else {
    p.side = OSIDE_BACK;
    for(i = 0; i < 3; ++i) N[i] = -N[i];
  }

  /* Monte Carlo weight (src/ssol_solver.c:283-299) */
  surface_sun_cos = ov_dot(N, p.dir);
  surface_sun0_cos = fabs(ov_dot(o->sun.dir, N));
  sun0_sun_cos = ov_dot(o->sun.dir, p.dir);
  proxy_cos = g->type == 0 ? 1 : fabs(ov_dot(Ns, N));
  w0 = s->sun.dni * o->sampled_area_proxy
     * fabs(surface_sun_cos / (proxy_cos * sun0_sun_cos));
  p.cos_factor = o->sampled_area_proxy / o->sampled_area
     * surface_sun0_cos / proxy_cos;
  p.initial_flux = w0;
  p.prev_out[0] = p.prev_out[1] = p.prev_out[2] = w0;
  ov_set(p.N, N);
  p.samp_slot = o->samp_slot[p.inst];
  c->nb_samples[p.samp_slot]++;

  /* Starting medium (src/ssol_solver.c:307-323) */
  p.mtl = point_material(o, &p);
  if(o->track) {
    /* sun point and start point (src/ssol_solver.c:909-921) */
    if(o->track_sun_len > 0) {
      double sp[3];
      for(i = 0; i < 3; ++i) sp[i] = p.pos[i] - p.dir[i] * o->track_sun_len;
      track_vertex(o, sp, s->sun.dni);
    }
    track_vertex(o, p.pos, p.initial_flux);
  }
  if(p.mtl->type == OM_DIELECTRIC || p.mtl->type == OM_THIN) {
    in_med = p.side == OSIDE_FRONT ? &p.mtl->medium_in : &p.mtl->medium_out;
  } else {
    in_med = &o->air;
  }
  out_med = in_med;

  /* Shadow ray toward the sun (src/ssol_solver.c:325-352) */
  ov_set(ray.org, p.pos);
  for(i = 0; i < 3; ++i) ray.dir[i] = -p.dir[i];
  ray.tmin = 0;
  ray.inst_from = p.inst; ray.sshape_from = p.sshape; ray.prim_from = p.prim;
  ray.side_from = p.side;
  ray.discard_virtual = 1;
  ray.reversed = 1;
  ray.any_hit = 1;
  otrace(o, &ray);
  c->segments++;
  is_lit = !ray.hit;

  if(!is_lit) {
    omc_add(c, c->samp + 2*p.samp_slot + 1, ir, p.initial_flux);
    omc_add(c, &c->global[G_SHADOWED], ir, p.initial_flux);
    c->path_type = P_SHADOW;
  } else {
    int hit_a_receiver = 0;
    ov_set(org, p.pos);
    ov_set(rdir, p.dir);
    hit_t = 0;
    for(;;) {
      const int in_atm = omedia_ceq(s, in_med, &o->air);
      const int hit_receiver = point_is_receiver(o, &p);
      const int hit_virtual = p.mtl->type == OM_VIRTUAL;
      int last_segment = 0, weight_is_zero = 0;
      double trans = 1;

      if(hit_t > 0) {
        const double k_ext = odata_value(&in_med->extinction, s, p.wl);
        if(k_ext > 0) trans = exp(-k_ext * hit_t);
      }
      p.in[0] = p.prev_out[0] * trans;
      p.in[1] = in_atm ? p.prev_out[1] : p.prev_out[1] * trans;
      p.in[2] = !in_atm ? p.prev_out[2] : p.prev_out[2] * trans;

      if(hit_virtual) {
        p.kabs = 0;
        p.out[0] = p.in[0]; p.out[1] = p.in[1]; p.out[2] = p.in[2];
        out_med = in_med;
      } else {
        if(point_shade(o, &p, in_med, &out_med, &rng)) return 5;
      }

      if(hit_receiver) {
        hit_a_receiver = 1;
        update_mc(o, c, &p, ir);
      } else {
        omc_add(c, &c->global[G_OTHER], ir, p.in[0] * p.kabs);
      }

      if(p.out[0] == 0) weight_is_zero = 1;

      if(!weight_is_zero) {
        if(hit_virtual) {
          tmin = (double)nextafterf((float)hit_t, FLT_MAX);
        } else {
          tmin = 0;
          ov_set(org, p.pos);
          ov_set(rdir, p.dir);
        }
        ov_set(ray.org, org);
        ov_set(ray.dir, rdir);
        ray.tmin = tmin;
        ray.inst_from = p.inst; ray.sshape_from = p.sshape; ray.prim_from = p.prim;
        ray.side_from = p.side;
        ray.discard_virtual = 0;
        ray.reversed = 0;
        ray.any_hit = 0;
        otrace(o, &ray);
        c->segments++;
        if(!ray.hit) {
          /* escape point (src/ssol_solver.c:1012-1019) */
          if(o->track && o->track_inf_len > 0) {
            double ep[3];
            for(i = 0; i < 3; ++i) ep[i] = p.pos[i] + rdir[i] * o->track_inf_len;
            track_vertex(o, ep, p.out[0]);
          }
          last_segment = 1;
          if(!omedia_ceq(s, out_med, &o->air)) return 5;
        }
      }

      if(weight_is_zero || last_segment || !hit_virtual) {
        const double absorbed = p.prev_out[0] - p.in[0];
        omc_add(c, &c->global[in_atm ? G_EXTINCT : G_OTHER], ir, absorbed);
        if(weight_is_zero || last_segment) break;
        p.prev_out[0] = p.out[0];
        p.prev_out[1] = p.out[1];
        p.prev_out[2] = p.out[2];
      }

      depth += !hit_virtual;
      if(depth > roulette_interval && depth % roulette_interval == 1) {
        if(orng_canonical(&rng) > 0.5) {
          p.survivor_score += 1;
          roulette_interval = typical_max_depth;
        } else {
          omc_apply_factor(c, ir, 0);
          killed = 1;
          break;
        }
      }

      /* point_update_from_hit (src/ssol_solver.c:360-411) */
      hit_t = ray.t;
      p.inst = ray.inst;
      p.sshape = ray.sshape;
      p.prim = ray.prim;
      g = o->shapes + oget_sshape(o, p.inst, p.sshape)->shape;
      if(g->type == 0) {
        ov_set(p.N, ray.Ng);
        ov_normalize(p.N);
        for(i = 0; i < 3; ++i) p.pos[i] = org[i] + rdir[i] * ray.t;
      } else {
        ov_set(p.N, ray.qN);
        ov_normalize(p.N);
        for(i = 0; i < 3; ++i) p.pos[i] = org[i] + p.dir[i] * ray.qdst;
      }
      if(ov_dot(p.dir, p.N) < 0) {
        p.side = OSIDE_FRONT;
      } else {
        p.side = OSIDE_BACK;
        for(i = 0; i < 3; ++i) p.N[i] = -p.N[i];
      }
      p.mtl = point_material(o, &p);
      if(o->track) track_vertex(o, p.pos, p.out[0]); /* src/ssol_solver.c:1080-1083 */
      in_med = out_med;
    }
    if(killed) {
      c->path_type = P_DROPPED;
      return 0;
    }
    omc_add(c, &c->global[G_MISSING], ir, p.out[0]);
    c->path_type = hit_a_receiver ? P_SUCCESS : P_MISSING;
  }
  omc_add(c, c->samp + 2*p.samp_slot + 0, ir, p.cos_factor);
  omc_add(c, &c->global[G_COS], ir, p.cos_factor);
  if(p.survivor_score) {
    omc_apply_factor(c, ir, (double)(1u << p.survivor_score));
  }
  return 0;
}

/*******************************************************************************
 * API
 ******************************************************************************/
struct ooracle*
oracle_create(const struct oscene* scn, int* err)
{
  struct ooracle* o = calloc(1, sizeof(*o));
  int res;
  if(!o) { if(err) *err = -1; return NULL; }
  o->scn = *scn;
  res = oscene_build(o);
  if(res) {
    if(err) *err = res;
    oscene_release(o);
    free(o);
    return NULL;
  }
  osun_state_setup(&o->sun, scn->sun.type, scn->sun.param, scn->sun.dir);
  o->air.extinction.real = 0;
  o->air.extinction.spectrum = -1;
  if(scn->has_atmosphere) o->air.extinction = scn->atmosphere_extinction;
  o->air.refractive_index.real = 1;
  o->air.refractive_index.spectrum = -1;
  if(err) *err = 0;
  return o;
}

static void free_results(struct ooracle* o)
{
  free(o->rcv); free(o->samp); free(o->nb_samples); free(o->sxr); free(o->prim);
  o->rcv = o->samp = o->sxr = o->prim = NULL;
  o->nb_samples = NULL;
}

static void free_tracked(struct ooracle* o)
{
  free(o->tv); free(o->tp_first); free(o->tp_count); free(o->tp_type);
  o->tv = NULL; o->tp_first = o->tp_count = NULL; o->tp_type = NULL;
  o->tv_n = o->tv_cap = o->tp_n = o->tp_cap = 0;
}

void
oracle_destroy(struct ooracle* o)
{
  if(!o) return;
  free_tracked(o);
  free_results(o);
  oscene_release(o);
  free(o);
}

static int tctx_init(struct otctx* c, const struct ooracle* o, int sxr)
{
  memset(c, 0, sizeof(*c));
  omc_init(c->global, 6);
  c->rcv = malloc((o->nrcv * 20 + 1) * sizeof(struct omc));
  c->samp = malloc((o->nsamp * 2 + 1) * sizeof(struct omc));
  c->nb_samples = calloc(o->nsamp + 1, sizeof(uint64_t));
  c->prim = malloc((o->nprim_slots * 20 + 1) * sizeof(struct omc));
  if(sxr) c->sxr = malloc((o->nsamp * o->nrcv * 20 + 1) * sizeof(struct omc));
  if(!c->rcv || !c->samp || !c->nb_samples || !c->prim || (sxr && !c->sxr))
    return -1;
  omc_init(c->rcv, o->nrcv * 20);
  omc_init(c->samp, o->nsamp * 2);
  omc_init(c->prim, o->nprim_slots * 20);
  if(sxr) omc_init(c->sxr, o->nsamp * o->nrcv * 20);
  return 0;
}

static void tctx_release(struct otctx* c)
{
  free(c->rcv); free(c->samp); free(c->nb_samples); free(c->sxr);
  free(c->prim); free(c->touched);
}

int
oracle_solve
  (struct ooracle* o, uint64_t rng_key, uint64_t rng_counter,
   uint64_t realisations, uint64_t max_failed, int nthreads)
{
  struct otctx* ctxs;
  const int sxr = o->scn.track_sampled_x_receiver;
  int64_t i, n = (int64_t)realisations;
  volatile int abort_flag = 0;
  uint64_t nfailures = 0;
  int t, res = 0;

  if(!realisations) return 1;
  if(nthreads <= 0) nthreads = omp_get_num_procs();
  if(o->track) nthreads = 1; /* paths are appended in realisation order */
  else free_tracked(o);       /* recorded paths belong to the last solve */
  free_results(o);
  ctxs = calloc((size_t)nthreads, sizeof(*ctxs));
  if(!ctxs) return 2;
  for(t = 0; t < nthreads; ++t) if(tctx_init(ctxs + t, o, sxr)) return 2;

  #pragma omp parallel for schedule(static) num_threads(nthreads)
  for(i = 0; i < n; ++i) {
    struct otctx* c = ctxs + omp_get_thread_num();
    int r;
    if(abort_flag) continue;
    r = trace_path(o, c, (uint64_t)i, rng_key, rng_counter);
    if(o->track) track_finish(o, c, r);
    if(r) {
      uint64_t nf;
      omc_apply_factor(c, (uint64_t)i, 0);
      #pragma omp atomic capture
      nf = ++nfailures;
      if(nf >= max_failed) abort_flag = 1;
      continue;
    }
    c->realisations++;
  }

  /* Merge (src/ssol_solver.c:735-797) */
  o->realisations = 0;
  o->segments = 0;
  omc_init(o->global, 6);
  o->rcv = calloc(o->nrcv * 20 + 1, sizeof(struct omc));
  o->samp = calloc(o->nsamp * 2 + 1, sizeof(struct omc));
  o->nb_samples = calloc(o->nsamp + 1, sizeof(uint64_t));
  o->prim = calloc(o->nprim_slots * 20 + 1, sizeof(struct omc));
  if(sxr) o->sxr = calloc(o->nsamp * o->nrcv * 20 + 1, sizeof(struct omc));
  for(t = 0; t < nthreads; ++t) {
    struct otctx* c = ctxs + t;
    size_t k;
    omc_merge(o->global, c->global, 6);
    omc_merge(o->rcv, c->rcv, o->nrcv * 20);
    omc_merge(o->samp, c->samp, o->nsamp * 2);
    omc_merge(o->prim, c->prim, o->nprim_slots * 20);
    if(sxr) omc_merge(o->sxr, c->sxr, o->nsamp * o->nrcv * 20);
    for(k = 0; k < o->nsamp; ++k) o->nb_samples[k] += c->nb_samples[k];
    o->realisations += c->realisations;
    o->segments += c->segments;
    tctx_release(c);
  }
  free(ctxs);
  o->failed = nfailures;
  if(abort_flag) res = 5;
  return res;
}

static void mc_result(const struct omc* d, double N, double out[3])
{
  const double w = d->w + d->tmp, w2 = d->w2 + d->tmp * d->tmp;
  out[0] = w / N;
  out[1] = w2 / N - out[0] * out[0];
  if(!(out[1] > 0)) out[1] = 0;
  out[2] = sqrt(out[1] / N);
}

void
oracle_get_global(struct ooracle* o, double out[6][3])
{
  int i;
  for(i = 0; i < 6; ++i) mc_result(o->global + i, (double)o->realisations, out[i]);
}

int
oracle_get_receiver(struct ooracle* o, size_t inst, int side, double out[10][3])
{
  int f, s;
  if(inst >= o->scn.ninstances || o->rcv_slot[inst] < 0) return 1;
  if(!(o->scn.instances[inst].receiver_mask & side)) return 1;
  s = side == OSIDE_FRONT ? 0 : 1;
  for(f = 0; f < 10; ++f)
    mc_result(o->rcv + ((size_t)o->rcv_slot[inst] * 2 + s) * 10 + f,
      (double)o->realisations, out[f]);
  return 0;
}

int
oracle_get_sampled(struct ooracle* o, size_t inst, double out[2][3],
  uint64_t* nb)
{
  int slot;
  if(inst >= o->scn.ninstances || (slot = o->samp_slot[inst]) < 0) return 1;
  *nb = o->nb_samples[slot];
  mc_result(o->samp + 2*slot + 0, (double)o->nb_samples[slot], out[0]);
  mc_result(o->samp + 2*slot + 1, (double)o->realisations, out[1]);
  return 0;
}

int
oracle_get_sampled_x_receiver(struct ooracle* o, size_t si, size_t ri,
  int side, double out[10][3])
{
  int f, s;
  if(!o->sxr || si >= o->scn.ninstances || ri >= o->scn.ninstances) return 1;
  if(o->samp_slot[si] < 0 || o->rcv_slot[ri] < 0) return 1;
  if(!(o->scn.instances[ri].receiver_mask & side)) return 1;
  s = side == OSIDE_FRONT ? 0 : 1;
  for(f = 0; f < 10; ++f) {
    const size_t b = (((size_t)o->samp_slot[si] * o->nrcv + o->rcv_slot[ri]) * 2 + s) * 10;
    mc_result(o->sxr + b + f, (double)o->realisations, out[f]);
  }
  return 0;
}

int
oracle_get_primitive(struct ooracle* o, size_t inst, size_t sshape,
  unsigned prim, int side, double out[10][3])
{
  const struct oshape_geo* g;
  size_t base;
  double area, e1[3], e2[3], n[3];
  int f, s, c;
  if(inst >= o->scn.ninstances || sshape >= o->max_sshapes) return 1;
  base = o->prim_base[inst * o->max_sshapes + sshape];
  if(base == SIZE_MAX) return 1;
  g = o->shapes + oget_sshape(o, (unsigned)inst, (unsigned)sshape)->shape;
  if(prim >= g->ntris) return 1;
  s = side == OSIDE_FRONT ? 0 : 1;
  {
    const double* a = g->rt_verts + 3 * g->rt_ids[3*prim];
    const double* b = g->rt_verts + 3 * g->rt_ids[3*prim+1];
    const double* cc = g->rt_verts + 3 * g->rt_ids[3*prim+2];
    for(c = 0; c < 3; ++c) { e1[c] = b[c] - a[c]; e2[c] = cc[c] - a[c]; }
    ov_cross(n, e1, e2);
    area = 0.5 * sqrt(ov_dot(n, n));
  }
  for(f = 0; f < 10; ++f) {
    mc_result(o->prim + ((base + prim) * 2 + s) * 10 + f,
      (double)o->realisations, out[f]);
    out[f][0] /= area;
    out[f][1] /= area * area;
    out[f][2] /= area;
  }
  return 0;
}

uint64_t oracle_realisation_count(const struct ooracle* o) { return o->realisations; }
uint64_t oracle_failed_count(const struct ooracle* o) { return o->failed; }
double oracle_sampled_area(const struct ooracle* o) { return o->sampled_area; }
double oracle_sampled_area_proxy(const struct ooracle* o) { return o->sampled_area_proxy; }
uint64_t oracle_segments(const struct ooracle* o) { return o->segments; }
unsigned oracle_shape_ntris(const struct ooracle* o, size_t s) { return o->shapes[s].ntris; }
double oracle_shape_rt_area(const struct ooracle* o, size_t s) { return o->shapes[s].rt_area; }
double oracle_shape_samp_area(const struct ooracle* o, size_t s) { return o->shapes[s].samp_area; }
unsigned oracle_shape_rt_vertex_count(const struct ooracle* o, size_t s)
{ return o->shapes[s].nverts; }
void oracle_shape_rt_vertex(const struct ooracle* o, size_t s, unsigned i, double pos[3])
{ ov_set(pos, o->shapes[s].rt_verts + 3*i); }

/* compute_infinite_path_segment_extend (src/ssol_solver.c:562-570) over the
 * world-space ray-traced triangles */
double
oracle_tracked_extend(const struct ooracle* o)
{
  double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  double e = 0;
  size_t t;
  int k, v;
  for(t = 0; t < o->ntris; ++t) {
    const struct owtri* tr = o->tris + t;
    for(v = 0; v < 3; ++v) {
      for(k = 0; k < 3; ++k) {
        const double x = tr->v0[k] + (v == 1 ? tr->e1[k] : 0) + (v == 2 ? tr->e2[k] : 0);
        if(x < lo[k]) lo[k] = x;
        if(x > hi[k]) hi[k] = x;
      }
    }
  }
  for(k = 0; k < 3; ++k) if(hi[k] >= lo[k] && hi[k] - lo[k] > e) e = hi[k] - lo[k];
  return e * 0.75;
}

int
oracle_solve_tracked
  (struct ooracle* o, uint64_t rng_key, uint64_t rng_counter,
   uint64_t realisations, uint64_t max_failed, double sun_ray_length,
   double infinite_ray_length)
{
  int res;
  free_tracked(o);
  if(sun_ray_length < 0 || infinite_ray_length < 0) {
    const double ext = oracle_tracked_extend(o);
    if(sun_ray_length < 0) sun_ray_length = ext;
    if(infinite_ray_length < 0) infinite_ray_length = ext;
  }
  o->track = 1;
  o->track_sun_len = sun_ray_length;
  o->track_inf_len = infinite_ray_length;
  res = oracle_solve(o, rng_key, rng_counter, realisations, max_failed, 1);
  o->track = 0;
  return res;
}

size_t oracle_tracked_count(const struct ooracle* o) { return o->tp_n; }

int
oracle_tracked_path
  (const struct ooracle* o, size_t i, int* type, size_t* n, const double** v)
{
  if(i >= o->tp_n) return 1;
  *type = o->tp_type[i];
  *n = o->tp_count[i];
  *v = o->tv + 4 * o->tp_first[i];
  return 0;
}
