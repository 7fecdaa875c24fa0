/* ORACLE -- test infrastructure only (see oracle.h).
 * Deterministic building blocks: RNG, quadric roots, sun and wavelength
 * samplers, BSDF sampling. */
#include "oracle_internal.h"

#include <float.h>
#include <math.h>
#include <string.h>

/*******************************************************************************
 * Threefry-4x64-20 (Salmon et al., SC'11; Random123 rotation constants)
 ******************************************************************************/
static inline uint64_t rotl64(uint64_t x, unsigned r)
{
  return (x << r) | (x >> (64 - r));
}

void
oracle_threefry4x64_20
  (const uint64_t ctr[4], const uint64_t key[4], uint64_t out[4])
{
  static const unsigned R[8][2] = {
    {14,16}, {52,57}, {23,40}, {5,37}, {25,33}, {46,12}, {58,22}, {32,32}
  };
  uint64_t ks[5];
  uint64_t x[4];
  int r, i;
  ks[4] = UINT64_C(0x1BD11BDAA9FC1A22);
  for(i = 0; i < 4; ++i) {
    ks[i] = key[i];
    ks[4] ^= key[i];
    x[i] = ctr[i] + key[i];
  }
  for(r = 0; r < 20; ++r) {
    const unsigned* rot = R[r % 8];
    if((r & 1) == 0) {
      x[0] += x[1]; x[1] = rotl64(x[1], rot[0]); x[1] ^= x[0];
      x[2] += x[3]; x[3] = rotl64(x[3], rot[1]); x[3] ^= x[2];
    } else {
      x[0] += x[3]; x[3] = rotl64(x[3], rot[0]); x[3] ^= x[0];
      x[2] += x[1]; x[1] = rotl64(x[1], rot[1]); x[1] ^= x[2];
    }
    if((r & 3) == 3) { /* key injection after every 4 rounds */
      const unsigned s = (unsigned)(r >> 2) + 1;
      for(i = 0; i < 4; ++i) x[i] += ks[(s + (unsigned)i) % 5];
      x[3] += s;
    }
  }
  memcpy(out, x, sizeof(x));
}

void
orng_init(struct orng* rng, uint64_t key, uint64_t path)
{
  rng->key[0] = key;
  rng->key[1] = ORACLE_TAG_PATH;
  rng->key[2] = 0;
  rng->key[3] = 0;
  rng->path = path;
  rng->block = 0;
  rng->avail = 0;
}

static void
orng_refill(struct orng* rng)
{
  uint64_t ctr[4];
  ctr[0] = rng->path;
  ctr[1] = rng->block++;
  ctr[2] = 0;
  ctr[3] = 0;
  oracle_threefry4x64_20(ctr, rng->key, rng->buf);
  rng->avail = 4;
}

double
orng_canonical(struct orng* rng)
{
  uint64_t x;
  if(!rng->avail) orng_refill(rng);
  x = rng->buf[4 - rng->avail--];
  return (double)(x >> 11) * (1.0 / 9007199254740992.0); /* [0, 1) */
}

/*******************************************************************************
 * Quadrics (reference src/ssol_shape.c:236-285, 864-1135, 1320-1350)
 ******************************************************************************/
/* Root of a.t^2 + b.t + c = 0 closest to `hint` (src/ssol_shape.c:864-896).
 * The reference never tests the discriminant; a negative one is treated
 * here as "no root", the case its NaN roots end up in (never stored as the
 * nearest hit, src/ssol_scene.c:529-533). */
int
oracle_solve_second(double a, double b, double c, double hint, double dist[2])
{
  if(a == 0) {
    if(b != 0) { dist[0] = -c / b; return 1; }
    return 0;
  } else {
    const double delta = b*b - 4*a*c;
    if(delta < 0) return 0;
    if(delta == 0) { dist[0] = -b / (2*a); return 1; }
    {
      const double sq = sqrt(delta);
      const double t1 = (-b - copysign(sq, b)) / (2*a);
      const double t2 = c / (a*t1);
      const int first = fabs(t1 - hint) < fabs(t2 - hint);
      dist[0] = first ? t1 : t2;
      dist[1] = first ? t2 : t1;
      return 2;
    }
  }
}

/* Private quadric data, as derived by src/ssol_shape.c:1260-1350 */
void
oquadric_setup(int type, const double qdata[2], double p[4])
{
  p[0] = p[1] = p[2] = p[3] = 0;
  switch(type) {
    case OQ_PLANE: break;
    case OQ_PARABOL:
    case OQ_PARABOLIC_CYLINDER:
      p[0] = qdata[0];
      p[1] = 1 / (4.0 * qdata[0]);
      break;
    case OQ_PARABOL2F:
      p[0] = 1 / (4.0 * qdata[0]);
      p[1] = 1 / (4.0 * qdata[1]);
      break;
    case OQ_HYPERBOL: {
      /* qdata = {img_focal, real_focal} */
      const double g = qdata[1] + qdata[0];
      const double f = qdata[1] / g;
      const double a2 = g * g * (f - f * f);
      const double abs_b = g * fabs(f - 0.5);
      p[0] = g * 0.5;              /* g_square */
      p[1] = a2 / (abs_b * abs_b); /* a2 / b2 */
      p[2] = 1 / a2;
      p[3] = abs_b;
      break;
    }
    case OQ_HEMISPHERE:
      p[0] = qdata[0];
      p[1] = qdata[0] * qdata[0];
      break;
    default: break;
  }
}

double
oquadric_z(int type, const double p[4], const double xy[2])
{
  const double x = xy[0], y = xy[1];
  switch(type) {
    case OQ_PLANE: return 0;
    case OQ_PARABOL: return (x*x + y*y) * p[1];
    case OQ_PARABOL2F: return x*x * p[0] + y*y * p[1];
    case OQ_PARABOLIC_CYLINDER: return y*y * p[1];
    case OQ_HYPERBOL: {
      const double z0 = p[0] + p[3];
      return p[3] * sqrt(1 + (x*x + y*y) * p[2]) + p[0] - z0;
    }
    case OQ_HEMISPHERE: {
      const double z2 = p[1] - (x*x + y*y);
      return z2 > 0 ? -sqrt(z2) + p[0] : p[0];
    }
    default: return 0;
  }
}

void
oquadric_grad(int type, const double p[4], const double pt[3], double g[3])
{
  switch(type) {
    case OQ_PLANE: g[0] = 0; g[1] = 0; g[2] = 1; break;
    case OQ_PARABOL: g[0] = -pt[0]; g[1] = -pt[1]; g[2] = 2 * p[0]; break;
    case OQ_PARABOL2F:
      g[0] = -pt[0] * (4.0 * p[0]);
      g[1] = -pt[1] * (4.0 * p[1]);
      g[2] = 2.0;
      break;
    case OQ_PARABOLIC_CYLINDER: g[0] = 0; g[1] = -pt[1]; g[2] = 2 * p[0]; break;
    case OQ_HYPERBOL: {
      const double z0 = p[0] + p[3];
      g[0] = pt[0];
      g[1] = pt[1];
      g[2] = -(pt[2] + z0 - p[0]) * p[1];
      break;
    }
    case OQ_HEMISPHERE: g[0] = -pt[0]; g[1] = -pt[1]; g[2] = p[0] - pt[2]; break;
    default: g[0] = g[1] = 0; g[2] = 1; break;
  }
}

/* Ray / quadric in the quadric frame (src/ssol_shape.c:970-1135) */
int
oquadric_intersect(int type, const double p[4], const double o[3],
  const double d[3], double hint, double* t, double grad[3])
{
  double a = 0, b = 0, c = 0, dst[2], hit[3];
  switch(type) {
    case OQ_PLANE: a = 0; b = d[2]; c = o[2]; break;
    case OQ_PARABOL:
      a = d[0]*d[0] + d[1]*d[1];
      b = 2*o[0]*d[0] + 2*o[1]*d[1] - 4*p[0]*d[2];
      c = o[0]*o[0] + o[1]*o[1] - 4*p[0]*o[2];
      break;
    case OQ_PARABOL2F:
      a = d[0]*d[0]*p[0] + d[1]*d[1]*p[1];
      b = 2.0*o[0]*d[0]*p[0] + 2.0*o[1]*d[1]*p[1] - d[2];
      c = o[0]*o[0]*p[0] + o[1]*o[1]*p[1] - o[2];
      break;
    case OQ_PARABOLIC_CYLINDER:
      a = d[1]*d[1];
      b = 2*o[1]*d[1] - 4*p[0]*d[2];
      c = o[1]*o[1] - 4*p[0]*o[2];
      break;
    case OQ_HYPERBOL: {
      const double b2 = p[3] * p[3];
      const double b2_a2 = b2 * p[2];
      const double z0 = p[0] + p[3];
      const double oz = o[2] + z0 - p[0];
      a = b2_a2 * (d[0]*d[0] + d[1]*d[1]) - d[2]*d[2];
      b = 2 * (b2_a2 * (o[0]*d[0] + o[1]*d[1]) - oz * d[2]);
      c = b2_a2 * (o[0]*o[0] + o[1]*o[1]) + b2 - oz * oz;
      break;
    }
    case OQ_HEMISPHERE: {
      const double z0 = -p[0];
      a = d[0]*d[0] + d[1]*d[1] + d[2]*d[2];
      b = 2 * (o[0]*d[0] + o[1]*d[1] + o[2]*d[2] + z0*d[2]);
      c = o[0]*o[0] + o[1]*o[1] + o[2]*o[2] - p[1] + 2*z0*o[2] + z0*z0;
      break;
    }
    default: return 0;
  }
  if(!oracle_solve_second(a, b, c, hint, dst)) { *t = INFINITY; return 0; }
  hit[0] = o[0] + d[0]*dst[0];
  hit[1] = o[1] + d[1]*dst[0];
  hit[2] = o[2] + d[2]*dst[0];
  oquadric_grad(type, p, hit, grad);
  *t = dst[0];
  return 1;
}

int
oracle_quadric_intersect_local
  (int quadric, const double qdata[2], const double org[3],
   const double dir[3], double hint, double* t, double grad[3])
{
  double p[4];
  oquadric_setup(quadric, qdata, p);
  return oquadric_intersect(quadric, p, org, dir, hint, t, grad);
}

double
oracle_quadric_z(int quadric, const double qdata[2], const double xy[2])
{
  double p[4];
  oquadric_setup(quadric, qdata, p);
  return oquadric_z(quadric, p, xy);
}

/*******************************************************************************
 * Frame and sun direction samplers (src/ssol_ranst_sun_dir.c)
 ******************************************************************************/
/* Orthonormal frame whose third column is `dir`. The reference uses RSys
 * d33_basis, whose exact x/y orientation is not vendored; every sampler
 * draws a uniform azimuth, so only the third column matters. */
void
oracle_basis(const double dir[3], double m[9])
{
  double z[3], x[3], y[3], h[3] = {0, 0, 0}, n;
  int i;
  n = sqrt(dir[0]*dir[0] + dir[1]*dir[1] + dir[2]*dir[2]);
  for(i = 0; i < 3; ++i) z[i] = dir[i] / n;
  if(fabs(z[0]) <= fabs(z[1]) && fabs(z[0]) <= fabs(z[2])) h[0] = 1;
  else if(fabs(z[1]) <= fabs(z[2])) h[1] = 1;
  else h[2] = 1;
  ov_cross(x, h, z);
  ov_normalize(x);
  ov_cross(y, z, x);
  for(i = 0; i < 3; ++i) { m[i] = x[i]; m[3+i] = y[i]; m[6+i] = z[i]; }
}

/* Buie sunshape (src/ssol_ranst_sun_dir.c:88-241) */
static double
buie_chi(double csr)
{
  if(csr > 0.145) {
    return -0.04419909985804843 + csr * (1.401323894233574 + csr
      * (-0.3639746714505299 + csr * (-0.9579768560161194
      + 1.1550475450828657 * csr)));
  }
  if(csr > 0.035) {
    return 0.022652077593662934 + csr * (0.5252380349996234 + csr
      * (2.5484334534423887 - 0.8763755326550412 * csr));
  }
  return 0.004733749294807862 + csr * (4.716738065192151 + csr
    * (-463.506669149804 + csr * (24745.88727411664 + csr
    * (-606122.7511711778 + 5521693.445014727 * csr))));
}

void
obuie_setup(double csr, struct obuie* s)
{
  const double thetaCS = 0.0436;
  const double integralA = 9.224724736098827E-6;
  double chi, k, g2, integralB, aR1, aR2;
  s->thetaSD = 0.00465;
  s->dThetaCSSD = thetaCS - s->thetaSD;
  chi = buie_chi(csr);
  k = 0.9 * log(13.5 * chi) * pow(chi, -0.3);
  s->gamma = 2.2 * log(0.52 * chi) * pow(chi, 0.43) - 0.1;
  s->ek1000g = exp(k) * pow(1000, s->gamma);
  g2 = s->gamma + 2.0;
  integralB = exp(k) * pow(1000, s->gamma) / g2
    * (pow(thetaCS, g2) - pow(s->thetaSD, g2));
  s->alpha = 1 / (integralA + integralB);
  s->h1 = 1.001 * obuie_pdf(s, 0.0038915695846209047);
  s->h2 = obuie_pdf(s, s->thetaSD);
  aR1 = s->thetaSD * s->h1;
  aR2 = s->dThetaCSSD * s->h2;
  s->p1 = aR1 / (aR1 + aR2);
}

double
obuie_pdf(const struct obuie* s, double theta)
{
  const double phi = theta < s->thetaSD
    ? cos(326 * theta) / cos(308 * theta)
    : s->ek1000g * pow(theta, s->gamma);
  return s->alpha * phi * sin(theta);
}

/* The 8 sampler constants in the device order {thetaSD, dThetaCSSD, gamma,
 * e^k 1000^gamma, alpha, h1, h2, p_rect1}, for fixed-input GPU tests */
void
oracle_buie_params(double csr, double out[8])
{
  struct obuie s;
  obuie_setup(csr, &s);
  out[0] = s.thetaSD; out[1] = s.dThetaCSSD; out[2] = s.gamma; out[3] = s.ek1000g;
  out[4] = s.alpha; out[5] = s.h1; out[6] = s.h2; out[7] = s.p1;
}

double
oracle_buie_pdf_theta(double csr, double theta)
{
  struct obuie s;
  obuie_setup(csr, &s);
  return obuie_pdf(&s, theta);
}

void
osun_sample(const struct osun_state* sun, double (*next)(void*), void* ctx,
  double dir[3])
{
  double pt[3], st, ct, phi;
  switch(sun->type) {
    case 0: ov_set(dir, sun->dir); return;
    case 1: { /* pillbox, src/ssol_ranst_sun_dir.c:246-267 */
      const double s2 = next(ctx) * sun->sin2_max;
      st = sqrt(s2);
      ct = sqrt(1 - s2);
      phi = next(ctx) * 2 * M_PI;
      pt[0] = cos(phi) * st; pt[1] = sin(phi) * st; pt[2] = ct;
      break;
    }
    case 2: /* gaussian, src/ssol_ranst_sun_dir.c:272-295 */
      st = sun->sigma * sqrt(-2 * log(next(ctx)));
      if(st > 1) st = 1;
      ct = sqrt(1 - st*st);
      phi = next(ctx) * 2 * M_PI;
      pt[0] = cos(phi) * st; pt[1] = sin(phi) * st; pt[2] = ct;
      break;
    case 3: { /* buie, src/ssol_ranst_sun_dir.c:170-241 */
      double theta, value;
      phi = next(ctx) * 2 * M_PI;
      do {
        if(next(ctx) < sun->buie.p1) {
          theta = next(ctx) * sun->buie.thetaSD;
          value = next(ctx) * sun->buie.h1;
        } else {
          theta = sun->buie.thetaSD + next(ctx) * sun->buie.dThetaCSSD;
          value = next(ctx) * sun->buie.h2;
        }
      } while(value > obuie_pdf(&sun->buie, theta));
      st = sin(theta); ct = cos(theta);
      pt[0] = st * sin(phi); pt[1] = st * cos(phi); pt[2] = ct;
      ov_mat_mul(dir, sun->basis, pt);
      return; /* the reference does not renormalise this one */
    }
    default: ov_set(dir, sun->dir); return;
  }
  ov_mat_mul(dir, sun->basis, pt);
  ov_normalize(dir);
}

void
osun_state_setup(struct osun_state* s, int type, double param,
  const double dir[3])
{
  memset(s, 0, sizeof(*s));
  s->type = type;
  ov_set(s->dir, dir);
  ov_normalize(s->dir);
  oracle_basis(s->dir, s->basis);
  if(type == 1) { const double si = sin(param); s->sin2_max = si * si; }
  if(type == 2) s->sigma = param;
  if(type == 3) obuie_setup(param, &s->buie);
}

struct uarray { const double* u; int n, i; };
static double uarray_next(void* ctx)
{
  struct uarray* a = ctx;
  return a->i < a->n ? a->u[a->i++] : 0.5;
}

int
oracle_sun_sample
  (int type, double param, const double dir[3], const double* u, int nu,
   double out[3])
{
  struct osun_state s;
  struct uarray a;
  a.u = u; a.n = nu; a.i = 0;
  osun_state_setup(&s, type, param, dir);
  osun_sample(&s, uarray_next, &a, out);
  return a.i;
}

/*******************************************************************************
 * Spectra (src/ssol_spectrum.c:57-89, Star-SP piecewise linear)
 ******************************************************************************/
double
oracle_spectrum_interpolate
  (const double* wl, const double* val, size_t n, double w)
{
  size_t lo = 0, hi = n;
  /* first index with wl[i] >= w */
  while(lo < hi) {
    const size_t mid = (lo + hi) / 2;
    if(wl[mid] < w) lo = mid + 1; else hi = mid;
  }
  if(lo == n) return val[n-1];
  if(lo == 0) return val[0];
  return val[lo-1] + (w - wl[lo-1])
    * ((val[lo] - val[lo-1]) / (wl[lo] - wl[lo-1]));
}

double
odata_value(const struct odata* d, const struct oscene* scn, double wl)
{
  const struct ospectrum* s;
  if(d->spectrum < 0) return d->real;
  s = scn->spectra + d->spectrum;
  return oracle_spectrum_interpolate(s->wavelengths, s->values, s->n, wl);
}

/* Inverse CDF of the pdf proportional to the linear interpolation of
 * (wl, val). A single sample is a Dirac (src/ssol_ranst_sun_wl.c:154-188). */
double
oracle_piecewise_linear_sample
  (const double* wl, const double* val, size_t n, double u)
{
  double total = 0, target, acc = 0;
  size_t i;
  if(n <= 1) return n ? wl[0] : -1;
  for(i = 0; i + 1 < n; ++i) total += 0.5 * (val[i] + val[i+1]) * (wl[i+1] - wl[i]);
  target = u * total;
  for(i = 0; i + 1 < n; ++i) {
    const double w = wl[i+1] - wl[i];
    const double area = 0.5 * (val[i] + val[i+1]) * w;
    if(acc + area >= target || i + 2 == n) {
      const double r = target - acc;
      const double a = val[i];
      const double b = (val[i+1] - val[i]) / w;
      double x;
      if(fabs(b) * w < 1e-12 * (a > 0 ? a : 1)) {
        x = a > 0 ? r / a : 0.5 * w;
      } else {
        const double disc = a*a + 2*b*r;
        x = (-a + sqrt(disc > 0 ? disc : 0)) / b;
      }
      if(x < 0) x = 0;
      if(x > w) x = w;
      return wl[i] + x;
    }
    acc += area;
  }
  return wl[n-1];
}

/*******************************************************************************
 * BSDFs, written from the physics (Star-SF 0.6 is absent)
 ******************************************************************************/
static double
fresnel_dielectric(double cos_i, double eta_i, double eta_t, double* cos_t)
{
  const double e = eta_i / eta_t;
  const double sin2_t = e * e * (1 - cos_i * cos_i);
  double rs, rp, ct;
  if(sin2_t >= 1) { *cos_t = 0; return 1; }
  ct = sqrt(1 - sin2_t);
  rs = (eta_i * cos_i - eta_t * ct) / (eta_i * cos_i + eta_t * ct);
  rp = (eta_t * cos_i - eta_i * ct) / (eta_t * cos_i + eta_i * ct);
  *cos_t = ct;
  return 0.5 * (rs * rs + rp * rp);
}

/* Sample a microfacet normal with density D(h) (h.N) in the frame of N */
static void
sample_microfacet(int distrib, double m, double u1, double u2,
  const double N[3], double h[3])
{
  double tan_t, cos_t, sin_t, phi, loc[3], B[9];
  if(distrib == 0) tan_t = m * sqrt(-log(1 - u1));   /* Beckmann */
  else tan_t = m * sqrt(u1);                          /* pillbox */
  cos_t = 1 / sqrt(1 + tan_t * tan_t);
  sin_t = tan_t * cos_t;
  phi = 2 * M_PI * u2;
  loc[0] = cos(phi) * sin_t; loc[1] = sin(phi) * sin_t; loc[2] = cos_t;
  oracle_basis(N, B);
  ov_mat_mul(h, B, loc);
}

double
obsdf_sample(const struct omaterial* m, double eta_out, double eta_in,
  double slab_eta, double slab_abs, const double wi[3], const double N[3],
  double (*next)(void*), void* ctx, double wo[3], int* type)
{
  const double cos_i = ov_dot(wi, N);
  int i;
  *type = 1;
  switch(m->type) {
    case 2: /* mirror */
      if(m->roughness == 0) { /* specular, Fresnel constant */
        for(i = 0; i < 3; ++i) wo[i] = 2 * cos_i * N[i] - wi[i];
        return m->reflectivity;
      } else {
        /* microfacet reflection with resampling until the reflected
         * direction leaves the macro surface */
        int k;
        for(k = 0; k < 256; ++k) {
          double h[3], c;
          const double u1 = next(ctx), u2 = next(ctx);
          sample_microfacet(m->distribution, m->roughness, u1, u2, N, h);
          c = ov_dot(wi, h);
          if(c <= 0) continue;
          for(i = 0; i < 3; ++i) wo[i] = 2 * c * h[i] - wi[i];
          if(ov_dot(wo, N) > 0) return m->reflectivity;
        }
        ov_set(wo, N);
        return 0;
      }
    case 1: { /* matte: Lambertian */
      const double u1 = next(ctx), u2 = next(ctx);
      const double r = sqrt(u1), phi = 2 * M_PI * u2;
      double loc[3], B[9];
      loc[0] = r * cos(phi); loc[1] = r * sin(phi); loc[2] = sqrt(1 - u1);
      oracle_basis(N, B);
      ov_mat_mul(wo, B, loc);
      return m->reflectivity;
    }
    case 0: { /* dielectric interface: reflect with probability F */
      double cos_t;
      const double F = fresnel_dielectric(cos_i, eta_out, eta_in, &cos_t);
      if(next(ctx) < F) {
        for(i = 0; i < 3; ++i) wo[i] = 2 * cos_i * N[i] - wi[i];
        return 1;
      } else {
        const double e = eta_out / eta_in;
        for(i = 0; i < 3; ++i) wo[i] = -e * wi[i] + (e * cos_i - cos_t) * N[i];
        ov_normalize(wo);
        *type = 2;
        return 1;
      }
    }
    case 3: { /* thin slab: incoherent sum of the internal reflections */
      double cos_t, R, T, A, den, Rt, Tt, w;
      R = fresnel_dielectric(cos_i, eta_out, slab_eta, &cos_t);
      if(R >= 1) {
        for(i = 0; i < 3; ++i) wo[i] = 2 * cos_i * N[i] - wi[i];
        return 1;
      }
      T = 1 - R;
      A = exp(-slab_abs * m->thickness / cos_t);
      den = 1 - R * R * A * A;
      Rt = R + T * T * R * A * A / den;
      Tt = T * T * A / den;
      w = Rt + Tt;
      if(next(ctx) * w < Rt) {
        for(i = 0; i < 3; ++i) wo[i] = 2 * cos_i * N[i] - wi[i];
      } else {
        for(i = 0; i < 3; ++i) wo[i] = -wi[i];
        *type = 2;
      }
      return w;
    }
    default:
      ov_set(wo, wi);
      return 0;
  }
}

double
oracle_bsdf_sample
  (const struct omaterial* m, double eta_out, double eta_in,
   double slab_eta, double slab_absorption, const double wi[3],
   const double N[3], const double* u, int nu, double wo[3], int* type)
{
  struct uarray a;
  a.u = u; a.n = nu; a.i = 0;
  return obsdf_sample(m, eta_out, eta_in, slab_eta, slab_absorption, wi, N,
    uarray_next, &a, wo, type);
}
