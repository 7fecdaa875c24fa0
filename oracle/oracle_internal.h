/* ORACLE -- test infrastructure only (see oracle.h). Internal declarations. */
#ifndef ORACLE_INTERNAL_H
#define ORACLE_INTERNAL_H

#ifndef _GNU_SOURCE
#define _GNU_SOURCE
#endif
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define ORACLE_TAG_PATH UINT64_C(0x5041544853534f4c)

enum { OQ_PLANE, OQ_PARABOL, OQ_PARABOL2F, OQ_HYPERBOL,
       OQ_PARABOLIC_CYLINDER, OQ_HEMISPHERE };
enum { OM_DIELECTRIC, OM_MATTE, OM_MIRROR, OM_THIN, OM_VIRTUAL };
enum { OSIDE_FRONT = 1, OSIDE_BACK = 2 };

/* Small vector helpers; 3x3 matrices are column major */
static inline void ov_set(double d[3], const double s[3])
{ d[0] = s[0]; d[1] = s[1]; d[2] = s[2]; }
static inline double ov_dot(const double a[3], const double b[3])
{ return a[0]*b[0] + a[1]*b[1] + a[2]*b[2]; }
static inline void ov_cross(double d[3], const double a[3], const double b[3])
{
  const double x = a[1]*b[2] - a[2]*b[1];
  const double y = a[2]*b[0] - a[0]*b[2];
  const double z = a[0]*b[1] - a[1]*b[0];
  d[0] = x; d[1] = y; d[2] = z;
}
static inline double ov_normalize(double v[3])
{
  const double n = sqrt(ov_dot(v, v));
  if(n > 0) { v[0] /= n; v[1] /= n; v[2] /= n; }
  return n;
}
static inline void ov_mat_mul(double d[3], const double m[9], const double v[3])
{
  const double x = m[0]*v[0] + m[3]*v[1] + m[6]*v[2];
  const double y = m[1]*v[0] + m[4]*v[1] + m[7]*v[2];
  const double z = m[2]*v[0] + m[5]*v[1] + m[8]*v[2];
  d[0] = x; d[1] = y; d[2] = z;
}
/* d = m^T v */
static inline void ov_mat_tmul(double d[3], const double m[9], const double v[3])
{
  const double x = m[0]*v[0] + m[1]*v[1] + m[2]*v[2];
  const double y = m[3]*v[0] + m[4]*v[1] + m[5]*v[2];
  const double z = m[6]*v[0] + m[7]*v[1] + m[8]*v[2];
  d[0] = x; d[1] = y; d[2] = z;
}
static inline void ov_mat_mat(double d[9], const double a[9], const double b[9])
{
  double t[9];
  int c;
  for(c = 0; c < 3; ++c) ov_mat_mul(t + 3*c, a, b + 3*c);
  memcpy(d, t, sizeof(t));
}
/* Inverse of a 3x3 matrix; returns the determinant */
static inline double ov_mat_inv(double d[9], const double m[9])
{
  double t[9], det;
  int i;
  t[0] = m[4]*m[8] - m[7]*m[5];
  t[1] = m[7]*m[2] - m[1]*m[8];
  t[2] = m[1]*m[5] - m[4]*m[2];
  t[3] = m[6]*m[5] - m[3]*m[8];
  t[4] = m[0]*m[8] - m[6]*m[2];
  t[5] = m[3]*m[2] - m[0]*m[5];
  t[6] = m[3]*m[7] - m[6]*m[4];
  t[7] = m[6]*m[1] - m[0]*m[7];
  t[8] = m[0]*m[4] - m[3]*m[1];
  det = m[0]*t[0] + m[3]*t[1] + m[6]*t[2];
  for(i = 0; i < 9; ++i) d[i] = t[i] / det;
  return det;
}
static inline void ov_transpose(double d[9], const double m[9])
{
  double t[9];
  int r, c;
  for(r = 0; r < 3; ++r) for(c = 0; c < 3; ++c) t[c*3+r] = m[r*3+c];
  memcpy(d, t, sizeof(t));
}

/* Counter-based path RNG: stream (key, PATH tag), first counter word = path */
struct orng {
  uint64_t key[4];
  uint64_t path;
  uint64_t block;
  uint64_t buf[4];
  int avail;
};
void orng_init(struct orng* rng, uint64_t key, uint64_t path);
double orng_canonical(struct orng* rng);

struct obuie {
  double thetaSD, dThetaCSSD, gamma, ek1000g, alpha, h1, h2, p1;
};
void obuie_setup(double csr, struct obuie* s);
double obuie_pdf(const struct obuie* s, double theta);

struct osun_state {
  int type;
  double dir[3];
  double basis[9];
  double sin2_max;
  double sigma;
  struct obuie buie;
};
void osun_state_setup(struct osun_state* s, int type, double param,
  const double dir[3]);
void osun_sample(const struct osun_state* sun, double (*next)(void*),
  void* ctx, double dir[3]);

void oquadric_setup(int type, const double qdata[2], double p[4]);
double oquadric_z(int type, const double p[4], const double xy[2]);
void oquadric_grad(int type, const double p[4], const double pt[3], double g[3]);
int oquadric_intersect(int type, const double p[4], const double o[3],
  const double d[3], double hint, double* t, double grad[3]);

double odata_value(const struct odata* d, const struct oscene* scn, double wl);

double obsdf_sample(const struct omaterial* m, double eta_out, double eta_in,
  double slab_eta, double slab_abs, const double wi[3], const double N[3],
  double (*next)(void*), void* ctx, double wo[3], int* type);

/* Tessellated shape */
struct oshape_geo {
  int type;
  unsigned nverts;     /* RT vertices */
  double* rt_verts;    /* [nverts][3] object space (float precision values) */
  unsigned ntris;
  unsigned* rt_ids;    /* [ntris][3] */
  double* samp_verts;  /* [nverts][3] sampling proxy (punched: flat) */
  double rt_area, samp_area;
  /* punched */
  int quadric;
  double qp[4];
  double qR[9], qT[3];
};

/* World-space triangle of the flattened RT scene */
struct owtri {
  double v0[3], e1[3], e2[3];
  double Ng[3];        /* world geometry normal (left-handed convention) */
  unsigned inst;
  unsigned sshape;
  unsigned prim;
};

struct obvh_node {
  double lo[3], hi[3];
  unsigned first;      /* leaf: first tri; inner: left child */
  unsigned count;      /* 0 <=> inner node, right child = first + 1 is not
                          assumed: stored in `right` */
  unsigned right;
};

/* Deferred MC accumulator (reference src/ssol_estimator_c.h:33-126) */
struct omc {
  uint64_t irealisation;
  double tmp;
  double w, w2;
};

struct ooracle {
  struct oscene scn;         /* shallow copy of the description */
  struct oshape_geo* shapes;
  /* per instance */
  double* inst_R;            /* [ninst][9] */
  double* inst_T;            /* [ninst][3] */
  double* inst_Rinvt;        /* [ninst][9] inverse transpose */
  int* rcv_slot;             /* [ninst] or -1 */
  int* samp_slot;            /* [ninst] or -1 */
  size_t* prim_base;         /* [ninst * max_sshapes] per-prim slot base */
  size_t max_sshapes;
  size_t nrcv, nsamp, nprim_slots;
  /* flattened RT scene */
  struct owtri* tris;
  size_t ntris;
  struct obvh_node* nodes;
  size_t nnodes;
  /* sampling */
  double* samp_cdf;          /* [nsamp_tris] cumulative area */
  unsigned* samp_inst;
  unsigned* samp_sshape;
  unsigned* samp_prim;
  size_t nsamp_tris;
  double sampled_area, sampled_area_proxy;
  struct osun_state sun;
  struct omedium air;
  /* results */
  uint64_t realisations, failed, segments;
  struct omc global[6];
  struct omc* rcv;           /* [nrcv][2][10] */
  struct omc* samp;          /* [nsamp][2] cos, shadowed */
  uint64_t* nb_samples;      /* [nsamp] */
  struct omc* sxr;           /* [nsamp][nrcv][2][10] or NULL */
  struct omc* prim;          /* [nprim_slots][2][10] */
  /* path recording (src/ssol_solver.c:898-1127), single-threaded solves */
  int track;
  double track_sun_len, track_inf_len;
  double* tv;                /* [n][4]: position, weight */
  size_t tv_n, tv_cap;
  size_t* tp_first;          /* per recorded path: first vertex, count, type */
  size_t* tp_count;
  int* tp_type;
  size_t tp_n, tp_cap;
};

int oscene_build(struct ooracle* o);
void oscene_release(struct ooracle* o);

/* Ray query with the reference's hit filter semantics */
struct oray {
  double org[3], dir[3];
  double tmin;
  /* filter data (reference struct ray_data, src/ssol_c.h:34-52) */
  unsigned inst_from, sshape_from, prim_from;
  int side_from;
  int discard_virtual;
  int reversed;
  int any_hit;
  /* results */
  int hit;
  double t;
  unsigned inst, sshape, prim;
  double Ng[3];
  double qdst;
  double qN[3];
};
void otrace(const struct ooracle* o, struct oray* ray);
void opunched_project(const struct ooracle* o, unsigned inst, unsigned sshape,
  const double pos[3], double out_pos[3], double out_N[3]);
const struct osshape* oget_sshape(const struct ooracle* o, unsigned inst,
  unsigned sshape);

#endif /* ORACLE_INTERNAL_H */
