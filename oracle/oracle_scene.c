/* ORACLE -- test infrastructure only (see oracle.h).
 * Scene build: punched-surface tessellation and carving
 * (src/ssol_shape.c:493-811, 1611-1699), areas, flattened world-space
 * triangles, binned-SAH BVH, and ray queries carrying the reference's hit
 * filter (src/ssol_scene.c:438-536) and quadric refinement
 * (src/ssol_shape.c:1449-1494). */
#include "oracle_internal.h"

#include <float.h>
#include <stdio.h>

/*******************************************************************************
 * Growable arrays
 ******************************************************************************/
struct dvec { double* p; size_t n, cap; };
struct uvec { unsigned* p; size_t n, cap; };

static int dvec_push(struct dvec* v, double x)
{
  if(v->n == v->cap) {
    size_t c = v->cap ? v->cap * 2 : 64;
    double* p = realloc(v->p, c * sizeof(double));
    if(!p) return -1;
    v->p = p; v->cap = c;
  }
  v->p[v->n++] = x;
  return 0;
}
static int uvec_push(struct uvec* v, unsigned x)
{
  if(v->n == v->cap) {
    size_t c = v->cap ? v->cap * 2 : 64;
    unsigned* p = realloc(v->p, c * sizeof(unsigned));
    if(!p) return -1;
    v->p = p; v->cap = c;
  }
  v->p[v->n++] = x;
  return 0;
}

/*******************************************************************************
 * 2D carving of the tessellation grid
 ******************************************************************************/
typedef struct { double x, y; } p2;

static double poly_area2(const p2* p, size_t n)
{
  double a = 0;
  size_t i;
  for(i = 0; i < n; ++i) {
    const p2 a0 = p[i], a1 = p[(i+1)%n];
    a += a0.x * a1.y - a1.x * a0.y;
  }
  return a;
}

static int poly_is_convex(const p2* p, size_t n)
{
  int sign = 0;
  size_t i;
  for(i = 0; i < n; ++i) {
    const p2 a = p[i], b = p[(i+1)%n], c = p[(i+2)%n];
    const double cr = (b.x-a.x)*(c.y-b.y) - (b.y-a.y)*(c.x-b.x);
    if(cr > 0) { if(sign < 0) return 0; sign = 1; }
    else if(cr < 0) { if(sign > 0) return 0; sign = -1; }
  }
  return 1;
}

/* Signed distance side of q w.r.t. the directed edge a->b (CCW polygon:
 * > 0 inside) */
static double edge_side(p2 a, p2 b, p2 q)
{
  return (b.x - a.x) * (q.y - a.y) - (b.y - a.y) * (q.x - a.x);
}

/* Clip polygon `in` (n vertices) by the half plane left of a->b (keep = +1)
 * or right of it (keep = -1). Output in `out`, returns its size. */
static size_t clip_half(const p2* in, size_t n, p2 a, p2 b, int keep, p2* out)
{
  size_t i, m = 0;
  for(i = 0; i < n; ++i) {
    const p2 s = in[i], e = in[(i+1)%n];
    const double ds = keep * edge_side(a, b, s);
    const double de = keep * edge_side(a, b, e);
    if(ds >= 0) out[m++] = s;
    if((ds >= 0) != (de >= 0)) {
      const double t = ds / (ds - de);
      p2 q;
      q.x = s.x + t * (e.x - s.x);
      q.y = s.y + t * (e.y - s.y);
      out[m++] = q;
    }
  }
  return m;
}

/* Append the fan triangulation of a convex polygon */
static int emit_fan(struct dvec* xy, const p2* p, size_t n)
{
  size_t i;
  if(n < 3) return 0;
  if(fabs(poly_area2(p, n)) < 1e-14) return 0;
  for(i = 1; i + 1 < n; ++i) {
    if(dvec_push(xy, p[0].x) || dvec_push(xy, p[0].y)
    || dvec_push(xy, p[i].x) || dvec_push(xy, p[i].y)
    || dvec_push(xy, p[i+1].x) || dvec_push(xy, p[i+1].y)) return -1;
  }
  return 0;
}

/* Apply one convex polygon (n vertices, orientation ccw = +1 or -1) to a
 * triangle soup (xy: 6 doubles per triangle); the result is appended to
 * `out`. AND keeps the part inside the polygon, SUB removes it. */
static int carve_convex(const struct dvec* soup, const p2* poly, size_t n,
  int ccw, int operation, struct dvec* out)
{
  /* a triangle clipped by n half planes keeps at most n + 3 vertices */
  p2* buf0 = malloc((n + 4) * sizeof(p2));
  p2* buf1 = malloc((n + 4) * sizeof(p2));
  size_t i, k;
  int res = 0;
  if(!buf0 || !buf1) { free(buf0); free(buf1); return -1; }
  for(i = 0; i + 6 <= soup->n && !res; i += 6) {
    p2 tri[3];
    tri[0].x = soup->p[i]; tri[0].y = soup->p[i+1];
    tri[1].x = soup->p[i+2]; tri[1].y = soup->p[i+3];
    tri[2].x = soup->p[i+4]; tri[2].y = soup->p[i+5];
    if(operation == 0) { /* AND */
      size_t m = 3;
      memcpy(buf0, tri, sizeof(tri));
      for(k = 0; k < n && m; ++k) {
        m = clip_half(buf0, m, poly[k], poly[(k+1)%n], ccw, buf1);
        memcpy(buf0, buf1, m * sizeof(p2));
      }
      if(emit_fan(out, buf0, m)) res = -1;
    } else { /* SUB: tri minus polygon = union over edges of the pieces
                outside edge k and inside edges 0..k-1 */
      for(k = 0; k < n && !res; ++k) {
        size_t m = 3, j;
        memcpy(buf0, tri, sizeof(tri));
        m = clip_half(buf0, m, poly[k], poly[(k+1)%n], -ccw, buf1);
        memcpy(buf0, buf1, m * sizeof(p2));
        for(j = 0; j < k && m; ++j) {
          m = clip_half(buf0, m, poly[j], poly[(j+1)%n], ccw, buf1);
          memcpy(buf0, buf1, m * sizeof(p2));
        }
        if(emit_fan(out, buf0, m)) res = -1;
      }
    }
  }
  free(buf0);
  free(buf1);
  return res;
}

/* Split a simple polygon into triangles by ear clipping: repeatedly cut off
 * a convex vertex whose triangle holds no other vertex (Meisters 1975).
 * `tris` receives 3 p2 per triangle, counter-clockwise; returns the count,
 * or -1 when no ear is found (self-intersecting input). */
static long ear_clip(const p2* in, size_t n, p2* tris)
{
  size_t* idx = malloc(n * sizeof(size_t));
  size_t m = n, i, ntri = 0;
  const int ccw = poly_area2(in, n) > 0;
  if(!idx) return -1;
  for(i = 0; i < n; ++i) idx[i] = ccw ? i : n - 1 - i;
  while(m > 3) {
    int cut = 0;
    for(i = 0; i < m && !cut; ++i) {
      const p2 a = in[idx[(i + m - 1) % m]], b = in[idx[i]], c = in[idx[(i + 1) % m]];
      size_t j;
      int empty = 1;
      if(edge_side(a, b, c) <= 0) continue; /* reflex or flat corner */
      for(j = 0; j < m && empty; ++j) {
        const p2 q = in[idx[j]];
        if(j == i || j == (i + 1) % m || j == (i + m - 1) % m) continue;
        if(edge_side(a, b, q) >= 0 && edge_side(b, c, q) >= 0 && edge_side(c, a, q) >= 0)
          empty = 0;
      }
      if(!empty) continue;
      tris[3*ntri] = a; tris[3*ntri+1] = b; tris[3*ntri+2] = c;
      ++ntri;
      memmove(idx + i, idx + i + 1, (m - i - 1) * sizeof(size_t));
      --m;
      cut = 1;
    }
    if(!cut) { free(idx); return -1; }
  }
  if(edge_side(in[idx[0]], in[idx[1]], in[idx[2]]) > 0) {
    tris[3*ntri] = in[idx[0]]; tris[3*ntri+1] = in[idx[1]]; tris[3*ntri+2] = in[idx[2]];
    ++ntri;
  }
  free(idx);
  return (long)ntri;
}

/* Apply one carving polygon to the soup (src/ssol_shape.c:579-652, where
 * Star-CPR clips). A non-convex polygon is split into triangles: AND keeps
 * the union of the parts inside each triangle (they do not overlap), SUB
 * removes the triangles one after the other. */
static int carve(struct dvec* soup, const struct ocarving* c)
{
  struct dvec out = {0};
  p2* poly;
  size_t i, n = c->nverts;
  int res = 0;
  if(n < 3) { if(c->operation == 0) soup->n = 0; return 0; }
  poly = malloc(n * sizeof(p2));
  if(!poly) return -1;
  for(i = 0; i < n; ++i) { poly[i].x = c->xy[2*i]; poly[i].y = c->xy[2*i+1]; }
  if(poly_is_convex(poly, n)) {
    res = carve_convex(soup, poly, n, poly_area2(poly, n) > 0 ? 1 : -1, c->operation, &out);
  } else {
    p2* tris = malloc(3 * n * sizeof(p2));
    long k, ntri = tris ? ear_clip(poly, n, tris) : -1;
    if(ntri < 0) res = -2;
    for(k = 0; k < ntri && !res; ++k) {
      if(c->operation == 0) {
        res = carve_convex(soup, tris + 3*k, 3, 1, 0, &out);
      } else {
        struct dvec rest = {0};
        res = carve_convex(soup, tris + 3*k, 3, 1, 1, &rest);
        free(soup->p);
        *soup = rest;
      }
    }
    free(tris);
    if(!res && c->operation != 0) { free(poly); return 0; } /* soup already carved */
  }
  free(poly);
  if(res) { free(out.p); return res; }
  free(soup->p);
  *soup = out;
  return 0;
}

/*******************************************************************************
 * Shapes
 ******************************************************************************/
static double tri_area(const double* a, const double* b, const double* c)
{
  double e0[3], e1[3], n[3];
  int i;
  for(i = 0; i < 3; ++i) { e0[i] = b[i] - a[i]; e1[i] = c[i] - a[i]; }
  ov_cross(n, e0, e1);
  return sqrt(ov_dot(n, n));
}

static int build_mesh(struct oshape_geo* g, const struct oshape* s)
{
  unsigned i;
  g->type = 0;
  g->nverts = s->nverts;
  g->ntris = s->ntris;
  g->rt_verts = malloc(3 * sizeof(double) * s->nverts);
  g->rt_ids = malloc(3 * sizeof(unsigned) * s->ntris);
  if(!g->rt_verts || !g->rt_ids) return -1;
  for(i = 0; i < 3 * s->nverts; ++i) g->rt_verts[i] = (double)s->positions[i];
  for(i = 0; i < 3 * s->ntris; ++i) {
    if(s->indices[i] >= s->nverts) return -2;
    g->rt_ids[i] = s->indices[i];
  }
  g->samp_verts = g->rt_verts;
  g->rt_area = 0;
  for(i = 0; i < s->ntris; ++i) {
    g->rt_area += tri_area(g->rt_verts + 3*g->rt_ids[3*i],
      g->rt_verts + 3*g->rt_ids[3*i+1], g->rt_verts + 3*g->rt_ids[3*i+2]);
  }
  g->rt_area *= 0.5;
  g->samp_area = g->rt_area;
  return 0;
}

/* Carve the 2D sheet, lift it onto the quadric and place it in object
 * space; the sampling proxy keeps the flat sheet */
static int carve_and_lift(struct oshape_geo* g, const struct oshape* s,
  struct dvec soup)
{
  size_t i, k;
  int res;
  /* Carvings, in order (src/ssol_shape.c:579-652) */
  for(k = 0; k < s->ncarvings; ++k) {
    res = carve(&soup, s->carvings + k);
    if(res) { free(soup.p); return res; }
  }

  /* Lift to the quadric and to object space, rounded to float like the
   * reference's Star-3D vertex buffers (src/ssol_shape.c:187-338) */
  g->ntris = (unsigned)(soup.n / 6);
  g->nverts = g->ntris * 3;
  g->rt_verts = malloc(sizeof(double) * 3 * g->nverts);
  g->samp_verts = malloc(sizeof(double) * 3 * g->nverts);
  g->rt_ids = malloc(sizeof(unsigned) * 3 * g->ntris);
  if(!g->rt_verts || !g->samp_verts || !g->rt_ids) { free(soup.p); return -1; }
  for(i = 0; i < g->nverts; ++i) {
    double p[3], q[3], xy[2];
    int c;
    xy[0] = soup.p[2*i];
    xy[1] = soup.p[2*i+1];
    if(s->quadric == OQ_PLANE) {
      p[0] = (float)xy[0]; p[1] = (float)xy[1]; p[2] = 0;
    } else {
      p[0] = xy[0]; p[1] = xy[1]; p[2] = oquadric_z(s->quadric, g->qp, xy);
    }
    ov_mat_mul(q, g->qR, p);
    for(c = 0; c < 3; ++c) g->rt_verts[3*i+c] = (float)(q[c] + g->qT[c]);
    p[0] = (float)xy[0]; p[1] = (float)xy[1]; p[2] = 0;
    ov_mat_mul(q, g->qR, p);
    for(c = 0; c < 3; ++c) g->samp_verts[3*i+c] = (float)(q[c] + g->qT[c]);
  }
  for(i = 0; i < 3 * (size_t)g->ntris; ++i) g->rt_ids[i] = (unsigned)i;
  free(soup.p);

  g->rt_area = g->samp_area = 0;
  for(i = 0; i < g->ntris; ++i) {
    g->rt_area += tri_area(g->rt_verts + 9*i, g->rt_verts + 9*i + 3,
      g->rt_verts + 9*i + 6);
    g->samp_area += tri_area(g->samp_verts + 9*i, g->samp_verts + 9*i + 3,
      g->samp_verts + 9*i + 6);
  }
  g->rt_area *= 0.5;
  g->samp_area *= 0.5;
  return 0;
}


/* Hemisphere (src/ssol_shape.c:1646-1671): the sheet is the xy projection
 * of a Star-3DUT hemisphere mesh (src/ssol_shape.c:437-489) of the carving
 * radius, with nslices sectors and nslices stacks; Star-3DUT 0.3.1 is not in
 * the image, so the mesh is rebuilt here from that description: a pole and
 * rings k = 1..n at polar angles k pi / (2n), i.e. radius r sin(k pi / 2n),
 * each of n vertices at azimuths 2 pi i / n */
static int build_hemisphere_disk(struct oshape_geo* g, const struct oshape* s)
{
  double radius = g->qp[0], r2 = -DBL_MAX;
  size_t n, k, i;
  struct dvec soup = {0};
  for(k = 0; k < s->ncarvings; ++k) {
    if(s->carvings[k].operation != 0) continue;
    for(i = 0; i < s->carvings[k].nverts; ++i) {
      const double x = s->carvings[k].xy[2*i], y = s->carvings[k].xy[2*i+1];
      if(x*x + y*y > r2) r2 = x*x + y*y;
    }
  }
  if(s->ncarvings && r2 >= 0 && sqrt(r2) < radius) radius = sqrt(r2);
  n = s->slices_hint != SIZE_MAX ? s->slices_hint
    : (size_t)(5.5 + 11 * radius / g->qp[0]);
  if(!n) return -2;
  for(k = 0; k < n; ++k) {
    const double ra = radius * sin((double)k * M_PI / (2.0 * (double)n));
    const double rb = radius * sin((double)(k+1) * M_PI / (2.0 * (double)n));
    for(i = 0; i < n; ++i) {
      const double a0 = 2 * M_PI * (double)i / (double)n;
      const double a1 = 2 * M_PI * (double)((i + 1) % n) / (double)n;
      const double x00 = ra*cos(a0), y00 = ra*sin(a0), x01 = ra*cos(a1), y01 = ra*sin(a1);
      const double x10 = rb*cos(a0), y10 = rb*sin(a0), x11 = rb*cos(a1), y11 = rb*sin(a1);
      /* clockwise seen from +Z, like the plane grid */
      dvec_push(&soup, x00); dvec_push(&soup, y00);
      dvec_push(&soup, x11); dvec_push(&soup, y11);
      dvec_push(&soup, x10); dvec_push(&soup, y10);
      if(k) {
        dvec_push(&soup, x00); dvec_push(&soup, y00);
        dvec_push(&soup, x01); dvec_push(&soup, y01);
        dvec_push(&soup, x11); dvec_push(&soup, y11);
      }
    }
  }
  return carve_and_lift(g, s, soup);
}

static int build_punched(struct oshape_geo* g, const struct oshape* s)
{
  double lower[2] = { DBL_MAX, DBL_MAX }, upper[2] = { -DBL_MAX, -DBL_MAX };
  double size[2], size_min, delta;
  size_t nslices, nsteps[2], nv[2], ix, iy, i, k;
  struct dvec soup = {0};

  g->type = 1;
  g->quadric = s->quadric;
  oquadric_setup(s->quadric, s->qdata, g->qp);
  memcpy(g->qR, s->qtransform, 9 * sizeof(double));
  memcpy(g->qT, s->qtransform + 9, 3 * sizeof(double));

  if(s->quadric == OQ_HEMISPHERE) return build_hemisphere_disk(g, s);
  if(!s->ncarvings) return -2;

  /* AABB of the AND carvings (src/ssol_shape.c:386-409) */
  for(k = 0; k < s->ncarvings; ++k) {
    if(s->carvings[k].operation != 0) continue;
    for(i = 0; i < s->carvings[k].nverts; ++i) {
      const double x = s->carvings[k].xy[2*i], y = s->carvings[k].xy[2*i+1];
      if(x < lower[0]) lower[0] = x;
      if(y < lower[1]) lower[1] = y;
      if(x > upper[0]) upper[0] = x;
      if(y > upper[1]) upper[1] = y;
    }
  }
  if(lower[0] >= upper[0] || lower[1] >= upper[1]) return -2;

  /* Number of slices (src/ssol_shape.c:1352-1390, 1656-1666) */
  if(s->slices_hint != SIZE_MAX) {
    nslices = s->slices_hint;
  } else if(s->quadric == OQ_PLANE) {
    nslices = 1;
  } else {
    const double zl = oquadric_z(s->quadric, g->qp, lower);
    const double zu = oquadric_z(s->quadric, g->qp, upper);
    const double max_z = zl > zu ? zl : zu;
    nslices = (size_t)(3 + sqrt(max_z) * 6);
    if(nslices > 50) nslices = 50;
  }
  if(!nslices) return -2;

  /* Regular grid (src/ssol_shape.c:493-577) */
  size[0] = upper[0] - lower[0];
  size[1] = upper[1] - lower[1];
  size_min = size[0] < size[1] ? size[0] : size[1];
  if(fabs(size_min) <= 1e-6) return -2;
  delta = size_min / (double)nslices;
  nsteps[0] = (size_t)ceil(size[0] / delta);
  nsteps[1] = (size_t)ceil(size[1] / delta);
  nv[0] = nsteps[0] + 1;
  nv[1] = nsteps[1] + 1;
  for(ix = 0; ix < nsteps[0]; ++ix) {
    for(iy = 0; iy < nsteps[1]; ++iy) {
      double x0 = lower[0] + (double)ix * delta;
      double x1 = lower[0] + (double)(ix+1) * delta;
      double y0 = lower[1] + (double)iy * delta;
      double y1 = lower[1] + (double)(iy+1) * delta;
      if(x0 > upper[0]) x0 = upper[0];
      if(x1 > upper[0]) x1 = upper[0];
      if(y0 > upper[1]) y0 = upper[1];
      if(y1 > upper[1]) y1 = upper[1];
      /* (id0, id3, id1) and (id0, id2, id3) with id0=(ix,iy), id1=(ix+1,iy),
       * id2=(ix,iy+1), id3=(ix+1,iy+1) */
      dvec_push(&soup, x0); dvec_push(&soup, y0);
      dvec_push(&soup, x1); dvec_push(&soup, y1);
      dvec_push(&soup, x1); dvec_push(&soup, y0);
      dvec_push(&soup, x0); dvec_push(&soup, y0);
      dvec_push(&soup, x0); dvec_push(&soup, y1);
      dvec_push(&soup, x1); dvec_push(&soup, y1);
    }
  }
  (void)nv;

  return carve_and_lift(g, s, soup);
}

/*******************************************************************************
 * BVH (binned SAH over world-space triangles)
 ******************************************************************************/
struct bref { double lo[3], hi[3], c[3]; };

static void box_grow(double lo[3], double hi[3], const double blo[3],
  const double bhi[3])
{
  int i;
  for(i = 0; i < 3; ++i) {
    if(blo[i] < lo[i]) lo[i] = blo[i];
    if(bhi[i] > hi[i]) hi[i] = bhi[i];
  }
}
static double box_area(const double lo[3], const double hi[3])
{
  const double dx = hi[0]-lo[0], dy = hi[1]-lo[1], dz = hi[2]-lo[2];
  if(dx < 0) return 0;
  return 2 * (dx*dy + dy*dz + dz*dx);
}

struct bvh_builder {
  struct ooracle* o;
  struct bref* refs;
  unsigned* idx;
  size_t nnodes, cap;
};

static size_t bvh_new_node(struct bvh_builder* b)
{
  if(b->nnodes == b->cap) {
    size_t c = b->cap ? b->cap * 2 : 1024;
    struct obvh_node* p = realloc(b->o->nodes, c * sizeof(struct obvh_node));
    if(!p) return (size_t)-1;
    b->o->nodes = p;
    b->cap = c;
  }
  return b->nnodes++;
}

#define NBINS 16
static int bvh_build_rec(struct bvh_builder* b, size_t node, size_t first,
  size_t count)
{
  double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX,-DBL_MAX,-DBL_MAX};
  double clo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, chi[3] = {-DBL_MAX,-DBL_MAX,-DBL_MAX};
  size_t i, best_axis = 0, best_split = 0, mid;
  double best_cost = DBL_MAX;
  int axis;
  for(i = first; i < first + count; ++i) {
    const struct bref* r = b->refs + b->idx[i];
    box_grow(lo, hi, r->lo, r->hi);
    box_grow(clo, chi, r->c, r->c);
  }
  memcpy(b->o->nodes[node].lo, lo, sizeof(lo));
  memcpy(b->o->nodes[node].hi, hi, sizeof(hi));
  if(count <= 4) {
    b->o->nodes[node].first = (unsigned)first;
    b->o->nodes[node].count = (unsigned)count;
    return 0;
  }
  for(axis = 0; axis < 3; ++axis) {
    double blo[NBINS][3], bhi[NBINS][3];
    size_t bcnt[NBINS] = {0};
    const double ext = chi[axis] - clo[axis];
    int k;
    if(ext <= 0) continue;
    for(k = 0; k < NBINS; ++k) {
      blo[k][0] = blo[k][1] = blo[k][2] = DBL_MAX;
      bhi[k][0] = bhi[k][1] = bhi[k][2] = -DBL_MAX;
    }
    for(i = first; i < first + count; ++i) {
      const struct bref* r = b->refs + b->idx[i];
      int bi = (int)((r->c[axis] - clo[axis]) / ext * NBINS);
      if(bi >= NBINS) bi = NBINS - 1;
      if(bi < 0) bi = 0;
      bcnt[bi]++;
      box_grow(blo[bi], bhi[bi], r->lo, r->hi);
    }
    for(k = 1; k < NBINS; ++k) {
      double llo[3] = {DBL_MAX,DBL_MAX,DBL_MAX}, lhi[3] = {-DBL_MAX,-DBL_MAX,-DBL_MAX};
      double rlo[3] = {DBL_MAX,DBL_MAX,DBL_MAX}, rhi[3] = {-DBL_MAX,-DBL_MAX,-DBL_MAX};
      size_t nl = 0, nr = 0;
      int j;
      double cost;
      for(j = 0; j < k; ++j) if(bcnt[j]) { box_grow(llo, lhi, blo[j], bhi[j]); nl += bcnt[j]; }
      for(j = k; j < NBINS; ++j) if(bcnt[j]) { box_grow(rlo, rhi, blo[j], bhi[j]); nr += bcnt[j]; }
      if(!nl || !nr) continue;
      cost = box_area(llo, lhi) * (double)nl + box_area(rlo, rhi) * (double)nr;
      if(cost < best_cost) { best_cost = cost; best_axis = (size_t)axis; best_split = (size_t)k; }
    }
  }
  if(best_cost == DBL_MAX) {
    mid = first + count / 2; /* degenerate centroids: median split */
  } else {
    const double ext = chi[best_axis] - clo[best_axis];
    size_t l = first, r = first + count;
    while(l < r) {
      const struct bref* rf = b->refs + b->idx[l];
      int bi = (int)((rf->c[best_axis] - clo[best_axis]) / ext * NBINS);
      if(bi >= NBINS) bi = NBINS - 1;
      if((size_t)bi < best_split) ++l;
      else { unsigned t = b->idx[l]; b->idx[l] = b->idx[--r]; b->idx[r] = t; }
    }
    mid = l;
    if(mid == first || mid == first + count) mid = first + count / 2;
  }
  {
    const size_t left = bvh_new_node(b);
    const size_t right = bvh_new_node(b);
    if(left == (size_t)-1 || right == (size_t)-1) return -1;
    b->o->nodes[node].first = (unsigned)left;
    b->o->nodes[node].right = (unsigned)right;
    b->o->nodes[node].count = 0;
    if(bvh_build_rec(b, left, first, mid - first)) return -1;
    if(bvh_build_rec(b, right, mid, first + count - mid)) return -1;
  }
  return 0;
}

static int bvh_build(struct ooracle* o)
{
  struct bvh_builder b;
  struct owtri* sorted;
  size_t i;
  memset(&b, 0, sizeof(b));
  b.o = o;
  o->nodes = NULL;
  o->nnodes = 0;
  if(!o->ntris) {
    o->nodes = calloc(1, sizeof(struct obvh_node));
    o->nodes[0].lo[0] = 1; o->nodes[0].hi[0] = -1;
    o->nnodes = 1;
    return 0;
  }
  b.refs = malloc(o->ntris * sizeof(struct bref));
  b.idx = malloc(o->ntris * sizeof(unsigned));
  if(!b.refs || !b.idx) return -1;
  for(i = 0; i < o->ntris; ++i) {
    const struct owtri* t = o->tris + i;
    int c;
    for(c = 0; c < 3; ++c) {
      const double a = t->v0[c], p1 = t->v0[c] + t->e1[c], p2 = t->v0[c] + t->e2[c];
      double mn = a < p1 ? a : p1, mx = a > p1 ? a : p1;
      if(p2 < mn) mn = p2;
      if(p2 > mx) mx = p2;
      b.refs[i].lo[c] = mn;
      b.refs[i].hi[c] = mx;
      b.refs[i].c[c] = 0.5 * (mn + mx);
    }
    b.idx[i] = (unsigned)i;
  }
  bvh_new_node(&b);
  if(bvh_build_rec(&b, 0, 0, o->ntris)) return -1;
  sorted = malloc(o->ntris * sizeof(struct owtri));
  if(!sorted) return -1;
  for(i = 0; i < o->ntris; ++i) sorted[i] = o->tris[b.idx[i]];
  free(o->tris);
  o->tris = sorted;
  o->nnodes = b.nnodes;
  free(b.refs);
  free(b.idx);
  return 0;
}

/*******************************************************************************
 * Build
 ******************************************************************************/
const struct osshape*
oget_sshape(const struct ooracle* o, unsigned inst, unsigned sshape)
{
  const struct oobject* obj = o->scn.objects + o->scn.instances[inst].object;
  return obj->sshapes + sshape;
}

int
oscene_build(struct ooracle* o)
{
  const struct oscene* scn = &o->scn;
  size_t i, j, k, nt = 0, ns = 0;
  int res;

  o->shapes = calloc(scn->nshapes, sizeof(struct oshape_geo));
  if(!o->shapes) return -1;
  for(i = 0; i < scn->nshapes; ++i) {
    res = scn->shapes[i].type == 0
      ? build_mesh(o->shapes + i, scn->shapes + i)
      : build_punched(o->shapes + i, scn->shapes + i);
    if(res) return res;
  }

  o->max_sshapes = 1;
  for(i = 0; i < scn->nobjects; ++i)
    if(scn->objects[i].nsshapes > o->max_sshapes)
      o->max_sshapes = scn->objects[i].nsshapes;

  o->inst_R = malloc(scn->ninstances * 9 * sizeof(double) + 1);
  o->inst_T = malloc(scn->ninstances * 3 * sizeof(double) + 1);
  o->inst_Rinvt = malloc(scn->ninstances * 9 * sizeof(double) + 1);
  o->rcv_slot = malloc(scn->ninstances * sizeof(int) + 1);
  o->samp_slot = malloc(scn->ninstances * sizeof(int) + 1);
  o->prim_base = malloc(scn->ninstances * o->max_sshapes * sizeof(size_t) + 1);
  if(!o->inst_R || !o->inst_T || !o->inst_Rinvt || !o->rcv_slot
  || !o->samp_slot || !o->prim_base) return -1;

  o->nrcv = o->nsamp = o->nprim_slots = 0;
  o->sampled_area = o->sampled_area_proxy = 0;
  for(i = 0; i < scn->ninstances; ++i) {
    const struct oinstance* in = scn->instances + i;
    const struct oobject* obj = scn->objects + in->object;
    double inv[9];
    memcpy(o->inst_R + 9*i, in->transform, 9 * sizeof(double));
    memcpy(o->inst_T + 3*i, in->transform + 9, 3 * sizeof(double));
    ov_mat_inv(inv, in->transform);
    ov_transpose(o->inst_Rinvt + 9*i, inv);
    o->rcv_slot[i] = in->receiver_mask ? (int)o->nrcv++ : -1;
    o->samp_slot[i] = in->sample ? (int)o->nsamp++ : -1;
    for(j = 0; j < o->max_sshapes; ++j) o->prim_base[i*o->max_sshapes + j] = SIZE_MAX;
    for(j = 0; j < obj->nsshapes; ++j) {
      const struct oshape_geo* g = o->shapes + obj->sshapes[j].shape;
      nt += g->ntris;
      if(in->sample) ns += g->ntris;
      if(in->receiver_mask && in->per_primitive) {
        o->prim_base[i*o->max_sshapes + j] = o->nprim_slots;
        o->nprim_slots += g->ntris;
      }
      if(in->sample) {
        o->sampled_area += g->rt_area;
        o->sampled_area_proxy += g->samp_area;
      }
    }
  }
  if(!o->nsamp) return -4; /* no instance to sample */

  /* Flattened world triangles and the sampling CDF */
  o->tris = malloc((nt ? nt : 1) * sizeof(struct owtri));
  o->samp_cdf = malloc((ns ? ns : 1) * sizeof(double));
  o->samp_inst = malloc((ns ? ns : 1) * sizeof(unsigned));
  o->samp_sshape = malloc((ns ? ns : 1) * sizeof(unsigned));
  o->samp_prim = malloc((ns ? ns : 1) * sizeof(unsigned));
  if(!o->tris || !o->samp_cdf || !o->samp_inst || !o->samp_sshape
  || !o->samp_prim) return -1;
  o->ntris = 0;
  o->nsamp_tris = 0;
  {
    double acc = 0;
    for(i = 0; i < scn->ninstances; ++i) {
      const struct oinstance* in = scn->instances + i;
      const struct oobject* obj = scn->objects + in->object;
      const double* R = o->inst_R + 9*i;
      const double* T = o->inst_T + 3*i;
      for(j = 0; j < obj->nsshapes; ++j) {
        const struct oshape_geo* g = o->shapes + obj->sshapes[j].shape;
        for(k = 0; k < g->ntris; ++k) {
          struct owtri* t = o->tris + o->ntris++;
          double w[3][3], e1[3], e2[3], n[3];
          int v, c;
          for(v = 0; v < 3; ++v) {
            const double* p = g->rt_verts + 3 * g->rt_ids[3*k + v];
            ov_mat_mul(w[v], R, p);
            for(c = 0; c < 3; ++c) w[v][c] += T[c];
          }
          for(c = 0; c < 3; ++c) {
            t->v0[c] = w[0][c];
            t->e1[c] = w[1][c] - w[0][c];
            t->e2[c] = w[2][c] - w[0][c];
          }
          /* object-space left-handed normal, to world with R^-T */
          {
            const double* a = g->rt_verts + 3 * g->rt_ids[3*k];
            const double* b = g->rt_verts + 3 * g->rt_ids[3*k+1];
            const double* cc = g->rt_verts + 3 * g->rt_ids[3*k+2];
            for(c = 0; c < 3; ++c) { e1[c] = b[c] - a[c]; e2[c] = cc[c] - a[c]; }
            ov_cross(n, e2, e1);
            ov_mat_mul(t->Ng, o->inst_Rinvt + 9*i, n);
          }
          t->inst = (unsigned)i;
          t->sshape = (unsigned)j;
          t->prim = (unsigned)k;
          if(in->sample) {
            const double* a = g->samp_verts + 3 * g->rt_ids[3*k];
            const double* b = g->samp_verts + 3 * g->rt_ids[3*k+1];
            const double* cc = g->samp_verts + 3 * g->rt_ids[3*k+2];
            acc += 0.5 * tri_area(a, b, cc);
            o->samp_cdf[o->nsamp_tris] = acc;
            o->samp_inst[o->nsamp_tris] = (unsigned)i;
            o->samp_sshape[o->nsamp_tris] = (unsigned)j;
            o->samp_prim[o->nsamp_tris] = (unsigned)k;
            o->nsamp_tris++;
          }
        }
      }
    }
  }
  if(bvh_build(o)) return -1;
  return 0;
}

void
oscene_release(struct ooracle* o)
{
  size_t i;
  if(o->shapes) {
    for(i = 0; i < o->scn.nshapes; ++i) {
      struct oshape_geo* g = o->shapes + i;
      if(g->samp_verts && g->samp_verts != g->rt_verts) free(g->samp_verts);
      free(g->rt_verts);
      free(g->rt_ids);
    }
  }
  free(o->shapes);
  free(o->inst_R); free(o->inst_T); free(o->inst_Rinvt);
  free(o->rcv_slot); free(o->samp_slot); free(o->prim_base);
  free(o->tris); free(o->nodes);
  free(o->samp_cdf); free(o->samp_inst); free(o->samp_sshape);
  free(o->samp_prim);
}

/*******************************************************************************
 * Queries
 ******************************************************************************/
/* World -> quadric frame of (inst, sshape): R = R_inst R_q, T = R_inst T_q +
 * T_inst (src/ssol_shape.c:1461-1466) */
static void
quadric_frame(const struct ooracle* o, unsigned inst, const struct oshape_geo* g,
  double R[9], double T[3], double Rinv[9], double Rinvt[9])
{
  const double* Ri = o->inst_R + 9*inst;
  const double* Ti = o->inst_T + 3*inst;
  int c;
  ov_mat_mat(R, Ri, g->qR);
  ov_mat_mul(T, Ri, g->qT);
  for(c = 0; c < 3; ++c) T[c] += Ti[c];
  ov_mat_inv(Rinv, R);
  ov_transpose(Rinvt, Rinv);
}

/* src/ssol_shape.c:1408-1447 */
void
opunched_project(const struct ooracle* o, unsigned inst, unsigned sshape,
  const double pos[3], double out_pos[3], double out_N[3])
{
  const struct oshape_geo* g = o->shapes + oget_sshape(o, inst, sshape)->shape;
  double R[9], T[3], Rinv[9], Rinvt[9], d[3], loc[3], n[3];
  int c;
  quadric_frame(o, inst, g, R, T, Rinv, Rinvt);
  for(c = 0; c < 3; ++c) d[c] = pos[c] - T[c];
  ov_mat_mul(loc, Rinv, d);
  loc[2] = oquadric_z(g->quadric, g->qp, loc);
  oquadric_grad(g->quadric, g->qp, loc, n);
  ov_mat_mul(out_pos, R, loc);
  for(c = 0; c < 3; ++c) out_pos[c] += T[c];
  ov_mat_mul(out_N, Rinvt, n);
  ov_normalize(out_N);
}

/* Quadric hit refined from the mesh hit distance (src/ssol_shape.c:1449-1494) */
static double
quadric_trace(const struct ooracle* o, unsigned inst, const struct oshape_geo* g,
  const double org[3], const double dir[3], double hint, double N[3])
{
  double R[9], T[3], Rinv[9], Rinvt[9], d[3], lo[3], ld[3], t, grad[3];
  int c;
  quadric_frame(o, inst, g, R, T, Rinv, Rinvt);
  for(c = 0; c < 3; ++c) d[c] = org[c] - T[c];
  ov_mat_mul(lo, Rinv, d);
  ov_mat_mul(ld, Rinv, dir);
  if(!oquadric_intersect(g->quadric, g->qp, lo, ld, hint, &t, grad))
    return INFINITY;
  ov_mat_mul(N, Rinvt, grad);
  ov_normalize(N);
  return t;
}

/* Hit filter (src/ssol_scene.c:438-536). Returns 1 to reject. */
static int
filter(const struct ooracle* o, struct oray* r, const struct owtri* t, double dist)
{
  const struct osshape* ss = oget_sshape(o, t->inst, t->sshape);
  const struct oshape_geo* g = o->shapes + ss->shape;
  const struct omaterial* m;
  int side;
  double N[3] = {0, 0, 0}, dst = INFINITY;
  if(dist < r->tmin) return 1;
  if(g->type == 0) {
    if(dist <= 1.e-6 || dist <= r->tmin
    || (t->prim == r->prim_from && t->sshape == r->sshape_from
      && t->inst == r->inst_from)) return 1;
    side = ov_dot(t->Ng, r->dir) < 0 ? OSIDE_FRONT : OSIDE_BACK;
  } else {
    dst = quadric_trace(o, t->inst, g, r->org, r->dir, dist, N);
    if(!(dst < (double)FLT_MAX)) return 1;
    if((float)dst <= r->tmin) return 1;
    side = ov_dot(r->dir, N) < 0 ? OSIDE_FRONT : OSIDE_BACK;
    if(t->inst == r->inst_from && t->sshape == r->sshape_from
    && side != r->side_from) return 1;
  }
  if(r->reversed) side = side == OSIDE_FRONT ? OSIDE_BACK : OSIDE_FRONT;
  m = o->scn.materials + (side == OSIDE_FRONT ? ss->mtl_front : ss->mtl_back);
  if(m->type == OM_VIRTUAL) {
    if(r->discard_virtual) return 1;
    if(!(o->scn.instances[t->inst].receiver_mask & side)) return 1;
  }
  if(g->type != 0 && r->qdst >= dst) {
    r->qdst = dst;
    ov_set(r->qN, N);
  }
  return 0;
}

static int
ray_box(const double org[3], const double inv[3], double tmin, double tmax,
  const double lo[3], const double hi[3])
{
  int i;
  for(i = 0; i < 3; ++i) {
    double t0 = (lo[i] - org[i]) * inv[i];
    double t1 = (hi[i] - org[i]) * inv[i];
    if(t0 > t1) { const double tmp = t0; t0 = t1; t1 = tmp; }
    if(t0 > tmin) tmin = t0;
    if(t1 < tmax) tmax = t1;
    if(tmin > tmax * (1 + 1e-12) + 1e-12) return 0;
  }
  return 1;
}

/* Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2013): edge
 * functions in a ray-aligned shear frame, so that a ray through a shared
 * edge or vertex hits at least one of the triangles. Returns the distance or
 * -1. */
static double
ray_tri(const double org[3], const double dir[3], const struct owtri* t)
{
  int kx, ky, kz;
  double Sx, Sy, Sz, A[3], B[3], C[3], Ax, Ay, Bx, By, Cx, Cy, U, V, W, det, T;
  const double ax = fabs(dir[0]), ay = fabs(dir[1]), az = fabs(dir[2]);
  int i;
  kz = ax > ay ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
  kx = (kz + 1) % 3;
  ky = (kx + 1) % 3;
  if(dir[kz] < 0) { const int tmp = kx; kx = ky; ky = tmp; }
  Sx = dir[kx] / dir[kz];
  Sy = dir[ky] / dir[kz];
  Sz = 1.0 / dir[kz];
  for(i = 0; i < 3; ++i) {
    A[i] = t->v0[i] - org[i];
    B[i] = t->v0[i] + t->e1[i] - org[i];
    C[i] = t->v0[i] + t->e2[i] - org[i];
  }
  Ax = A[kx] - Sx * A[kz]; Ay = A[ky] - Sy * A[kz];
  Bx = B[kx] - Sx * B[kz]; By = B[ky] - Sy * B[kz];
  Cx = C[kx] - Sx * C[kz]; Cy = C[ky] - Sy * C[kz];
  U = Cx * By - Cy * Bx;
  V = Ax * Cy - Ay * Cx;
  W = Bx * Ay - By * Ax;
  if((U < 0 || V < 0 || W < 0) && (U > 0 || V > 0 || W > 0)) return -1;
  det = U + V + W;
  if(det == 0) return -1;
  T = U * Sz * A[kz] + V * Sz * B[kz] + W * Sz * C[kz];
  return T / det;
}

void
otrace(const struct ooracle* o, struct oray* r)
{
  unsigned stack[128];
  int sp = 0;
  double inv[3], tfar = INFINITY;
  int i;
  r->hit = 0;
  r->qdst = INFINITY;
  for(i = 0; i < 3; ++i) inv[i] = 1.0 / r->dir[i];
  stack[sp++] = 0;
  while(sp) {
    const struct obvh_node* n = o->nodes + stack[--sp];
    if(!ray_box(r->org, inv, r->tmin, tfar, n->lo, n->hi)) continue;
    if(n->count) {
      unsigned k;
      for(k = n->first; k < n->first + n->count; ++k) {
        const struct owtri* t = o->tris + k;
        const double d = ray_tri(r->org, r->dir, t);
        if(!(d >= r->tmin && d < tfar)) continue;
        if(filter(o, r, t, d)) continue;
        tfar = d;
        r->hit = 1;
        r->t = d;
        r->inst = t->inst;
        r->sshape = t->sshape;
        r->prim = t->prim;
        ov_set(r->Ng, t->Ng);
        if(r->any_hit) return;
      }
    } else {
      if(sp + 2 > 128) continue;
      stack[sp++] = n->right;
      stack[sp++] = n->first;
    }
  }
}
