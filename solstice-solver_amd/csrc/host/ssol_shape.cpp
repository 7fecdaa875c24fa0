// Punched quadric surfaces: tessellation of the carving AABB, carving of
// the 2D sheet by the AND/SUB polygons, lift onto the quadric, and the flat
// sampling proxy. Follows src/ssol_shape.c:386-811 and 1611-1699 for the
// grid, slice counts and areas; the reference clips with Star-CPR (Clipper),
// here convex carvings clip by half planes and non-convex ones are first
// split into triangles by ear clipping.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>

#include "ssol_internal.hpp"

using namespace ssol;

namespace {

struct P2 {
  double x, y;
};

using Poly = std::vector<P2>;

double area2(const Poly& p) {
  double a = 0;
  for (size_t i = 0, n = p.size(); i < n; ++i) {
    const P2 &u = p[i], &v = p[(i + 1) % n];
    a += u.x * v.y - v.x * u.y;
  }
  return a;
}

double orient(P2 a, P2 b, P2 q) { return (b.x - a.x) * (q.y - a.y) - (b.y - a.y) * (q.x - a.x); }

bool convex(const Poly& p) {
  int sign = 0;
  for (size_t i = 0, n = p.size(); i < n; ++i) {
    const double c = orient(p[i], p[(i + 1) % n], p[(i + 2) % n]);
    if (c > 0) { if (sign < 0) return false; sign = 1; }
    else if (c < 0) { if (sign > 0) return false; sign = -1; }
  }
  return true;
}

// Keep the part of `in` on the inner side (orient * s >= 0) of edge a->b
Poly half_plane(const Poly& in, P2 a, P2 b, double s) {
  Poly out;
  out.reserve(in.size() + 2);
  for (size_t i = 0, n = in.size(); i < n; ++i) {
    const P2 p = in[i], q = in[(i + 1) % n];
    const double dp = s * orient(a, b, p), dq = s * orient(a, b, q);
    if (dp >= 0) out.push_back(p);
    if ((dp >= 0) != (dq >= 0)) {
      const double t = dp / (dp - dq);
      out.push_back({p.x + t * (q.x - p.x), p.y + t * (q.y - p.y)});
    }
  }
  return out;
}

// Convex clip polygon, counter-clockwise
Poly clip_convex(Poly subject, const Poly& ccw) {
  for (size_t k = 0, n = ccw.size(); k < n && subject.size() >= 3; ++k)
    subject = half_plane(subject, ccw[k], ccw[(k + 1) % n], 1.0);
  return subject;
}

// subject minus a convex CCW polygon, as convex pieces
void subtract_convex(const Poly& subject, const Poly& ccw, std::vector<Poly>& out) {
  Poly rest = subject;
  for (size_t k = 0, n = ccw.size(); k < n && rest.size() >= 3; ++k) {
    Poly outside = half_plane(rest, ccw[k], ccw[(k + 1) % n], -1.0);
    if (outside.size() >= 3 && std::fabs(area2(outside)) > 1e-14) out.push_back(outside);
    rest = half_plane(rest, ccw[k], ccw[(k + 1) % n], 1.0);
  }
}

bool point_in_tri(P2 p, P2 a, P2 b, P2 c) {
  const double d1 = orient(a, b, p), d2 = orient(b, c, p), d3 = orient(c, a, p);
  return d1 >= 0 && d2 >= 0 && d3 >= 0;
}

// Ear clipping of a simple CCW polygon into CCW triangles
std::vector<Poly> ear_clip(Poly p) {
  std::vector<Poly> tris;
  size_t guard = 0;
  while (p.size() > 3 && guard++ < 100000) {
    bool cut = false;
    for (size_t i = 0, n = p.size(); i < n; ++i) {
      const P2 a = p[(i + n - 1) % n], b = p[i], c = p[(i + 1) % n];
      if (orient(a, b, c) <= 0) continue;
      bool empty = true;
      for (size_t j = 0; j < n && empty; ++j) {
        if (j == i || j == (i + 1) % n || j == (i + n - 1) % n) continue;
        if (point_in_tri(p[j], a, b, c)) empty = false;
      }
      if (!empty) continue;
      tris.push_back({a, b, c});
      p.erase(p.begin() + (long)i);
      cut = true;
      break;
    }
    if (!cut) break;  // degenerate input
  }
  if (p.size() == 3 && orient(p[0], p[1], p[2]) > 0) tris.push_back(p);
  return tris;
}

// Triangle soup carving by one polygon
void carve(std::vector<Poly>& soup, const ssol_carving& c) {
  Poly poly(c.nb_vertices);
  for (size_t i = 0; i < c.nb_vertices; ++i) {
    double xy[2];
    c.get(i, xy, c.context);
    poly[i] = {xy[0], xy[1]};
  }
  if (poly.size() < 3) {
    if (c.operation == SSOL_AND) soup.clear();
    return;
  }
  if (area2(poly) < 0) std::reverse(poly.begin(), poly.end());
  const bool is_convex = convex(poly);
  std::vector<Poly> pieces = is_convex ? std::vector<Poly>{poly} : ear_clip(poly);
  std::vector<Poly> out;
  out.reserve(soup.size());
  for (const Poly& t : soup) {
    if (c.operation == SSOL_AND) {
      for (const Poly& pc : pieces) {
        Poly r = clip_convex(t, pc);
        if (r.size() >= 3 && std::fabs(area2(r)) > 1e-14) out.push_back(r);
      }
    } else {
      std::vector<Poly> cur{t};
      for (const Poly& pc : pieces) {
        std::vector<Poly> next;
        for (const Poly& q : cur) subtract_convex(q, pc, next);
        cur.swap(next);
      }
      for (Poly& q : cur) out.push_back(std::move(q));
    }
  }
  soup.swap(out);
}

double quadric_z(int q, const double* p, double x, double y) {
  switch (q) {
    case SSOL_QUADRIC_PLANE: return 0;
    case SSOL_QUADRIC_PARABOL: return (x * x + y * y) * p[1];
    case SSOL_QUADRIC_PARABOL2F: return x * x * p[0] + y * y * p[1];
    case SSOL_QUADRIC_PARABOLIC_CYLINDER: return y * y * p[1];
    case SSOL_QUADRIC_HYPERBOL: {
      const double z0 = p[0] + p[3];
      return p[3] * std::sqrt(1 + (x * x + y * y) * p[2]) + p[0] - z0;
    }
    case SSOL_QUADRIC_HEMISPHERE: {
      const double z2 = p[1] - (x * x + y * y);
      return z2 > 0 ? -std::sqrt(z2) + p[0] : p[0];
    }
    default: return 0;
  }
}

}  // namespace

namespace ssol {

res_T punched_setup(ssol_shape* shape, const ssol_punched_surface* ps) {
  if (!ps || !ps->quadric) return RES_BAD_ARG;
  const ssol_quadric& q = *ps->quadric;
  if (ps->nb_carvings == 0 && q.type != SSOL_QUADRIC_HEMISPHERE) return RES_BAD_ARG;
  if (ps->nb_carvings && !ps->carvings) return RES_BAD_ARG;
  for (size_t i = 0; i < ps->nb_carvings; ++i) {
    const ssol_carving& c = ps->carvings[i];
    if (!c.get || !c.nb_vertices || (unsigned)c.operation >= SSOL_CLIPPING_OPS_COUNT__)
      return RES_BAD_ARG;
  }
  // quadric parameters (src/ssol_shape.c:65-118, 1260-1350)
  double p[4] = {0, 0, 0, 0};
  switch (q.type) {
    case SSOL_QUADRIC_PLANE: break;
    case SSOL_QUADRIC_PARABOL:
      if (!(q.data.parabol.focal > 0)) return RES_BAD_ARG;
      p[0] = q.data.parabol.focal;
      p[1] = 1 / (4.0 * q.data.parabol.focal);
      break;
    case SSOL_QUADRIC_PARABOL2F:
      if (!(q.data.parabol2f.focal_x > 0 && q.data.parabol2f.focal_y > 0)) return RES_BAD_ARG;
      p[0] = 1 / (4.0 * q.data.parabol2f.focal_x);
      p[1] = 1 / (4.0 * q.data.parabol2f.focal_y);
      break;
    case SSOL_QUADRIC_HYPERBOL: {
      const double img = q.data.hyperbol.img_focal, real = q.data.hyperbol.real_focal;
      if (!(img > 0 && real > 0)) return RES_BAD_ARG;
      const double g = real + img, f = real / g, a2 = g * g * (f - f * f);
      const double abs_b = g * std::fabs(f - 0.5);
      p[0] = g * 0.5;
      p[1] = a2 / (abs_b * abs_b);
      p[2] = 1 / a2;
      p[3] = abs_b;
      break;
    }
    case SSOL_QUADRIC_PARABOLIC_CYLINDER:
      if (!(q.data.parabolic_cylinder.focal > 0)) return RES_BAD_ARG;
      p[0] = q.data.parabolic_cylinder.focal;
      p[1] = 1 / (4.0 * q.data.parabolic_cylinder.focal);
      break;
    case SSOL_QUADRIC_HEMISPHERE:
      if (!(q.data.hemisphere.radius > 0)) return RES_BAD_ARG;
      p[0] = q.data.hemisphere.radius;
      p[1] = p[0] * p[0];
      break;
    default: return RES_BAD_ARG;
  }

  std::vector<Poly> soup;
  double lower[2], upper[2];
  size_t nslices = 0;
  if (q.type == SSOL_QUADRIC_HEMISPHERE) {
    // disc of the carving radius (AND polygons), at most the sphere radius
    double r2 = -DBL_MAX;
    for (size_t i = 0; i < ps->nb_carvings; ++i) {
      const ssol_carving& c = ps->carvings[i];
      if (c.operation == SSOL_SUB) continue;
      for (size_t v = 0; v < c.nb_vertices; ++v) {
        double xy[2];
        c.get(v, xy, c.context);
        r2 = std::max(r2, xy[0] * xy[0] + xy[1] * xy[1]);
      }
    }
    double radius = (ps->nb_carvings && r2 >= 0) ? std::sqrt(r2) : DBL_MAX;
    radius = std::min(radius, p[0]);
    lower[0] = lower[1] = -radius;
    upper[0] = upper[1] = radius;
    nslices = q.slices_count_hint != SIZE_MAX ? q.slices_count_hint
                                             : (size_t)(5.5 + 11 * radius / p[0]);
    if (!nslices) return RES_BAD_ARG;
    // Triangulated disk: the xy projection of a hemisphere mesh of the
    // carving radius with nslices sectors around the pole and nslices
    // stacks from the pole to the equator (Star-3DUT's hemisphere, which
    // the reference projects, src/ssol_shape.c:437-489): ring k at radius
    // r sin(k pi / (2 nslices)), so the outer ring is the regular
    // nslices-gon inscribed in the carving circle
    const size_t n = nslices;
    auto ring_pt = [&](size_t k, size_t i) {
      const double rr = radius * std::sin((double)k * M_PI / (2.0 * (double)n));
      const double a = 2 * M_PI * (double)(i % n) / (double)n;
      return P2{rr * std::cos(a), rr * std::sin(a)};
    };
    for (size_t k = 0; k < n; ++k) {
      for (size_t i = 0; i < n; ++i) {
        const P2 p00 = ring_pt(k, i), p01 = ring_pt(k, i + 1);
        const P2 p10 = ring_pt(k + 1, i), p11 = ring_pt(k + 1, i + 1);
        // clockwise seen from +Z, like the plane grid
        if (k == 0) {
          soup.push_back({p00, p11, p10});
        } else {
          soup.push_back({p00, p11, p10});
          soup.push_back({p00, p01, p11});
        }
      }
    }
  } else {
    lower[0] = lower[1] = DBL_MAX;
    upper[0] = upper[1] = -DBL_MAX;
    for (size_t i = 0; i < ps->nb_carvings; ++i) {
      const ssol_carving& c = ps->carvings[i];
      if (c.operation == SSOL_SUB) continue;
      for (size_t v = 0; v < c.nb_vertices; ++v) {
        double xy[2];
        c.get(v, xy, c.context);
        lower[0] = std::min(lower[0], xy[0]);
        lower[1] = std::min(lower[1], xy[1]);
        upper[0] = std::max(upper[0], xy[0]);
        upper[1] = std::max(upper[1], xy[1]);
      }
    }
    if (lower[0] >= upper[0] || lower[1] >= upper[1]) {
      log_error(shape->dev, "ssol_punched_surface_setup: infinite or null punched surface.\n");
      return RES_BAD_ARG;
    }
    if (q.slices_count_hint != SIZE_MAX) {
      nslices = q.slices_count_hint;
    } else if (q.type == SSOL_QUADRIC_PLANE) {
      nslices = 1;
    } else {
      const double max_z = std::max(quadric_z(q.type, p, lower[0], lower[1]),
                                    quadric_z(q.type, p, upper[0], upper[1]));
      nslices = std::min<size_t>(50, (size_t)(3 + std::sqrt(max_z) * 6));
    }
    if (!nslices) return RES_BAD_ARG;
    // regular grid of square cells (src/ssol_shape.c:493-577)
    const double sx = upper[0] - lower[0], sy = upper[1] - lower[1];
    const double smin = std::min(sx, sy);
    if (std::fabs(smin) <= 1e-6) return RES_BAD_ARG;
    const double delta = smin / (double)nslices;
    const size_t nx = (size_t)std::ceil(sx / delta), ny = (size_t)std::ceil(sy / delta);
    auto X = [&](size_t i) { return std::min(lower[0] + (double)i * delta, upper[0]); };
    auto Y = [&](size_t j) { return std::min(lower[1] + (double)j * delta, upper[1]); };
    soup.reserve(nx * ny * 2);
    for (size_t ix = 0; ix < nx; ++ix) {
      for (size_t iy = 0; iy < ny; ++iy) {
        const P2 v0{X(ix), Y(iy)}, v1{X(ix + 1), Y(iy)}, v2{X(ix), Y(iy + 1)}, v3{X(ix + 1), Y(iy + 1)};
        soup.push_back({v0, v3, v1});
        soup.push_back({v0, v2, v3});
      }
    }
  }
  for (size_t i = 0; i < ps->nb_carvings; ++i) {
    // carving polygons keep the orientation of the grid triangles: clip
    // preserves each subject polygon's winding
    carve(soup, ps->carvings[i]);
  }

  // triangulate the carved polygons (fan) and lift them
  std::vector<float> pos, samp;
  std::vector<unsigned> ids;
  const double* T = q.transform;
  auto push = [&](std::vector<float>& dst, double x, double y, double z) {
    for (int k = 0; k < 3; ++k)
      dst.push_back((float)(T[k] * x + T[3 + k] * y + T[6 + k] * z + T[9 + k]));
  };
  auto add_vertex = [&](P2 v) {
    if (q.type == SSOL_QUADRIC_PLANE) {
      push(pos, (float)v.x, (float)v.y, 0);
    } else {
      push(pos, v.x, v.y, quadric_z(q.type, p, v.x, v.y));
    }
    push(samp, (float)v.x, (float)v.y, 0);
  };
  for (const Poly& poly : soup) {
    for (size_t k = 1; k + 1 < poly.size(); ++k) {
      const unsigned base = (unsigned)(pos.size() / 3);
      add_vertex(poly[0]);
      add_vertex(poly[k]);
      add_vertex(poly[k + 1]);
      ids.push_back(base);
      ids.push_back(base + 1);
      ids.push_back(base + 2);
    }
  }
  double rt = 0, sa = 0;
  for (size_t t = 0; t < ids.size() / 3; ++t) {
    auto area = [&](const std::vector<float>& v) {
      const float* a = &v[3 * ids[3 * t]];
      const float* b = &v[3 * ids[3 * t + 1]];
      const float* c = &v[3 * ids[3 * t + 2]];
      const double e0[3] = {(double)b[0] - a[0], (double)b[1] - a[1], (double)b[2] - a[2]};
      const double e1[3] = {(double)c[0] - a[0], (double)c[1] - a[1], (double)c[2] - a[2]};
      const double n[3] = {e0[1] * e1[2] - e0[2] * e1[1], e0[2] * e1[0] - e0[0] * e1[2],
                           e0[0] * e1[1] - e0[1] * e1[0]};
      return std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    };
    rt += area(pos);
    sa += area(samp);
  }
  shape->pos.swap(pos);
  shape->samp_pos.swap(samp);
  shape->ids.swap(ids);
  shape->normals.clear();
  shape->uvs.clear();
  shape->rt_area = rt * 0.5;
  shape->samp_area = sa * 0.5;
  shape->quadric = q.type;
  memcpy(shape->qparam, p, sizeof(p));
  memcpy(shape->qtransform, q.transform, sizeof(shape->qtransform));
  shape->ready = true;
  return RES_OK;
}

}  // namespace ssol
