// ssol_solve and the estimator getters.
//
// ssol_solve (reference src/ssol_solver.c:613-842) checks its arguments and
// the scene (src/ssol_scene.c:538-558), flattens the scene into the plain
// arrays of ssolh.h (shader baking, media deduplication, per-object BLAS,
// instance TLAS, area-weighted alias table over the sampling view, tally
// layout), and runs the realisations on the GPU through ssolh_*. There is no
// CPU path. The estimator getters restate src/ssol_estimator.c:112-268 and
// src/ssol_mc_receiver.c:32-173 over the downloaded tally arrays.
#include <chrono>
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <unordered_map>

#include "ssol_internal.hpp"

using namespace ssol;

namespace {

// A TLAS refit for moved instances is kept while its summed child-box
// surface area stays within this factor of the tree's as built
constexpr double REFIT_LOOSENESS = 1.5;

// deepest TLAS + BLAS descent accepted: the GPU traversal keeps its first
// stack entries in LDS and the rest, for deep trees, in device memory
constexpr int MAX_BVH_LEVELS = 200;

struct DataKey {
  int spectrum;
  double real;
  uint64_t c0, c1;
};

// Flat scene under construction; owns every array handed to ssolh.
struct Flat {
  std::vector<ssolh_spectrum> spectra;
  std::vector<double> wl, val, cdf;
  std::map<const ssol_spectrum*, int> spec_index;
  std::vector<ssolh_medium> media;
  std::vector<std::pair<DataKey, DataKey>> media_keys;
  std::vector<ssolh_material> materials;
  std::map<const ssol_material*, uint32_t> mtl_index;
  std::vector<ssolh_shape> shapes;
  std::map<const ssol_shape*, uint32_t> shape_index;
  std::vector<ssolh_sshape> sshapes;
  std::vector<ssolh_object> objects;
  std::map<const ssol_object*, uint32_t> obj_index;
  std::vector<unsigned> obj_depth;
  std::vector<aabb> obj_box;  // object-space bounds of each BLAS
  std::vector<ssolh_bvh_node> nodes;
  std::vector<ssolh_triangle> tris;
  std::vector<ssolh_instance> instances;
  std::vector<uint32_t> tlas_inst;
  std::vector<ssolh_inst_sshape> isshapes;
  std::vector<ssolh_sample_tri> stris;
  std::vector<ssolh_alias> alias;
  size_t blas_nodes = 0;      // nodes of the BLASes; the TLAS follows them
  ssolh_scene scn{};
};

int spectrum_id(Flat& f, const ssol_spectrum* s) {
  auto it = f.spec_index.find(s);
  if (it != f.spec_index.end()) return it->second;
  const int id = (int)f.spectra.size();
  ssolh_spectrum d;
  d.offset = (uint32_t)f.wl.size();
  d.count = (uint32_t)s->wavelengths.size();
  f.spectra.push_back(d);
  double acc = 0;
  for (size_t i = 0; i < s->wavelengths.size(); ++i) {
    if (i) acc += 0.5 * (s->intensities[i - 1] + s->intensities[i]) *
                  (s->wavelengths[i] - s->wavelengths[i - 1]);
    f.wl.push_back(s->wavelengths[i]);
    f.val.push_back(s->intensities[i]);
    f.cdf.push_back(acc);
  }
  f.spec_index[s] = id;
  return id;
}

ssolh_data flat_data(Flat& f, const ssol_data& d, DataKey& key) {
  ssolh_data o{};
  if (d.type == SSOL_DATA_REAL) {
    o.real = d.value.real;
    o.spectrum = -1;
    key = DataKey{-1, d.value.real, 0, 0};
  } else {
    o.real = 0;
    o.spectrum = spectrum_id(f, d.value.spectrum);
    key = DataKey{1, 0, d.value.spectrum->checksum[0], d.value.spectrum->checksum[1]};
  }
  return o;
}

bool key_eq(const DataKey& a, const DataKey& b) {
  if (a.spectrum != b.spectrum) return false;
  if (a.spectrum < 0) return a.real == b.real;
  return a.c0 == b.c0 && a.c1 == b.c1;
}

// Media are identified with the reference's media_ceq rule
uint32_t medium_id(Flat& f, const ssol_medium& m) {
  DataKey ke, kr;
  ssolh_medium md;
  md.extinction = flat_data(f, m.extinction, ke);
  md.refractive_index = flat_data(f, m.refractive_index, kr);
  for (size_t i = 0; i < f.media.size(); ++i)
    if (key_eq(f.media_keys[i].first, ke) && key_eq(f.media_keys[i].second, kr)) return (uint32_t)i;
  f.media.push_back(md);
  f.media_keys.push_back({ke, kr});
  return (uint32_t)(f.media.size() - 1);
}

// The k-th of the three surface fragments shaders are baked on
ssol_surface_fragment probe_fragment(int k) {
  ssol_surface_fragment fr = SSOL_SURFACE_FRAGMENT_NULL;
  const double a = 0.3 + 0.7 * k;
  fr.dir[0] = 0.1 * k; fr.dir[1] = -0.2; fr.dir[2] = -1;
  fr.P[0] = 1.5 * k; fr.P[1] = -2.0 * k; fr.P[2] = 0.25 * k;
  fr.Ng[0] = 0; fr.Ng[1] = std::sin(a); fr.Ng[2] = std::cos(a);
  fr.Ns[0] = std::sin(a); fr.Ns[1] = 0; fr.Ns[2] = std::cos(a);
  fr.uv[0] = 0.25 * k; fr.uv[1] = 1 - 0.25 * k;
  fr.dPdu[0] = 1; fr.dPdv[1] = 1;
  return fr;
}

// The wavelengths shaders are baked at: the sun spectrum's, or its first,
// middle and last when it has more than three
std::vector<double> probe_wavelengths(const ssol_spectrum* sp) {
  std::vector<double> wls = sp->wavelengths;
  if (wls.size() > 3) wls = {wls.front(), wls[wls.size() / 2], wls.back()};
  return wls;
}

// Evaluate a shader on probe fragments; require a constant value.
bool bake_scalar(ssol_device* dev, const ssol_material* m, ssol_shader_getter_T fn,
                 const std::vector<double>& wls, double* out, const char* what) {
  double first = 0;
  bool have = false;
  for (int k = 0; k < 3; ++k) {
    const ssol_surface_fragment fr = probe_fragment(k);
    for (double wl : wls) {
      double v = 0;
      fn(dev, m->buf, wl, &fr, &v);
      if (!have) { first = v; have = true; }
      else if (v != first) {
        log_error(dev, "ssol_solve: the %s shader of a material varies with the wavelength or "
                       "the surface fragment; only constant shaders run on the GPU.\n", what);
        return false;
      }
    }
  }
  *out = first;
  return true;
}

bool bake_normal(ssol_device* dev, const ssol_material* m, int* from_ng) {
  ssol_surface_fragment fr = SSOL_SURFACE_FRAGMENT_NULL;
  fr.dir[2] = -1;
  fr.Ng[0] = 0; fr.Ng[1] = 0.6; fr.Ng[2] = 0.8;
  fr.Ns[0] = 0.8; fr.Ns[1] = 0; fr.Ns[2] = 0.6;
  double v[3] = {0, 0, 0};
  m->normal(dev, m->buf, 500, &fr, v);
  if (v[0] == fr.Ns[0] && v[1] == fr.Ns[1] && v[2] == fr.Ns[2]) { *from_ng = 0; return true; }
  if (v[0] == fr.Ng[0] && v[1] == fr.Ng[1] && v[2] == fr.Ng[2]) { *from_ng = 1; return true; }
  log_error(dev, "ssol_solve: a normal shader returns neither the shading nor the geometry "
                 "normal; it cannot be baked for the GPU.\n");
  return false;
}

res_T material_id(Flat& f, ssol_device* dev, const ssol_material* m,
                  const std::vector<double>& wls, uint32_t* out) {
  auto it = f.mtl_index.find(m);
  if (it != f.mtl_index.end()) { *out = it->second; return RES_OK; }
  ssolh_material d{};
  d.type = (int32_t)m->type;
  d.medium_in = d.medium_out = -1;
  d.slab_eta.spectrum = d.slab_absorption.spectrum = -1;
  switch (m->type) {
    case SSOL_MATERIAL_MIRROR:
      if (!bake_normal(dev, m, &d.normal_from_ng)) return RES_BAD_ARG;
      if (!bake_scalar(dev, m, m->reflectivity, wls, &d.reflectivity, "reflectivity")) return RES_BAD_ARG;
      if (!bake_scalar(dev, m, m->roughness, wls, &d.roughness, "roughness")) return RES_BAD_ARG;
      d.distribution = (int32_t)m->distrib;
      break;
    case SSOL_MATERIAL_MATTE:
      if (!bake_normal(dev, m, &d.normal_from_ng)) return RES_BAD_ARG;
      if (!bake_scalar(dev, m, m->reflectivity, wls, &d.reflectivity, "reflectivity")) return RES_BAD_ARG;
      break;
    case SSOL_MATERIAL_DIELECTRIC:
      if (!bake_normal(dev, m, &d.normal_from_ng)) return RES_BAD_ARG;
      d.medium_in = (int32_t)medium_id(f, m->in_medium);
      d.medium_out = (int32_t)medium_id(f, m->out_medium);
      break;
    case SSOL_MATERIAL_THIN_DIELECTRIC: {
      if (!bake_normal(dev, m, &d.normal_from_ng)) return RES_BAD_ARG;
      d.medium_in = (int32_t)medium_id(f, m->in_medium);
      d.medium_out = (int32_t)medium_id(f, m->out_medium);
      DataKey k;
      d.slab_eta = flat_data(f, m->slab_medium.refractive_index, k);
      d.slab_absorption = flat_data(f, m->slab_medium.extinction, k);
      d.thickness = m->thickness;
      break;
    }
    case SSOL_MATERIAL_VIRTUAL: break;
    default: return RES_BAD_ARG;
  }
  const uint32_t id = (uint32_t)f.materials.size();
  f.materials.push_back(d);
  f.mtl_index[m] = id;
  *out = id;
  return RES_OK;
}

void mat_inv(const double* m, double* inv) {
  double t[9];
  t[0] = m[4] * m[8] - m[7] * m[5];
  t[1] = m[7] * m[2] - m[1] * m[8];
  t[2] = m[1] * m[5] - m[4] * m[2];
  t[3] = m[6] * m[5] - m[3] * m[8];
  t[4] = m[0] * m[8] - m[6] * m[2];
  t[5] = m[3] * m[2] - m[0] * m[5];
  t[6] = m[3] * m[7] - m[6] * m[4];
  t[7] = m[6] * m[1] - m[0] * m[7];
  t[8] = m[0] * m[4] - m[3] * m[1];
  const double det = m[0] * t[0] + m[3] * t[1] + m[6] * t[2];
  for (int i = 0; i < 9; ++i) inv[i] = t[i] / det;
}

void mat_mul(const double* a, const double* b, double* out) {
  double t[9];
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r)
      t[c * 3 + r] = a[r] * b[c * 3] + a[3 + r] * b[c * 3 + 1] + a[6 + r] * b[c * 3 + 2];
  memcpy(out, t, sizeof(t));
}

void inflate(aabb& b, float rel, float abs_) {
  float m = 0;
  for (int k = 0; k < 3; ++k) m = std::max(m, std::max(std::fabs(b.lo[k]), std::fabs(b.hi[k])));
  const float e = m * rel + abs_;
  for (int k = 0; k < 3; ++k) { b.lo[k] -= e; b.hi[k] += e; }
}

// Whether every shape of `o` is a mesh and all its vertices lie in one plane
// (to 1e-12 of the object's size, i.e. exactly for float data such as a
// facet grid in z = 0).
bool object_is_planar(const ssol_object* o) {
  double p0[3] = {0, 0, 0}, n[3] = {0, 0, 0}, best = 0, lo[3], hi[3];
  bool first = true;
  for (int k = 0; k < 3; ++k) { lo[k] = DBL_MAX; hi[k] = -DBL_MAX; }
  for (const auto& ss : o->sshapes) {
    if (ss.shape->kind != shape_kind::mesh) return false;
    const std::vector<float>& P = ss.shape->pos;
    const std::vector<unsigned>& I = ss.shape->ids;
    for (size_t v = 0; v + 2 < P.size(); v += 3) {
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(lo[k], (double)P[v + k]);
        hi[k] = std::max(hi[k], (double)P[v + k]);
      }
      if (first) { for (int k = 0; k < 3; ++k) p0[k] = P[v + k]; first = false; }
    }
    for (size_t t = 0; t + 2 < I.size(); t += 3) {  // normal of the largest triangle
      const float* a = &P[3 * I[t]];
      const float* b = &P[3 * I[t + 1]];
      const float* c = &P[3 * I[t + 2]];
      const double e1[3] = {(double)b[0] - a[0], (double)b[1] - a[1], (double)b[2] - a[2]};
      const double e2[3] = {(double)c[0] - a[0], (double)c[1] - a[1], (double)c[2] - a[2]};
      const double m[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                           e1[0] * e2[1] - e1[1] * e2[0]};
      const double a2 = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
      if (a2 > best) { best = a2; for (int k = 0; k < 3; ++k) n[k] = m[k]; }
    }
  }
  if (first || !(best > 0)) return false;
  const double len = std::sqrt(best);
  double ext = 0;
  for (int k = 0; k < 3; ++k) { n[k] /= len; ext = std::max(ext, hi[k] - lo[k]); }
  for (const auto& ss : o->sshapes) {
    const std::vector<float>& P = ss.shape->pos;
    for (size_t v = 0; v + 2 < P.size(); v += 3) {
      const double dist = n[0] * (P[v] - p0[0]) + n[1] * (P[v + 1] - p0[1]) + n[2] * (P[v + 2] - p0[2]);
      if (std::fabs(dist) > 1e-12 * ext) return false;
    }
  }
  return true;
}

res_T object_id(Flat& f, const ssol_object* o, uint32_t* out) {
  auto it = f.obj_index.find(o);
  if (it != f.obj_index.end()) { *out = it->second; return RES_OK; }
  ssolh_object d{};
  d.sshape_first = (uint32_t)f.sshapes.size();
  d.sshape_count = (uint32_t)o->sshapes.size();
  std::vector<aabb> boxes;
  std::vector<ssolh_triangle> tris;
  aabb all;
  for (int k = 0; k < 3; ++k) { all.lo[k] = FLT_MAX; all.hi[k] = -FLT_MAX; }
  for (size_t s = 0; s < o->sshapes.size(); ++s) {
    const ssol_shape* sh = o->sshapes[s].shape;
    for (size_t t = 0; t < sh->ids.size() / 3; ++t) {
      ssolh_triangle tr{};
      aabb b;
      for (int k = 0; k < 3; ++k) { b.lo[k] = FLT_MAX; b.hi[k] = -FLT_MAX; }
      for (int v = 0; v < 3; ++v) {
        const float* p = &sh->pos[3 * sh->ids[3 * t + v]];
        float* dst = v == 0 ? tr.v0 : (v == 1 ? tr.v1 : tr.v2);
        for (int k = 0; k < 3; ++k) {
          dst[k] = p[k];
          b.lo[k] = std::min(b.lo[k], p[k]);
          b.hi[k] = std::max(b.hi[k], p[k]);
        }
      }
      inflate(b, 1e-6f, 1e-7f);
      for (int k = 0; k < 3; ++k) { all.lo[k] = std::min(all.lo[k], b.lo[k]); all.hi[k] = std::max(all.hi[k], b.hi[k]); }
      tr.sshape = (uint32_t)s;
      tr.prim = (uint32_t)t;
      tris.push_back(tr);
      boxes.push_back(b);
    }
  }
  if (object_is_planar(o)) d.flags |= SSOLH_OBJECT_PLANAR;
  std::vector<uint32_t> order;
  const size_t node0 = f.nodes.size();
  unsigned depth = 0;
  // leaves of at most 2 triangles: a leaf's triangles are FP64 watertight
  // tests, dearer than the FP32 node visits a deeper tree adds (10k field:
  // k_trace 35.5 -> 32.4 ms per step against leaves of 4, 33.4 with 1)
  d.blas_root = bvh_build(boxes, 2, f.nodes, order, &depth);
  const uint32_t base = (uint32_t)f.tris.size();
  for (size_t n = node0; n < f.nodes.size(); ++n) {
    for (int c = 0; c < 2; ++c) {
      if (f.nodes[n].count[c] != 0 && f.nodes[n].count[c] != SSOLH_BVH_EMPTY)
        f.nodes[n].child[c] += base;
    }
  }
  for (uint32_t i : order) f.tris.push_back(tris[i]);
  if (f.tris.size() >= (1u << 28)) return RES_MEM_ERR;
  *out = (uint32_t)f.objects.size();
  f.objects.push_back(d);
  f.obj_depth.push_back(depth);
  f.obj_box.push_back(all);
  f.obj_index[o] = *out;
  for (const shaded_shape& ss : o->sshapes) {
    ssolh_sshape e{};
    auto si = f.shape_index.find(ss.shape);
    if (si == f.shape_index.end()) {
      ssolh_shape s{};
      s.type = ss.shape->kind == shape_kind::mesh ? SSOLH_SHAPE_MESH : SSOLH_SHAPE_PUNCHED;
      s.quadric = ss.shape->quadric;
      s.ntris = (uint32_t)(ss.shape->ids.size() / 3);
      memcpy(s.qparam, ss.shape->qparam, sizeof(s.qparam));
      memcpy(s.qtransform, ss.shape->qtransform, sizeof(s.qtransform));
      f.shape_index[ss.shape] = (uint32_t)f.shapes.size();
      f.shapes.push_back(s);
      si = f.shape_index.find(ss.shape);
    }
    e.shape = si->second;
    f.sshapes.push_back(e);  // materials filled by the caller
  }
  return RES_OK;
}

// Vose alias table over non-negative weights
void build_alias(const std::vector<double>& w, std::vector<ssolh_alias>& out) {
  const size_t n = w.size();
  out.assign(n, ssolh_alias{1.0, 0, 0});
  double total = 0;
  for (double x : w) total += x;
  std::vector<double> p(n);
  std::vector<uint32_t> small, large;
  for (size_t i = 0; i < n; ++i) {
    p[i] = total > 0 ? w[i] * (double)n / total : 1.0;
    (p[i] < 1 ? small : large).push_back((uint32_t)i);
  }
  while (!small.empty() && !large.empty()) {
    const uint32_t s = small.back(); small.pop_back();
    const uint32_t l = large.back();
    out[s].prob = p[s];
    out[s].alias = l;
    p[l] = (p[l] + p[s]) - 1.0;
    if (p[l] < 1) { large.pop_back(); small.push_back(l); }
  }
  for (uint32_t l : large) { out[l].prob = 1.0; out[l].alias = l; }
  for (uint32_t s : small) { out[s].prob = 1.0; out[s].alias = s; }
}

// Buie sunshape state (src/ssol_ranst_sun_dir.c:88-222)
void buie_params(double csr, double* p) {
  double chi;
  if (csr > 0.145)
    chi = -0.04419909985804843 + csr * (1.401323894233574 + csr * (-0.3639746714505299 +
          csr * (-0.9579768560161194 + 1.1550475450828657 * csr)));
  else if (csr > 0.035)
    chi = 0.022652077593662934 + csr * (0.5252380349996234 + csr * (2.5484334534423887 -
          0.8763755326550412 * csr));
  else
    chi = 0.004733749294807862 + csr * (4.716738065192151 + csr * (-463.506669149804 +
          csr * (24745.88727411664 + csr * (-606122.7511711778 + 5521693.445014727 * csr))));
  const double thetaSD = 0.00465, thetaCS = 0.0436, integralA = 9.224724736098827E-6;
  const double k = 0.9 * std::log(13.5 * chi) * std::pow(chi, -0.3);
  const double gamma = 2.2 * std::log(0.52 * chi) * std::pow(chi, 0.43) - 0.1;
  const double g2 = gamma + 2.0;
  const double ek = std::exp(k) * std::pow(1000, gamma);
  const double integralB = ek / g2 * (std::pow(thetaCS, g2) - std::pow(thetaSD, g2));
  const double alpha = 1 / (integralA + integralB);
  auto pdf = [&](double th) {
    const double phi = th < thetaSD ? std::cos(326 * th) / std::cos(308 * th) : ek * std::pow(th, gamma);
    return alpha * phi * std::sin(th);
  };
  const double h1 = 1.001 * pdf(0.0038915695846209047);
  const double h2 = pdf(thetaSD);
  const double a1 = thetaSD * h1, a2 = (thetaCS - thetaSD) * h2;
  p[0] = thetaSD; p[1] = thetaCS - thetaSD; p[2] = gamma; p[3] = ek;
  p[4] = alpha; p[5] = h1; p[6] = h2; p[7] = a1 / (a1 + a2);
}

void frame(const double* d, double* m) {
  const double ax = std::fabs(d[0]), ay = std::fabs(d[1]), az = std::fabs(d[2]);
  double h[3] = {0, 0, 0};
  if (ax <= ay && ax <= az) h[0] = 1; else if (ay <= az) h[1] = 1; else h[2] = 1;
  double x[3] = {h[1] * d[2] - h[2] * d[1], h[2] * d[0] - h[0] * d[2], h[0] * d[1] - h[1] * d[0]};
  const double n = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  for (double& v : x) v /= n;
  const double y[3] = {d[1] * x[2] - d[2] * x[1], d[2] * x[0] - d[0] * x[2], d[0] * x[1] - d[1] * x[0]};
  for (int k = 0; k < 3; ++k) { m[k] = x[k]; m[3 + k] = y[k]; m[6 + k] = d[k]; }
}

struct Layout {
  std::vector<int> rcv_slot, samp_slot;
  std::map<std::pair<uint32_t, uint32_t>, size_t> prim_first;
  size_t nrcv = 0, nsamp = 0, nprim = 0;
};

void push_data_sig(std::vector<uint64_t>& sig, const ssol_data& d) {
  sig.push_back((uint64_t)d.type);
  if (d.type == SSOL_DATA_REAL) {
    uint64_t b;
    memcpy(&b, &d.value.real, 8);
    sig.push_back(b);
  } else {
    sig.push_back((uint64_t)(uintptr_t)d.value.spectrum);
    sig.push_back(d.value.spectrum->gen);
  }
}

void push_medium_sig(std::vector<uint64_t>& sig, const ssol_medium& m) {
  push_data_sig(sig, m.extinction);
  push_data_sig(sig, m.refractive_index);
}

// What a material's shaders return on the probes they are baked on: the
// shaders read their parameters from the material's parameter buffer (or
// anywhere else), which a client may rewrite between solves without any
// setter call, so the baked constants are checked against a fresh
// evaluation on every solve
void push_shader_sig(std::vector<uint64_t>& sig, const ssol_material* m,
                     const std::vector<double>& wls) {
  auto push = [&](double v) {
    uint64_t b;
    memcpy(&b, &v, 8);
    sig.push_back(b);
  };
  for (int k = 0; k < 3; ++k) {
    const ssol_surface_fragment fr = probe_fragment(k);
    if (m->normal) {
      double v[3] = {0, 0, 0};
      m->normal(m->dev, m->buf, wls.empty() ? 500 : wls.front(), &fr, v);
      push(v[0]); push(v[1]); push(v[2]);
    }
    for (ssol_shader_getter_T fn : {m->reflectivity, m->roughness}) {
      if (!fn) continue;
      for (double wl : wls) {
        double v = 0;
        fn(m->dev, m->buf, wl, &fr, &v);
        push(v);
      }
    }
  }
}

// Everything flatten() reads, so that an unchanged scene is not flattened
// and uploaded again
// Instance transforms go to *xf, the rest to the returned signature: scenes
// that differ only in *xf share their objects, sampling and tally layout.
std::vector<uint64_t> scene_signature(const ssol_scene* scn, std::vector<uint64_t>* xf) {
  std::vector<uint64_t> sig;
  xf->clear();
  auto P = [&](const void* p) { sig.push_back((uint64_t)(uintptr_t)p); };
  P(scn);
  P(scn->sun);
  sig.push_back(scn->sun->gen);
  P(scn->sun->spectrum);
  sig.push_back(scn->sun->spectrum->gen);
  P(scn->atmosphere);
  if (scn->atmosphere) {
    sig.push_back(scn->atmosphere->gen);
    push_data_sig(sig, scn->atmosphere->extinction);
  }
  std::vector<const ssol_material*> mats;
  for (const ssol_instance* in : scn->instances) {
    P(in);
    sig.push_back(in->id);
    for (int k = 0; k < 12; ++k) {
      uint64_t b;
      memcpy(&b, &in->transform[k], 8);
      xf->push_back(b);
    }
    sig.push_back((uint64_t)in->receiver_mask);
    sig.push_back((uint64_t)in->per_primitive);
    sig.push_back((uint64_t)in->sample);
    P(in->object);
    sig.push_back(in->object->sshapes.size());
    for (const shaded_shape& ss : in->object->sshapes) {
      P(ss.shape);
      sig.push_back(ss.shape->gen);
      for (const ssol_material* m : {ss.front, ss.back}) {
        P(m);
        sig.push_back(m->gen);
        push_medium_sig(sig, m->in_medium);
        push_medium_sig(sig, m->out_medium);
        push_medium_sig(sig, m->slab_medium);
        mats.push_back(m);
      }
    }
  }
  std::sort(mats.begin(), mats.end());
  mats.erase(std::unique(mats.begin(), mats.end()), mats.end());
  if (scn->sun && scn->sun->spectrum) {
    const std::vector<double> wls = probe_wavelengths(scn->sun->spectrum);
    for (const ssol_material* m : mats)
      if (m->type != SSOL_MATERIAL_VIRTUAL) push_shader_sig(sig, m, wls);
  }
  return sig;
}

// Device form of a sun (src/ssol_sun.c:214-279). With add_spectrum false
// the sun's spectrum must already be in the flat scene's pool.
res_T flat_sun(Flat& f, const ssol_sun* sun, bool add_spectrum, ssolh_sun& su) {
  memset(&su, 0, sizeof(su));
  if (!sun->spectrum || sun->spectrum->wavelengths.empty() || !(sun->dni > 0)) return RES_BAD_ARG;
  for (int k = 0; k < 3; ++k) su.dir[k] = sun->direction[k];
  if (su.dir[0] == 0 && su.dir[1] == 0 && su.dir[2] == 0) return RES_BAD_ARG;
  frame(su.dir, su.basis);
  su.dni = sun->dni;
  if (add_spectrum) {
    su.wl_spectrum = spectrum_id(f, sun->spectrum);
  } else {
    auto it = f.spec_index.find(sun->spectrum);
    if (it == f.spec_index.end()) return RES_BAD_ARG;
    su.wl_spectrum = it->second;
  }
  switch (sun->kind) {
    case sun_kind::directional: su.type = SSOLH_SUN_DIRECTIONAL; break;
    case sun_kind::pillbox:
      if (!(sun->param > 0)) return RES_BAD_ARG;
      su.type = SSOLH_SUN_PILLBOX;
      su.param[0] = std::sin(sun->param) * std::sin(sun->param);
      break;
    case sun_kind::gaussian:
      if (!(sun->param > 0)) return RES_BAD_ARG;
      su.type = SSOLH_SUN_GAUSSIAN;
      su.param[0] = sun->param;
      break;
    case sun_kind::buie:
      if (sun->param < 1e-6 || sun->param > 0.849) return RES_BAD_ARG;
      su.type = SSOLH_SUN_BUIE;
      buie_params(sun->param, su.param);
      break;
  }
  const ssolh_spectrum& sp = f.spectra[su.wl_spectrum];
  su.wl_cdf_total = sp.count ? f.cdf[sp.offset + sp.count - 1] : 0;
  return RES_OK;
}

// World transform of an instance and, per shaded shape, the world <->
// quadric frames of punched shapes (src/ssol_shape.c:1461-1466)
void instance_frames(const ssol_instance* in, ssolh_instance& d, ssolh_inst_sshape* iss) {
  memcpy(d.transform, in->transform, sizeof(d.transform));
  mat_inv(in->transform, d.inv_linear);
  for (size_t s = 0; s < in->object->sshapes.size(); ++s) {
    const ssol_shape* sh = in->object->sshapes[s].shape;
    ssolh_inst_sshape& is = iss[s];
    if (sh->kind == shape_kind::punched) {
      mat_mul(in->transform, sh->qtransform, is.R);
      for (int k = 0; k < 3; ++k)
        is.T[k] = in->transform[k] * sh->qtransform[9] + in->transform[3 + k] * sh->qtransform[10] +
                  in->transform[6 + k] * sh->qtransform[11] + in->transform[9 + k];
      mat_inv(is.R, is.R_inv);
    } else {
      memset(is.R, 0, sizeof(is.R));
      memset(is.T, 0, sizeof(is.T));
      memset(is.R_inv, 0, sizeof(is.R_inv));
      is.R[0] = is.R[4] = is.R[8] = 1;
      is.R_inv[0] = is.R_inv[4] = is.R_inv[8] = 1;
    }
  }
}

// TLAS over the world boxes of the instances, after the BLAS nodes
res_T build_tlas(ssol_device* dev, Flat& f) {
  f.nodes.resize(f.blas_nodes);
  f.tlas_inst.clear();
  std::vector<aabb> boxes(f.instances.size());
  unsigned max_blas = 0;
  for (size_t i = 0; i < f.instances.size(); ++i) {
    const ssolh_instance& in = f.instances[i];
    const aabb& ob = f.obj_box[in.object];
    max_blas = std::max(max_blas, f.obj_depth[in.object]);
    aabb& b = boxes[i];
    for (int k = 0; k < 3; ++k) { b.lo[k] = FLT_MAX; b.hi[k] = -FLT_MAX; }
    for (int c = 0; c < 8; ++c) {
      const double p[3] = {(c & 1) ? ob.hi[0] : ob.lo[0], (c & 2) ? ob.hi[1] : ob.lo[1],
                           (c & 4) ? ob.hi[2] : ob.lo[2]};
      for (int k = 0; k < 3; ++k) {
        const double w = in.transform[k] * p[0] + in.transform[3 + k] * p[1] +
                         in.transform[6 + k] * p[2] + in.transform[9 + k];
        b.lo[k] = std::min(b.lo[k], (float)w);
        b.hi[k] = std::max(b.hi[k], (float)w);
      }
    }
    if (ob.lo[0] > ob.hi[0]) {  // empty object: box that no ray reaches
      for (int k = 0; k < 3; ++k) { b.lo[k] = 0; b.hi[k] = 0; }
    }
    inflate(b, 2e-6f, 1e-6f);
  }
  unsigned tdepth = 0;
  f.scn.tlas_root = bvh_build(boxes, 1, f.nodes, f.tlas_inst, &tdepth);
  if (tdepth + max_blas + 2 > (unsigned)MAX_BVH_LEVELS) {
    log_error(dev, "ssol_solve: BVH too deep (%u + %u levels).\n",
              tdepth, max_blas);
    return RES_MEM_ERR;
  }
  f.scn.stack_depth = tdepth + max_blas + 2;
  f.scn.nnodes = (uint32_t)f.nodes.size();
  f.scn.nodes = f.nodes.data();
  f.scn.tlas_instances = f.tlas_inst.data();
  log_warning(dev, "ssol_solve: %zu instances, %zu triangles, TLAS depth %u, BLAS depth %u.\n",
              f.instances.size(), f.tris.size(), tdepth, max_blas);
  return RES_OK;
}

// Instances moved, nothing else changed: new instance frames over the same
// objects, sampling tables and tally layout (the sampled areas are
// object-space areas, which rigid motions keep).
res_T refresh_frames(ssol_scene* scn, Flat& f) {
  if (f.instances.size() != scn->instances.size()) return RES_BAD_ARG;
  for (size_t i = 0; i < scn->instances.size(); ++i) {
    ssolh_instance& d = f.instances[i];
    instance_frames(scn->instances[i], d, f.isshapes.data() + d.isshape_first);
  }
  return RES_OK;
}

res_T flatten(ssol_scene* scn, Flat& f, Layout& L) {
  ssol_device* dev = scn->dev;
  ssol_sun* sun = scn->sun;
  // wavelengths at which shaders are probed
  const std::vector<double> wls = probe_wavelengths(sun->spectrum);

  // air (src/ssol_solver.c:660-664)
  ssol_medium air = SSOL_MEDIUM_VACUUM;
  if (scn->atmosphere) air.extinction = scn->atmosphere->extinction;
  const uint32_t air_id = medium_id(f, air);

  double S = 0, S_proxy = 0;
  bool has_sampled = false, has_receiver = false;
  std::vector<double> weights;
  for (size_t i = 0; i < scn->instances.size(); ++i) {
    ssol_instance* in = scn->instances[i];
    for (const shaded_shape& ss : in->object->sshapes)
      if (!ss.shape->ready) {
        log_error(dev, "ssol_solve: a shape of the scene is not set up.\n");
        return RES_BAD_ARG;
      }
    uint32_t oid;
    res_T r = object_id(f, in->object, &oid);
    if (r != RES_OK) return r;
    const ssolh_object& obj = f.objects[oid];
    for (size_t s = 0; s < in->object->sshapes.size(); ++s) {
      uint32_t mf, mb;
      if ((r = material_id(f, dev, in->object->sshapes[s].front, wls, &mf)) != RES_OK) return r;
      if ((r = material_id(f, dev, in->object->sshapes[s].back, wls, &mb)) != RES_OK) return r;
      f.sshapes[obj.sshape_first + s].mtl_front = mf;
      f.sshapes[obj.sshape_first + s].mtl_back = mb;
    }
    ssolh_instance d{};
    d.object = oid;
    d.receiver_mask = (uint32_t)in->receiver_mask;
    d.receiver = in->receiver_mask ? (int32_t)L.nrcv++ : -1;
    d.sampled = in->sample ? (int32_t)L.nsamp++ : -1;
    d.isshape_first = (uint32_t)f.isshapes.size();
    d.user_id = in->id;
    L.rcv_slot.push_back(d.receiver);
    L.samp_slot.push_back(d.sampled);
    if (in->receiver_mask) has_receiver = true;
    std::vector<ssolh_inst_sshape> frames(in->object->sshapes.size());
    instance_frames(in, d, frames.data());
    for (size_t s = 0; s < in->object->sshapes.size(); ++s) {
      const ssol_shape* sh = in->object->sshapes[s].shape;
      ssolh_inst_sshape is = frames[s];
      is.prim_tally = -1;
      if (in->receiver_mask && in->per_primitive) {
        is.prim_tally = (int64_t)L.nprim;
        L.prim_first[{in->id, sh->uid}] = L.nprim;
        L.nprim += sh->ids.size() / 3;
      }
      f.isshapes.push_back(is);
      if (in->sample) {
        const std::vector<float>& sp = sh->kind == shape_kind::punched ? sh->samp_pos : sh->pos;
        for (size_t t = 0; t < sh->ids.size() / 3; ++t) {
          ssolh_sample_tri st{};
          for (int v = 0; v < 3; ++v) {
            const float* p = &sp[3 * sh->ids[3 * t + v]];
            float* dst = v == 0 ? st.v0 : (v == 1 ? st.v1 : st.v2);
            for (int k = 0; k < 3; ++k) dst[k] = p[k];
          }
          st.instance = (uint32_t)i;
          st.sshape = (uint32_t)s;
          st.prim = (uint32_t)t;
          const double e0[3] = {(double)st.v1[0] - st.v0[0], (double)st.v1[1] - st.v0[1], (double)st.v1[2] - st.v0[2]};
          const double e1[3] = {(double)st.v2[0] - st.v0[0], (double)st.v2[1] - st.v0[1], (double)st.v2[2] - st.v0[2]};
          const double n[3] = {e0[1] * e1[2] - e0[2] * e1[1], e0[2] * e1[0] - e0[0] * e1[2],
                               e0[0] * e1[1] - e0[1] * e1[0]};
          weights.push_back(0.5 * std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]));
          f.stris.push_back(st);
        }
      }
    }
    if (in->sample) {
      has_sampled = true;
      S += in->object->rt_area;
      S_proxy += in->object->samp_area;
    }
    f.instances.push_back(d);
  }
  if (!has_sampled || f.stris.empty()) {
    log_error(dev, "No solstice instance to sample.\n");
    return RES_BAD_ARG;
  }
  if (!has_receiver) log_warning(dev, "No receiver is defined.\n");
  build_alias(weights, f.alias);

  f.blas_nodes = f.nodes.size();
  if (res_T r = build_tlas(dev, f)) return r;

  // sun (src/ssol_sun.c:214-279)
  {
    const res_T rs = flat_sun(f, sun, true, f.scn.sun);
    if (rs != RES_OK) return rs;
  }

  ssolh_scene& s = f.scn;
  s.nspectra = (uint32_t)f.spectra.size();
  s.spectra_len = (uint32_t)f.wl.size();
  s.spectra = f.spectra.data();
  s.spectra_wl = f.wl.data();
  s.spectra_val = f.val.data();
  s.spectra_cdf = f.cdf.data();
  s.nmedia = (uint32_t)f.media.size();
  s.air_medium = air_id;
  s.media = f.media.data();
  s.nmaterials = (uint32_t)f.materials.size();
  s.materials = f.materials.data();
  s.nshapes = (uint32_t)f.shapes.size();
  s.shapes = f.shapes.data();
  s.nsshapes = (uint32_t)f.sshapes.size();
  s.sshapes = f.sshapes.data();
  s.nobjects = (uint32_t)f.objects.size();
  s.objects = f.objects.data();
  static_assert(sizeof(aabb) == 6 * sizeof(float), "aabb is lo xyz, hi xyz");
  s.object_boxes = (const float*)f.obj_box.data();
  s.ntriangles = (uint32_t)f.tris.size();
  s.triangles = f.tris.data();
  s.ninstances = (uint32_t)f.instances.size();
  s.instances = f.instances.data();
  s.ninst_sshapes = (uint32_t)f.isshapes.size();
  s.inst_sshapes = f.isshapes.data();
  s.nsample_tris = (uint32_t)f.stris.size();
  s.sample_tris = f.stris.data();
  s.sample_alias = f.alias.data();
  s.sampled_area = S;
  s.sampled_area_proxy = S_proxy;
  s.nreceivers = (uint32_t)L.nrcv;
  s.nsampled = (uint32_t)L.nsamp;
  s.nprim_tallies = L.nprim;
  return RES_OK;
}

void mc_result(double w, double w2, double N, ssol_mc_result* r) {
  r->E = w / N;
  r->V = w2 / N - r->E * r->E;
  r->V = r->V > 0 ? r->V : 0;
  r->SE = std::sqrt(r->V / N);
}

void estimator_release(ssol_estimator* e) {
  if (e->rng) ssp_rng_ref_put(e->rng);
  ssol_device* d = e->dev;
  delete e;
  device_ref_put_(d);
}

void fill_receiver(const double* base, double N, ssol_mc_receiver* r) {
  ssol_mc_result* fields[10] = {&r->incoming_flux, &r->incoming_if_no_atm_loss,
                                &r->incoming_if_no_field_loss, &r->incoming_lost_in_field,
                                &r->incoming_lost_in_atmosphere, &r->absorbed_flux,
                                &r->absorbed_if_no_atm_loss, &r->absorbed_if_no_field_loss,
                                &r->absorbed_lost_in_field, &r->absorbed_lost_in_atmosphere};
  for (int f = 0; f < 10; ++f) mc_result(base[2 * f], base[2 * f + 1], N, fields[f]);
}

}  // namespace

struct solve_cache {
  std::vector<uint64_t> sig, xf;
  Flat flat;
  Layout layout;
};

namespace ssol {
void solve_cache_free(ssol_device* d) {
  delete d->cache;
  d->cache = nullptr;
}
}  // namespace ssol

namespace {

// Flattened scene of `scn` on the device: reuses the cached copy when no
// setter ran since it was built (the cache key covers every object's
// generation stamp), creates the GPU context at the first solve.
res_T prepare_scene(ssol_scene* scn) {
  ssol_device* dev = scn->dev;
  std::vector<uint64_t> xf;
  std::vector<uint64_t> sig = scene_signature(scn, &xf);
  const bool same = dev->cache && dev->cache->sig == sig;
  const bool cached = same && dev->cache->xf == xf;
  if (!same) {
    solve_cache_free(dev);
    dev->cache = new solve_cache();
    res_T r = flatten(scn, dev->cache->flat, dev->cache->layout);
    if (r != RES_OK) {
      solve_cache_free(dev);
      return r;
    }
  } else if (!cached) {
    // only instances moved: new frames, and the TLAS refit on the GPU while
    // it stays within REFIT_LOOSENESS of the tree as built, else built anew
    // on the GPU; a full upload only if either fails
    Flat& f = dev->cache->flat;
    res_T r = refresh_frames(scn, f);
    if (r == RES_OK && dev->gpu) {
      double ratio = 0;
      int e = ssolh_scene_move_instances(dev->gpu, (uint32_t)f.instances.size(),
                                         f.instances.data(), (uint32_t)f.isshapes.size(),
                                         f.isshapes.data(), &ratio);
      if (!e && ratio > REFIT_LOOSENESS) e = ssolh_scene_build_tlas(dev->gpu);
      if (!e) {
        dev->cache->xf.swap(xf);
        return RES_OK;
      }
    }
    if (r == RES_OK) r = build_tlas(dev, f);
    if (r != RES_OK) {
      solve_cache_free(dev);
      return r;
    }
  }
  if (!dev->gpu) {
    const int e = ssolh_create(0, &dev->gpu);
    if (e) {
      log_error(dev, "ssol_solve: %s\n", ssolh_last_error());
      fprintf(stderr, "ssol_solve: GPU unavailable: %s\n", ssolh_last_error());
      dev->gpu = nullptr;
      solve_cache_free(dev);
      return RES_UNKNOWN_ERR;
    }
  }
  if (!cached) {
    const int e = ssolh_scene_upload(dev->gpu, &dev->cache->flat.scn);
    if (e) {
      fprintf(stderr, "ssol_solve: scene upload failed: %s\n", ssolh_last_error());
      solve_cache_free(dev);
      return e == SSOLH_ERR_ARG ? RES_BAD_ARG : RES_MEM_ERR;
    }
    dev->cache->sig.swap(sig);
    dev->cache->xf.swap(xf);
  }
  return RES_OK;
}

res_T check_sun(ssol_scene* scn, const char* who) {
  ssol_device* dev = scn->dev;
  if (!scn->sun) { log_error(dev, "%s: no sun attached.\n", who); return RES_BAD_ARG; }
  if (!scn->sun->spectrum || scn->sun->spectrum->wavelengths.empty()) {
    log_error(dev, "%s: sun's spectrum undefined.\n", who);
    return RES_BAD_ARG;
  }
  if (scn->sun->dni <= 0) { log_error(dev, "%s: sun's DNI undefined.\n", who); return RES_BAD_ARG; }
  return RES_OK;
}

// Largest extent of the world box of the ray-traced triangles, times 0.75:
// the default length of the tracked sun and escape segments
// (compute_infinite_path_segment_extend, src/ssol_solver.c:562-570).
double tracked_extend(const ssol_scene* scn) {
  double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  for (const ssol_instance* in : scn->instances) {
    if (!in->object) continue;
    const double* t = in->transform;
    for (const auto& ss : in->object->sshapes) {
      const std::vector<float>& pos = ss.shape->pos;
      for (size_t v = 0; v + 2 < pos.size(); v += 3) {
        for (int k = 0; k < 3; ++k) {
          const double w = t[k] * pos[v] + t[3 + k] * pos[v + 1] + t[6 + k] * pos[v + 2] + t[9 + k];
          lo[k] = std::min(lo[k], w);
          hi[k] = std::max(hi[k], w);
        }
      }
    }
  }
  double e = 0;
  for (int k = 0; k < 3; ++k) if (hi[k] >= lo[k]) e = std::max(e, hi[k] - lo[k]);
  return e * 0.75;
}

// Solve that records every path (FP64 form): runs once with a buffer sized
// for a few vertices per path, and again with the exact size if the paths
// produced more (the realisations are the same), then orders the vertices
// by realisation and index. Paths cancelled by the roulette are dropped as
// in src/ssol_solver.c:1111-1117.
res_T solve_tracked(ssol_scene* scn, const ssolh_solve_args& a0, const ssol_path_tracker* tracker,
                    ssol_estimator* est) {
  ssol_device* dev = scn->dev;
  double sun_len = tracker->sun_ray_length, inf_len = tracker->infinite_ray_length;
  if (sun_len < 0 || inf_len < 0) {
    const double ext = tracked_extend(scn);
    if (sun_len < 0) sun_len = ext;
    if (inf_len < 0) inf_len = ext;
  }
  ssolh_solve_args a = a0;
  a.precision = SSOLH_FP64;
  a.graph_iterations = 0;
  ssolh_track_args t{sun_len, inf_len, 0};
  t.max_vertices = std::max<uint64_t>(4096, a.realisations * 8);
  uint64_t nv = 0;
  int e = ssolh_solve_tracked(dev->gpu, &a, &t, &nv);
  if (!e && nv > t.max_vertices) {
    t.max_vertices = nv;
    e = ssolh_solve_tracked(dev->gpu, &a, &t, &nv);
  }
  if (e) {
    fprintf(stderr, "ssol_solve: GPU solve failed: %s\n", ssolh_last_error());
    return e == SSOLH_ERR_ARG ? RES_BAD_ARG : RES_UNKNOWN_ERR;
  }
  std::vector<ssolh_track_vertex> v(nv);
  std::vector<uint8_t> types(a.realisations);
  e = ssolh_track_download(dev->gpu, v.data(), nv, types.data(), a.realisations);
  if (e) {
    fprintf(stderr, "ssol_solve: path download failed: %s\n", ssolh_last_error());
    return RES_UNKNOWN_ERR;
  }
  std::sort(v.begin(), v.end(), [](const ssolh_track_vertex& x, const ssolh_track_vertex& y) {
    return x.path != y.path ? x.path < y.path : x.index < y.index;
  });
  size_t nkept = 0, npaths = 0;
  for (const auto& x : v) nkept += types[x.path] != SSOLH_PATH_DROPPED;
  for (uint8_t ty : types) npaths += ty != SSOLH_PATH_DROPPED;
  est->path_vertices.reserve(nkept);
  est->paths.reserve(npaths);
  std::vector<size_t> at;  // first vertex of each kept path
  at.reserve(npaths);
  size_t i = 0;
  for (uint64_t p = 0; p < a.realisations; ++p) {
    const size_t first = i;
    while (i < v.size() && v[i].path == p) ++i;
    if (types[p] == SSOLH_PATH_DROPPED) continue;
    at.push_back(est->path_vertices.size());
    for (size_t k = first; k < i; ++k) {
      ssol_path_vertex pv;
      for (int c = 0; c < 3; ++c) pv.pos[c] = v[k].pos[c];
      pv.weight = v[k].weight;
      est->path_vertices.push_back(pv);
    }
    est->paths.push_back({(int)types[p], nullptr, i - first});
  }
  for (size_t k = 0; k < est->paths.size(); ++k) est->paths[k].v = est->path_vertices.data() + at[k];
  return RES_OK;
}

}  // namespace

extern "C" {

res_T ssol_solve(struct ssol_scene* scn, const struct ssp_rng* rng, const size_t N,
                 const size_t max_failed, const struct ssol_path_tracker* tracker,
                 struct ssol_estimator** out) {
  if (!scn || !rng || !N || !out) return RES_BAD_ARG;
  *out = nullptr;
  if (N > (size_t)INT64_MAX || max_failed > (size_t)INT64_MAX) return RES_BAD_ARG;
  ssol_device* dev = scn->dev;
  if (res_T rs = check_sun(scn, "ssol_solve")) return rs;
  res_T pr = prepare_scene(scn);
  if (pr != RES_OK) return pr;
  Flat& f = dev->cache->flat;
  Layout& L = dev->cache->layout;
  int e = 0;
  ssolh_solve_args a{};
  a.realisations = N;
  a.max_failed = max_failed;
  a.rng_key = rng->key;
  a.rng_counter = rng->counter;
  a.precision = dev->precision == SSOL_PRECISION_FP32 ? SSOLH_FP32 : SSOLH_FP64;
  a.path_slots = (uint32_t)std::min<size_t>(dev->path_slots, UINT32_MAX);
  a.pending_entries = 4;
  a.graph_iterations = dev->graph_iterations;
  ssol_estimator* est = new ssol_estimator();
  est->dev = dev;
  ref_get(dev);
  if (tracker) {
    const res_T tr = solve_tracked(scn, a, tracker, est);
    if (tr != RES_OK) {
      estimator_release(est);
      return tr;
    }
  } else {
    e = ssolh_solve(dev->gpu, &a);
    if (e) {
      fprintf(stderr, "ssol_solve: GPU solve failed: %s\n", ssolh_last_error());
      estimator_release(est);
      return e == SSOLH_ERR_ARG ? RES_BAD_ARG : RES_UNKNOWN_ERR;
    }
  }
  est->nrcv = L.nrcv;
  est->nsamp = L.nsamp;
  est->rcv.assign(L.nrcv * 40, 0);
  est->samp.assign(L.nsamp * 4, 0);
  est->samp_count.assign(L.nsamp, 0);
  est->sxr.assign(L.nsamp * L.nrcv * 40, 0);
  est->prim.assign(L.nprim * 40, 0);
  ssolh_tallies t{};
  t.receivers = est->rcv.data();
  t.sampled = est->samp.data();
  t.sampled_count = (uint64_t*)est->samp_count.data();
  t.sampled_x_receivers = est->sxr.data();
  t.primitives = est->prim.data();
  e = ssolh_tallies_download(dev->gpu, &t);
  if (e) {
    fprintf(stderr, "ssol_solve: tally download failed: %s\n", ssolh_last_error());
    estimator_release(est);
    return RES_UNKNOWN_ERR;
  }
  memcpy(est->global, t.global, sizeof(est->global));
  est->realisation_count = t.realisation_count;
  est->failed_count = t.failed_count;
  est->sampled_area = f.scn.sampled_area;
  for (size_t i = 0; i < scn->instances.size(); ++i) {
    const ssol_instance* in = scn->instances[i];
    if (L.rcv_slot[i] >= 0) est->rcv_slot[in->id] = L.rcv_slot[i];
    if (L.samp_slot[i] >= 0) est->samp_slot[in->id] = L.samp_slot[i];
  }
  est->prim_first = L.prim_first;
  est->stats.segments = t.segments;
  est->stats.paths = N;
  est->stats.pending_spills = t.pending_spills;
  est->stats.kernel_ms = t.kernel_ms;
  est->stats.trace_ms = t.trace_ms;
  est->stats.shade_ms = t.shade_ms;
  est->stats.iterations = t.iterations;
  est->stats.wavefronts = t.wavefronts;
  est->stats.refilled = t.refilled;
  // RNG state after the run: the next realisations' counters
  ssp_rng* nr = new ssp_rng();
  nr->type = rng->type;
  nr->key = rng->key;
  nr->counter = rng->counter + N;
  est->rng = nr;
  *out = est;
  if (t.aborted) {
    log_error(dev, "Too many unexpected radiative paths.\n");
    return RES_BAD_OP;
  }
  return RES_OK;
}

res_T ssol_estimator_ref_get(struct ssol_estimator* e) {
  if (!e) return RES_BAD_ARG;
  ref_get(e);
  return RES_OK;
}

res_T ssol_estimator_ref_put(struct ssol_estimator* e) {
  if (!e) return RES_BAD_ARG;
  if (--e->ref == 0) estimator_release(e);
  return RES_OK;
}

res_T ssol_estimator_get_mc_global(struct ssol_estimator* e, struct ssol_mc_global* g) {
  if (!e || !g) return RES_BAD_ARG;
  const double N = (double)e->realisation_count;
  ssol_mc_result* fields[6] = {&g->cos_factor, &g->absorbed_by_receivers, &g->shadowed,
                               &g->missing, &g->extinguished_by_atmosphere, &g->other_absorbed};
  for (int i = 0; i < 6; ++i) mc_result(e->global[i][0], e->global[i][1], N, fields[i]);
  return RES_OK;
}

res_T ssol_estimator_get_mc_receiver(struct ssol_estimator* e, const struct ssol_instance* inst,
                                     const enum ssol_side_flag side, struct ssol_mc_receiver* rcv) {
  if (!e || !inst || !rcv || !(inst->receiver_mask & (int)side)) return RES_BAD_ARG;
  memset(rcv, 0, sizeof(*rcv));
  auto it = e->rcv_slot.find(inst->id);
  if (it == e->rcv_slot.end()) return RES_OK;
  const int s = side == SSOL_FRONT ? 0 : 1;
  fill_receiver(&e->rcv[((size_t)it->second * 2 + s) * 20], (double)e->realisation_count, rcv);
  e->refs.push_back(std::unique_ptr<ssol_estimator::rcv_ref>(
      new ssol_estimator::rcv_ref{e, it->second, s, inst}));
  rcv->mc__ = e->refs.back().get();
  rcv->N__ = e->realisation_count;
  rcv->instance__ = inst;
  return RES_OK;
}

res_T ssol_estimator_get_mc_sampled_x_receiver(struct ssol_estimator* e,
                                               const struct ssol_instance* samp,
                                               const struct ssol_instance* ri,
                                               const enum ssol_side_flag side,
                                               struct ssol_mc_receiver* rcv) {
  if (!e || !samp || !ri || !rcv || (side != SSOL_BACK && side != SSOL_FRONT) || !samp->sample ||
      !(ri->receiver_mask & (int)side))
    return RES_BAD_ARG;
  memset(rcv, 0, sizeof(*rcv));
  auto si = e->samp_slot.find(samp->id);
  if (si == e->samp_slot.end()) return RES_BAD_ARG;
  auto it = e->rcv_slot.find(ri->id);
  if (it == e->rcv_slot.end()) return RES_OK;
  const int s = side == SSOL_FRONT ? 0 : 1;
  const size_t key = ((size_t)si->second * e->nrcv + (size_t)it->second) * 2 + (size_t)s;
  fill_receiver(&e->sxr[key * 20], (double)e->realisation_count, rcv);
  e->refs.push_back(std::unique_ptr<ssol_estimator::rcv_ref>(
      new ssol_estimator::rcv_ref{e, -1, s, ri}));
  rcv->mc__ = e->refs.back().get();
  rcv->N__ = e->samp_count[si->second];
  return RES_OK;
}

res_T ssol_estimator_get_realisation_count(const struct ssol_estimator* e, size_t* n) {
  if (!e || !n) return RES_BAD_ARG;
  *n = e->realisation_count;
  return RES_OK;
}

res_T ssol_estimator_get_failed_count(const struct ssol_estimator* e, size_t* n) {
  if (!e || !n) return RES_BAD_ARG;
  *n = e->failed_count;
  return RES_OK;
}

res_T ssol_estimator_get_sampled_area(const struct ssol_estimator* e, double* area) {
  if (!e || !area) return RES_BAD_ARG;
  *area = e->sampled_area;
  return RES_OK;
}

res_T ssol_estimator_get_sampled_count(const struct ssol_estimator* e, size_t* n) {
  if (!e || !n) return RES_BAD_ARG;
  *n = e->nsamp;
  return RES_OK;
}

res_T ssol_estimator_get_mc_sampled(struct ssol_estimator* e, const struct ssol_instance* inst,
                                    struct ssol_mc_sampled* s) {
  if (!e || !inst || !s) return RES_BAD_ARG;
  auto it = e->samp_slot.find(inst->id);
  if (it == e->samp_slot.end()) return RES_BAD_ARG;
  const double* b = &e->samp[(size_t)it->second * 4];
  s->nb_samples = e->samp_count[it->second];
  mc_result(b[0], b[1], (double)s->nb_samples, &s->cos_factor);
  mc_result(b[2], b[3], (double)e->realisation_count, &s->shadowed);
  return RES_OK;
}

res_T ssol_estimator_get_rng_state(const struct ssol_estimator* e, const struct ssp_rng** r) {
  if (!e || !r) return RES_BAD_ARG;
  *r = e->rng;
  return RES_OK;
}

// Tracked paths (src/ssol_estimator.c:270-322)
res_T ssol_estimator_get_tracked_paths_count(const struct ssol_estimator* e, size_t* n) {
  if (!e || !n) return RES_BAD_ARG;
  *n = e->paths.size();
  return RES_OK;
}

res_T ssol_estimator_get_tracked_path(const struct ssol_estimator* e, const size_t i,
                                      struct ssol_path* p) {
  if (!e || !p || i >= e->paths.size()) return RES_BAD_ARG;
  p->path__ = &e->paths[i];
  return RES_OK;
}

res_T ssol_path_get_vertices_count(const struct ssol_path* p, size_t* n) {
  if (!p || !n || !p->path__) return RES_BAD_ARG;
  *n = static_cast<const ssol_estimator::tracked_path*>(p->path__)->n;
  return RES_OK;
}

res_T ssol_path_get_vertex(const struct ssol_path* p, const size_t i, struct ssol_path_vertex* v) {
  if (!p || !v || !p->path__) return RES_BAD_ARG;
  const auto* tp = static_cast<const ssol_estimator::tracked_path*>(p->path__);
  if (i >= tp->n) return RES_BAD_ARG;
  *v = tp->v[i];
  return RES_OK;
}

res_T ssol_path_get_type(const struct ssol_path* p, enum ssol_path_type* t) {
  if (!p || !t || !p->path__) return RES_BAD_ARG;
  *t = (enum ssol_path_type) static_cast<const ssol_estimator::tracked_path*>(p->path__)->type;
  return RES_OK;
}

res_T ssol_mc_receiver_get_mc_shape(struct ssol_mc_receiver* rcv, const struct ssol_shape* shape,
                                    struct ssol_mc_shape* mc) {
  if (!rcv || !shape || !mc || !rcv->mc__ || !rcv->instance__) return RES_BAD_ARG;
  const ssol_instance* inst = rcv->instance__;
  bool owns = false;
  for (const shaded_shape& ss : inst->object->sshapes) owns |= ss.shape == shape;
  if (!owns) return RES_BAD_ARG;
  auto* ref = (ssol_estimator::rcv_ref*)rcv->mc__;
  ssol_estimator* e = ref->est;
  mc->N__ = rcv->N__;
  mc->shape__ = shape;
  mc->mc__ = nullptr;
  auto it = e->prim_first.find({inst->id, shape->uid});
  if (it != e->prim_first.end() && ref->slot >= 0) {
    e->shape_refs.push_back(std::unique_ptr<ssol_estimator::shape_ref>(
        new ssol_estimator::shape_ref{e, it->second * 2 + (size_t)ref->side, shape}));
    mc->mc__ = e->shape_refs.back().get();
  }
  return RES_OK;
}

res_T ssol_mc_shape_get_mc_primitive(struct ssol_mc_shape* sh, const unsigned i,
                                     struct ssol_mc_primitive* prim) {
  if (!sh || !prim || !sh->shape__) return RES_BAD_ARG;
  const ssol_shape* shape = sh->shape__;
  if (i >= shape->ids.size() / 3) return RES_BAD_ARG;
  ssol_mc_result* fields[10] = {&prim->incoming_flux, &prim->incoming_if_no_atm_loss,
                                &prim->incoming_if_no_field_loss, &prim->incoming_lost_in_field,
                                &prim->incoming_lost_in_atmosphere, &prim->absorbed_flux,
                                &prim->absorbed_if_no_atm_loss, &prim->absorbed_if_no_field_loss,
                                &prim->absorbed_lost_in_field, &prim->absorbed_lost_in_atmosphere};
  if (!sh->mc__) {
    for (auto* f : fields) *f = SSOL_MC_RESULT_NULL;
    return RES_OK;
  }
  auto* ref = (ssol_estimator::shape_ref*)sh->mc__;
  // first = slot*2 + side of triangle 0; triangle i is (slot + i)*2 + side
  const size_t key = ref->first + 2 * (size_t)i;
  const double* b = &ref->est->prim[key * 20];
  const float* v0 = &shape->pos[3 * shape->ids[3 * i]];
  const float* v1 = &shape->pos[3 * shape->ids[3 * i + 1]];
  const float* v2 = &shape->pos[3 * shape->ids[3 * i + 2]];
  const double e0[3] = {(double)v1[0] - v0[0], (double)v1[1] - v0[1], (double)v1[2] - v0[2]};
  const double e1[3] = {(double)v2[0] - v0[0], (double)v2[1] - v0[1], (double)v2[2] - v0[2]};
  const double n[3] = {e0[1] * e1[2] - e0[2] * e1[1], e0[2] * e1[0] - e0[0] * e1[2],
                       e0[0] * e1[1] - e0[1] * e1[0]};
  const double area = 0.5 * std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  for (int f = 0; f < 10; ++f) {
    mc_result(b[2 * f], b[2 * f + 1], (double)sh->N__, fields[f]);
    fields[f]->E /= area;
    fields[f]->V /= area * area;
    fields[f]->SE /= area;
  }
  return RES_OK;
}

// Checks a sun batch, prepares the scene and flattens the suns; the
// caller's realisation counts in `nr`
static res_T batch_setup(struct ssol_scene* scn, const size_t nsuns, struct ssol_sun* const* suns,
                  const size_t* realisations, const char* who, std::vector<ssolh_sun>& dsuns,
                  std::vector<uint64_t>& nr) {
  ssol_device* dev = scn->dev;
  if (res_T rs = check_sun(scn, who)) return rs;
  if (nsuns > (1u << 24)) return RES_BAD_ARG;
  uint64_t total = 0;
  for (size_t j = 0; j < nsuns; ++j) {
    if (!suns[j] || realisations[j] > (size_t)INT64_MAX) return RES_BAD_ARG;
    total += realisations[j];
    if (total > (uint64_t)INT64_MAX) return RES_BAD_ARG;
  }
  if (!total) return RES_BAD_ARG;
  res_T pr = prepare_scene(scn);
  if (pr != RES_OK) return pr;
  Flat& f = dev->cache->flat;
  dsuns.resize(nsuns);
  for (size_t j = 0; j < nsuns; ++j) {
    if (flat_sun(f, suns[j], false, dsuns[j]) != RES_OK) {
      log_error(dev, "%s: sun %zu is invalid or its spectrum is not the scene sun's.\n", who, j);
      return RES_BAD_ARG;
    }
  }
  nr.assign(realisations, realisations + nsuns);
  return RES_OK;
}

static ssolh_batch_args batch_args(const ssol_device* dev, std::vector<ssolh_sun>& dsuns,
                            std::vector<uint64_t>& nr, uint64_t key, uint64_t counter,
                            uint32_t slots, uint32_t graph) {
  ssolh_batch_args a{};
  a.nsuns = (uint32_t)dsuns.size();
  a.suns = dsuns.data();
  a.realisations = nr.data();
  a.max_failed = 0;
  a.rng_key = key;
  a.rng_counter = counter;
  a.precision = dev->precision == SSOL_PRECISION_FP32 ? SSOLH_FP32 : SSOLH_FP64;
  a.path_slots = slots;
  a.pending_entries = 4;
  a.graph_iterations = graph;
  return a;
}

res_T ssol_solve_sun_batch(struct ssol_scene* scn, const struct ssp_rng* rng, const size_t nsuns,
                           struct ssol_sun* const* suns, const size_t* realisations,
                           struct ssol_mc_global* globals, size_t* counts) {
  if (!scn || !rng || !nsuns || !suns || !realisations || !globals) return RES_BAD_ARG;
  ssol_device* dev = scn->dev;
  std::vector<ssolh_sun> dsuns;
  std::vector<uint64_t> nr;
  if (res_T r = batch_setup(scn, nsuns, suns, realisations, "ssol_solve_sun_batch", dsuns, nr))
    return r;
  ssolh_batch_args a = batch_args(dev, dsuns, nr, rng->key, rng->counter,
                                  (uint32_t)std::min<size_t>(dev->path_slots, UINT32_MAX),
                                  dev->graph_iterations);
  std::vector<ssolh_batch_result> res(nsuns);
  const int e = ssolh_solve_batch(dev->gpu, &a, res.data());
  if (e) {
    fprintf(stderr, "ssol_solve_sun_batch: GPU solve failed: %s\n", ssolh_last_error());
    return e == SSOLH_ERR_ARG ? RES_BAD_ARG : RES_UNKNOWN_ERR;
  }
  for (size_t j = 0; j < nsuns; ++j) {
    const double N = (double)res[j].realisation_count;
    if (counts) counts[j] = (size_t)res[j].realisation_count;
    if (N == 0) {  // a sun given no realisation: zero estimates
      memset(&globals[j], 0, sizeof(globals[j]));
      continue;
    }
    ssol_mc_result* fields[6] = {&globals[j].cos_factor, &globals[j].absorbed_by_receivers,
                                 &globals[j].shadowed, &globals[j].missing,
                                 &globals[j].extinguished_by_atmosphere,
                                 &globals[j].other_absorbed};
    for (int k = 0; k < 6; ++k) mc_result(res[j].global[k][0], res[j].global[k][1], N, fields[k]);
  }
  return RES_OK;
}

res_T ssol_solve_adjoint(struct ssol_scene* scn, const struct ssp_rng* rng, const size_t N,
                         const size_t nparams, struct ssol_instance* const* heliostats,
                         const double* axes, const double d_absorbed,
                         const struct ssol_instance* map_receiver,
                         const struct ssol_shape* map_shape, const double* d_map,
                         double* grad) {
  if (!scn || !rng || !N || !nparams || !heliostats || !axes || !grad) return RES_BAD_ARG;
  if (N > (size_t)INT64_MAX || nparams > (1u << 24)) return RES_BAD_ARG;
  if (d_map && (!map_receiver || !map_shape)) return RES_BAD_ARG;
  ssol_device* dev = scn->dev;
  if (res_T rs = check_sun(scn, "ssol_solve_adjoint")) return rs;
  if (dev->precision != SSOL_PRECISION_FP64) {
    log_error(dev, "ssol_solve_adjoint: the adjoint runs in the FP64 form.\n");
    return RES_BAD_ARG;
  }
  res_T pr = prepare_scene(scn);
  if (pr != RES_OK) return pr;
  Layout& L = dev->cache->layout;
  // parameter slot of every scene instance (flat instances follow the
  // scene's instance order)
  std::vector<int32_t> inst_param(scn->instances.size(), -1);
  for (size_t p = 0; p < nparams; ++p) {
    const ssol_instance* h = heliostats[p];
    size_t i = 0;
    while (i < scn->instances.size() && scn->instances[i] != h) ++i;
    if (i == scn->instances.size() || inst_param[i] >= 0) return RES_BAD_ARG;
    inst_param[i] = (int32_t)p;
  }
  std::vector<double> gmap;
  if (d_map) {
    auto it = L.prim_first.find({map_receiver->id, map_shape->uid});
    if (it == L.prim_first.end()) {
      log_error(dev, "ssol_solve_adjoint: the map shape has no per-primitive tallies.\n");
      return RES_BAD_ARG;
    }
    gmap.assign(L.nprim * 2, 0.0);
    const size_t ntris = map_shape->ids.size() / 3;
    for (size_t t = 0; t < ntris; ++t) {
      // the map is in W/m2 (the getter divides by the triangle's area)
      const float* v0 = &map_shape->pos[3 * map_shape->ids[3 * t]];
      const float* v1 = &map_shape->pos[3 * map_shape->ids[3 * t + 1]];
      const float* v2 = &map_shape->pos[3 * map_shape->ids[3 * t + 2]];
      const double e0[3] = {(double)v1[0] - v0[0], (double)v1[1] - v0[1], (double)v1[2] - v0[2]};
      const double e1[3] = {(double)v2[0] - v0[0], (double)v2[1] - v0[1], (double)v2[2] - v0[2]};
      const double n[3] = {e0[1] * e1[2] - e0[2] * e1[1], e0[2] * e1[0] - e0[0] * e1[2],
                           e0[0] * e1[1] - e0[1] * e1[0]};
      const double area = 0.5 * std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
      gmap[(it->second + t) * 2] = area > 0 ? d_map[t] / area : 0.0;  // front side
    }
  }
  ssolh_adjoint_args a{};
  a.solve.realisations = N;
  a.solve.rng_key = rng->key;
  a.solve.rng_counter = rng->counter;
  a.solve.precision = SSOLH_FP64;
  a.solve.path_slots = (uint32_t)std::min<size_t>(dev->path_slots, UINT32_MAX);
  a.solve.pending_entries = 4;
  a.solve.graph_iterations = dev->graph_iterations;
  a.nparams = (uint32_t)nparams;
  a.inst_param = inst_param.data();
  a.axes = axes;
  a.gmap = gmap.empty() ? nullptr : gmap.data();
  a.d_absorbed = d_absorbed;
  const int e = ssolh_solve_adjoint(dev->gpu, &a, grad);
  if (e) {
    fprintf(stderr, "ssol_solve_adjoint: GPU solve failed: %s\n", ssolh_last_error());
    return e == SSOLH_ERR_ARG ? RES_BAD_ARG : RES_UNKNOWN_ERR;
  }
  return RES_OK;
}

res_T ssol_device_autotune(struct ssol_device* dev, struct ssol_scene* scn, const size_t probe,
                           struct ssol_autotune_result* out) {
  if (!dev || !scn || scn->dev != dev || probe < 1024 || !out) return RES_BAD_ARG;
  if (res_T rs = check_sun(scn, "ssol_device_autotune")) return rs;
  res_T pr = prepare_scene(scn);
  if (pr != RES_OK) return pr;
  // candidates: path slots (the wavefront width) x iterations per graph
  // launch (0: host-driven iterations); each timed on a probe solve after
  // one warm-up, the fastest kept on the device
  // (0: the memory-sized default width)
  static const uint32_t slot_choices[] = {0, 1u << 20, 1u << 22, 1u << 24};
  static const uint32_t graph_choices[] = {0, 8, 32};
  double best = 1e300;
  uint32_t best_slots = 0, best_graph = 0;
  unsigned ntried = 0;
  for (uint32_t sl : slot_choices) {
    if (sl >= probe) continue;  // the default already holds the whole probe
    for (uint32_t gi : graph_choices) {
      ssolh_solve_args a{};
      a.realisations = probe;
      a.rng_key = 0x6175746f74756e65ull;  // fixed key: the probe is only timed
      a.precision = dev->precision == SSOL_PRECISION_FP32 ? SSOLH_FP32 : SSOLH_FP64;
      a.path_slots = sl;
      a.pending_entries = 4;
      a.graph_iterations = gi;
      double t = 0;
      for (int rep = 0; rep < 2; ++rep) {
        const auto t0 = std::chrono::steady_clock::now();
        const int e = ssolh_solve(dev->gpu, &a);
        const auto t1 = std::chrono::steady_clock::now();
        if (e) {
          fprintf(stderr, "ssol_device_autotune: probe failed: %s\n", ssolh_last_error());
          return RES_UNKNOWN_ERR;
        }
        t = std::chrono::duration<double>(t1 - t0).count();  // keep the second run
      }
      ++ntried;
      if (t < best) { best = t; best_slots = sl; best_graph = gi; }
    }
  }
  dev->path_slots = best_slots;
  dev->graph_iterations = best_graph;
  out->path_slots = best_slots;
  out->graph_iterations = best_graph;
  out->paths_per_second = (double)probe / best;
  out->candidates = ntried;
  return RES_OK;
}

}  // extern "C"

namespace ssol {
res_T scene_for_gpu(ssol_scene* scn, const char* who) {
  if (!scn) return RES_BAD_ARG;
  if (res_T rs = check_sun(scn, who)) return rs;
  return prepare_scene(scn);
}
}  // namespace ssol

extern "C" {

res_T ssol_device_autotune_batch(struct ssol_device* dev, struct ssol_scene* scn,
                                 const size_t nsuns, struct ssol_sun* const* suns,
                                 const size_t* realisations, struct ssol_autotune_result* out) {
  if (!dev || !scn || scn->dev != dev || !nsuns || !suns || !realisations || !out)
    return RES_BAD_ARG;
  std::vector<ssolh_sun> dsuns;
  std::vector<uint64_t> nr;
  if (res_T r = batch_setup(scn, nsuns, suns, realisations, "ssol_device_autotune_batch", dsuns,
                            nr))
    return r;
  uint64_t total = 0;
  for (uint64_t x : nr) total += x;
  // candidates: wavefront widths (0: the memory-sized default) x iterations
  // per graph launch (0: host-driven); each timed on the probe batch after
  // one warm-up, the fastest kept on the device
  static const uint32_t slot_choices[] = {0, 1u << 22, 1u << 24, 1u << 26};
  static const uint32_t graph_choices[] = {0, 16, 32};
  std::vector<ssolh_batch_result> res(nsuns);
  double best = 1e300;
  uint32_t best_slots = 0, best_graph = 0;
  unsigned ntried = 0;
  for (uint32_t sl : slot_choices) {
    if (sl && sl >= total) continue;  // the default already holds the whole probe
    for (uint32_t gi : graph_choices) {
      ssolh_batch_args a = batch_args(dev, dsuns, nr, 0x6175746f74756e65ull, 0, sl, gi);
      double t = 0;
      for (int rep = 0; rep < 2; ++rep) {
        const auto t0 = std::chrono::steady_clock::now();
        const int e = ssolh_solve_batch(dev->gpu, &a, res.data());
        const auto t1 = std::chrono::steady_clock::now();
        if (e) {
          fprintf(stderr, "ssol_device_autotune_batch: probe failed: %s\n", ssolh_last_error());
          return RES_UNKNOWN_ERR;
        }
        t = std::chrono::duration<double>(t1 - t0).count();  // keep the second run
      }
      ++ntried;
      if (t < best) { best = t; best_slots = sl; best_graph = gi; }
    }
  }
  dev->path_slots = best_slots;
  dev->graph_iterations = best_graph;
  out->path_slots = best_slots;
  out->graph_iterations = best_graph;
  out->paths_per_second = (double)total / best;
  out->candidates = ntried;
  return RES_OK;
}

res_T ssol_device_set_graph_iterations(struct ssol_device* dev, const unsigned n) {
  if (!dev || n > 1024) return RES_BAD_ARG;
  dev->graph_iterations = n;
  return RES_OK;
}

res_T ssol_estimator_get_solve_stats(const struct ssol_estimator* e, struct ssol_solve_stats* s) {
  if (!e || !s) return RES_BAD_ARG;
  *s = e->stats;
  return RES_OK;
}

}  // extern "C"
