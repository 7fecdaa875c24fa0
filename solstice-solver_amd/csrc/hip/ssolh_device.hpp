// Device-side building blocks of the flux integrator on gfx950: vector math,
// Threefry-4x64-20 path streams, sun and wavelength samplers, the baked
// BSDFs, quadric roots. Everything here is __device__ code shared by the
// wavefront kernels in ssolh.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ssolh.h"

namespace ssolh {

#define SD __device__ __forceinline__

constexpr uint64_t TAG_PATH = 0x5041544853534f4cull;  // same tag as ssp.h
constexpr uint32_t SSOL_FRONT_ = 1u, SSOL_BACK_ = 2u;  // ssol_side_flag
constexpr double PI = 3.14159265358979323846;

struct d3 { double x, y, z; };

SD d3 mk(double x, double y, double z) { return d3{x, y, z}; }
SD d3 operator+(d3 a, d3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
SD d3 operator-(d3 a, d3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
SD d3 operator*(d3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
SD d3 operator-(d3 a) { return mk(-a.x, -a.y, -a.z); }
SD double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
SD d3 cross(d3 a, d3 b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
SD d3 normalize(d3 a) {
  const double n = sqrt(dot(a, a));
  return n > 0 ? a * (1.0 / n) : a;
}
// column-major 3x3 times vector
SD d3 mul(const double* m, d3 v) {
  return mk(m[0] * v.x + m[3] * v.y + m[6] * v.z,
            m[1] * v.x + m[4] * v.y + m[7] * v.z,
            m[2] * v.x + m[5] * v.y + m[8] * v.z);
}
// transpose(m) times vector
SD d3 tmul(const double* m, d3 v) {
  return mk(m[0] * v.x + m[1] * v.y + m[2] * v.z,
            m[3] * v.x + m[4] * v.y + m[5] * v.z,
            m[6] * v.x + m[7] * v.y + m[8] * v.z);
}
SD d3 ld3(const float* p) { return mk(p[0], p[1], p[2]); }
SD d3 ld3(const double* p) { return mk(p[0], p[1], p[2]); }

// Threefry-4x64-20 (Salmon et al., SC'11), Random123 rotation schedule
SD uint64_t rotl(uint64_t x, unsigned r) { return (x << r) | (x >> (64 - r)); }

SD void threefry4x64_20(const uint64_t c[4], const uint64_t k[4], uint64_t o[4]) {
  uint64_t ks[5];
  ks[4] = 0x1BD11BDAA9FC1A22ull;
#pragma unroll
  for (int i = 0; i < 4; ++i) { ks[i] = k[i]; ks[4] ^= k[i]; }
  uint64_t x0 = c[0] + k[0], x1 = c[1] + k[1], x2 = c[2] + k[2], x3 = c[3] + k[3];
  constexpr unsigned R[8][2] = {{14, 16}, {52, 57}, {23, 40}, {5, 37},
                                {25, 33}, {46, 12}, {58, 22}, {32, 32}};
#pragma unroll
  for (int r = 0; r < 20; ++r) {
    if ((r & 1) == 0) {
      x0 += x1; x1 = rotl(x1, R[r % 8][0]); x1 ^= x0;
      x2 += x3; x3 = rotl(x3, R[r % 8][1]); x3 ^= x2;
    } else {
      x0 += x3; x3 = rotl(x3, R[r % 8][0]); x3 ^= x0;
      x2 += x1; x1 = rotl(x1, R[r % 8][1]); x1 ^= x2;
    }
    if ((r & 3) == 3) {
      const unsigned s = (unsigned)(r >> 2) + 1;
      x0 += ks[(s + 0) % 5];
      x1 += ks[(s + 1) % 5];
      x2 += ks[(s + 2) % 5];
      x3 += ks[(s + 3) % 5] + s;
    }
  }
  o[0] = x0; o[1] = x1; o[2] = x2; o[3] = x3;
}

// Sequential draws of one realisation: block b of path p is Threefry(key,
// (p, b, 0, 0)); each block gives four 53-bit uniforms in [0, 1).
// The four words of a block are consumed in order; they are kept in named
// registers (a dynamically indexed array would live in scratch or LDS).
struct PathRng {
  uint64_t key, path;
  uint32_t block;
  uint32_t avail;
  uint64_t b0, b1, b2, b3;

  SD void init(uint64_t k, uint64_t p, uint32_t b) { key = k; path = p; block = b; avail = 0; }
  SD double next() {
    if (!avail) {
      const uint64_t c[4] = {path, (uint64_t)block, 0, 0};
      const uint64_t kk[4] = {key, TAG_PATH, 0, 0};
      uint64_t o[4];
      threefry4x64_20(c, kk, o);
      b0 = o[0]; b1 = o[1]; b2 = o[2]; b3 = o[3];
      ++block;
      avail = 4;
    }
    const uint64_t x = avail == 4 ? b0 : (avail == 3 ? b1 : (avail == 2 ? b2 : b3));
    --avail;
    return (double)(x >> 11) * (1.0 / 9007199254740992.0);
  }
};

// Spectra ------------------------------------------------------------------
struct Spectra {
  const ssolh_spectrum* sp;
  const double* wl;
  const double* val;
  const double* cdf;
};

// Reference src/ssol_spectrum.c:57-89 (lower-bound search, clamped)
SD double spectrum_at(const Spectra& s, int idx, double w) {
  const ssolh_spectrum d = s.sp[idx];
  const double* wl = s.wl + d.offset;
  const double* va = s.val + d.offset;
  uint32_t lo = 0, hi = d.count;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (wl[mid] < w) lo = mid + 1; else hi = mid;
  }
  if (lo == d.count) return va[d.count - 1];
  if (lo == 0) return va[0];
  return va[lo - 1] + (w - wl[lo - 1]) * ((va[lo] - va[lo - 1]) / (wl[lo] - wl[lo - 1]));
}

SD double data_value(const Spectra& s, const ssolh_data& d, double w) {
  return d.spectrum < 0 ? d.real : spectrum_at(s, d.spectrum, w);
}

// Inverse CDF of the piecewise-linear density through (wl, val); cdf holds
// the cumulative trapezoid areas (cdf[0] = 0).
SD double sample_wavelength(const Spectra& s, int idx, double u) {
  const ssolh_spectrum d = s.sp[idx];
  const double* wl = s.wl + d.offset;
  const double* va = s.val + d.offset;
  const double* cdf = s.cdf + d.offset;
  if (d.count <= 1) return d.count ? wl[0] : -1.0;
  const double target = u * cdf[d.count - 1];
  uint32_t lo = 0, hi = d.count - 1;  // segment i in [0, count-2]
  while (lo + 1 < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (cdf[mid] <= target) lo = mid; else hi = mid;
  }
  const uint32_t i = lo;
  const double w = wl[i + 1] - wl[i];
  const double r = target - cdf[i];
  const double a = va[i];
  const double b = (va[i + 1] - va[i]) / w;
  double x;
  if (fabs(b) * w < 1e-12 * (a > 0 ? a : 1.0)) {
    x = a > 0 ? r / a : 0.5 * w;
  } else {
    const double disc = a * a + 2 * b * r;
    x = (-a + sqrt(disc > 0 ? disc : 0.0)) / b;
  }
  x = x < 0 ? 0 : (x > w ? w : x);
  return wl[i] + x;
}

// Sun directions (reference src/ssol_ranst_sun_dir.c) ------------------------
SD double buie_pdf(const double* p, double theta) {
  // p = {thetaSD, dThetaCSSD, gamma, e^k 1000^gamma, alpha, h1, h2, p1}
  const double phi = theta < p[0] ? cos(326 * theta) / cos(308 * theta)
                                  : p[3] * pow(theta, p[2]);
  return p[4] * phi * sin(theta);
}

template <class Rng>
SD d3 sample_sun(const ssolh_sun& sun, Rng& rng) {
  double st, ct, phi;
  d3 loc;
  switch (sun.type) {
    case SSOLH_SUN_PILLBOX: {
      const double s2 = rng.next() * sun.param[0];
      st = sqrt(s2);
      ct = sqrt(1 - s2);
      phi = rng.next() * 2 * PI;
      loc = mk(cos(phi) * st, sin(phi) * st, ct);
      return normalize(mul(sun.basis, loc));
    }
    case SSOLH_SUN_GAUSSIAN: {
      st = sun.param[0] * sqrt(-2 * log(rng.next()));
      st = st > 1 ? 1 : st;
      ct = sqrt(1 - st * st);
      phi = rng.next() * 2 * PI;
      loc = mk(cos(phi) * st, sin(phi) * st, ct);
      return normalize(mul(sun.basis, loc));
    }
    case SSOLH_SUN_BUIE: {
      const double* p = sun.param;
      phi = rng.next() * 2 * PI;
      double theta, value;
      do {
        if (rng.next() < p[7]) {
          theta = rng.next() * p[0];
          value = rng.next() * p[5];
        } else {
          theta = p[0] + rng.next() * p[1];
          value = rng.next() * p[6];
        }
      } while (value > buie_pdf(p, theta));
      st = sin(theta);
      ct = cos(theta);
      loc = mk(st * sin(phi), st * cos(phi), ct);
      return mul(sun.basis, loc);
    }
    default:
      return ld3(sun.dir);
  }
}

// Frame whose z axis is n (unit); the azimuth of every sampler is uniform so
// the x/y orientation is irrelevant to the distribution.
SD void basis(d3 n, d3& x, d3& y) {
  d3 h;
  const double ax = fabs(n.x), ay = fabs(n.y), az = fabs(n.z);
  if (ax <= ay && ax <= az) h = mk(1, 0, 0);
  else if (ay <= az) h = mk(0, 1, 0);
  else h = mk(0, 0, 1);
  x = normalize(cross(h, n));
  y = cross(n, x);
}

// Quadrics (reference src/ssol_shape.c:236-285, 864-1135) ------------------
// The real roots of a t^2 + b t + c (0, 1 or 2 of them)
SD int quadric_roots(double a, double b, double c, double t[2]) {
  if (a == 0) {
    if (b != 0) { t[0] = -c / b; return 1; }
    return 0;
  }
  const double delta = b * b - 4 * a * c;
  if (delta < 0) return 0;  // the reference's NaN roots; never kept as a hit
  if (delta == 0) { t[0] = -b / (2 * a); return 1; }
  const double sq = sqrt(delta);
  t[0] = (-b - copysign(sq, b)) / (2 * a);
  t[1] = c / (a * t[0]);
  return 2;
}

// The root nearest the hint
SD int solve_second(double a, double b, double c, double hint, double& t) {
  double r[2];
  const int n = quadric_roots(a, b, c, r);
  if (n == 0) return 0;
  t = n == 1 || fabs(r[0] - hint) < fabs(r[1] - hint) ? r[0] : r[1];
  return n;
}

SD double quadric_z(int q, const double* p, double x, double y) {
  switch (q) {
    case SSOLH_Q_PARABOL: return (x * x + y * y) * p[1];
    case SSOLH_Q_PARABOL2F: return x * x * p[0] + y * y * p[1];
    case SSOLH_Q_PARABOLIC_CYLINDER: return y * y * p[1];
    case SSOLH_Q_HYPERBOL: {
      const double z0 = p[0] + p[3];
      return p[3] * sqrt(1 + (x * x + y * y) * p[2]) + p[0] - z0;
    }
    case SSOLH_Q_HEMISPHERE: {
      const double z2 = p[1] - (x * x + y * y);
      return z2 > 0 ? -sqrt(z2) + p[0] : p[0];
    }
    default: return 0;
  }
}

SD d3 quadric_grad(int q, const double* p, d3 pt) {
  switch (q) {
    case SSOLH_Q_PARABOL: return mk(-pt.x, -pt.y, 2 * p[0]);
    case SSOLH_Q_PARABOL2F: return mk(-pt.x * (4.0 * p[0]), -pt.y * (4.0 * p[1]), 2.0);
    case SSOLH_Q_PARABOLIC_CYLINDER: return mk(0, -pt.y, 2 * p[0]);
    case SSOLH_Q_HYPERBOL: {
      const double z0 = p[0] + p[3];
      return mk(pt.x, pt.y, -(pt.z + z0 - p[0]) * p[1]);
    }
    case SSOLH_Q_HEMISPHERE: return mk(-pt.x, -pt.y, p[0] - pt.z);
    default: return mk(0, 0, 1);
  }
}

// Coefficients of the ray / quadric equation a t^2 + b t + c = 0 in the
// quadric frame; false for no quadric.
SD bool quadric_coeffs(int q, const double* p, d3 o, d3 d, double& a, double& b, double& c) {
  switch (q) {
    case SSOLH_Q_PLANE: a = 0; b = d.z; c = o.z; break;
    case SSOLH_Q_PARABOL:
      a = d.x * d.x + d.y * d.y;
      b = 2 * o.x * d.x + 2 * o.y * d.y - 4 * p[0] * d.z;
      c = o.x * o.x + o.y * o.y - 4 * p[0] * o.z;
      break;
    case SSOLH_Q_PARABOL2F:
      a = d.x * d.x * p[0] + d.y * d.y * p[1];
      b = 2.0 * o.x * d.x * p[0] + 2.0 * o.y * d.y * p[1] - d.z;
      c = o.x * o.x * p[0] + o.y * o.y * p[1] - o.z;
      break;
    case SSOLH_Q_PARABOLIC_CYLINDER:
      a = d.y * d.y;
      b = 2 * o.y * d.y - 4 * p[0] * d.z;
      c = o.y * o.y - 4 * p[0] * o.z;
      break;
    case SSOLH_Q_HYPERBOL: {
      const double b2 = p[3] * p[3];
      const double b2_a2 = b2 * p[2];
      const double oz = o.z + p[3];  // o.z + z0 - g/2
      a = b2_a2 * (d.x * d.x + d.y * d.y) - d.z * d.z;
      b = 2 * (b2_a2 * (o.x * d.x + o.y * d.y) - oz * d.z);
      c = b2_a2 * (o.x * o.x + o.y * o.y) + b2 - oz * oz;
      break;
    }
    case SSOLH_Q_HEMISPHERE: {
      const double z0 = -p[0];
      a = d.x * d.x + d.y * d.y + d.z * d.z;
      b = 2 * (o.x * d.x + o.y * d.y + o.z * d.z + z0 * d.z);
      c = o.x * o.x + o.y * o.y + o.z * o.z - p[1] + 2 * z0 * o.z + z0 * z0;
      break;
    }
    default: return false;
  }
  return true;
}

// Ray / quadric in the quadric frame: the root nearest the hint and the
// gradient there; returns 0 when there is no root.
SD int quadric_hit(int q, const double* p, d3 o, d3 d, double hint, double& t, d3& grad) {
  double a, b, c;
  if (!quadric_coeffs(q, p, o, d, a, b, c)) return 0;
  if (!solve_second(a, b, c, hint, t)) return 0;
  grad = quadric_grad(q, p, o + d * t);
  return 1;
}

// BSDFs, from the physics (Star-SF is not vendored) -------------------------
enum { LOBE_NONE = 0, LOBE_REFLECT = 1, LOBE_TRANSMIT = 2 };

SD double fresnel_dielectric(double cos_i, double eta_i, double eta_t, double& cos_t) {
  const double e = eta_i / eta_t;
  const double s2 = e * e * (1 - cos_i * cos_i);
  if (s2 >= 1) { cos_t = 0; return 1; }
  const double ct = sqrt(1 - s2);
  const double rs = (eta_i * cos_i - eta_t * ct) / (eta_i * cos_i + eta_t * ct);
  const double rp = (eta_t * cos_i - eta_i * ct) / (eta_t * cos_i + eta_i * ct);
  cos_t = ct;
  return 0.5 * (rs * rs + rp * rp);
}

// Microfacet normal with density D(h)(h.n): Beckmann tan2 = -m2 ln(1-u),
// pillbox tan = m sqrt(u); uniform azimuth.
SD d3 sample_microfacet(int distrib, double m, double u1, double u2, d3 n) {
  const double tan_t = distrib == 0 ? m * sqrt(-log(1 - u1)) : m * sqrt(u1);
  const double cos_t = 1 / sqrt(1 + tan_t * tan_t);
  const double sin_t = tan_t * cos_t;
  const double phi = 2 * PI * u2;
  d3 x, y;
  basis(n, x, y);
  return x * (cos(phi) * sin_t) + y * (sin(phi) * sin_t) + n * cos_t;
}

SD d3 reflect(d3 wi, d3 n, double cos_i) { return n * (2 * cos_i) - wi; }

// wi points away from the surface on n's side. Returns the MC weight.
// h_out (optional): the microfacet normal of a glossy reflection.
template <class Rng>
SD double bsdf_sample(const ssolh_material& m, double eta_out, double eta_in,
                      double slab_eta, double slab_abs, d3 wi, d3 n,
                      Rng& rng, d3& wo, int& lobe, d3* h_out = nullptr) {
  const double cos_i = dot(wi, n);
  lobe = LOBE_REFLECT;
  switch (m.type) {
    case SSOLH_MTL_MIRROR:
      if (m.roughness == 0) {
        wo = reflect(wi, n, cos_i);
        return m.reflectivity;
      }
      for (int k = 0; k < 256; ++k) {
        const double u1 = rng.next(), u2 = rng.next();
        const d3 h = sample_microfacet(m.distribution, m.roughness, u1, u2, n);
        const double c = dot(wi, h);
        if (c <= 0) continue;
        wo = reflect(wi, h, c);
        if (dot(wo, n) > 0) {
          if (h_out) *h_out = h;
          return m.reflectivity;
        }
      }
      wo = n;
      return 0;
    case SSOLH_MTL_MATTE: {
      const double u1 = rng.next(), u2 = rng.next();
      const double r = sqrt(u1), phi = 2 * PI * u2;
      d3 x, y;
      basis(n, x, y);
      wo = x * (r * cos(phi)) + y * (r * sin(phi)) + n * sqrt(1 - u1);
      return m.reflectivity;
    }
    case SSOLH_MTL_DIELECTRIC: {
      double cos_t;
      const double F = fresnel_dielectric(cos_i, eta_out, eta_in, cos_t);
      if (rng.next() < F) {
        wo = reflect(wi, n, cos_i);
      } else {
        const double e = eta_out / eta_in;
        wo = normalize(wi * (-e) + n * (e * cos_i - cos_t));
        lobe = LOBE_TRANSMIT;
      }
      return 1;
    }
    case SSOLH_MTL_THIN_DIELECTRIC: {
      double cos_t;
      const double R = fresnel_dielectric(cos_i, eta_out, slab_eta, cos_t);
      if (R >= 1) {
        wo = reflect(wi, n, cos_i);
        return 1;
      }
      const double T = 1 - R;
      const double A = exp(-slab_abs * m.thickness / cos_t);
      const double den = 1 - R * R * A * A;
      const double Rt = R + T * T * R * A * A / den;
      const double Tt = T * T * A / den;
      const double w = Rt + Tt;
      if (rng.next() * w < Rt) {
        wo = reflect(wi, n, cos_i);
      } else {
        wo = -wi;
        lobe = LOBE_TRANSMIT;
      }
      return w;
    }
    default:
      wo = -wi;
      lobe = LOBE_NONE;
      return 0;
  }
}

}  // namespace ssolh
