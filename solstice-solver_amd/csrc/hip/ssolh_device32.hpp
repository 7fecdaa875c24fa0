// FP32 building blocks of the single-precision form of the flux integrator:
// 3-vectors, the path random stream as 24-bit uniforms, and the sun and BSDF
// samplers of ssolh_device.hpp restated in float (same formulas, same draw
// order). Quadric roots stay in ssolh_device.hpp (FP64).
#pragma once

#include "ssolh_device.hpp"

namespace ssolh {

struct f3 { float x, y, z; };

SD f3 mkf(float x, float y, float z) { return f3{x, y, z}; }
SD f3 operator+(f3 a, f3 b) { return mkf(a.x + b.x, a.y + b.y, a.z + b.z); }
SD f3 operator-(f3 a, f3 b) { return mkf(a.x - b.x, a.y - b.y, a.z - b.z); }
SD f3 operator*(f3 a, float s) { return mkf(a.x * s, a.y * s, a.z * s); }
SD f3 operator-(f3 a) { return mkf(-a.x, -a.y, -a.z); }
SD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
SD f3 cross(f3 a, f3 b) {
  return mkf(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
SD f3 normalize(f3 a) {
  const float n2 = dot(a, a);
  return n2 > 0 ? a * rsqrtf(n2) : a;
}
SD f3 tof(d3 a) { return mkf((float)a.x, (float)a.y, (float)a.z); }
SD d3 tod(f3 a) { return mk(a.x, a.y, a.z); }
SD f3 ldf(const float* p) { return mkf(p[0], p[1], p[2]); }
// column-major 3x3 (double) times float vector, in float
SD f3 mulf(const double* m, f3 v) {
  return mkf((float)m[0] * v.x + (float)m[3] * v.y + (float)m[6] * v.z,
             (float)m[1] * v.x + (float)m[4] * v.y + (float)m[7] * v.z,
             (float)m[2] * v.x + (float)m[5] * v.y + (float)m[8] * v.z);
}
SD f3 tmulf(const double* m, f3 v) {
  return mkf((float)m[0] * v.x + (float)m[1] * v.y + (float)m[2] * v.z,
             (float)m[3] * v.x + (float)m[4] * v.y + (float)m[5] * v.z,
             (float)m[6] * v.x + (float)m[7] * v.y + (float)m[8] * v.z);
}

// The realisation's Threefry stream (same key, counter and block layout as
// PathRng) cut into eight 24-bit uniforms per block, in the open interval
// (0, 1): (k + 0.5) 2^-24.
struct PathRng32 {
  uint64_t key, path;
  uint32_t block;
  uint32_t avail;
  uint32_t w[8];

  SD void init(uint64_t k, uint64_t p, uint32_t b) { key = k; path = p; block = b; avail = 0; }
  SD float next() {
    if (!avail) {
      const uint64_t c[4] = {path, (uint64_t)block, 0, 0};
      const uint64_t kk[4] = {key, TAG_PATH, 0, 0};
      uint64_t o[4];
      threefry4x64_20(c, kk, o);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        w[2 * i] = (uint32_t)o[i];
        w[2 * i + 1] = (uint32_t)(o[i] >> 32);
      }
      ++block;
      avail = 8;
    }
    uint32_t x = w[0];
#pragma unroll
    for (int i = 1; i < 8; ++i) x = (8u - avail == (uint32_t)i) ? w[i] : x;
    --avail;
    return ((float)(x >> 8) + 0.5f) * (1.0f / 16777216.0f);
  }
};

template <class Rng>
SD f3 sample_sun32(const ssolh_sun& sun, Rng& rng) {
  float st, ct, phi;
  f3 loc;
  switch (sun.type) {
    case SSOLH_SUN_PILLBOX: {
      const float s2 = rng.next() * (float)sun.param[0];
      st = sqrtf(s2);
      ct = sqrtf(1 - s2);
      phi = rng.next() * (float)(2 * PI);
      loc = mkf(cosf(phi) * st, sinf(phi) * st, ct);
      return normalize(mulf(sun.basis, loc));
    }
    case SSOLH_SUN_GAUSSIAN: {
      st = (float)sun.param[0] * sqrtf(-2 * logf(rng.next()));
      st = st > 1 ? 1 : st;
      ct = sqrtf(1 - st * st);
      phi = rng.next() * (float)(2 * PI);
      loc = mkf(cosf(phi) * st, sinf(phi) * st, ct);
      return normalize(mulf(sun.basis, loc));
    }
    case SSOLH_SUN_BUIE: {
      // the rejection test runs on the FP64 density
      const double* p = sun.param;
      phi = rng.next() * (float)(2 * PI);
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
      st = (float)sin(theta);
      ct = (float)cos(theta);
      loc = mkf(st * sinf(phi), st * cosf(phi), ct);
      return mulf(sun.basis, loc);
    }
    default:
      return mkf((float)sun.dir[0], (float)sun.dir[1], (float)sun.dir[2]);
  }
}

SD void basis32(f3 n, f3& x, f3& y) {
  f3 h;
  const float ax = fabsf(n.x), ay = fabsf(n.y), az = fabsf(n.z);
  if (ax <= ay && ax <= az) h = mkf(1, 0, 0);
  else if (ay <= az) h = mkf(0, 1, 0);
  else h = mkf(0, 0, 1);
  x = normalize(cross(h, n));
  y = cross(n, x);
}

SD float fresnel_dielectric32(float cos_i, float eta_i, float eta_t, float& cos_t) {
  const float e = eta_i / eta_t;
  const float s2 = e * e * (1 - cos_i * cos_i);
  if (s2 >= 1) { cos_t = 0; return 1; }
  const float ct = sqrtf(1 - s2);
  const float rs = (eta_i * cos_i - eta_t * ct) / (eta_i * cos_i + eta_t * ct);
  const float rp = (eta_t * cos_i - eta_i * ct) / (eta_t * cos_i + eta_i * ct);
  cos_t = ct;
  return 0.5f * (rs * rs + rp * rp);
}

SD f3 sample_microfacet32(int distrib, float m, float u1, float u2, f3 n) {
  const float tan_t = distrib == 0 ? m * sqrtf(-logf(1 - u1)) : m * sqrtf(u1);
  const float cos_t = rsqrtf(1 + tan_t * tan_t);
  const float sin_t = tan_t * cos_t;
  const float phi = (float)(2 * PI) * u2;
  f3 x, y;
  basis32(n, x, y);
  return x * (cosf(phi) * sin_t) + y * (sinf(phi) * sin_t) + n * cos_t;
}

SD f3 reflect32(f3 wi, f3 n, float cos_i) { return n * (2 * cos_i) - wi; }

// Same models and draw order as bsdf_sample (ssolh_device.hpp), in float
template <class Rng>
SD float bsdf_sample32(const ssolh_material& m, float eta_out, float eta_in, float slab_eta,
                       float slab_abs, f3 wi, f3 n, Rng& rng, f3& wo, int& lobe) {
  const float cos_i = dot(wi, n);
  lobe = LOBE_REFLECT;
  switch (m.type) {
    case SSOLH_MTL_MIRROR:
      if (m.roughness == 0) {
        wo = reflect32(wi, n, cos_i);
        return (float)m.reflectivity;
      }
      for (int k = 0; k < 256; ++k) {
        const float u1 = rng.next(), u2 = rng.next();
        const f3 h = sample_microfacet32(m.distribution, (float)m.roughness, u1, u2, n);
        const float c = dot(wi, h);
        if (c <= 0) continue;
        wo = reflect32(wi, h, c);
        if (dot(wo, n) > 0) return (float)m.reflectivity;
      }
      wo = n;
      return 0;
    case SSOLH_MTL_MATTE: {
      const float u1 = rng.next(), u2 = rng.next();
      const float r = sqrtf(u1), phi = (float)(2 * PI) * u2;
      f3 x, y;
      basis32(n, x, y);
      wo = x * (r * cosf(phi)) + y * (r * sinf(phi)) + n * sqrtf(1 - u1);
      return (float)m.reflectivity;
    }
    case SSOLH_MTL_DIELECTRIC: {
      float cos_t;
      const float F = fresnel_dielectric32(cos_i, eta_out, eta_in, cos_t);
      if (rng.next() < F) {
        wo = reflect32(wi, n, cos_i);
      } else {
        const float e = eta_out / eta_in;
        wo = normalize(wi * (-e) + n * (e * cos_i - cos_t));
        lobe = LOBE_TRANSMIT;
      }
      return 1;
    }
    case SSOLH_MTL_THIN_DIELECTRIC: {
      float cos_t;
      const float R = fresnel_dielectric32(cos_i, eta_out, slab_eta, cos_t);
      if (R >= 1) {
        wo = reflect32(wi, n, cos_i);
        return 1;
      }
      const float T = 1 - R;
      const float A = expf(-slab_abs * (float)m.thickness / cos_t);
      const float den = 1 - R * R * A * A;
      const float Rt = R + T * T * R * A * A / den;
      const float Tt = T * T * A / den;
      const float w = Rt + Tt;
      if (rng.next() * w < Rt) {
        wo = reflect32(wi, n, cos_i);
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
