// ssp_rng: the Threefry-4x64-20 counter state behind ssol_solve's rng
// argument and ssol_estimator_get_rng_state (see ssp.h).
#include <cstring>

#include "ssol_internal.hpp"

namespace {

inline uint64_t rotl(uint64_t x, unsigned r) { return (x << r) | (x >> (64 - r)); }

void refill(ssp_rng* r) {
  const uint64_t ctr[4] = {r->counter, 0, 0, 0};
  const uint64_t key[4] = {r->key, SSP_THREEFRY_TAG_HOST, 0, 0};
  ssp_threefry4x64_20(ctr, key, r->buf);
  r->counter += 1;
  r->avail = 4;
}

}  // namespace

extern "C" {

void ssp_threefry4x64_20(const uint64_t c[4], const uint64_t k[4], uint64_t o[4]) {
  static const unsigned R[8][2] = {{14, 16}, {52, 57}, {23, 40}, {5, 37},
                                   {25, 33}, {46, 12}, {58, 22}, {32, 32}};
  uint64_t ks[5];
  ks[4] = 0x1BD11BDAA9FC1A22ull;
  uint64_t x[4];
  for (int i = 0; i < 4; ++i) {
    ks[i] = k[i];
    ks[4] ^= k[i];
    x[i] = c[i] + k[i];
  }
  for (int r = 0; r < 20; ++r) {
    const unsigned* rot = R[r % 8];
    if ((r & 1) == 0) {
      x[0] += x[1]; x[1] = rotl(x[1], rot[0]); x[1] ^= x[0];
      x[2] += x[3]; x[3] = rotl(x[3], rot[1]); x[3] ^= x[2];
    } else {
      x[0] += x[3]; x[3] = rotl(x[3], rot[0]); x[3] ^= x[0];
      x[2] += x[1]; x[1] = rotl(x[1], rot[1]); x[1] ^= x[2];
    }
    if ((r & 3) == 3) {
      const unsigned s = (unsigned)(r >> 2) + 1;
      for (int i = 0; i < 4; ++i) x[i] += ks[(s + (unsigned)i) % 5];
      x[3] += s;
    }
  }
  memcpy(o, x, sizeof(x));
}

res_T ssp_rng_create(struct mem_allocator*, const enum ssp_rng_type type, struct ssp_rng** out) {
  if (!out) return RES_BAD_ARG;
  if (type != SSP_RNG_THREEFRY) return RES_BAD_ARG;
  *out = new ssp_rng();
  return RES_OK;
}

res_T ssp_rng_ref_get(struct ssp_rng* r) {
  if (!r) return RES_BAD_ARG;
  ++r->ref;
  return RES_OK;
}

res_T ssp_rng_ref_put(struct ssp_rng* r) {
  if (!r) return RES_BAD_ARG;
  if (--r->ref == 0) delete r;
  return RES_OK;
}

res_T ssp_rng_get_type(const struct ssp_rng* r, enum ssp_rng_type* t) {
  if (!r || !t) return RES_BAD_ARG;
  *t = r->type;
  return RES_OK;
}

res_T ssp_rng_set(struct ssp_rng* r, const uint64_t seed) {
  if (!r) return RES_BAD_ARG;
  r->key = seed;
  r->counter = 0;
  r->avail = 0;
  return RES_OK;
}

uint64_t ssp_rng_get(struct ssp_rng* r) {
  if (!r->avail) refill(r);
  return r->buf[4 - r->avail--];
}

double ssp_rng_canonical(struct ssp_rng* r) {
  return (double)(ssp_rng_get(r) >> 11) * (1.0 / 9007199254740992.0);
}

float ssp_rng_canonical_float(struct ssp_rng* r) {
  return (float)(ssp_rng_get(r) >> 40) * (1.0f / 16777216.0f);
}

res_T ssp_rng_discard(struct ssp_rng* r, uint64_t n) {
  if (!r) return RES_BAD_ARG;
  while (n--) (void)ssp_rng_get(r);
  return RES_OK;
}

res_T ssp_rng_write(const struct ssp_rng* r, FILE* stream) {
  if (!r || !stream) return RES_BAD_ARG;
  if (fprintf(stream, "threefry4x64_20 %llu %llu %d %llu %llu %llu %llu\n",
              (unsigned long long)r->key, (unsigned long long)r->counter, r->avail,
              (unsigned long long)r->buf[0], (unsigned long long)r->buf[1],
              (unsigned long long)r->buf[2], (unsigned long long)r->buf[3]) < 0)
    return RES_IO_ERR;
  return RES_OK;
}

res_T ssp_rng_read(struct ssp_rng* r, FILE* stream) {
  if (!r || !stream) return RES_BAD_ARG;
  char name[32];
  unsigned long long k, c, b0, b1, b2, b3;
  int avail;
  if (fscanf(stream, "%31s %llu %llu %d %llu %llu %llu %llu", name, &k, &c, &avail, &b0, &b1, &b2,
             &b3) != 8)
    return RES_IO_ERR;
  if (strcmp(name, "threefry4x64_20") != 0 || avail < 0 || avail > 4) return RES_BAD_ARG;
  r->key = k;
  r->counter = c;
  r->avail = avail;
  r->buf[0] = b0; r->buf[1] = b1; r->buf[2] = b2; r->buf[3] = b3;
  return RES_OK;
}

res_T ssp_rng_get_counter_state(const struct ssp_rng* r, uint64_t* key, uint64_t* counter) {
  if (!r || !key || !counter) return RES_BAD_ARG;
  *key = r->key;
  *counter = r->counter;
  return RES_OK;
}

res_T ssp_rng_set_counter_state(struct ssp_rng* r, const uint64_t key, const uint64_t counter) {
  if (!r) return RES_BAD_ARG;
  r->key = key;
  r->counter = counter;
  r->avail = 0;
  return RES_OK;
}

}  // extern "C"
