/* Random number generator state used at the ssol_solve boundary.
 *
 * ssol_solve takes a `const struct ssp_rng*` and the estimator hands one back
 * (ssol_estimator_get_rng_state). In the reference both are Star-SP 0.12
 * objects (un-vendored; call sites src/ssol_solver.c:675-677 and
 * src/ssol_estimator.c:368-418). This header gives that type a home with the
 * Star-SP function names a client uses around ssol_solve.
 *
 * Generator: Threefry-4x64-20 (Salmon, Moraes, Dror, Shaw, "Parallel random
 * numbers: as easy as 1, 2, 3", SC'11), the counter-based family behind
 * Star-SP's SSP_RNG_THREEFRY. The state is a 64-bit key and a 64-bit
 * counter. ssol_solve gives realisation i of a solve the independent stream
 * keyed by (key, PATH tag) with first counter word counter+i, and the
 * estimator's state comes back with counter advanced by the number of
 * realisations, so chained solves draw disjoint streams. Host draws
 * (ssp_rng_get & co) use a separate key tag and advance the counter by one per
 * 256-bit block. Bit streams differ from Star-SP's (its proxy partitioning
 * is not reproduced); only SSP_RNG_THREEFRY is implemented, the other type
 * names exist for source compatibility and ssp_rng_create rejects them with
 * RES_BAD_ARG. */
#ifndef SSP_H
#define SSP_H

#include <stdint.h>
#include <stdio.h>
#include "ssol.h"

#ifdef __cplusplus
extern "C" {
#endif

enum ssp_rng_type {
  SSP_RNG_KISS,
  SSP_RNG_MT19937_64,
  SSP_RNG_RANLUX48,
  SSP_RNG_RANDOM_DEVICE,
  SSP_RNG_THREEFRY,
  SSP_RNG_AES,
  SSP_RNG_TYPES_COUNT__
};

/* Key tags of the two stream families (second key word of Threefry) */
#define SSP_THREEFRY_TAG_PATH UINT64_C(0x5041544853534f4c) /* realisations */
#define SSP_THREEFRY_TAG_HOST UINT64_C(0x484f535453534f4c) /* ssp_rng_get */

SSOL_API res_T ssp_rng_create
  (struct mem_allocator* allocator, const enum ssp_rng_type type,
   struct ssp_rng** rng);
SSOL_API res_T ssp_rng_ref_get(struct ssp_rng* rng);
SSOL_API res_T ssp_rng_ref_put(struct ssp_rng* rng);
SSOL_API res_T ssp_rng_get_type(const struct ssp_rng* rng, enum ssp_rng_type* type);
SSOL_API res_T ssp_rng_set(struct ssp_rng* rng, const uint64_t seed);
SSOL_API uint64_t ssp_rng_get(struct ssp_rng* rng);
SSOL_API double ssp_rng_canonical(struct ssp_rng* rng);       /* [0, 1) */
SSOL_API float ssp_rng_canonical_float(struct ssp_rng* rng); /* [0, 1) */
SSOL_API res_T ssp_rng_discard(struct ssp_rng* rng, uint64_t n);
SSOL_API res_T ssp_rng_write(const struct ssp_rng* rng, FILE* stream);
SSOL_API res_T ssp_rng_read(struct ssp_rng* rng, FILE* stream);

/* Extension: the (key, counter) pair ssol_solve derives its streams from */
SSOL_API res_T ssp_rng_get_counter_state
  (const struct ssp_rng* rng, uint64_t* key, uint64_t* counter);
SSOL_API res_T ssp_rng_set_counter_state
  (struct ssp_rng* rng, const uint64_t key, const uint64_t counter);

/* Extension: one Threefry-4x64-20 block, exported for known-answer tests */
SSOL_API void ssp_threefry4x64_20
  (const uint64_t ctr[4], const uint64_t key[4], uint64_t out[4]);

#ifdef __cplusplus
}
#endif

#endif /* SSP_H */
