/*
 * orc_rng.h -- OpenMC 0.13 random-number streams, restated for the oracle.
 *
 * TEST INFRASTRUCTURE: part of the oracle, used only by tests and by the
 * CPU-baseline leg of the benchmark, never by the product library.
 *
 * Upstream: openmc-dev/openmc v0.13, src/random_lcg.cpp (not vendored in the
 * reference; its call sites are src/raykernels/CollisionKernel.C:150 and
 * src/userobjects/OpenMCStudy.C:326,480-482,536).
 *   state   s <- M s + C (mod 2^64), M = 6364136223846793005,
 *           C = 1442695040888963407            (Knuth's MMIX LCG)
 *   output  PCG-RXS-M-XS-64 permutation of the new state (O'Neill 2014,
 *           "PCG: A Family of Simple Fast Space-Efficient Statistically Good
 *           Algorithms for Random Number Generation", section 6.3.4),
 *           mapped to [0,1] by ldexp(out, -64)
 *   skip    future_seed(n, s) = M^n s + C (M^n - 1)/(M - 1) by binary
 *           exponentiation (F. Brown, "Random Number Generation with
 *           Arbitrary Strides", Trans. ANS 71, 1994)
 *   streams init_seed(id, k) = future_seed(id * 152917, master + k),
 *           k = 0 tracking, 1 source, 2 URR
 *
 * Pinned by the reference's gold trace (test/tests/study/gold/
 * openmc_study_rays.e) for the state step, the seeding formula and the
 * one-draw-per-segment rule; the permutation is pinned by the published
 * PCG constant and by statistical tests.
 */
#ifndef ORC_RNG_H
#define ORC_RNG_H

#include <stdint.h>

#define ORC_PRN_MULT 6364136223846793005ULL
#define ORC_PRN_ADD 1442695040888963407ULL
#define ORC_PRN_STRIDE 152917ULL

static inline uint64_t orc_step(uint64_t s) { return ORC_PRN_MULT * s + ORC_PRN_ADD; }

static inline double orc_u64_to_unit(uint64_t r)
{
  /* (double)r is round-to-nearest; times 2^-64 is exact */
  return (double)r * 0x1p-64;
}

static inline double orc_prn(uint64_t *seed)
{
  uint64_t s = orc_step(*seed);
  *seed = s;
  uint64_t word = ((s >> ((s >> 59u) + 5u)) ^ s) * 12605985483714917081ULL;
  uint64_t out = (word >> 43u) ^ word;
  return orc_u64_to_unit(out);
}

static inline uint64_t orc_future_seed(uint64_t n, uint64_t seed)
{
  uint64_t g = ORC_PRN_MULT, c = ORC_PRN_ADD;
  uint64_t g_new = 1, c_new = 0;
  while (n > 0) {
    if (n & 1u) {
      g_new *= g;
      c_new = c_new * g + c;
    }
    c = (g + 1) * c;
    g *= g;
    n >>= 1;
  }
  return g_new * seed + c_new;
}

static inline uint64_t orc_init_seed(int64_t id, int stream, uint64_t master)
{
  return orc_future_seed((uint64_t)id * ORC_PRN_STRIDE, master + (uint64_t)stream);
}

/* MaCaw keeps seeds in Real ray aux data: a uint64 stored into a double and
 * read back (OpenMCStudy.C:345-347, CollisionKernel.C:132-134).  A value that
 * rounds to 2^64 wraps to 0 (the conversion is otherwise undefined). */
static inline uint64_t orc_seed_roundtrip(uint64_t s)
{
  double d = (double)s;
  if (d >= 18446744073709551616.0) return 0;
  return (uint64_t)d;
}

#endif
