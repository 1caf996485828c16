/*
 * oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU oracle for mc-mpi's Layer::simulate: a line-for-line restatement of
 * src/layer.cpp and src/random.cpp (reference paths), built as liboracle.so and
 * used by tests/, __graft_entry__.smoke() and bench.py's CPU baseline to check
 * the HIP product.  The product never links, loads or calls it.
 *
 * Two instances of orc_impl.inc:
 *   orc32_*  real_t = float  (the reference as shipped; pinned byte-exactly
 *            against data/test_layer_target_WA.out and against the reference
 *            sources compiled into oracle/_ref);
 *   orc64_*  real_t = double (the FP64 restatement the GPU must match
 *            bit-exactly per history).
 *
 * FP64 log/exp: correctly rounded (orc_math.c) by default, or glibc.
 * FP64 tallies: the reference's real_t summation, plus an exact fixed-point
 * sum (signed 128-bit, 88 fractional bits) of the same terms, which is what
 * the GPU computes; see orc_fx_add.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "orc_math.h"

#define ORC_CAT2(a, b) a##b
#define ORC_CAT(a, b) ORC_CAT2(a, b)
#define ORC_MIN(a, b) (((a) < (b)) ? (a) : (b))
#define ORC_MAX(a, b) (((a) > (b)) ? (a) : (b))

/* src/layer.cpp:15 */
#define MAX_PARTICLES_VECTOR 1000000

/* src/random.cpp:8-10 */
#define RNG_G ((uint64_t)(6364136223846793005ull))
#define RNG_C ((uint64_t)(1442695040888963407ull))
#define RNG_P (((uint64_t)(1)) << 63)
/* src/random.cpp:20-22 */
#define RNGS_G ((uint64_t)(5177284530976225183ull))
#define RNGS_C ((uint64_t)(2096348467109453893ull))
#define RNGS_P (((uint64_t)(1)) << 63)

/* src/random.cpp:24-29 */
static inline uint64_t orc_rnd_seed(uint64_t *seed) {
  *seed = (RNGS_G * *seed + RNGS_C) % RNGS_P;
  return *seed;
}

uint64_t orc_rnd_seed_c(uint64_t *seed) { return orc_rnd_seed(seed); }

/* ---------------------------------------------------- exact fixed-point tally
 * A tally term v (double) maps to round_half_up(v * 2^88) as a signed 128-bit
 * integer; terms with |v| >= 2^39 set flag 8, NaN flag 1, +inf flag 2,
 * -inf flag 4.  Sums of such integers are exact and independent of order. */
typedef unsigned __int128 orc_fx_t;
#define ORC_FX_FRAC 88

static inline orc_fx_t orc_fx_from_double(double v, uint32_t *flags) {
  if (v != v) {
    *flags |= 1;
    return 0;
  }
  if (isinf(v)) {
    *flags |= v > 0 ? 2 : 4;
    return 0;
  }
  double a = fabs(v);
  if (a >= 0x1p39) {
    *flags |= 8;
    return 0;
  }
  if (a == 0.0) return 0;
  int e;
  double mfrac = frexp(a, &e); /* a = mfrac * 2^e, mfrac in [0.5, 1) */
  uint64_t M = (uint64_t)ldexp(mfrac, 53);
  int s = e - 53 + ORC_FX_FRAC;
  orc_fx_t f;
  if (s >= 0)
    f = (orc_fx_t)M << s;
  else if (s > -64)
    f = (orc_fx_t)(((M >> (-s - 1)) + 1) >> 1);
  else
    f = 0;
  return signbit(v) ? (orc_fx_t)0 - f : f;
}

static inline void orc_fx_add(orc_fx_t *acc, uint32_t *flags, double v) {
  *acc += orc_fx_from_double(v, flags);
}

static inline double orc_fx_to_double(orc_fx_t v, uint32_t flags) {
  if ((flags & 1) || (flags & 8) || ((flags & 6) == 6)) return NAN;
  if (flags & 2) return INFINITY;
  if (flags & 4) return -INFINITY;
  return (double)(__int128)v * 0x1p-88;
}

/* ctypes helpers: fixed point as (lo, hi) u64 pairs */
void orc_fx_convert(const double *v, int64_t n, uint64_t *out, uint32_t *flags) {
  for (int64_t i = 0; i < n; ++i) {
    orc_fx_t f = orc_fx_from_double(v[i], &flags[i]);
    out[2 * i] = (uint64_t)f;
    out[2 * i + 1] = (uint64_t)(f >> 64);
  }
}
void orc_fx_finalize(const uint64_t *fx, const uint32_t *flags, int64_t n, double *out) {
  for (int64_t i = 0; i < n; ++i) {
    orc_fx_t f = ((orc_fx_t)fx[2 * i + 1] << 64) | fx[2 * i];
    out[i] = orc_fx_to_double(f, flags[i]);
  }
}

/* ------------------------------------------------------------- FP32 instance */
#define ORC_P 32
#define REAL float
#define ORC_PFX orc32_
#define ORC_LOG logf /* `log(float)` resolves to logf (src/layer.cpp:137) */
#define ORC_EXP expf /* src/layer.cpp:175 calls expf; src/layer.cpp:59 exp(float) */
#define ORC_FABS fabsf
#define EPS_PRECISION 1e-4F /* include/types/types.hpp:18 */
#define MAXREAL FLT_MAX     /* include/types/types.hpp:17 */
#include "orc_impl.inc"
#undef ORC_P
#undef REAL
#undef ORC_PFX
#undef ORC_LOG
#undef ORC_EXP
#undef ORC_FABS
#undef EPS_PRECISION
#undef MAXREAL

/* ------------------------------------------------------------- FP64 instance */
static int orc_math_mode; /* 0: correctly rounded, 1: glibc */
void orc_set_math_mode(int mode) { orc_math_mode = mode; }
int orc_get_math_mode(void) { return orc_math_mode; }
static inline double orc64_log(double x) { return orc_math_mode ? log(x) : orc_log_cr(x); }
static inline double orc64_exp(double x) { return orc_math_mode ? exp(x) : orc_exp_cr(x); }

#define ORC_P 64
#define REAL double
#define ORC_PFX orc64_
#define ORC_LOG orc64_log
#define ORC_EXP orc64_exp
#define ORC_FABS fabs
#define EPS_PRECISION 1e-4
#define MAXREAL DBL_MAX
#include "orc_impl.inc"

/* FP64 bank simulation with the fixed-point tally as (lo, hi) u64 pairs */
int orc64_simulate_bank_fx(const orc64_params *L, const double *sigs, const double *ar,
                           orc64_particle *particles, int64_t n, double *weights,
                           uint64_t *fx_pairs, uint32_t *flags, int32_t *result,
                           uint64_t *events, uint64_t *collisions, int nthread) {
  orc_fx_t *fx = NULL;
  if (fx_pairs) {
    fx = (orc_fx_t *)aligned_alloc(16, sizeof(orc_fx_t) * (size_t)(L->m > 0 ? L->m : 1));
    if (!fx) return -1;
    for (int j = 0; j < L->m; ++j)
      fx[j] = ((orc_fx_t)fx_pairs[2 * j + 1] << 64) | fx_pairs[2 * j];
  }
  int rc = orc64_simulate_bank(L, sigs, ar, particles, n, weights, fx, flags, result, events,
                               collisions, nthread);
  if (fx) {
    for (int j = 0; j < L->m; ++j) {
      fx_pairs[2 * j] = (uint64_t)fx[j];
      fx_pairs[2 * j + 1] = (uint64_t)(fx[j] >> 64);
    }
    free(fx);
  }
  return rc;
}

/* RNG known-answer helpers (src/random.cpp) for tests */
void orc_rnd_seed_chain(uint64_t seed, int64_t n, uint64_t *out) {
  for (int64_t i = 0; i < n; ++i) out[i] = orc_rnd_seed(&seed);
}
void orc32_rnd_real_step(uint64_t *seeds, int64_t n, float *out) {
  for (int64_t i = 0; i < n; ++i) out[i] = orc32_rnd_real(&seeds[i]);
}
void orc64_rnd_real_step(uint64_t *seeds, int64_t n, double *out) {
  for (int64_t i = 0; i < n; ++i) out[i] = orc64_rnd_real(&seeds[i]);
}
void orc64_math_eval(int which, const double *x, int64_t n, double *out) {
  for (int64_t i = 0; i < n; ++i) out[i] = which == 0 ? orc64_log(x[i]) : orc64_exp(x[i]);
}
