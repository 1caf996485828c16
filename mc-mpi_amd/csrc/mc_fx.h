/*
 * mc_fx.h -- the exact fixed-point arithmetic of the tallies: conversion of
 * a double term to a signed 128-bit integer with 2^-88 resolution and the
 * correctly rounded conversion of a 128-bit sum back to double.
 *
 * Integer code only (plus one ldexp), so it compiles as HIP device code and
 * as plain host C/C++ (gcc) for CPU-side tests; the device kernels use this
 * exact source.
 */
#ifndef MC_FX_H
#define MC_FX_H

#include <stdint.h>

#include "mc_math.h" /* mc_asu64, MC_FN, MC_FN_COLD */

#if defined(__HIPCC__)
#define MC_FN_HD __host__ __device__ __forceinline__ static
#else
#define MC_FN_HD static inline
#endif

/* ------------------------------------------------- exact fixed-point tally
 * A term v maps to the signed 128-bit integer round_half_up(v * 2^88),
 * stored as (lo, hi) two's complement.  Non-finite terms and |v| >= 2^39 set
 * flag bits instead (1 NaN, 2 +inf, 4 -inf, 8 out of range).  Integer sums
 * are exact, so the tally does not depend on the order in which lanes, waves
 * and blocks add their terms. */
#define MC_FX_FRAC 88

typedef struct mc_fx {
  uint64_t lo;
  uint64_t hi;
} mc_fx;

/* The conversion for every term (zero, subnormal, negative, large and
   non-finite ones included), with its flags. */
typedef struct mc_fx_flagged {
  mc_fx v;
  uint32_t flags;
} mc_fx_flagged;

MC_FN_COLD mc_fx_flagged mc_fx_from_double_general(double v) {
  mc_fx_flagged out;
  out.flags = 0;
  uint64_t b = mc_asu64(v);
  uint32_t E = (uint32_t)(b >> 52) & 0x7FFu;
  mc_fx r = {0, 0};
  out.v = r;
  if (E == 0x7FFu) {
    out.flags = (b & 0x000FFFFFFFFFFFFFULL) ? 1u : ((b >> 63) ? 4u : 2u);
    return out;
  }
  if (E >= 1023u + 39u) {
    out.flags = 8u;
    return out;
  }
  if (E == 0u) return out; /* zero or subnormal: below 2^-89 -> 0 */
  uint64_t M = (b & 0x000FFFFFFFFFFFFFULL) | 0x0010000000000000ULL;
  int s = (int)E - 987; /* v = M * 2^(E-1075); v*2^88 = M * 2^(E-987) */
  if (s >= 64) {
    r.hi = M << (s - 64);
  } else if (s > 0) {
    r.lo = M << s;
    r.hi = M >> (64 - s);
  } else if (s == 0) {
    r.lo = M;
  } else if (s > -64) {
    r.lo = ((M >> (-s - 1)) + 1) >> 1; /* round half up */
  }
  if (b >> 63) { /* negate (two's complement) */
    r.lo = ~r.lo + 1;
    r.hi = ~r.hi + (r.lo == 0 ? 1 : 0);
  }
  out.v = r;
  return out;
}

/* Fast path for 2^-35 <= |v| < 2^28: |v| * 2^88 = M * 2^s with 0 < s < 64,
   so the 128-bit integer is two shifts of the 53-bit significand (negated for
   v < 0), exactly as in the general path.  Returns 0 and leaves r alone
   outside that range.  POSITIVE_ONLY sends negative terms to the general
   path too (the transport's deposits are never negative; the adjoint's terms
   are). */
#define MC_FX_FAST(POSITIVE_ONLY, v, r, taken)                               \
  do {                                                                      \
    const uint64_t b_ = mc_asu64(v);                                        \
    const int s_ = (POSITIVE_ONLY) ? (int)((uint32_t)(b_ >> 52)) - 987      \
                                   : (int)((uint32_t)(b_ >> 52) & 0x7FFu) - 987; \
    (taken) = (unsigned)(s_ - 1) < 63u;                                     \
    if (taken) {                                                            \
      const uint64_t M_ = (b_ & 0x000FFFFFFFFFFFFFULL) | 0x0010000000000000ULL; \
      (r).lo = M_ << s_;                                                    \
      (r).hi = M_ >> (64 - s_);                                             \
      if (!(POSITIVE_ONLY) && (b_ >> 63)) { /* two's complement */          \
        (r).lo = ~(r).lo + 1;                                               \
        (r).hi = ~(r).hi + ((r).lo == 0 ? 1 : 0);                           \
      }                                                                     \
    }                                                                       \
  } while (0)

/* round_half_up(v * 2^88) as a signed 128-bit integer; flags as above.
   positive_only selects the fast path's sign handling (results identical). */
MC_FN mc_fx mc_fx_from_double_p(double v, uint32_t *flags, int positive_only) {
  mc_fx r;
  int taken;
  if (positive_only) {
    MC_FX_FAST(1, v, r, taken);
  } else {
    MC_FX_FAST(0, v, r, taken);
  }
  if (taken) return r;
  const mc_fx_flagged g = mc_fx_from_double_general(v);
  *flags |= g.flags;
  return g.v;
}

MC_FN mc_fx mc_fx_from_double(double v, uint32_t *flags) {
  return mc_fx_from_double_p(v, flags, 0);
}

/* Correctly rounded conversion of the 128-bit sum back to double. */
MC_FN_HD double mc_fx_to_double(uint64_t lo, uint64_t hi) {
  int neg = (int64_t)hi < 0;
  if (neg) {
    lo = ~lo + 1;
    hi = ~hi + (lo == 0 ? 1 : 0);
  }
  double d;
  if (hi == 0) {
    d = (double)lo * 0x1p-88;
  } else {
    int lz = __builtin_clzll(hi);
    uint64_t top = lz ? ((hi << lz) | (lo >> (64 - lz))) : hi;
    uint64_t rest = lz ? (lo << lz) : lo;
    top |= (rest != 0) ? 1u : 0u; /* sticky bit below the rounding position */
    d = __builtin_ldexp((double)top, 64 - lz - MC_FX_FRAC);
  }
  return neg ? -d : d;
}

MC_FN_HD double mc_fx_finalize(uint64_t lo, uint64_t hi, uint32_t flags) {
  if ((flags & 1u) || (flags & 8u) || ((flags & 6u) == 6u)) return __builtin_nan("");
  if (flags & 2u) return __builtin_inf();
  if (flags & 4u) return -__builtin_inf();
  return mc_fx_to_double(lo, hi);
}

#endif /* MC_FX_H */
