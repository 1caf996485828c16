/*
 * orc_math.c -- TEST INFRASTRUCTURE ONLY (part of the CPU oracle).
 *
 * Independent FP64 log/exp for the oracle.  The reference computes
 * `log(h)` (src/layer.cpp:137) and `expf(-sig_a * di)` (src/layer.cpp:175)
 * in float; its FP64 restatement needs a double log/exp whose value does not
 * depend on the libm.  The oracle defines that value as the correctly rounded
 * one and obtains it without sharing any code with the product:
 *
 *   1. evaluate in x87 extended precision (logl/expl, 64-bit significand,
 *      error of a few units of 2^-63 relative);
 *   2. if rounding that value to double is unambiguous given a margin of
 *      2^-60 relative, return it;
 *   3. otherwise evaluate in binary128 with libquadmath (logq/expq, about
 *      2^-112 relative) and round that; the margin there is 2^-108.
 *   Cases still ambiguous after step 3 are counted in orc_math_unresolved.
 *
 * The glibc mode calls the system libm's log/exp (the literal calls a
 * `real_t = double` build of the reference would make); it is used to time
 * the CPU baseline and to measure how often glibc differs from correct
 * rounding.
 */
#include <math.h>
#include <quadmath.h>
#include <stdint.h>

#include "orc_math.h"

static unsigned long long orc_unresolved_total;

unsigned long long orc_math_unresolved(void) {
  return __atomic_load_n(&orc_unresolved_total, __ATOMIC_RELAXED);
}

static int ld_unambiguous(long double v, double *out) {
  double d = (double)v;
  long double e = fabsl(v) * 0x1p-60L;
  if ((double)(v + e) == d && (double)(v - e) == d) {
    *out = d;
    return 1;
  }
  return 0;
}

static double q_round(__float128 q) {
  double d = (double)q;
  __float128 e = fabsq(q) * 0x1p-108Q;
  if (!((double)(q + e) == d && (double)(q - e) == d))
    __atomic_fetch_add(&orc_unresolved_total, 1, __ATOMIC_RELAXED);
  return d;
}

double orc_log_cr(double x) {
  if (!(x > 0.0) || isinf(x)) return log(x); /* 0, <0, NaN, inf: exact specials */
  if (x == 1.0) return 0.0;
  double d;
  if (ld_unambiguous(logl((long double)x), &d)) return d;
  return q_round(logq((__float128)x));
}

double orc_exp_cr(double x) {
  if (isnan(x)) return x;
  if (x == 0.0) return 1.0;
  if (x < -746.0) return 0.0;
  if (x > 710.0) return INFINITY;
  double d;
  if (ld_unambiguous(expl((long double)x), &d)) return d;
  if (fabs(x) < 0x1p-20) {
    /* near 1: 1 + expm1(x) in binary128 is exact to about 2^-112 absolute,
       which separates cases such as exp(2^-53) = 1 + 2^-53 + 2^-107 from the
       rounding boundary 1 + 2^-53 that a relative margin cannot. */
    __float128 q = 1 + expm1q((__float128)x);
    double dq = (double)q;
    __float128 e = 0x1p-110Q;
    if (!((double)(q + e) == dq && (double)(q - e) == dq))
      __atomic_fetch_add(&orc_unresolved_total, 1, __ATOMIC_RELAXED);
    return dq;
  }
  return q_round(expq((__float128)x));
}

double orc_log_glibc(double x) { return log(x); }
double orc_exp_glibc(double x) { return exp(x); }
