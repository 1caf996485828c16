/*
 * mc_math.h -- correctly rounded FP64 natural log and exp.
 *
 * The transport step (reference: src/layer.cpp:137 `-log(h)` and
 * src/layer.cpp:175 `expf(-sig_a * di)`, restated in FP64) needs log and exp
 * whose results do not depend on which libm or GPU math library evaluates
 * them.  The only such definition is correct rounding: these functions return
 * the double nearest to the exact value (ties to even), so the GPU kernels and
 * any correctly rounded CPU evaluation agree bit for bit.
 *
 * Method (own implementation; no libm code or coefficients are used):
 *   - Tang's table-driven argument reduction (P. T. P. Tang, "Table-driven
 *     implementation of the exponential function" / "... of the logarithm
 *     function", ACM TOMS 15(2) 1989 and 16(4) 1990), with tables computed from
 *     their definitions by gen_math_tables.py;
 *   - Taylor polynomials with exact rational coefficients 1/n and 1/n!;
 *   - double-double arithmetic (Dekker 1971; Knuth's TwoSum) with explicit
 *     fma, so no step depends on compiler contraction;
 *   - Ziv's strategy: a fast evaluation with a bounded relative error, a
 *     one-fma rounding test valid for errors below 2^-68 (mc_rounds_to_hi),
 *     and, when the test cannot decide, an all-double-double evaluation
 *     (relative error about 2^-100).  The fast log is below 2^-68.6: a bound
 *     gen_math_tables.py computes for every table row and asserts; the
 *     largest rows are next to x = 1, where log x is small.  The fast exp is
 *     below 2^-72 (a 1,024-entry table, argument reduction below 2^-76,
 *     each rounding below 2^-75.5, series truncation below 2^-78).  An fma
 *     in place of a multiply and an add only removes a rounding.  The result can differ from the correctly
 *     rounded one only when the exact value lies within about 2^-100
 *     (relative) of a rounding boundary.
 *
 * The header compiles as HIP device code (hipcc) or as plain host C/C++ (gcc)
 * for CPU-side tests; the arithmetic is the same IEEE-754 binary64 sequence in
 * both.  Build with -ffp-contract=off.
 */
#ifndef MC_MATH_H
#define MC_MATH_H

#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MC_FN __device__ __forceinline__ static
#define MC_FN_COLD __device__ __noinline__ static
#define MC_TAB __device__ static const
#else
#define MC_FN static inline
#define MC_FN_COLD static
#define MC_TAB static const
#endif

#include "mc_math_tables.h"

/* ---------------------------------------------------------------- basics */

MC_FN uint64_t mc_asu64(double x) {
  uint64_t u;
  memcpy(&u, &x, 8);
  return u;
}
MC_FN double mc_asf64(uint64_t u) {
  double x;
  memcpy(&x, &u, 8);
  return x;
}
MC_FN double mc_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }

/* s + t = a + b exactly (Knuth TwoSum, no magnitude condition). */
MC_FN void mc_two_sum(double a, double b, double *s, double *t) {
  double sum = a + b;
  double bb = sum - a;
  *t = (a - (sum - bb)) + (b - bb);
  *s = sum;
}
/* s + t = a + b exactly when |a| >= |b| or a == 0 (Dekker FastTwoSum). */
MC_FN void mc_fast_two_sum(double a, double b, double *s, double *t) {
  double sum = a + b;
  *t = b - (sum - a);
  *s = sum;
}
/* (rh, rl) ~ (ah + al) * (bh + bl) */
MC_FN void mc_dd_mul(double ah, double al, double bh, double bl, double *rh, double *rl) {
  double p = ah * bh;
  double e = mc_fma(ah, bh, -p);
  e = e + (ah * bl + al * bh);
  mc_fast_two_sum(p, e, rh, rl);
}
/* (rh, rl) ~ (ah + al) + (bh + bl), accurate version */
MC_FN void mc_dd_add(double ah, double al, double bh, double bl, double *rh, double *rl) {
  double s, t;
  mc_two_sum(ah, bh, &s, &t);
  t = t + (al + bl);
  mc_fast_two_sum(s, t, rh, rl);
}

/* 1 when hi + lo, known to within err (absolute), rounds to hi regardless of
   where in [lo - err, lo + err] the exact tail lies. */
MC_FN int mc_round_ok(double hi, double lo, double err) {
  return (hi + (lo + err) == hi) && (hi + (lo - err) == hi);
}

/* The same certificate in one fma, for hi + lo (hi = RN(hi + lo)) within
   relative error eps of the exact value Y, eps < 2^-68 (the fast paths are
   below 2^-68.6 for log and 2^-72 for exp).  Let b be the distance from hi to its rounding boundary on
   the side of lo: b >= 2^-54 |hi| (half an ulp of hi, or a quarter below a
   power of two).  RN(hi + e lo) == hi with e = 1 + 2^-13 (one rounding, at
   that boundary, ties included) gives |lo| <= b/e, so on lo's side
   |Y - hi| <= b/e + eps |Y| < b, because b (1 - 1/e) > 2^-68 |hi|; on the
   other side |Y - hi| <= eps |Y| is far below 2^-54 |hi|.  Y thus lies
   strictly inside hi's rounding interval and rounds to hi. */
#define MC_ZIV_E (1.0 + 0x1p-13)
MC_FN int mc_rounds_to_hi(double hi, double lo) { return mc_fma(lo, MC_ZIV_E, hi) == hi; }

/* The fast paths' polynomial coefficients and the certificate's factor e.
   mc_log and mc_exp use them as constants.  Kernels pass a copy whose values
   the compiler cannot fold (made once at kernel entry, mc_coef_live in
   mc_device.h) to mc_log_k / mc_exp_k: each value then stays in one 64-bit
   register across the particle loop and each Horner step compiles to one
   fma, where folded constants are rebuilt from 32-bit halves on every use
   and copied into the accumulator first.  Values and results are the same
   either way. */
typedef struct {
  double l8, l7, l6, l5, l4, l3; /* -1/8, 1/7, -1/6, 1/5, -1/4, 1/3 */
  double e5, e4, e3;             /* 1/5!, 1/4!, 1/3! */
  double ziv;                    /* MC_ZIV_E */
} mc_coef;

MC_FN mc_coef mc_coef_default(void) {
  mc_coef K;
  K.l8 = -mc_inv_n[8][0]; K.l7 = mc_inv_n[7][0]; K.l6 = -mc_inv_n[6][0];
  K.l5 = mc_inv_n[5][0];  K.l4 = -mc_inv_n[4][0]; K.l3 = mc_inv_n[3][0];
  K.e5 = mc_inv_fact[5][0];
  K.e4 = mc_inv_fact[4][0]; K.e3 = mc_inv_fact[3][0];
  K.ziv = MC_ZIV_E;
  return K;
}

/* a / b correctly rounded from y = RN(1/b) (an IEEE division):
   q0 = RN(a y) is within 1.5 ulp of a/b; q1 = RN(q0 + r0 y) with the exact
   remainder r0 = a - b q0 is then within half an ulp plus 2^-104 relative,
   hence within 1 ulp; and with r1 = a - b q1 exact, RN(q1 + r1 y) is the
   correctly rounded quotient by Markstein's theorem (y within half an ulp of
   1/b, q1 within 1 ulp of a/b).  For |b| >= 2^-60 the fmas stay in the
   normal range while 2^-900 <= |q0| <= 2^900; otherwise (a = 0 among them)
   the IEEE division is used.  (mc_event uses the quotient only for
   |mu| > EPS_PRECISION, far above 2^-60.) */
MC_FN double mc_div_by_rcp(double a, double b, double y) {
  const double q0 = a * y;
  const double aq = __builtin_fabs(q0);
  if (!(aq >= 0x1p-900 && aq <= 0x1p900)) return a / b;
  const double r0 = mc_fma(-b, q0, a);
  const double q1 = mc_fma(r0, y, q0);
  const double r1 = mc_fma(-b, q1, a);
  return mc_fma(r1, y, q1);
}

/* ------------------------------------------------------------------- log */

/* log(x) as double-double, all steps in double-double (about 2^-100 rel). */
MC_FN_COLD double mc_log_accurate(int e, double r, int idx) {
  /* log1p(r), |r| < 2^-8: Horner over 1/n in double-double for n <= 6, the
     tail r^7/7 - ... + r^14/14 in double (below 2^-56 * r, error < 2^-108). */
  double t = 0.0;
  for (int n = 14; n >= 7; --n) {
    double c = (n & 1) ? mc_inv_n[n][0] : -mc_inv_n[n][0];
    t = mc_fma(t, r, c);
  }
  /* y = t*r + (-1/6); then Horner down to 1/1 in double-double */
  double yh, yl;
  {
    double ph = t * r, pl = mc_fma(t, r, -ph);
    mc_dd_add(ph, pl, -mc_inv_n[6][0], -mc_inv_n[6][1], &yh, &yl);
  }
  for (int n = 5; n >= 1; --n) {
    double ph, pl;
    mc_dd_mul(yh, yl, r, 0.0, &ph, &pl);
    double ch = (n & 1) ? mc_inv_n[n][0] : -mc_inv_n[n][0];
    double cl = (n & 1) ? mc_inv_n[n][1] : -mc_inv_n[n][1];
    mc_dd_add(ph, pl, ch, cl, &yh, &yl);
  }
  double lh, ll; /* log1p(r) = y * r */
  mc_dd_mul(yh, yl, r, 0.0, &lh, &ll);
  /* + (-log c) = head + mid + lo (|head| >= |mid| or head == 0) */
  double th, tl;
  mc_fast_two_sum(mc_log_tab[idx][1], mc_log_tab[idx][2], &th, &tl);
  tl = tl + mc_log_tab[idx][3];
  double sh, sl;
  mc_dd_add(lh, ll, th, tl, &sh, &sl);
  /* + e*ln2 (three-piece ln2; e*MC_LN2_HI is exact) */
  double ed = (double)e;
  double kh = ed * MC_LN2_HI;
  double mh = ed * MC_LN2_MID, ml = mc_fma(ed, MC_LN2_MID, -mh);
  ml = ml + ed * MC_LN2_LO;
  double k2h, k2l;
  mc_dd_add(kh, 0.0, mh, ml, &k2h, &k2l);
  double rh, rl;
  mc_dd_add(sh, sl, k2h, k2l, &rh, &rl);
  return rh + rl;
}

/* log in two halves.  mc_log_lookup splits x (bits ix, times 2^eadj) and
   loads its table row; mc_log_eval finishes from the row.  A caller with
   independent work can put it between the two, so that the row's loads are
   in flight meanwhile. */
typedef struct {
  double m, c, head, mid;
  int e;
  unsigned idx;
} mc_log_row;

MC_FN mc_log_row mc_log_lookup(uint64_t ix, int eadj) {
  mc_log_row R;
  const uint64_t tmp = ix - MC_LOG_OFF;
  /* e = tmp >> 52 (arithmetic), m = x / 2^e in [0.7075, 1.4150);
     table row: bucket j = bits 43..51 of tmp (512 buckets of m) */
#if defined(__HIP_DEVICE_COMPILE__)
  /* the same on the high words, in 32-bit operations (a 64-bit shift is
     otherwise turned into a 64-bit integer-to-double conversion) */
  const uint32_t thi = (uint32_t)(tmp >> 32);
  R.e = __builtin_amdgcn_sbfe((int)thi, 20, 12) + eadj;
  const uint32_t mhi = (uint32_t)(ix >> 32) - (thi & 0xFFF00000u);
  R.m = mc_asf64(((uint64_t)mhi << 32) | (uint32_t)ix);
  R.idx = __builtin_amdgcn_ubfe(thi, 11, 9);
#else
  const int e = (int)((int64_t)tmp >> 52);
  R.e = e + eadj;
  R.m = mc_asf64(ix - ((uint64_t)(int64_t)e << 52));
  R.idx = (unsigned)(tmp >> 43) & 511u;
#endif
  R.c = mc_log_tab[R.idx][0];
  R.head = mc_log_tab[R.idx][1];
  R.mid = mc_log_tab[R.idx][2];
  return R;
}

MC_FN double mc_log_eval(const mc_log_row *R, const mc_coef *K) {
  const double r = mc_fma(R->m, R->c, -1.0); /* exact, |r| < 2^-8.67 */

  /* log1p(r) = r - r^2/2 + r^3 * P(r), P = 1/3 - r/4 + ... - r^5/8; the
     omitted tail (r^9/9 - ...) is below R^9/(9(1 - R)) for the table's
     largest |r|, R = 2^-8.67.  Near x = 1 the rows have c = 1 (r = m - 1,
     log x ~ r); gen_math_tables.py bounds the relative error of this whole
     fast path per row (all terms, roundings included). */
  double r2h = r * r;
  double r2l = mc_fma(r, r, -r2h);
  double p = mc_fma(r, K->l8, K->l7);
  p = mc_fma(r, p, K->l6);
  p = mc_fma(r, p, K->l5);
  p = mc_fma(r, p, K->l4);
  p = mc_fma(r, p, K->l3);
  double t = (r2h * r) * p;

  /* e*ln2 + (-log c) + r - r^2/2 + t, accumulated in double-double.
     s = e*ln2_hi + head is exact (both on the 2^-42 grid, |s| < 2^10); the
     next two sums are FastTwoSums: |s| >= |r| unless s == 0 (e != 0 gives
     |s| > 0.34; e == 0 is checked per table row by gen_math_tables.py), and
     s + r then dominates r^2/2. */
  double ed = (double)R->e;
  double s2, serr2, s3, serr3;
  const double s = ed * MC_LN2_HI + R->head;
  double lo = mc_fma(ed, MC_LN2_MID, R->mid);
  mc_fast_two_sum(s, r, &s2, &serr2);
  lo = lo + serr2;
  mc_fast_two_sum(s2, -0.5 * r2h, &s3, &serr3);
  lo = lo + (serr3 + mc_fma(-0.5, r2l, t));
  double hi;
  mc_fast_two_sum(s3, lo, &hi, &lo);
  if (mc_fma(lo, K->ziv, hi) == hi) return hi; /* mc_rounds_to_hi */
  return mc_log_accurate(R->e, r, (int)R->idx);
}

/* x not a positive normal finite number (kept out of line) */
MC_FN_COLD double mc_log_special(double x) {
  uint64_t ix = mc_asu64(x);
  if (x != x) return x + x;                       /* NaN */
  if ((ix << 1) == 0) return -1.0 / 0.0;           /* +-0 -> -inf */
  if (ix >> 63) return (x - x) / (x - x);          /* x < 0 -> NaN */
  if (ix == 0x7FF0000000000000ULL) return x;       /* +inf */
  /* positive subnormal: scaled into the normal range */
  const mc_coef K = mc_coef_default();
  const mc_log_row R = mc_log_lookup(mc_asu64(x * 0x1p54), -54);
  return mc_log_eval(&R, &K);
}

/* x not a positive normal finite number; the subtrahend's low word is zero,
   so the test needs only the high word */
MC_FN int mc_log_is_special(double x) {
  return (uint32_t)(mc_asu64(x) >> 32) - 0x00100000u >= 0x7FF00000u - 0x00100000u;
}

/* The table row of x.  For a special x the row is meaningless but valid
   (the index is a 9-bit field of the bits); mc_log_finish replaces the value,
   and the transport step sends such x to its exact path. */
MC_FN mc_log_row mc_log_start(double x) { return mc_log_lookup(mc_asu64(x), 0); }

MC_FN double mc_log_finish(double x, const mc_log_row *R, const mc_coef *K) {
  const double y = mc_log_eval(R, K);
  return mc_log_is_special(x) ? mc_log_special(x) : y;
}

MC_FN double mc_log_k(double x, const mc_coef *K) {
  const mc_log_row R = mc_log_start(x);
  return mc_log_finish(x, &R, K);
}

MC_FN double mc_log(double x) {
  const mc_coef K = mc_coef_default();
  return mc_log_k(x, &K);
}

/* ------------------------------------------------------------------- exp */

/* exp(xh + xl) as (yh, yl) * 2^k, all steps in double-double; |xh| < 800. */
MC_FN_COLD void mc_exp_dd(double xh, double xl, double *yh, double *yl, int *kexp) {
  double kd = __builtin_rint(xh * MC_INV_LN2_128);
  int k = (int)kd;
  /* r = x - k*ln2/128 with ln2/128 in three pieces */
  double t = xh - kd * MC_LN2_128_1; /* exact */
  double ph = kd * MC_LN2_128_2, pl = mc_fma(kd, MC_LN2_128_2, -ph);
  double rh, rl;
  mc_two_sum(t, -ph, &rh, &rl);
  rl = rl + ((xl - pl) - kd * MC_LN2_128_3);
  mc_fast_two_sum(rh, rl, &rh, &rl);
  /* exp(r) = sum r^n/n!, |r| < 2^-8.4: n >= 6 in double, n <= 5 double-double */
  double tail = 0.0;
  for (int n = 13; n >= 6; --n) tail = mc_fma(tail, rh, mc_inv_fact[n][0]);
  double eh, el;
  {
    double qh = tail * rh, ql = mc_fma(tail, rh, -qh);
    mc_dd_add(qh, ql, mc_inv_fact[5][0], mc_inv_fact[5][1], &eh, &el);
  }
  for (int n = 4; n >= 0; --n) {
    double qh, ql;
    mc_dd_mul(eh, el, rh, rl, &qh, &ql);
    mc_dd_add(qh, ql, mc_inv_fact[n][0], mc_inv_fact[n][1], &eh, &el);
  }
  int j = k & 127;
  mc_dd_mul(mc_exp_tab[j][0], mc_exp_tab[j][1], eh, el, yh, yl);
  *kexp = (k - j) / 128;
}

/* exp(x) for 2^-54 <= |x| < 2^-30.  Near 1 a double-double of the result
   cannot hold cases such as exp(2^-53) = 1 + 2^-53 + 2^-107 + ..., which
   lie 2^-107 above a rounding boundary.  Instead compute u = exp(x) - 1 =
   x + x^2/2 + x^3/6 + x^4/24 in double-double relative to u (error below
   2^-100 |u|, truncation below 2^-120 |u|) and round 1 + u on the grid of
   doubles near 1 (spacing 2^-52 above 1, 2^-53 below). */
MC_FN_COLD double mc_exp_tiny(double x) {
  double x2h = x * x, x2l = mc_fma(x, x, -x2h);
  double x3h = x2h * x, x3l = mc_fma(x2h, x, -x3h) + x2l * x;
  double ch, cl;
  mc_dd_mul(x3h, x3l, mc_inv_fact[3][0], mc_inv_fact[3][1], &ch, &cl);
  double d4 = (x2h * x2h) * mc_inv_fact[4][0];
  double uh, ul;
  mc_dd_add(0.5 * x2h, 0.5 * x2l, ch, cl + d4, &uh, &ul);
  mc_dd_add(x, 0.0, uh, ul, &uh, &ul);
  /* v = |u| in units of the grid spacing, as vh + vl; n = round(v) */
  int up = uh > 0.0;
  double sc = up ? 0x1p52 : -0x1p53;
  double vh = uh * sc, vl = ul * sc;
  double n = __builtin_rint(vh);
  double f = vh - n; /* exact, |f| <= 1/2 */
  if (f - 0.5 > -vl) n += 1.0;        /* f + vl > 1/2 */
  else if (f + 0.5 < -vl) n -= 1.0;   /* f + vl < -1/2 */
  return up ? 1.0 + n * 0x1p-52 : 1.0 - n * 0x1p-53;
}

/* 2^n as a double for n in [-1022, 1023] */
MC_FN double mc_pow2(int n) { return mc_asf64((uint64_t)(n + 1023) << 52); }

MC_FN_COLD double mc_exp_special(double x) {
  if (x != x) return x + x;          /* NaN */
  if (x < -745.2) return 0.0;        /* exp(x) < 2^-1075: rounds to 0 (also -inf) */
  if (x > 709.782712893384) return 1.0 / 0.0; /* overflow (also +inf) */
  double yh, yl;
  int k;
  if (x > 0.0) {
    /* 708.3 .. 709.78: result near DBL_MAX; scale in two steps */
    mc_exp_dd(x, 0.0, &yh, &yl, &k);
    double y = yh + yl;
    return (y * mc_pow2(k - 1)) * 2.0;
  }
  /* subnormal range: exp(x) = z * 2^-1074 with z = exp(x + 1074 ln2);
     round z to the nearest integer (ties to even). */
  double th = 1074.0 * MC_LN2_HI; /* exact */
  double mh = 1074.0 * MC_LN2_MID, ml = mc_fma(1074.0, MC_LN2_MID, -mh);
  ml = ml + 1074.0 * MC_LN2_LO;
  double sh, sl;
  mc_two_sum(x, th, &sh, &sl);
  double s2h, s2l;
  mc_dd_add(sh, sl, mh, ml, &s2h, &s2l);
  mc_exp_dd(s2h, s2l, &yh, &yl, &k);
  /* z = (yh + yl) * 2^k, k <= 53, exact scaling */
  double zs = mc_pow2(k);
  double zh = yh * zs, zl = yl * zs;
  double n = __builtin_rint(zh);
  double f = (zh - n) + zl;
  if (f > 0.5) n += 1.0;
  else if (f < -0.5) n -= 1.0;
  else if (f == 0.5 || f == -0.5) {
    double nn = n + (f > 0 ? 1.0 : -1.0);
    if (__builtin_fmod(nn, 2.0) == 0.0) n = nn;
  }
  return n * 0x1p-1074;
}

MC_FN double mc_exp_k(double x, const mc_coef *K) {
  /* |x| >= 708.3 (and NaN) out of line; the band 708.3 .. 709.7 is exact
     there too */
  if (!(__builtin_fabs(x) < 708.3)) return mc_exp_special(x);
  /* Tiny |x| (zero and subnormals included) need no test of their own:
     kd = 0, r = x, and the fast path's error stays relative to exp(x) ~ 1
     (a failed rounding test goes to mc_exp_tiny below). */
  /* k = the integer nearest x*1024/ln2, from one fma against 1.5*2^52: the
     sum is rounded to an integer (|x*1024/ln2| < 2^20), whose value sits in
     the low bits of the double; kd = k exactly.  (Any integer within 1/2 of
     x*1024/ln2 serves: it only chooses the reduction.) */
  const double shifted = mc_fma(x, MC_INV_LN2_1024, 0x1.8p52);
  const double kd = shifted - 0x1.8p52;
  const int k = (int)(uint32_t)mc_asu64(shifted);
  const int j = k & 1023;
  const double th = mc_exp_tab10[j][0], tl = mc_exp_tab10[j][1];
#if defined(__HIP_DEVICE_COMPILE__)
  /* keep the table row's load ahead of the series below (the scheduler
     would otherwise issue it just before its use and wait for it there) */
  __builtin_amdgcn_sched_barrier(0);
#endif
  /* r = x - kd*ln2/1024 as rh + rl, |kd| < 2^20, ln2/1024 = A + B (+ below
     2^-97), |r| < 2^-11.5.  t = x - kd*A is exact (kd*A has at most 53 bits;
     t is a multiple of min(ulp(x), 2^-43) below 2^-11 in magnitude).
     rh = RN(t - kd*B); t - rh = kd*B - (rounding of rh) is below 2^-24 +
     2^-65, so its rounding errs by less than 2^-77; rl = RN((t - rh) - kd*B)
     then leaves |rh + rl - r| < 2^-76. */
  const double t = mc_fma(-kd, MC_LN2_1024_A, x);
  const double rh = mc_fma(-kd, MC_LN2_1024_B, t);
  const double rl = mc_fma(-kd, MC_LN2_1024_B, t - rh);
  /* exp(r) = 1 + r + b, b = r^2 (1/2 + r/6 + r^2/24 + r^3/120), |b| < 2^-24,
     evaluated in plain double; the head 1 + rh by FastTwoSum.  Errors,
     relative to the result: series tail r^6/720 < 2^-78.7; the roundings
     of b, of rl + b and of el below 2^-76.5 each; rl * (exp(rh) - 1)
     dropped, below 2^-76.5; the product by 2^(j/1024) = th + tl (fma
     remainders, th * el rounded) below 2^-75.5; the reduction below 2^-76.
     In all below 2^-72. */
  double c = mc_fma(rh, K->e5, K->e4);
  c = mc_fma(rh, c, K->e3);
  c = mc_fma(rh, c, 0.5);
  const double b = (rh * rh) * c;
  double eh, el;
  mc_fast_two_sum(1.0, rh, &eh, &el);
  el = el + (rl + b);
  double yh = th * eh;
  double yl = mc_fma(th, eh, -yh);
  yl = yl + mc_fma(th, el, tl * eh);
  mc_fast_two_sum(yh, yl, &yh, &yl);
  /* times 2^((k - j)/1024) = 2^(k >> 10): exact, the result being normal for
     |x| < 708.3 */
  if (mc_fma(yl, K->ziv, yh) == yh) return __builtin_ldexp(yh, k >> 10); /* mc_rounds_to_hi */
  if (__builtin_fabs(x) < 0x1p-30) return mc_exp_tiny(x);
  int kk;
  mc_exp_dd(x, 0.0, &yh, &yl, &kk);
  return (yh + yl) * mc_pow2(kk);
}

MC_FN double mc_exp(double x) {
  const mc_coef K = mc_coef_default();
  return mc_exp_k(x, &K);
}

/* mc_exp_k in two halves, for kernels that load the table row one step
   before they evaluate: mc_exp_lookup(x) reads the row mc_exp_k reads for x
   (any x, NaN included, gives a row of the table), and
   mc_exp_k_row(x, row, K) == mc_exp_k(x, K) bit for bit: the same
   operations, the row taken from the argument. */
typedef struct mc_exp_row {
  double th, tl;
} mc_exp_row;

MC_FN mc_exp_row mc_exp_lookup(double x) {
  const double shifted = mc_fma(x, MC_INV_LN2_1024, 0x1.8p52);
  const int j = (int)(uint32_t)mc_asu64(shifted) & 1023;
  mc_exp_row R;
  R.th = mc_exp_tab10[j][0];
  R.tl = mc_exp_tab10[j][1];
  return R;
}

MC_FN double mc_exp_k_row(double x, const mc_exp_row *R, const mc_coef *K) {
  if (!(__builtin_fabs(x) < 708.3)) return mc_exp_special(x);
  const double shifted = mc_fma(x, MC_INV_LN2_1024, 0x1.8p52);
  const double kd = shifted - 0x1.8p52;
  const int k = (int)(uint32_t)mc_asu64(shifted);
  const double th = R->th, tl = R->tl;
  /* from here as mc_exp_k (the error analysis there applies unchanged) */
  const double t = mc_fma(-kd, MC_LN2_1024_A, x);
  const double rh = mc_fma(-kd, MC_LN2_1024_B, t);
  const double rl = mc_fma(-kd, MC_LN2_1024_B, t - rh);
  double c = mc_fma(rh, K->e5, K->e4);
  c = mc_fma(rh, c, K->e3);
  c = mc_fma(rh, c, 0.5);
  const double b = (rh * rh) * c;
  double eh, el;
  mc_fast_two_sum(1.0, rh, &eh, &el);
  el = el + (rl + b);
  double yh = th * eh;
  double yl = mc_fma(th, eh, -yh);
  yl = yl + mc_fma(th, el, tl * eh);
  mc_fast_two_sum(yh, yl, &yh, &yl);
  if (mc_fma(yl, K->ziv, yh) == yh) return __builtin_ldexp(yh, k >> 10);
  if (__builtin_fabs(x) < 0x1p-30) return mc_exp_tiny(x);
  int kk;
  mc_exp_dd(x, 0.0, &yh, &yl, &kk);
  return (yh + yl) * mc_pow2(kk);
}

#endif /* MC_MATH_H */
