/*
 * oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU oracle for mc-mpi's Layer::simulate: an operation-by-operation C
 * restatement of src/layer.cpp and src/random.cpp (reference paths), built as liboracle.so and
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
#ifdef _OPENMP
#include <omp.h>
#endif

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

/* ============================================================ adjoint
 * Gradient of a weighted tally T = sum_c g_c * tally_c with respect to the
 * per-cell cross sections the step uses, sig_a_c and sig_i_c (src/layer.cpp
 * :131-133), by replaying each history from its initial particle.
 *
 * For event j of a history (cell c_j, flight length d_j, weight w_j before
 * and w_{j+1} = w_j - dw_j after), with P_j = sum_{k<j} g_{c_k} dw_k and the
 * history total T_h = sum_k g_{c_k} dw_k:
 *   pathwise (sig_a enters only through exp(-sig_a d)):
 *     dT/dsig_a_c += d_j * (g_{c_j} w_{j+1} - (T_h - P_j - g_{c_j} dw_j))
 *   score function (sig_i sets the flight-length density; the term of event
 *   j multiplies only the tally from j on, T_h - P_j):
 *     dT/dsig_i_c += ([collision_j] / sig_i - d_j) * (T_h - P_j)
 *   for events with sig_i > EPS_PRECISION (otherwise the flight is not
 *   sampled and the term is 0).
 * The weight kill (src/layer.cpp:195) is not differentiated.
 * Terms are summed exactly in the same fixed point as the tallies.
 */

/* one event with the cell values given directly (same arithmetic as
   orc64_particle_step after its sig_a / sig_i products) */
static inline double orc64_event_cells(double sig_a, double sig_i, double dx, orc64_particle *p,
                                       double *di_out, int *coll_out) {
  const double h = orc64_rnd_real(&p->seed);
  double di = sig_i > 1e-4 ? -orc64_log(h) / sig_i : DBL_MAX;
  int index_new;
  double x_new_edge;
  if (p->mu < 0) {
    index_new = p->index - 1;
    x_new_edge = p->index * dx;
  } else {
    index_new = p->index + 1;
    x_new_edge = (p->index + 1) * dx;
  }
  double di_edge = DBL_MAX;
  if (p->mu < -1e-4 || 1e-4 < p->mu) di_edge = (x_new_edge - p->x) / p->mu;
  int coll = di < di_edge;
  if (coll) {
    index_new = p->index;
    p->x += di * p->mu;
    p->mu = 2 * orc64_rnd_real(&p->seed) - 1;
  } else {
    di = di_edge;
    p->x = x_new_edge;
  }
  const double dw = (1 - orc64_exp(-sig_a * di)) * p->wmc;
  p->wmc -= dw;
  p->index = index_new;
  *di_out = di;
  *coll_out = coll;
  return dw;
}

static inline int orc64_alive(const orc64_params *L, const orc64_particle *p) {
  return (p->wmc >= L->min_weight) && (p->index < L->index_start + L->m) &&
         (p->index >= L->index_start);
}

/* Forward with direct cell values; optional g -> per-history score. */
int orc64_forward_cells(const orc64_params *L, const double *sig_a, const double *sig_i,
                        orc64_particle *particles, int64_t n, uint64_t *fx_pairs, uint32_t *flags,
                        const double *g, double *score, int32_t *result, int nthread) {
  const int m = L->m;
  orc_fx_t *fx = (orc_fx_t *)aligned_alloc(16, sizeof(orc_fx_t) * (size_t)(m > 0 ? m : 1) * (nthread + 1));
  uint32_t *fl = (uint32_t *)calloc((size_t)m * (nthread + 1), sizeof(uint32_t));
  if (!fx || !fl) return -1;
  memset(fx, 0, sizeof(orc_fx_t) * (size_t)m * (nthread + 1));
#pragma omp parallel for num_threads(nthread) schedule(dynamic, 256)
  for (int64_t p = 0; p < n; ++p) {
    int t = 0;
#ifdef _OPENMP
    t = omp_get_thread_num();
#endif
    orc_fx_t *tf = fx + (size_t)t * m;
    uint32_t *tg = fl + (size_t)t * m;
    orc64_particle *q = &particles[p];
    double T = 0.0;
    while (orc64_alive(L, q)) {
      const int c = q->index - L->index_start;
      double di;
      int coll;
      const double dw = orc64_event_cells(sig_a[c], sig_i[c], L->dx, q, &di, &coll);
      orc_fx_add(&tf[c], &tg[c], dw);
      if (g) T = T + g[c] * dw;
    }
    if (score) score[p] = T;
    if (result) {
      int r;
      if (q->index == L->index_start - 1) r = -1;
      else if (q->index == L->index_start + L->m) r = 1;
      else if (q->wmc < L->min_weight) r = 0;
      else r = -1;
      result[p] = r;
    }
  }
  for (int t = 0; t < nthread; ++t)
    for (int c = 0; c < m; ++c) {
      orc_fx_t v = ((orc_fx_t)fx_pairs[2 * c + 1] << 64) | fx_pairs[2 * c];
      v += fx[(size_t)t * m + c];
      fx_pairs[2 * c] = (uint64_t)v;
      fx_pairs[2 * c + 1] = (uint64_t)(v >> 64);
      flags[c] |= fl[(size_t)t * m + c];
    }
  free(fx);
  free(fl);
  return 0;
}

/* Adjoint replay; score may be NULL (then computed by a first replay). */
int orc64_adjoint_cells(const orc64_params *L, const double *sig_a, const double *sig_i,
                        const orc64_particle *initial, int64_t n, const double *g,
                        const double *score, uint64_t *ga_pairs, uint64_t *gi_pairs,
                        uint32_t *flags, int nthread) {
  const int m = L->m;
  const size_t cells = (size_t)(m > 0 ? m : 1) * (nthread + 1);
  orc_fx_t *ga = (orc_fx_t *)aligned_alloc(16, sizeof(orc_fx_t) * cells);
  orc_fx_t *gi = (orc_fx_t *)aligned_alloc(16, sizeof(orc_fx_t) * cells);
  uint32_t *fl = (uint32_t *)calloc(cells, sizeof(uint32_t));
  if (!ga || !gi || !fl) return -1;
  memset(ga, 0, sizeof(orc_fx_t) * cells);
  memset(gi, 0, sizeof(orc_fx_t) * cells);
#pragma omp parallel for num_threads(nthread) schedule(dynamic, 256)
  for (int64_t p = 0; p < n; ++p) {
    int t = 0;
#ifdef _OPENMP
    t = omp_get_thread_num();
#endif
    orc_fx_t *ta = ga + (size_t)t * m, *ti = gi + (size_t)t * m;
    uint32_t *tg = fl + (size_t)t * m;
    double T = 0.0;
    if (score) {
      T = score[p];
    } else {
      orc64_particle q = initial[p];
      while (orc64_alive(L, &q)) {
        const int c = q.index - L->index_start;
        double di;
        int coll;
        const double dw = orc64_event_cells(sig_a[c], sig_i[c], L->dx, &q, &di, &coll);
        T = T + g[c] * dw;
      }
    }
    orc64_particle q = initial[p];
    double P = 0.0;
    while (orc64_alive(L, &q)) {
      const int c = q.index - L->index_start;
      double di;
      int coll;
      const double si = sig_i[c];
      const double dw = orc64_event_cells(sig_a[c], si, L->dx, &q, &di, &coll);
      const double gdw = g[c] * dw;
      const double t_from = T - P;
      const double a_after = t_from - gdw;
#if ORC_MUTANT == 6 /* planted mistake: pathwise term without the downstream tally */
      const double term_a = di * (g[c] * q.wmc);
      (void)a_after;
#else
      const double term_a = di * (g[c] * q.wmc - a_after);
#endif
      orc_fx_add(&ta[c], &tg[c], term_a);
      if (si > 1e-4) {
#if ORC_MUTANT == 7 /* planted mistake: sign of the flight-length score */
        const double s = coll ? (1.0 / si) + di : di;
#elif ORC_MUTANT == 8 /* planted mistake: collision density term dropped */
        const double s = -di;
        (void)coll;
#else
        const double s = coll ? (1.0 / si) - di : -di;
#endif
        orc_fx_add(&ti[c], &tg[c], s * t_from);
      }
      P = P + gdw;
    }
  }
  for (int t = 0; t < nthread; ++t)
    for (int c = 0; c < m; ++c) {
      orc_fx_t v = ((orc_fx_t)ga_pairs[2 * c + 1] << 64) | ga_pairs[2 * c];
      v += ga[(size_t)t * m + c];
      ga_pairs[2 * c] = (uint64_t)v;
      ga_pairs[2 * c + 1] = (uint64_t)(v >> 64);
      v = ((orc_fx_t)gi_pairs[2 * c + 1] << 64) | gi_pairs[2 * c];
      v += gi[(size_t)t * m + c];
      gi_pairs[2 * c] = (uint64_t)v;
      gi_pairs[2 * c + 1] = (uint64_t)(v >> 64);
      flags[c] |= fl[(size_t)t * m + c];
    }
  free(ga);
  free(gi);
  free(fl);
  return 0;
}
