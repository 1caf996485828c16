/*
 * mc_device.h -- device building blocks of the transport kernels:
 * the reference's two LCGs, the per-cell cross-section table, the exact
 * fixed-point tally arithmetic and one transport event.
 *
 * Everything here is FP64 and must be compiled with -ffp-contract=off so the
 * arithmetic is exactly the sequence of IEEE operations the reference's
 * expressions specify (no fused multiply-adds that the reference does not do).
 */
#ifndef MC_DEVICE_H
#define MC_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mc_math.h"

/* ------------------------------------------------------------------- RNG */
/* src/random.cpp:8-16: s <- (G*s + C) mod 2^63; real = s * 2^-63 */
#define MC_RNG_G 6364136223846793005ULL
#define MC_RNG_C 1442695040888963407ULL
/* src/random.cpp:20-29: seed chain */
#define MC_RNGS_G 5177284530976225183ULL
#define MC_RNGS_C 2096348467109453893ULL
#define MC_MASK63 0x7FFFFFFFFFFFFFFFULL

__device__ __forceinline__ double mc_rnd_real(uint64_t &seed) {
  seed = (MC_RNG_G * seed + MC_RNG_C) & MC_MASK63;
  /* (double)seed * 2^-63, the conversion rounded to nearest and the scaling
     exact, formed as one rounding of hi*2^-31 + lo*2^-63 (both exact).  The
     high word is taken with a bit-field extract: a plain shift and truncation
     is widened by the compiler into a 64-bit conversion with an extra add. */
  const double hi = (double)__builtin_amdgcn_ubfe((uint32_t)(seed >> 32), 0, 31);
  const double lo = (double)(uint32_t)seed;
  return __builtin_fma(hi, 0x1p-31, lo * 0x1p-63);
}

__device__ __forceinline__ uint64_t mc_rnd_seed(uint64_t &seed) {
  seed = (MC_RNGS_G * seed + MC_RNGS_C) & MC_MASK63;
  return seed;
}

/* The k-th output of rnd_seed from `seed` (src/random.cpp:24-29), k >= 1,
   via the affine map f(s) = G s + C raised to the power k by squaring; the
   map is taken mod 2^64 and reduced mod 2^63 at the end, which is exact
   because 2^63 divides 2^64. */
__host__ __device__ __forceinline__ uint64_t mc_seed_chain_at(uint64_t seed, uint64_t k) {
  uint64_t A = MC_RNGS_G, B = MC_RNGS_C; /* f^(2^j) */
  uint64_t RA = 1, RB = 0;               /* accumulated power */
  while (k) {
    if (k & 1) {
      RA = A * RA;
      RB = A * RB + B;
    }
    B = A * B + B;
    A = A * A;
    k >>= 1;
  }
  return (RA * seed + RB) & MC_MASK63;
}

/* ------------------------------------------------------ cross sections */
/* Per cell, the two products particle_step forms (src/layer.cpp:131-133):
   sig_a = sigs*ar and sig_i = sigs*(1.0 - ar), computed once with the same
   operations so every event reads the bit-identical values. */
struct mc_cell {
  double sig_a;
  double sig_i;
};
static_assert(sizeof(mc_cell) == 16, "cell offsets are formed as c << 4");

__device__ __forceinline__ mc_cell mc_make_cell(double sigs, double ar) {
  const double interaction_rate = 1.0 - ar;
  mc_cell c;
  c.sig_a = sigs * ar;
  c.sig_i = sigs * interaction_rate;
  return c;
}

#include "mc_fx.h"

/* 128-bit add into two 64-bit words with 64-bit atomics (LDS or global):
   the carry out of the low word is detected from the returned old value. */
__device__ __forceinline__ void mc_fx_atomic_add(unsigned long long *lo_word,
                                                 unsigned long long *hi_word, mc_fx v) {
  if ((v.lo | v.hi) == 0) return;
  unsigned long long old = atomicAdd(lo_word, (unsigned long long)v.lo);
  unsigned long long carry = (old + v.lo < old) ? 1ULL : 0ULL;
  unsigned long long h = v.hi + carry;
  if (h) atomicAdd(hi_word, h);
}

/* The same without the test of the whole term: both words are added even
   when zero (adding zero changes nothing; the transport's deposits are
   almost never zero, and the tests and their branches cost more than the
   rare extra atomic).  HI_COND: the high word only when it changes -- for
   particle weights below 2^-24, whose deposits add nothing there but their
   carries; then the skipped atomic costs more than the test. */
template <bool HI_COND>
__device__ __forceinline__ void mc_fx_atomic_add_always(unsigned long long *lo_word,
                                                        unsigned long long *hi_word, mc_fx v) {
  const unsigned long long old = atomicAdd(lo_word, (unsigned long long)v.lo);
  const unsigned long long carry = (old + v.lo < old) ? 1ULL : 0ULL;
  const unsigned long long h = (unsigned long long)v.hi + carry;
  if (!HI_COND || h) atomicAdd(hi_word, h);
}

/* --------------------------------------------------------- particle state */
struct mc_state {
  uint64_t seed;
  double x;
  double mu;
  double wmc;
  int32_t index;
  double inv_mu; /* RN(1/mu), kept by the kernels that divide by reciprocal */
};

__device__ __forceinline__ void mc_refresh_rcp(mc_state &p) { p.inv_mu = 1.0 / p.mu; }

/* Layer constants read by the step (include/layer/layer.hpp:465-485) */
struct mc_layer_k {
  int32_t index_start;
  int32_t m;
  double dx;
  double min_weight;
};

/* EPS_PRECISION and MAXREAL of include/types/types.hpp:6-22, real_t = double */
#define MC_EPS 1e-4
#define MC_MAXREAL 1.7976931348623157e308

/* the log/exp coefficients as values the compiler cannot fold; made once at
   kernel entry and passed to every mc_event of the kernel */
__device__ __forceinline__ mc_coef mc_coef_live() {
  mc_coef K = mc_coef_default();
  asm volatile("" : "+v"(K.l8), "+v"(K.l7), "+v"(K.l6), "+v"(K.l5), "+v"(K.l4), "+v"(K.l3));
  asm volatile("" : "+v"(K.e5), "+v"(K.e4), "+v"(K.e3), "+v"(K.ziv));
  return K;
}

/* One transport event: Layer::particle_step (src/layer.cpp:123-190) in FP64
   with correctly rounded log/exp, on cell c (cells[c]).  Returns the
   absorbed weight dw (tallied by the caller to cell c); `collided` reports
   which branch was taken and di_out the flight length.
   The reference draws the flight length di = -log(h) / sig_i and the edge
   distance di_edge, and the particle collides iff di < di_edge (both
   correctly rounded quotients here).  On fine meshes most events cross the
   cell edge, and then di itself is never used.  For h in [0, 1],
   -log(h) >= 1 - h, so
     1 - h > sig_i * di_edge * (1 + 2^-40)  (computed in floating point)
   implies the exact -log(h) / sig_i > di_edge, hence RN(di) >= di_edge and
   the particle crosses: such an event needs no log at all.  Only when that
   is not certain (h near 1, a collision, a thick cell, |mu| <= EPS) are the
   correctly rounded log and the division formed and compared exactly, as the
   reference does.  Results are identical either way.
   RCP: the edge distance divides by the particle's stored RN(1/mu)
   (mc_div_by_rcp; same result), refreshed when mu changes, for kernels that
   keep the state in registers. */
template <bool RCP = false, typename Count = uint64_t>
__device__ __forceinline__ double mc_event(mc_state &p, const mc_cell *cells, int c,
                                           const double dx, const mc_coef &K, bool &collided,
                                           double &di_out, mc_cell *cell_out = nullptr,
                                           Count *col_count = nullptr) {
  /* The edge quotient is formed unconditionally and then selected; when
     |mu| <= EPS the discarded quotient may be inf or NaN and is never used.
     *cell_out (if given) receives the cell; *col_count (if given) is
     incremented on a collision, inside the collision's own branch. */
  /* an unsigned 32-bit byte offset (c is in [0, m)): the load takes the
     uniform base in scalar registers and the offset as is */
  const mc_cell cell =
      *reinterpret_cast<const mc_cell *>(reinterpret_cast<const char *>(cells) + ((uint32_t)c << 4));
  if (cell_out) *cell_out = cell;
  const double h = mc_rnd_real(p.seed);

  const bool back = p.mu < 0;
  int index_new = back ? p.index - 1 : p.index + 1;
  const double x_new_edge = (back ? p.index : p.index + 1) * dx;
  const double q_edge = RCP ? mc_div_by_rcp(x_new_edge - p.x, p.mu, p.inv_mu)
                            : (x_new_edge - p.x) / p.mu;
  const double di_edge = (p.mu < -MC_EPS || MC_EPS < p.mu) ? q_edge : MC_MAXREAL;

  const double reach = cell.sig_i * di_edge;
  /* bitwise, not short-circuit: one test, no branch */
  const int sure_cross =
      (int)!(cell.sig_i > MC_EPS) | (int)((1.0 - h) > mc_fma(reach, 0x1p-40, reach));
  double di = di_edge;
  collided = false;
  if (!sure_cross) {
    const double nlog = -mc_log_k(h, &K);
    const double dq = cell.sig_i > MC_EPS ? nlog / cell.sig_i : MC_MAXREAL;
    collided = dq < di_edge;
    if (collided) di = dq;
  }
  if (collided) {
    index_new = p.index;
    p.x += di * p.mu;
    p.mu = 2 * mc_rnd_real(p.seed) - 1;
    if (RCP) mc_refresh_rcp(p);
    if (col_count) ++*col_count;
  } else {
    p.x = x_new_edge;
  }

  const double dw = (1 - mc_exp_k(-cell.sig_a * di, &K)) * p.wmc;
  p.wmc -= dw;
  p.index = index_new;
  di_out = di;
  return dw;
}

/* mc_event in two halves, for a kernel that loads the exp table row one
   step before it evaluates (the event kernel's pipeline).  mc_event_decide
   is mc_event up to the exp argument: the draw, the edge distance (IEEE
   division), the collision decision (certified crossing, else the correctly
   rounded log and division), the collision's new x and mu, or the crossing's
   x; it returns -sig_a * di and the new cell index.  The caller then forms
   dw = (1 - mc_exp_k_row(arg, row, K)) * wmc, wmc -= dw, index = index_new,
   as mc_event does.  The operations and their order are mc_event's. */
__device__ __forceinline__ double mc_event_decide(mc_state &p, const mc_cell cell, const double dx,
                                                  const mc_coef &K, bool &collided,
                                                  int &index_new) {
  const double h = mc_rnd_real(p.seed);
  const bool back = p.mu < 0;
  index_new = back ? p.index - 1 : p.index + 1;
  const double x_new_edge = (back ? p.index : p.index + 1) * dx;
  const double q_edge = (x_new_edge - p.x) / p.mu;
  const double di_edge = (p.mu < -MC_EPS || MC_EPS < p.mu) ? q_edge : MC_MAXREAL;
  const double reach = cell.sig_i * di_edge;
  const int sure_cross =
      (int)!(cell.sig_i > MC_EPS) | (int)((1.0 - h) > mc_fma(reach, 0x1p-40, reach));
  double di = di_edge;
  collided = false;
  if (!sure_cross) {
    const double nlog = -mc_log_k(h, &K);
    const double dq = cell.sig_i > MC_EPS ? nlog / cell.sig_i : MC_MAXREAL;
    collided = dq < di_edge;
    if (collided) di = dq;
  }
  if (collided) {
    index_new = p.index;
    p.x += di * p.mu;
    p.mu = 2 * mc_rnd_real(p.seed) - 1;
  } else {
    p.x = x_new_edge;
  }
  return -cell.sig_a * di;
}

/* src/layer.cpp:195-197: the history continues while this holds */
/* index_start <= index < index_start + m as one unsigned compare (the
   difference wraps to a large value below index_start) */
__device__ __forceinline__ bool mc_alive(const mc_state &p, const mc_layer_k &L) {
  return (p.wmc >= L.min_weight) &&
         ((unsigned)p.index - (unsigned)L.index_start < (unsigned)L.m);
}

/* src/layer.cpp:202-217: exit code of a finished history */
__device__ __forceinline__ int mc_exit_code(const mc_state &p, const mc_layer_k &L) {
  if (p.index == L.index_start - 1) return -1;
  if (p.index == L.index_start + L.m) return +1;
  if (p.wmc < L.min_weight) return 0;
  return -1;
}

/* -------------------------------------------------------------- wave utils */
/* value of v in lane `lane` (wave-uniform lane index) */
__device__ __forceinline__ unsigned long long mc_readlane64(unsigned long long v, int lane) {
  unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)v, lane);
  unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(v >> 32), lane);
  return ((unsigned long long)hi << 32) | lo;
}

/* number of set bits of `mask` in lanes below this one (prefix count) */
__device__ __forceinline__ int mc_mbcnt64(unsigned long long mask) {
  return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                        __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

/* the largest v over the wave, as a wave-uniform (scalar) value; every lane
   of the wave must take part */
__device__ __forceinline__ unsigned long long mc_wave_max_u64(unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(v, off);
    v = o > v ? o : v;
  }
  return mc_readlane64(v, 0);
}

__device__ __forceinline__ unsigned long long mc_wave_sum(unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

/* add one tally term (exact fixed point) to cell c of (lo, hi); non-finite or
   out-of-range terms set the cell's flag word and the call's status */
/* SIGNED: terms of either sign are common (adjoint); otherwise negative terms
   take the general conversion (same result, fewer instructions for the
   transport's non-negative deposits). */
/* on_flags(g) records the flag word g of a term the general conversion
   flagged (never called for the common terms) */
template <bool SIGNED = false, bool HI_COND = false, typename OnFlags>
__device__ __forceinline__ void mc_tally_with(unsigned long long *lo, unsigned long long *hi, int c,
                                              double dw, OnFlags on_flags) {
  mc_fx v;
  int fast;
  MC_FX_FAST(!SIGNED, dw, v, fast);
  if (!fast) { /* flags can only come from the general conversion */
    const mc_fx_flagged g = mc_fx_from_double_general(dw);
    v = g.v;
    if (g.flags) on_flags(g.flags);
  }
  if (SIGNED)
    mc_fx_atomic_add(&lo[c], &hi[c], v);
  else
    mc_fx_atomic_add_always<HI_COND>(&lo[c], &hi[c], v);
}

template <bool SIGNED = false, bool HI_COND = false>
__device__ __forceinline__ void mc_tally(unsigned long long *lo, unsigned long long *hi,
                                         uint32_t *flags, unsigned int *status, int c, double dw) {
  mc_tally_with<SIGNED, HI_COND>(lo, hi, c, dw, [&](uint32_t g) {
    atomicOr(&flags[c], g);
    if (g & 8u) atomicOr(status, 2u);
  });
}

/* Gradient terms of one replayed event for the target T = sum g_c tally_c
   (cell c, flight length di, collision flag, deposit dw, weight wmc after the
   event, history total T and partial sum P = sum of g dw before the event):
     pathwise, sig_a:       di * (g_c * wmc - (T - P - g_c * dw))
     score function, sig_i: ([coll] / sig_i - di) * (T - P), only where
                            sig_i > EPS_PRECISION (sampled flight length).
   Returns whether the sig_i term exists. */
__device__ __forceinline__ bool mc_adjoint_terms(const mc_cell cell, bool coll, double di,
                                                 double dw, double gc, double wmc, double T,
                                                 double P, double &term_a, double &term_i) {
  const double gdw = gc * dw;
  const double t_from = T - P;
  const double a_after = t_from - gdw;
  term_a = di * (gc * wmc - a_after);
  if (cell.sig_i > MC_EPS) {
    const double s = coll ? (1.0 / cell.sig_i) - di : -di;
    term_i = s * t_from;
    return true;
  }
  term_i = 0.0;
  return false;
}

#endif /* MC_DEVICE_H */
