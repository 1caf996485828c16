/*
 * adjoint.hip -- gradient of a weighted tally by history replay.
 *
 * Target: T = sum_c g_c * tally_c, where tally_c is the absorbed weight of
 * cell c (src/layer.cpp:174-179).  Parameters: the per-cell cross sections
 * the step uses, sig_a_c and sig_i_c (src/layer.cpp:131-133).
 *
 * The forward pass (history kernel with target weights) stores, per history,
 * T_h = sum_j g_{c_j} dw_j.  The adjoint kernel replays every history from
 * its initial particle -- the particle's seed is its RNG state, so the replay
 * recomputes the path instead of storing it -- and for each event j (cell
 * c, flight length d_j, weight w_{j+1} after the event, collision flag,
 * P_j = sum_{k<j} g_{c_k} dw_k) adds
 *   pathwise, sig_a:      d_j * (g_c * w_{j+1} - (T_h - P_j - g_c * dw_j))
 *   score function, sig_i: ([collision] / sig_i - d_j) * (T_h - P_j)
 *                          (only where sig_i > EPS_PRECISION: there the flight
 *                          length is sampled with density sig_i e^{-sig_i d})
 * to two exact fixed-point per-cell tallies (same arithmetic as the absorbed
 * weight tally).  The kill at the minimum weight is not differentiated.
 * When T_h was not stored, each lane first replays the history once to
 * compute it, then replays it again for the gradient.
 */
#include "internal.h"

namespace {

struct adj_args {
  mc_layer_k L;
  int64_t n;
  mcgpu_bank init; /* read only */
  const mc_cell *cells;
  const double *g;
  const double *score; /* may be null */
  unsigned long long *fx; /* ga lo [0,m), ga hi [m,2m), gi lo [2m,3m), gi hi [3m,4m) */
  uint32_t *flags;        /* ga flags [0,m), gi flags [m,2m) */
  mc_counters *ctr;
  uint64_t max_events;
};

template <bool LDS, bool HAVE_SCORE>
__global__ __launch_bounds__(1024) void k_adjoint(adj_args a, int refill_k) {
  extern __shared__ __align__(16) unsigned char smem[];
  const mc_coef K = mc_coef_live();
  const int m = a.L.m;
  unsigned long long *tl = a.fx; /* 4m words */
  const mc_cell *cells = a.cells;
  const double *g = a.g;
  if (LDS) {
    tl = reinterpret_cast<unsigned long long *>(smem);
    mc_cell *lc = reinterpret_cast<mc_cell *>(smem + 32 * (size_t)m);
    double *lg = reinterpret_cast<double *>(smem + 48 * (size_t)m);
    for (int c = threadIdx.x; c < m; c += blockDim.x) {
      tl[c] = 0;
      tl[m + c] = 0;
      tl[2 * m + c] = 0;
      tl[3 * m + c] = 0;
      lc[c] = a.cells[c];
      lg[c] = a.g[c];
    }
    cells = lc;
    g = lg;
  }
  __syncthreads();

  const int lane = __lane_id();
  const uint32_t max_ev = a.max_events > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)a.max_events;
  mc_state p;
  p.index = 0;
  int64_t idx = 0;
  bool has = false, done = false, capped = false;
  int phase = 1;       /* 0: score replay, 1: gradient replay */
  uint32_t ev = 0;
  double T = 0.0, P = 0.0;
  unsigned long long tot_ev = 0;

  /* events between refills: refill_k if positive, else adapted per wave as
     in the history kernel (halved when many lanes were idle at a refill,
     doubled when almost none were) */
  const bool adapt = refill_k <= 0;
  int kk = adapt ? 128 : refill_k;
  for (;;) {
    if (!done) {
      const unsigned long long need = __ballot(!has);
      if (adapt) {
        const int idle = __popcll(need);
        kk = idle > 16 ? (kk > 8 ? kk >> 1 : 8) : (idle < 4 ? (kk < 256 ? kk << 1 : 256) : kk);
      }
      if (need) {
        const int cnt = __popcll(need);
        const int leader = __ffsll((long long)need) - 1;
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd(&a.ctr->next, (unsigned long long)cnt);
        base = mc_readlane64(base, leader);
        if (!has) {
          const int64_t i = (int64_t)(base + (unsigned long long)mc_mbcnt64(need));
          if (i < a.n) {
            p.seed = a.init.seed[i];
            p.x = a.init.x[i];
            p.mu = a.init.mu[i];
            mc_refresh_rcp(p);
            p.wmc = a.init.wmc[i];
            p.index = a.init.index[i];
            idx = i;
            has = true;
            ev = 0;
            P = 0.0;
            if (HAVE_SCORE) {
              T = a.score[i];
              phase = 1;
            } else {
              T = 0.0;
              phase = 0;
            }
          }
        }
        if (base + (unsigned long long)cnt >= (unsigned long long)a.n) done = true;
      }
    }
    if (__ballot(has) == 0) break;
    if (has) {
      for (int k = 0; k < kk; ++k) {
        if (!mc_alive(p, a.L) || ev >= max_ev) break;
        const int c = p.index - a.L.index_start;
        mc_cell cell;
        bool coll;
        double di;
        const double dw = mc_event<true>(p, cells, c, a.L.dx, K, coll, di, &cell);
        ++ev;
        if (!HAVE_SCORE && phase == 0) {
          T = T + g[c] * dw;
          continue;
        }
        const double gc = g[c];
        double term_a, term_i;
        const bool has_i = mc_adjoint_terms(cell, coll, di, dw, gc, p.wmc, T, P, term_a, term_i);
        mc_tally<true>(tl, tl + m, a.flags, &a.ctr->status, c, term_a);
        if (has_i) mc_tally<true>(tl + 2 * m, tl + 3 * m, a.flags + m, &a.ctr->status, c, term_i);
        P = P + gc * dw;
      }
      const bool alive = mc_alive(p, a.L);
      if (!alive || ev >= max_ev) {
        if (alive) capped = true;
        tot_ev += ev;
        if (!HAVE_SCORE && phase == 0 && !alive) {
          /* score known: replay the same history for the gradient */
          p.seed = a.init.seed[idx];
          p.x = a.init.x[idx];
          p.mu = a.init.mu[idx];
          mc_refresh_rcp(p);
          p.wmc = a.init.wmc[idx];
          p.index = a.init.index[idx];
          ev = 0;
          P = 0.0;
          phase = 1;
        } else {
          has = false;
        }
      }
    }
  }

  tot_ev = mc_wave_sum(tot_ev);
  if (lane == 0) atomicAdd(&a.ctr->events, tot_ev);
  if (capped) atomicOr(&a.ctr->status, 1u);
  if (LDS) {
    __syncthreads();
    for (int c = threadIdx.x; c < m; c += blockDim.x) {
      mc_fx v;
      v.lo = tl[c];
      v.hi = tl[m + c];
      mc_fx_atomic_add(&a.fx[c], &a.fx[m + c], v);
      v.lo = tl[2 * m + c];
      v.hi = tl[3 * m + c];
      mc_fx_atomic_add(&a.fx[2 * m + c], &a.fx[3 * m + c], v);
    }
  }
}

__global__ void k_setup_cells(int m, const double *sig_a, const double *sig_i, mc_cell *cells,
                              unsigned long long *fx, int fx_words_per_cell, uint32_t *flags,
                              int flag_words_per_cell, mc_counters *ctr) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < m) {
    mc_cell k;
    k.sig_a = sig_a[c];
    k.sig_i = sig_i[c];
    cells[c] = k;
    for (int w = 0; w < fx_words_per_cell; ++w) fx[(size_t)w * m + c] = 0;
    for (int w = 0; w < flag_words_per_cell; ++w) flags[(size_t)w * m + c] = 0;
  }
  if (c == 0) {
    ctr->next = 0;
    ctr->events = 0;
    ctr->collisions = 0;
    ctr->n_left = 0;
    ctr->n_right = 0;
    ctr->n_killed = 0;
    ctr->status = 0;
  }
}

__global__ void k_finalize2(int m, const unsigned long long *fx, const uint32_t *flags,
                            double *ga, double *gi) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < m) {
    ga[c] = ga[c] + mc_fx_finalize(fx[c], fx[m + c], flags[c]);
    gi[c] = gi[c] + mc_fx_finalize(fx[2 * m + c], fx[3 * m + c], flags[m + c]);
  }
}

__global__ void k_finalize1(int m, const unsigned long long *fx, const uint32_t *flags,
                            double *w) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < m) w[c] = w[c] + mc_fx_finalize(fx[c], fx[m + c], flags[c]);
}

int read_counters(mcgpu_context *ctx, hipStream_t s, mcgpu_stats *stats) {
  MC_HIP(ctx, hipMemcpyAsync(ctx->host_counters, ctx->counters, sizeof(mc_counters),
                             hipMemcpyDeviceToHost, s));
  MC_HIP(ctx, hipStreamSynchronize(s));
  const mc_counters *hc = (const mc_counters *)ctx->host_counters;
  if (stats) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    stats->events = hc->events;
    stats->collisions = hc->collisions;
    stats->n_left = (int64_t)hc->n_left;
    stats->n_right = (int64_t)hc->n_right;
    stats->n_killed = (int64_t)hc->n_killed;
    stats->kernel_ms = ms;
  }
  if (hc->status & 1u)
    return mc_fail(ctx, MCGPU_ERR_EVENT_CAP, "a history exceeded max_events_per_history");
  if (hc->status & 2u) return mc_fail(ctx, MCGPU_ERR_RANGE, "a tally term was outside +-2^39");
  return MCGPU_OK;
}

int cells_setup(mcgpu_context *ctx, int m, const double *sig_a, const double *sig_i, int fx_words,
                int flag_words, hipStream_t s) {
  int rc;
  if ((rc = mc_ensure(ctx, &ctx->cells, &ctx->cells_bytes, sizeof(mc_cell) * (size_t)m))) return rc;
  if ((rc = mc_ensure(ctx, &ctx->fx, &ctx->fx_bytes, 8 * (size_t)fx_words * m))) return rc;
  if ((rc = mc_ensure(ctx, &ctx->flags, &ctx->flags_bytes, 4 * (size_t)flag_words * m))) return rc;
  hipLaunchKernelGGL(k_setup_cells, dim3((m + 255) / 256), dim3(256), 0, s, m, sig_a, sig_i,
                     (mc_cell *)ctx->cells, (unsigned long long *)ctx->fx, fx_words,
                     (uint32_t *)ctx->flags, flag_words, (mc_counters *)ctx->counters);
  MC_HIP(ctx, hipGetLastError());
  return MCGPU_OK;
}

template <bool LDS, bool S>
int launch_adj(mcgpu_context *ctx, const adj_args &a, const mcgpu_options &o, hipStream_t s,
               size_t lds) {
  auto kern = k_adjoint<LDS, S>;
  if (lds > 65536)
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  int per_cu = 0;
  int block = o.block_size;
  if (block <= 0) {
    block = mc_auto_block(kern, lds, &per_cu);
  } else if (block > 1024 || block % 64) {
    return mc_fail(ctx, MCGPU_ERR_ARG, "block_size must be a multiple of 64 and <= 1024");
  } else if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, block, lds) !=
                 hipSuccess || per_cu < 1) {
    per_cu = 1;
  }
  int grid = o.grid_blocks;
  if (grid <= 0) {
    grid = per_cu * ctx->num_cus;
    const int64_t blocks_needed = (a.n + block - 1) / block;
    if (blocks_needed < grid) grid = (int)(blocks_needed > 0 ? blocks_needed : 1);
  }
  const int k = o.events_per_refill > 0 ? o.events_per_refill : 0; /* 0: adaptive */
  hipLaunchKernelGGL(kern, dim3(grid), dim3(block), lds, s, a, k);
  MC_HIP(ctx, hipGetLastError());
  return MCGPU_OK;
}

} // namespace

extern "C" {

int mcgpu_forward_cells(mcgpu_context *ctx, int64_t n, const mcgpu_bank *bank,
                        const double *sig_a, const double *sig_i, double *weights_absorbed,
                        int32_t min_index, int32_t max_index, double dx,
                        double particle_min_weight, const double *g, double *history_score,
                        int8_t *result, const mcgpu_options *opt, mcgpu_stats *stats,
                        void *stream) {
  if (!ctx) return MCGPU_ERR_ARG;
  ctx->err.clear();
  if (n < 0 || !bank || !sig_a || !sig_i || !weights_absorbed || max_index <= min_index ||
      (history_score && !g))
    return mc_fail(ctx, MCGPU_ERR_ARG, "forward_cells: bad arguments");
  MC_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const int m = max_index - min_index;
  int rc = cells_setup(ctx, m, sig_a, sig_i, 2, 1, s);
  if (rc) return rc;
  mcgpu_options o;
  mcgpu_default_options(&o);
  if (opt) o = *opt;
  if (history_score && o.kernel != MCGPU_KERNEL_HISTORY)
    return mc_fail(ctx, MCGPU_ERR_ARG, "per-history scores need the history kernel");
  mc_transport_args a;
  memset(&a, 0, sizeof a);
  a.L.index_start = min_index;
  a.L.m = m;
  a.L.dx = dx;
  a.L.min_weight = particle_min_weight;
  a.n = n;
  a.soa = *bank;
  a.result8 = result;
  a.cells = (const mc_cell *)ctx->cells;
  a.fx = (unsigned long long *)ctx->fx;
  a.flags = (uint32_t *)ctx->flags;
  a.ctr = (mc_counters *)ctx->counters;
  a.max_events = o.max_events_per_history ? o.max_events_per_history : (1ULL << 27);
  a.g = g;
  a.score = history_score;
  MC_HIP(ctx, hipEventRecord(ctx->ev0, s));
  if (n > 0) {
    rc = o.kernel == MCGPU_KERNEL_EVENT ? mc_launch_event(ctx, a, o, s) : mc_launch_history(ctx, a, o, s);
    if (rc) return rc;
  }
  MC_HIP(ctx, hipEventRecord(ctx->ev1, s));
  hipLaunchKernelGGL(k_finalize1, dim3((m + 255) / 256), dim3(256), 0, s, m,
                     (const unsigned long long *)ctx->fx, (const uint32_t *)ctx->flags,
                     weights_absorbed);
  MC_HIP(ctx, hipGetLastError());
  return read_counters(ctx, s, stats);
}

int mcgpu_adjoint_cells(mcgpu_context *ctx, int64_t n, const mcgpu_bank *initial,
                        const double *sig_a, const double *sig_i, int32_t min_index,
                        int32_t max_index, double dx, double particle_min_weight,
                        const double *g, const double *history_score, double *grad_sig_a,
                        double *grad_sig_i, const mcgpu_options *opt, mcgpu_stats *stats,
                        void *stream) {
  if (!ctx) return MCGPU_ERR_ARG;
  ctx->err.clear();
  if (n < 0 || !initial || !sig_a || !sig_i || !g || !grad_sig_a || !grad_sig_i ||
      max_index <= min_index)
    return mc_fail(ctx, MCGPU_ERR_ARG, "adjoint_cells: bad arguments");
  if (n > 0 && (!initial->seed || !initial->x || !initial->mu || !initial->wmc || !initial->index))
    return mc_fail(ctx, MCGPU_ERR_ARG, "adjoint_cells: bank has a null array");
  MC_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const int m = max_index - min_index;
  int rc = cells_setup(ctx, m, sig_a, sig_i, 4, 2, s);
  if (rc) return rc;
  mcgpu_options o;
  mcgpu_default_options(&o);
  if (opt) o = *opt;
  adj_args a;
  memset(&a, 0, sizeof a);
  a.L.index_start = min_index;
  a.L.m = m;
  a.L.dx = dx;
  a.L.min_weight = particle_min_weight;
  a.n = n;
  a.init = *initial;
  a.cells = (const mc_cell *)ctx->cells;
  a.g = g;
  a.score = history_score;
  a.fx = (unsigned long long *)ctx->fx;
  a.flags = (uint32_t *)ctx->flags;
  a.ctr = (mc_counters *)ctx->counters;
  a.max_events = o.max_events_per_history ? o.max_events_per_history : (1ULL << 27);
  const size_t lds = 56 * (size_t)m; /* two tallies, cells, target weights */
  const bool in_lds = lds <= ctx->lds_per_block;
  MC_HIP(ctx, hipEventRecord(ctx->ev0, s));
  if (n > 0) {
    if (history_score)
      rc = in_lds ? launch_adj<true, true>(ctx, a, o, s, lds) : launch_adj<false, true>(ctx, a, o, s, 0);
    else
      rc = in_lds ? launch_adj<true, false>(ctx, a, o, s, lds) : launch_adj<false, false>(ctx, a, o, s, 0);
    if (rc) return rc;
  }
  MC_HIP(ctx, hipEventRecord(ctx->ev1, s));
  hipLaunchKernelGGL(k_finalize2, dim3((m + 255) / 256), dim3(256), 0, s, m,
                     (const unsigned long long *)ctx->fx, (const uint32_t *)ctx->flags,
                     grad_sig_a, grad_sig_i);
  MC_HIP(ctx, hipGetLastError());
  return read_counters(ctx, s, stats);
}

} /* extern "C" */
