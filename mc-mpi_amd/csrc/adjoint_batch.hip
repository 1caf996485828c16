/*
 * adjoint_batch.hip -- forward-with-scores and adjoint over many independent
 * slabs, one persistent launch each.
 *
 * The single-slab entry points (mcgpu_forward_cells, mcgpu_adjoint_cells)
 * fill the GPU when a slab has millions of histories; a batch of slabs shares
 * one launch instead, as the batched transport does:
 *  - the host cuts every slab's particles into tiles of particles_per_lane x
 *    block_size histories and orders them by estimated work (histories x
 *    cells), largest first;
 *  - a persistent block takes the next tile from a global counter, stages the
 *    slab's cells, target weights and (zeroed) fixed-point accumulators in
 *    LDS when they fit (otherwise it adds straight into the global ones),
 *    runs the tile's histories with lanes refilled from an LDS counter, and
 *    adds its accumulators to the slab's global cells once per tile.
 * Per history the arithmetic is that of the single-slab kernels (the same
 * step, the same score and gradient terms in event order, exact sums), so
 * every slab's outputs are bit-identical to running it alone.
 */
#include <algorithm>
#include <vector>

#include "internal.h"

namespace {

struct bslab_k {
  int32_t index_start;
  int32_t m;
  double dx;
  double min_weight;
  int64_t cell_off;
};

struct btile_k {
  int32_t slab;
  int32_t pad_;
  int64_t begin; /* absolute particle indices [begin, end) */
  int64_t end;
};

struct bcell_args {
  const bslab_k *slabs;
  const btile_k *tiles;
  int32_t n_tiles;
  int32_t lds_max_m;      /* slabs with m <= this stage their cells in LDS */
  mcgpu_bank bank;        /* forward: in/out; adjoint: initial bank, read only */
  const mc_cell *cells;   /* [total] */
  const double *g;        /* [total] */
  const double *score_in; /* adjoint: per-history scores, may be null */
  double *score_out;      /* forward: per-history scores, may be null */
  int8_t *result8;        /* forward: exit codes, may be null */
  unsigned long long *fx; /* forward: 2 words per cell; adjoint: 4 (ga lo, ga hi, gi lo, gi hi) */
  int64_t total;
  uint32_t *flags;        /* forward: [total]; adjoint: [2 total] */
  mc_counters *ctr;
  uint64_t max_events;
};

/* block-uniform: claim the next tile */
__device__ __forceinline__ int next_tile(mc_counters *ctr, int *s_tile) {
  if (threadIdx.x == 0) *s_tile = (int)atomicAdd(&ctr->next, 1ull);
  __syncthreads();
  return *s_tile;
}

/* Per-lane counters of one block, carried across its tiles. */
struct blane {
  bool capped;
  unsigned long long ev, col;
};

/* One tile of the forward: its histories, lanes refilled from the LDS queue
   s_next.  Called with LDS or global accumulator/cell pointers from two
   places, so each inlined copy uses the matching memory instructions. */
template <bool SCORE>
__device__ __forceinline__ void fwd_tile(const bcell_args &a, const btile_k &tile,
                                         const bslab_k &sl, const mc_layer_k &L,
                                         unsigned long long *t_lo, unsigned long long *t_hi,
                                         const mc_cell *cells, const double *g,
                                         unsigned long long *s_next, unsigned long long *s_tot,
                                         uint32_t max_ev, int refill_k, const mc_coef &K,
                                         blane &bl) {
  const int lane = __lane_id();
  mc_state p;
  p.index = 0;
  int64_t idx = 0;
  bool has = false, done = false;
  uint32_t ev = 0, col = 0;
  double score = 0.0;
  for (;;) {
    if (!done) {
      const unsigned long long need = __ballot(!has);
      if (need) {
        const int cnt = __popcll(need);
        const int leader = __ffsll((long long)need) - 1;
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd(s_next, (unsigned long long)cnt);
        base = mc_readlane64(base, leader);
        if (!has) {
          const int64_t i = (int64_t)(base + (unsigned long long)mc_mbcnt64(need));
          if (i < tile.end) {
            p.seed = a.bank.seed[i];
            p.x = a.bank.x[i];
            p.mu = a.bank.mu[i];
            mc_refresh_rcp(p);
            p.wmc = a.bank.wmc[i];
            p.index = a.bank.index[i];
            idx = i;
            has = true;
            ev = 0;
            col = 0;
            score = 0.0;
          }
        }
        if (base + (unsigned long long)cnt >= (unsigned long long)tile.end) done = true;
      }
    }
    if (__ballot(has) == 0) break;
    if (has) {
      for (int k = 0; k < refill_k; ++k) {
        if (!mc_alive(p, L) || ev >= max_ev) break;
        const int c = p.index - L.index_start;
        bool coll;
        double di;
        const double dw = mc_event<true>(p, cells, c, L.dx, K, coll, di);
        ++ev;
        col += coll ? 1u : 0u;
        mc_tally(t_lo, t_hi, a.flags + sl.cell_off, &a.ctr->status, c, dw);
        if (SCORE) score = score + g[c] * dw;
      }
      const bool alive = mc_alive(p, L);
      if (!alive || ev >= max_ev) {
        int code = mc_exit_code(p, L);
        if (alive) {
          bl.capped = true;
          code = 2;
        }
        a.bank.seed[idx] = p.seed;
        a.bank.x[idx] = p.x;
        a.bank.mu[idx] = p.mu;
        a.bank.wmc[idx] = p.wmc;
        a.bank.index[idx] = p.index;
        if (a.result8) a.result8[idx] = (int8_t)code;
        if (SCORE) a.score_out[idx] = score;
        bl.ev += ev;
        bl.col += col;
        atomicAdd(&s_tot[code == -1 ? 2 : (code == 1 ? 3 : 4)], code == 2 ? 0ull : 1ull);
        has = false;
      }
    }
  }
}

/* Forward: mcgpu_forward_cells per slab (history kernel arithmetic). */
template <bool SCORE>
__global__ __launch_bounds__(1024) void k_forward_cells_batch(bcell_args a, int refill_k) {
  extern __shared__ __align__(16) unsigned char smem[];
  const mc_coef K = mc_coef_live();
  __shared__ int s_tile;
  __shared__ unsigned long long s_next;
  __shared__ unsigned long long s_tot[5];
  const int lane = __lane_id();
  const int B = blockDim.x, tid = threadIdx.x;
  if (tid < 5) s_tot[tid] = 0;
  blane bl;
  bl.capped = false;
  bl.ev = bl.col = 0;
  const uint32_t max_ev = a.max_events > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)a.max_events;
  unsigned long long *t_lds = reinterpret_cast<unsigned long long *>(smem);

  for (;;) {
    const int t = next_tile(a.ctr, &s_tile);
    if (t >= a.n_tiles) break;
    const btile_k tile = a.tiles[t];
    const bslab_k sl = a.slabs[tile.slab];
    const int m = sl.m;
    mc_layer_k L;
    L.index_start = sl.index_start;
    L.m = m;
    L.dx = sl.dx;
    L.min_weight = sl.min_weight;
    const bool in_lds = m <= a.lds_max_m; /* block-uniform */
    mc_cell *lc = reinterpret_cast<mc_cell *>(smem + (size_t)16 * m);
    double *lg = reinterpret_cast<double *>(smem + (size_t)32 * m);
    if (in_lds) {
      for (int c = tid; c < m; c += B) {
        t_lds[c] = 0;
        t_lds[m + c] = 0;
        lc[c] = a.cells[sl.cell_off + c];
        if (SCORE) lg[c] = a.g[sl.cell_off + c];
      }
    }
    if (tid == 0) s_next = (unsigned long long)tile.begin;
    __syncthreads();

    if (in_lds)
      fwd_tile<SCORE>(a, tile, sl, L, t_lds, t_lds + m, lc, lg, &s_next, s_tot, max_ev,
                      refill_k, K, bl);
    else
      fwd_tile<SCORE>(a, tile, sl, L, a.fx + sl.cell_off, a.fx + a.total + sl.cell_off,
                      a.cells + sl.cell_off, a.g ? a.g + sl.cell_off : nullptr, &s_next, s_tot,
                      max_ev, refill_k, K, bl);

    __syncthreads();
    if (in_lds) {
      for (int c = tid; c < m; c += B) {
        mc_fx v;
        v.lo = t_lds[c];
        v.hi = t_lds[m + c];
        mc_fx_atomic_add(&a.fx[sl.cell_off + c], &a.fx[a.total + sl.cell_off + c], v);
      }
    }
    __syncthreads();
  }

  {
    const unsigned long long e = mc_wave_sum(bl.ev), cl = mc_wave_sum(bl.col);
    if (lane == 0) {
      atomicAdd(&s_tot[0], e);
      atomicAdd(&s_tot[1], cl);
    }
  }
  if (bl.capped) atomicOr(&a.ctr->status, 1u);
  __syncthreads();
  if (tid == 0) {
    atomicAdd(&a.ctr->events, s_tot[0]);
    atomicAdd(&a.ctr->collisions, s_tot[1]);
    atomicAdd(&a.ctr->n_left, s_tot[2]);
    atomicAdd(&a.ctr->n_right, s_tot[3]);
    atomicAdd(&a.ctr->n_killed, s_tot[4]);
  }
}

/* One tile of the adjoint (replay from the initial bank); pointers as in
   fwd_tile. */
template <bool HAVE_SCORE>
__device__ __forceinline__ void adj_tile(const bcell_args &a, const btile_k &tile,
                                         const bslab_k &sl, const mc_layer_k &L,
                                         unsigned long long *ga_lo, unsigned long long *ga_hi,
                                         unsigned long long *gi_lo, unsigned long long *gi_hi,
                                         const mc_cell *cells, const double *g,
                                         unsigned long long *s_next, uint32_t max_ev,
                                         int refill_k, const mc_coef &K, blane &bl) {
  const int lane = __lane_id();
  uint32_t *fa = a.flags + sl.cell_off, *fi = a.flags + a.total + sl.cell_off;
  mc_state p;
  p.index = 0;
  int64_t idx = 0;
  bool has = false, done = false;
  int phase = 1; /* 0: score replay, 1: gradient replay */
  uint32_t ev = 0;
  double T = 0.0, P = 0.0;
  for (;;) {
    if (!done) {
      const unsigned long long need = __ballot(!has);
      if (need) {
        const int cnt = __popcll(need);
        const int leader = __ffsll((long long)need) - 1;
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd(s_next, (unsigned long long)cnt);
        base = mc_readlane64(base, leader);
        if (!has) {
          const int64_t i = (int64_t)(base + (unsigned long long)mc_mbcnt64(need));
          if (i < tile.end) {
            p.seed = a.bank.seed[i];
            p.x = a.bank.x[i];
            p.mu = a.bank.mu[i];
            mc_refresh_rcp(p);
            p.wmc = a.bank.wmc[i];
            p.index = a.bank.index[i];
            idx = i;
            has = true;
            ev = 0;
            P = 0.0;
            if (HAVE_SCORE) {
              T = a.score_in[i];
              phase = 1;
            } else {
              T = 0.0;
              phase = 0;
            }
          }
        }
        if (base + (unsigned long long)cnt >= (unsigned long long)tile.end) done = true;
      }
    }
    if (__ballot(has) == 0) break;
    if (has) {
      for (int k = 0; k < refill_k; ++k) {
        if (!mc_alive(p, L) || ev >= max_ev) break;
        const int c = p.index - L.index_start;
        mc_cell cell;
        bool coll;
        double di;
        const double dw = mc_event<true>(p, cells, c, L.dx, K, coll, di, &cell);
        ++ev;
        if (!HAVE_SCORE && phase == 0) {
          T = T + g[c] * dw;
          continue;
        }
        const double gc = g[c];
        double term_a, term_i;
        const bool has_i = mc_adjoint_terms(cell, coll, di, dw, gc, p.wmc, T, P, term_a, term_i);
        mc_tally<true>(ga_lo, ga_hi, fa, &a.ctr->status, c, term_a);
        if (has_i) mc_tally<true>(gi_lo, gi_hi, fi, &a.ctr->status, c, term_i);
        P = P + gc * dw;
      }
      const bool alive = mc_alive(p, L);
      if (!alive || ev >= max_ev) {
        if (alive) bl.capped = true;
        bl.ev += ev;
        if (!HAVE_SCORE && phase == 0 && !alive) {
          p.seed = a.bank.seed[idx];
          p.x = a.bank.x[idx];
          p.mu = a.bank.mu[idx];
          mc_refresh_rcp(p);
          p.wmc = a.bank.wmc[idx];
          p.index = a.bank.index[idx];
          ev = 0;
          P = 0.0;
          phase = 1;
        } else {
          has = false;
        }
      }
    }
  }
}

/* Adjoint: mcgpu_adjoint_cells per slab (replay from the initial bank). */
template <bool HAVE_SCORE>
__global__ __launch_bounds__(1024) void k_adjoint_cells_batch(bcell_args a, int refill_k) {
  extern __shared__ __align__(16) unsigned char smem[];
  const mc_coef K = mc_coef_live();
  __shared__ int s_tile;
  __shared__ unsigned long long s_next;
  const int lane = __lane_id();
  const int B = blockDim.x, tid = threadIdx.x;
  blane bl;
  bl.capped = false;
  bl.ev = bl.col = 0;
  const uint32_t max_ev = a.max_events > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)a.max_events;
  const int64_t T4 = a.total;
  unsigned long long *w_lds = reinterpret_cast<unsigned long long *>(smem);

  for (;;) {
    const int t = next_tile(a.ctr, &s_tile);
    if (t >= a.n_tiles) break;
    const btile_k tile = a.tiles[t];
    const bslab_k sl = a.slabs[tile.slab];
    const int m = sl.m;
    mc_layer_k L;
    L.index_start = sl.index_start;
    L.m = m;
    L.dx = sl.dx;
    L.min_weight = sl.min_weight;
    const bool in_lds = m <= a.lds_max_m; /* block-uniform */
    mc_cell *lc = reinterpret_cast<mc_cell *>(smem + (size_t)32 * m);
    double *lg = reinterpret_cast<double *>(smem + (size_t)48 * m);
    if (in_lds) {
      for (int c = tid; c < m; c += B) {
        w_lds[c] = 0;
        w_lds[m + c] = 0;
        w_lds[2 * m + c] = 0;
        w_lds[3 * m + c] = 0;
        lc[c] = a.cells[sl.cell_off + c];
        lg[c] = a.g[sl.cell_off + c];
      }
    }
    if (tid == 0) s_next = (unsigned long long)tile.begin;
    __syncthreads();

    if (in_lds)
      adj_tile<HAVE_SCORE>(a, tile, sl, L, w_lds, w_lds + m, w_lds + 2 * m, w_lds + 3 * m, lc, lg,
                           &s_next, max_ev, refill_k, K, bl);
    else
      adj_tile<HAVE_SCORE>(a, tile, sl, L, a.fx + sl.cell_off, a.fx + T4 + sl.cell_off,
                           a.fx + 2 * T4 + sl.cell_off, a.fx + 3 * T4 + sl.cell_off,
                           a.cells + sl.cell_off, a.g + sl.cell_off, &s_next, max_ev, refill_k,
                           K, bl);

    __syncthreads();
    if (in_lds) {
      for (int c = tid; c < m; c += B) {
        mc_fx v;
        v.lo = w_lds[c];
        v.hi = w_lds[m + c];
        mc_fx_atomic_add(&a.fx[sl.cell_off + c], &a.fx[T4 + sl.cell_off + c], v);
        v.lo = w_lds[2 * m + c];
        v.hi = w_lds[3 * m + c];
        mc_fx_atomic_add(&a.fx[2 * T4 + sl.cell_off + c], &a.fx[3 * T4 + sl.cell_off + c], v);
      }
    }
    __syncthreads();
  }
  const unsigned long long e = mc_wave_sum(bl.ev);
  if (lane == 0) atomicAdd(&a.ctr->events, e);
  if (bl.capped) atomicOr(&a.ctr->status, 1u);
}

/* cells from (sig_a, sig_i); zeroed accumulators (words per cell) and counters */
__global__ void k_batch_cells_setup(int64_t total, const double *sig_a, const double *sig_i,
                                    mc_cell *cells, unsigned long long *fx, int fx_words,
                                    uint32_t *flags, int flag_words, mc_counters *ctr) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < total) {
    mc_cell k;
    k.sig_a = sig_a[c];
    k.sig_i = sig_i[c];
    cells[c] = k;
    for (int w = 0; w < fx_words; ++w) fx[(size_t)w * total + c] = 0;
    for (int w = 0; w < flag_words; ++w) flags[(size_t)w * total + c] = 0;
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

/* out_k[c] += the k-th accumulator (k < n_out) */
__global__ void k_batch_cells_finalize(int64_t total, const unsigned long long *fx,
                                       const uint32_t *flags, double *out0, double *out1) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= total) return;
  out0[c] = out0[c] + mc_fx_finalize(fx[c], fx[total + c], flags[c]);
  if (out1) out1[c] = out1[c] + mc_fx_finalize(fx[2 * total + c], fx[3 * total + c], flags[total + c]);
}

/* Validates the slab table and builds the device slab and tile tables in
   ctx->tables; returns the largest slab's cell count in *max_m. */
int build_tables(mcgpu_context *ctx, int32_t n_slabs, const mcgpu_slab *slabs,
                 int64_t total_cells, int64_t tile_size, const bslab_k **d_slabs,
                 const btile_k **d_tiles, int32_t *n_tiles, int *max_m, int64_t *max_particle) {
  std::vector<bslab_k> ks(n_slabs);
  std::vector<btile_k> tiles;
  std::vector<double> work;
  *max_m = 0;
  *max_particle = 0;
  for (int s = 0; s < n_slabs; ++s) {
    const mcgpu_slab &h = slabs[s];
    const int m = h.max_index - h.min_index;
    if (m <= 0 || h.n < 0 || h.cell_offset < 0 || h.cell_offset + m > total_cells ||
        h.particle_offset < 0)
      return mc_fail(ctx, MCGPU_ERR_ARG, "batch: invalid slab descriptor");
    ks[s].index_start = h.min_index;
    ks[s].m = m;
    ks[s].dx = h.dx;
    ks[s].min_weight = h.particle_min_weight;
    ks[s].cell_off = h.cell_offset;
    *max_m = std::max(*max_m, m);
    *max_particle = std::max(*max_particle, h.particle_offset + h.n);
    for (int64_t p0 = 0; p0 < h.n; p0 += tile_size) {
      btile_k t;
      t.slab = s;
      t.pad_ = 0;
      t.begin = h.particle_offset + p0;
      t.end = h.particle_offset + std::min(h.n, p0 + tile_size);
      tiles.push_back(t);
      work.push_back((double)(t.end - t.begin) * (double)m);
    }
  }
  std::vector<size_t> order(tiles.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return work[x] > work[y]; });
  std::vector<btile_k> sorted(tiles.size());
  for (size_t i = 0; i < order.size(); ++i) sorted[i] = tiles[order[i]];
  const size_t slab_bytes = (sizeof(bslab_k) * ks.size() + 255) / 256 * 256;
  const size_t bytes = slab_bytes + sizeof(btile_k) * (sorted.size() + 1);
  int rc = mc_ensure(ctx, &ctx->tables, &ctx->tables_bytes, bytes);
  if (rc) return rc;
  unsigned char *base = (unsigned char *)ctx->tables;
  MC_HIP(ctx, hipMemcpy(base, ks.data(), sizeof(bslab_k) * ks.size(), hipMemcpyHostToDevice));
  if (!sorted.empty())
    MC_HIP(ctx, hipMemcpy(base + slab_bytes, sorted.data(), sizeof(btile_k) * sorted.size(),
                          hipMemcpyHostToDevice));
  *d_slabs = (const bslab_k *)base;
  *d_tiles = (const btile_k *)(base + slab_bytes);
  *n_tiles = (int32_t)sorted.size();
  return MCGPU_OK;
}

int read_batch_counters(mcgpu_context *ctx, hipStream_t s, mcgpu_stats *stats) {
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

template <typename K>
int launch_persistent(mcgpu_context *ctx, K kern, const bcell_args &a, const mcgpu_options &o,
                      size_t lds, hipStream_t s) {
  if (lds > 65536)
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  const int block = o.block_size > 0 ? o.block_size : 1024;
  int grid = o.grid_blocks;
  if (grid <= 0) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, block, lds) != hipSuccess ||
        per_cu < 1)
      per_cu = 1;
    grid = per_cu * ctx->num_cus;
  }
  const int k = o.events_per_refill > 0 ? o.events_per_refill : 8;
  hipLaunchKernelGGL(kern, dim3(grid), dim3(block), lds, s, a, k);
  MC_HIP(ctx, hipGetLastError());
  return MCGPU_OK;
}

/* Common set-up of both batched calls: options, tables, cells, accumulators,
   LDS size.  fx_words / flag_words: accumulator words per cell. */
int batch_prepare(mcgpu_context *ctx, int32_t n_slabs, const mcgpu_slab *slabs,
                  int64_t total_cells, const mcgpu_bank *bank, const double *sig_a,
                  const double *sig_i, int fx_words, int flag_words, size_t lds_bytes_per_cell,
                  const mcgpu_options *opt, hipStream_t s, mcgpu_options *o, bcell_args *a,
                  size_t *lds) {
  mcgpu_default_options(o);
  if (opt) *o = *opt;
  const int block = o->block_size > 0 ? o->block_size : 1024;
  if (block > 1024 || block % 64)
    return mc_fail(ctx, MCGPU_ERR_ARG, "block_size must be a multiple of 64 and <= 1024");
  o->block_size = block;
  const int ppl = o->particles_per_lane > 0 ? o->particles_per_lane : 16;
  memset(a, 0, sizeof *a);
  int max_m = 0;
  int64_t max_particle = 0;
  int rc = build_tables(ctx, n_slabs, slabs, total_cells, (int64_t)ppl * block, &a->slabs,
                        &a->tiles, &a->n_tiles, &max_m, &max_particle);
  if (rc) return rc;
  if (max_particle > 0 && (!bank->seed || !bank->x || !bank->mu || !bank->wmc || !bank->index))
    return mc_fail(ctx, MCGPU_ERR_ARG, "batch: bank has a null array");
  if ((rc = mc_ensure(ctx, &ctx->cells, &ctx->cells_bytes, sizeof(mc_cell) * (size_t)total_cells)))
    return rc;
  if ((rc = mc_ensure(ctx, &ctx->fx, &ctx->fx_bytes, 8 * (size_t)fx_words * total_cells))) return rc;
  if ((rc = mc_ensure(ctx, &ctx->flags, &ctx->flags_bytes, 4 * (size_t)flag_words * total_cells)))
    return rc;
  hipLaunchKernelGGL(k_batch_cells_setup, dim3((unsigned)((total_cells + 255) / 256)), dim3(256), 0,
                     s, total_cells, sig_a, sig_i, (mc_cell *)ctx->cells,
                     (unsigned long long *)ctx->fx, fx_words, (uint32_t *)ctx->flags, flag_words,
                     (mc_counters *)ctx->counters);
  MC_HIP(ctx, hipGetLastError());
  /* LDS for the largest slab that fits; larger slabs use the global accumulators */
  const size_t budget = ctx->lds_per_block - 128; /* the kernels' static LDS */
  int lds_max_m = (int)std::min<size_t>((size_t)max_m, budget / lds_bytes_per_cell);
  *lds = lds_bytes_per_cell * (size_t)lds_max_m;
  a->lds_max_m = lds_max_m;
  a->bank = *bank;
  a->cells = (const mc_cell *)ctx->cells;
  a->fx = (unsigned long long *)ctx->fx;
  a->total = total_cells;
  a->flags = (uint32_t *)ctx->flags;
  a->ctr = (mc_counters *)ctx->counters;
  a->max_events = o->max_events_per_history ? o->max_events_per_history : (1ULL << 27);
  return MCGPU_OK;
}

} // namespace

extern "C" {

int mcgpu_forward_cells_batch(mcgpu_context *ctx, int32_t n_slabs, const mcgpu_slab *slabs,
                              int64_t total_cells, const mcgpu_bank *bank, const double *sig_a,
                              const double *sig_i, double *weights_absorbed, const double *g,
                              double *history_score, int8_t *result, const mcgpu_options *opt,
                              mcgpu_stats *stats, void *stream) {
  if (!ctx) return MCGPU_ERR_ARG;
  ctx->err.clear();
  if (n_slabs <= 0 || !slabs || total_cells <= 0 || !bank || !sig_a || !sig_i ||
      !weights_absorbed || (history_score && !g))
    return mc_fail(ctx, MCGPU_ERR_ARG, "forward_cells_batch: bad arguments");
  MC_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  mcgpu_options o;
  bcell_args a;
  size_t lds = 0;
  const bool score = history_score != nullptr;
  int rc = batch_prepare(ctx, n_slabs, slabs, total_cells, bank, sig_a, sig_i, 2, 1,
                         score ? 40 : 32, opt, s, &o, &a, &lds);
  if (rc) return rc;
  a.g = g;
  a.score_out = history_score;
  a.result8 = result;
  MC_HIP(ctx, hipEventRecord(ctx->ev0, s));
  if (a.n_tiles > 0) {
    rc = score ? launch_persistent(ctx, k_forward_cells_batch<true>, a, o, lds, s)
               : launch_persistent(ctx, k_forward_cells_batch<false>, a, o, lds, s);
    if (rc) return rc;
  }
  MC_HIP(ctx, hipEventRecord(ctx->ev1, s));
  hipLaunchKernelGGL(k_batch_cells_finalize, dim3((unsigned)((total_cells + 255) / 256)), dim3(256),
                     0, s, total_cells, (const unsigned long long *)ctx->fx,
                     (const uint32_t *)ctx->flags, weights_absorbed, (double *)nullptr);
  MC_HIP(ctx, hipGetLastError());
  return read_batch_counters(ctx, s, stats);
}

int mcgpu_adjoint_cells_batch(mcgpu_context *ctx, int32_t n_slabs, const mcgpu_slab *slabs,
                              int64_t total_cells, const mcgpu_bank *initial,
                              const double *sig_a, const double *sig_i, const double *g,
                              const double *history_score, double *grad_sig_a,
                              double *grad_sig_i, const mcgpu_options *opt, mcgpu_stats *stats,
                              void *stream) {
  if (!ctx) return MCGPU_ERR_ARG;
  ctx->err.clear();
  if (n_slabs <= 0 || !slabs || total_cells <= 0 || !initial || !sig_a || !sig_i || !g ||
      !grad_sig_a || !grad_sig_i)
    return mc_fail(ctx, MCGPU_ERR_ARG, "adjoint_cells_batch: bad arguments");
  MC_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  mcgpu_options o;
  bcell_args a;
  size_t lds = 0;
  int rc = batch_prepare(ctx, n_slabs, slabs, total_cells, initial, sig_a, sig_i, 4, 2, 56, opt,
                         s, &o, &a, &lds);
  if (rc) return rc;
  a.g = g;
  a.score_in = history_score;
  MC_HIP(ctx, hipEventRecord(ctx->ev0, s));
  if (a.n_tiles > 0) {
    rc = history_score ? launch_persistent(ctx, k_adjoint_cells_batch<true>, a, o, lds, s)
                       : launch_persistent(ctx, k_adjoint_cells_batch<false>, a, o, lds, s);
    if (rc) return rc;
  }
  MC_HIP(ctx, hipEventRecord(ctx->ev1, s));
  hipLaunchKernelGGL(k_batch_cells_finalize, dim3((unsigned)((total_cells + 255) / 256)), dim3(256),
                     0, s, total_cells, (const unsigned long long *)ctx->fx,
                     (const uint32_t *)ctx->flags, grad_sig_a, grad_sig_i);
  MC_HIP(ctx, hipGetLastError());
  return read_batch_counters(ctx, s, stats);
}

} /* extern "C" */
