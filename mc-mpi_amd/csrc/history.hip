/*
 * history.hip -- history-based transport kernel.
 *
 * One history per lane, run to termination with the loop condition of
 * Layer::simulate_particle (src/layer.cpp:195-200) around the event of
 * Layer::particle_step (src/layer.cpp:123-190).  Unlike the reference CUDA
 * kernel (src/culayer_kernel.cu:19-113: one launch per 325 steps, host
 * round trips and re-sorting between launches), the grid is persistent:
 * a lane whose history ends takes the next particle index from a global
 * work queue (one atomic per wave per refill), so lanes stay busy until the
 * queue is empty and there is one launch per call.
 *
 * Tallies: each block accumulates the exact fixed-point tally of its cells
 * in LDS (two 64-bit words per cell, 64-bit LDS atomics with carry) and adds
 * it to the global 128-bit tally once at the end.  Integer sums are exact, so
 * the result is independent of which lane, wave or block handled which
 * history.  For layers too large for LDS the same atomics go to the global
 * (L2-resident) tally.  The per-cell cross-section pairs are staged in LDS
 * when they fit next to the tally, otherwise read through L2.
 */
#include "internal.h"

namespace {

template <bool LDS_TALLY, bool LDS_CELLS, bool SCORE>
__global__ __launch_bounds__(1024) void k_history(mc_transport_args a, int refill_k) {
  extern __shared__ __align__(16) unsigned char smem[];
  const mc_coef K = mc_coef_live();
  const int m = a.L.m;
  unsigned long long *t_lo, *t_hi;
  size_t off = 0;
  if (LDS_TALLY) {
    t_lo = reinterpret_cast<unsigned long long *>(smem);
    t_hi = t_lo + m;
    off = (size_t)16 * m;
    for (int c = threadIdx.x; c < m; c += blockDim.x) {
      t_lo[c] = 0;
      t_hi[c] = 0;
    }
  } else {
    t_lo = a.fx;
    t_hi = a.fx + m;
  }
  const mc_cell *cells = a.cells;
  const double *g = a.g;
  if (LDS_CELLS) {
    mc_cell *lc = reinterpret_cast<mc_cell *>(smem + off);
    for (int c = threadIdx.x; c < m; c += blockDim.x) lc[c] = a.cells[c];
    cells = lc;
    if (SCORE) {
      double *lg = reinterpret_cast<double *>(smem + off + sizeof(mc_cell) * (size_t)m);
      for (int c = threadIdx.x; c < m; c += blockDim.x) lg[c] = a.g[c];
      g = lg;
    }
  }
  __syncthreads();

  const int lane = __lane_id();
  mc_state p;
  p.index = 0;
  int64_t idx = 0;
  bool has = false, done = false, capped = false;
  uint64_t ev = 0, col = 0;
  double score = 0.0;
  unsigned long long tot_ev = 0, tot_col = 0, nl = 0, nr = 0, nk = 0;

  /* events between refills: refill_k if positive, else adapted per wave
     from the lanes found idle at each refill (many idle: the histories are
     short against the interval, halve it; almost none: double it, since
     each refill costs a queue atomic and a pass through this loop) */
  const bool adapt = refill_k <= 0;
  int kk = adapt ? 128 : refill_k;
  /* wave-uniform bound on every lane's event count: while it stays a whole
     interval below the cap, no history can reach the cap in the next
     interval, and the events run without the cap test; recomputed exactly
     (the largest count in the wave) when it nears the cap */
  uint64_t ev_bound = 0;
  for (;;) {
    if (!done) { /* wave-uniform */
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
            load_state(a, i, p);
            mc_refresh_rcp(p);
            idx = i;
            has = true;
            ev = 0;
            col = 0;
            score = 0.0;
          }
        }
        if (base + (unsigned long long)cnt >= (unsigned long long)a.n) done = true;
      }
    }
    if (__ballot(has) == 0) break; /* wave-uniform */
    if (ev_bound + (uint64_t)kk >= a.max_events)
      ev_bound = mc_wave_max_u64(has ? ev : 0ull);
    const bool below_cap = ev_bound + (uint64_t)kk < a.max_events; /* wave-uniform */
    ev_bound += (uint64_t)kk;
    if (has) {
      /* one event of this lane's history: Layer::particle_step */
      auto step = [&]() {
        const int c = p.index - a.L.index_start;
        bool coll;
        double di;
        const double dw = mc_event<true>(p, cells, c, a.L.dx, K, coll, di, nullptr, &col);
        mc_tally(t_lo, t_hi, a.flags, &a.ctr->status, c, dw);
        if (SCORE) score = score + g[c] * dw;
      };
      /* the same loop twice, so that the common one carries no cap test
         (the compiler keeps a shared test in the loop even when the flag
         guarding it is uniform) */
      if (below_cap) {
        for (int k = 0; k < kk; ++k) {
          if (!mc_alive(p, a.L)) break;
          step();
          ++ev;
        }
      } else {
        for (int k = 0; k < kk; ++k) {
          if (!mc_alive(p, a.L) || ev >= a.max_events) break;
          step();
          ++ev;
        }
      }
      const bool alive = mc_alive(p, a.L);
      if (!alive || ev >= a.max_events) {
        int code = mc_exit_code(p, a.L);
        if (alive) { /* runaway history stopped by the guard */
          capped = true;
          code = 2;
        }
        store_state(a, idx, p, code);
        if (a.events) a.events[idx] = ev;
        if (a.collisions) a.collisions[idx] = col;
        if (SCORE) a.score[idx] = score;
        tot_ev += ev;
        tot_col += col;
        nl += code == -1;
        nr += code == 1;
        nk += code == 0;
        has = false;
      }
    }
  }

  tot_ev = mc_wave_sum(tot_ev);
  tot_col = mc_wave_sum(tot_col);
  nl = mc_wave_sum(nl);
  nr = mc_wave_sum(nr);
  nk = mc_wave_sum(nk);
  if (lane == 0) {
    atomicAdd(&a.ctr->events, tot_ev);
    atomicAdd(&a.ctr->collisions, tot_col);
    atomicAdd(&a.ctr->n_left, nl);
    atomicAdd(&a.ctr->n_right, nr);
    atomicAdd(&a.ctr->n_killed, nk);
  }
  if (capped) atomicOr(&a.ctr->status, 1u);

  if (LDS_TALLY) {
    __syncthreads();
    for (int c = threadIdx.x; c < m; c += blockDim.x) {
      mc_fx v;
      v.lo = t_lo[c];
      v.hi = t_hi[c];
      mc_fx_atomic_add(&a.fx[c], &a.fx[m + c], v);
    }
  }
}

template <bool T, bool C, bool S>
int launch_one(mcgpu_context *ctx, const mc_transport_args &a, const mcgpu_options &o,
               hipStream_t s, size_t lds) {
  auto kern = k_history<T, C, S>;
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
    /* no more blocks than there are waves of work */
    int64_t waves = (a.n + 63) / 64;
    int64_t blocks_needed = (waves * 64 + block - 1) / block;
    if (blocks_needed < grid) grid = (int)(blocks_needed > 0 ? blocks_needed : 1);
  }
  int k = o.events_per_refill > 0 ? o.events_per_refill : 0; /* 0: adaptive */
  hipLaunchKernelGGL(kern, dim3(grid), dim3(block), lds, s, a, k);
  MC_HIP(ctx, hipGetLastError());
  return MCGPU_OK;
}

} // namespace

int mc_launch_history(mcgpu_context *ctx, const mc_transport_args &a, const mcgpu_options &o,
                      hipStream_t s) {
  const size_t m = (size_t)a.L.m;
  const size_t tally = 16 * m, cells = sizeof(mc_cell) * m;
  const bool lds_tally = tally <= ctx->lds_per_block;
  if (a.score) { /* per-history score of a weighted tally (forward of the adjoint) */
    if (!a.g) return mc_fail(ctx, MCGPU_ERR_ARG, "score output needs target weights g");
    const size_t extra = 8 * m;
    const bool lds_cells = lds_tally && tally + cells + extra <= 65536;
    if (lds_tally && lds_cells) return launch_one<true, true, true>(ctx, a, o, s, tally + cells + extra);
    if (lds_tally) return launch_one<true, false, true>(ctx, a, o, s, tally);
    return launch_one<false, false, true>(ctx, a, o, s, 0);
  }
  const bool lds_cells = lds_tally && tally + cells <= 65536;
  if (lds_tally && lds_cells) return launch_one<true, true, false>(ctx, a, o, s, tally + cells);
  if (lds_tally) return launch_one<true, false, false>(ctx, a, o, s, tally);
  return launch_one<false, false, false>(ctx, a, o, s, 0);
}
