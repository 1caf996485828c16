/*
 * event.hip -- event-based transport kernel.
 *
 * The reference advances each particle to the end of its history in one go
 * (src/layer.cpp:195-200).  Here a pass applies exactly ONE event
 * (Layer::particle_step, src/layer.cpp:123-190) to every live particle of a
 * bank, and the bank is compacted between passes, so all lanes of a wave
 * always work on the same step of different particles.
 *
 * Layout and schedule (one persistent block per CU-slot, no launch per pass):
 *  - each block owns a particle pool in HBM: two halves (ping-pong) of P
 *    slots in structure-of-arrays form (seed, x, mu, wmc, index, id, event
 *    and collision counters), so every field is a coalesced 512-byte wave
 *    access;
 *  - a pass streams the live part of the current half: load, one event,
 *    tally, then the survivors are compacted into the other half -- a wave
 *    ballot + mbcnt prefix gives each survivor its rank, one LDS atomic per
 *    wave reserves the wave's range; finished histories are written to their
 *    original bank slot (the particle id), so compaction order is free;
 *  - with compaction interval c > 1, c-1 passes out of c write survivors back
 *    in place and mark finished slots dead instead of compacting;
 *  - when the pool has room for a block-width of particles, it is refilled
 *    from the call's bank through the global work queue (one atomic per
 *    refill), so blocks stay full until the bank is exhausted;
 *  - the exact fixed-point tally lives in LDS for the block's lifetime and is
 *    added to the global tally once (same arithmetic as the history kernel).
 * Results are bit-identical to the history kernel and the oracle: histories
 * are independent and the tally is an exact integer sum.
 */
#include "internal.h"

namespace {

#define DEAD_ID 0xFFFFFFFFu

/* one half of a block's pool, structure of arrays */
struct pool_half {
  uint64_t *seed;
  double *x;
  double *mu;
  double *wmc;
  int32_t *index;
  uint32_t *id;
  uint32_t *ev;
  uint32_t *col;
};

constexpr size_t SLOT_BYTES = 8 * 4 + 4 * 4; /* 48 bytes per slot */

__device__ __forceinline__ pool_half pool_at(unsigned char *base, int P) {
  pool_half h;
  h.seed = reinterpret_cast<uint64_t *>(base);
  h.x = reinterpret_cast<double *>(base + 8 * (size_t)P);
  h.mu = reinterpret_cast<double *>(base + 16 * (size_t)P);
  h.wmc = reinterpret_cast<double *>(base + 24 * (size_t)P);
  h.index = reinterpret_cast<int32_t *>(base + 32 * (size_t)P);
  h.id = reinterpret_cast<uint32_t *>(base + 36 * (size_t)P);
  h.ev = reinterpret_cast<uint32_t *>(base + 40 * (size_t)P);
  h.col = reinterpret_cast<uint32_t *>(base + 44 * (size_t)P);
  return h;
}

__device__ __forceinline__ void put(const pool_half &h, int j, const mc_state &p, uint32_t id,
                                    uint32_t ev, uint32_t col) {
  h.seed[j] = p.seed;
  h.x[j] = p.x;
  h.mu[j] = p.mu;
  h.wmc[j] = p.wmc;
  h.index[j] = p.index;
  h.id[j] = id;
  h.ev[j] = ev;
  h.col[j] = col;
}

template <bool LDS_TALLY, bool LDS_CELLS>
__global__ __launch_bounds__(1024) void k_event(mc_transport_args a, unsigned char *pools, int P,
                                                int compact_every) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ unsigned int s_cnt[3];
  __shared__ unsigned long long s_base;
  __shared__ int s_take;

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
  if (LDS_CELLS) {
    mc_cell *lc = reinterpret_cast<mc_cell *>(smem + off);
    for (int c = threadIdx.x; c < m; c += blockDim.x) lc[c] = a.cells[c];
    cells = lc;
  }
  if (threadIdx.x == 0) {
    s_cnt[0] = 0;
    s_cnt[1] = 0;
    s_cnt[2] = 0;
  }

  unsigned char *mine = pools + (size_t)blockIdx.x * 2 * SLOT_BYTES * (size_t)P;
  pool_half half[2] = {pool_at(mine, P), pool_at(mine + SLOT_BYTES * (size_t)P, P)};
  const int B = blockDim.x;
  const int tid = threadIdx.x;
  const int lane = __lane_id();
  const uint32_t max_ev = a.max_events > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)a.max_events;

  int cur = 0;
  int n_pool = 0; /* slots in use in half[cur] (block-uniform) */
  bool queue_done = false;
  bool capped = false;
  unsigned long long tot_ev = 0, tot_col = 0, nl = 0, nr = 0, nk = 0;
  __syncthreads();

  for (unsigned pass = 0;; ++pass) {
    /* ---- refill from the work queue (block-uniform condition) */
    if (!queue_done && n_pool <= P - B) {
      if (tid == 0) {
        const int want = P - n_pool;
        unsigned long long base = atomicAdd(&a.ctr->next, (unsigned long long)want);
        int take = 0;
        if (base < (unsigned long long)a.n) {
          unsigned long long left = (unsigned long long)a.n - base;
          take = left < (unsigned long long)want ? (int)left : want;
        }
        s_base = base;
        s_take = take;
      }
      __syncthreads();
      const unsigned long long base = s_base;
      const int take = s_take;
      if (take < P - n_pool) queue_done = true;
      for (int j = tid; j < take; j += B) {
        mc_state p;
        load_state(a, (int64_t)(base + j), p);
        put(half[cur], n_pool + j, p, (uint32_t)(base + j), 0u, 0u);
      }
      n_pool += take;
      __syncthreads();
    }
    if (n_pool == 0) {
      if (queue_done) break;
      continue;
    }

    /* ---- one event per live particle */
    const bool compact = (pass % (unsigned)compact_every) == 0;
    unsigned int *cnt = &s_cnt[pass % 3];
    if (tid == 0) s_cnt[(pass + 1) % 3] = 0;
    const pool_half src = half[cur];
    const pool_half dst = half[cur ^ 1];
    for (int j0 = 0; j0 < n_pool; j0 += B) { /* block-uniform trip count */
      const int j = j0 + tid;
      bool live = false, keep = false;
      mc_state p;
      uint32_t id = DEAD_ID, ev = 0, col = 0;
      if (j < n_pool) {
        id = src.id[j];
        if (id != DEAD_ID) {
          p.seed = src.seed[j];
          p.x = src.x[j];
          p.mu = src.mu[j];
          p.wmc = src.wmc[j];
          p.index = src.index[j];
          ev = src.ev[j];
          col = src.col[j];
          live = true;
        }
      }
      if (live) {
        if (mc_alive(p, a.L) && ev < max_ev) {
          const int c = p.index - a.L.index_start;
          bool coll;
          double di;
          const double dw = mc_event(p, cells[c], a.L.dx, coll, di);
          ++ev;
          col += coll ? 1u : 0u;
          mc_tally(t_lo, t_hi, a.flags, &a.ctr->status, c, dw);
        }
        keep = mc_alive(p, a.L) && ev < max_ev;
        if (!keep) { /* history finished: write it to its bank slot */
          int code = mc_exit_code(p, a.L);
          if (mc_alive(p, a.L)) {
            capped = true;
            code = 2;
          }
          store_state(a, (int64_t)id, p, code);
          if (a.events) a.events[id] = ev;
          if (a.collisions) a.collisions[id] = col;
          tot_ev += ev;
          tot_col += col;
          nl += code == -1;
          nr += code == 1;
          nk += code == 0;
        }
      }
      if (compact) {
        const unsigned long long mask = __ballot(keep);
        if (mask) {
          const int leader = __ffsll((long long)mask) - 1;
          unsigned int wbase = 0;
          if (lane == leader) wbase = atomicAdd(cnt, (unsigned int)__popcll(mask));
          wbase = (unsigned int)__builtin_amdgcn_readlane((int)wbase, leader);
          if (keep) put(dst, (int)wbase + mc_mbcnt64(mask), p, id, ev, col);
        }
      } else if (live) {
        if (keep)
          put(src, j, p, id, ev, col);
        else
          src.id[j] = DEAD_ID;
      }
    }
    __syncthreads();
    if (compact) {
      n_pool = (int)*cnt;
      cur ^= 1;
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
    for (int c = tid; c < m; c += B) {
      mc_fx v;
      v.lo = t_lo[c];
      v.hi = t_hi[c];
      mc_fx_atomic_add(&a.fx[c], &a.fx[m + c], v);
    }
  }
}

template <bool T, bool C>
int launch_one(mcgpu_context *ctx, const mc_transport_args &a, const mcgpu_options &o,
               hipStream_t s, size_t lds) {
  auto kern = k_event<T, C>;
  if (lds > 65536)
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  int block = o.block_size > 0 ? o.block_size : (lds > 81920 ? 1024 : 256);
  if (block > 1024 || block % 64)
    return mc_fail(ctx, MCGPU_ERR_ARG, "block_size must be a multiple of 64 and <= 1024");
  int grid = o.grid_blocks;
  if (grid <= 0) {
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, block, lds);
    if (e != hipSuccess || per_cu < 1) per_cu = 1;
    grid = per_cu * ctx->num_cus;
  }
  /* particles per lane: slots of the pool per thread */
  const int ppl = o.particles_per_lane > 0 ? o.particles_per_lane : 4;
  const int P = ppl * block;
  int64_t per_block = (a.n + grid - 1) / grid;
  if (grid > 1 && per_block < block) { /* small banks: fewer blocks */
    grid = (int)((a.n + block - 1) / block);
    if (grid < 1) grid = 1;
  }
  const int c = o.compaction_interval > 0 ? o.compaction_interval : 1;
  int rc = mc_ensure(ctx, &ctx->work, &ctx->work_bytes, (size_t)grid * 2 * SLOT_BYTES * (size_t)P);
  if (rc) return rc;
  hipLaunchKernelGGL(kern, dim3(grid), dim3(block), lds, s, a, (unsigned char *)ctx->work, P, c);
  MC_HIP(ctx, hipGetLastError());
  return MCGPU_OK;
}

} // namespace

int mc_launch_event(mcgpu_context *ctx, const mc_transport_args &a, const mcgpu_options &o,
                    hipStream_t s) {
  if (a.n >= (int64_t)DEAD_ID)
    return mc_fail(ctx, MCGPU_ERR_ARG, "event kernel: bank larger than 2^32-2 particles");
  const size_t m = (size_t)a.L.m;
  const size_t tally = 16 * m, cells = sizeof(mc_cell) * m;
  const size_t statics = 64; /* s_cnt, s_base, s_take */
  const bool lds_tally = tally + statics <= ctx->lds_per_block;
  const bool lds_cells = lds_tally && tally + cells <= 65536;
  if (lds_tally && lds_cells) return launch_one<true, true>(ctx, a, o, s, tally + cells);
  if (lds_tally) return launch_one<true, false>(ctx, a, o, s, tally);
  return launch_one<false, false>(ctx, a, o, s, 0);
}
