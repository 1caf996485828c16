/*
 * batched.hip -- many independent slab problems in one persistent launch.
 *
 * The reference runs one Layer per MPI rank and transports each layer's bank
 * with Layer::simulate (src/layer.cpp:239-361).  Here thousands of
 * independent slabs (each one Layer's worth of work) share a single launch:
 *  - the host cuts every slab's bank into tiles of particles_per_lane *
 *    block_size particles and orders the tiles by estimated work (particles x
 *    cells), largest first, so the long tiles start early;
 *  - each persistent block takes the next tile from a global counter, zeroes
 *    its LDS tally for that slab's cells (and stages the slab's cross sections
 *    in LDS when both fit), runs the tile's histories with lanes refilled
 *    from an LDS counter (as in the history kernel), and adds the exact
 *    fixed-point tally to the slab's global cells once per tile;
 *  - setup (cross-section table, zeroed tallies and counters), transport and
 *    finalize (tally -> weights_absorbed) are captured once into a hipGraph
 *    and replayed; zeroing is done by kernels (no memset nodes).
 * Per history the arithmetic is the single-slab kernels' arithmetic, so each
 * slab's results are bit-identical to transporting it alone.
 */
#include <algorithm>
#include <vector>

#include "internal.h"

namespace {

struct batch_args {
  const mc_slab_k *slabs;
  const mc_tile_k *tiles;
  int32_t n_tiles;
  int32_t lds_cells_max_m; /* slabs with m <= this stage cross sections in LDS */
  int32_t lds_tally_max_m; /* slabs with m <= this keep their tally in LDS; larger
                              ones add straight into the global tally */
  mcgpu_bank soa;
  const mc_cell *cells;
  unsigned long long *fx; /* lo words [0, total), hi words [total, 2 total) */
  int64_t total_cells;
  uint32_t *flags;
  int8_t *result8;
  mc_counters *ctr;
  uint64_t max_events;
};

/* Per-lane counters of one block, carried across its tiles. */
struct batch_lane {
  bool capped;
  unsigned int ev_pending, col_pending;
};

/* One tile: its histories, lanes refilled from the LDS queue s_next.  The
   tally and cell pointers are LDS or global according to the template
   arguments, so every access compiles to the matching instruction. */
template <bool LDS_TALLY, bool LDS_CELLS, bool HI_COND>
__device__ __forceinline__ void batch_tile(const batch_args &a, const mc_tile_k &tile,
                                           const mc_slab_k &sl, const mc_layer_k &L,
                                           unsigned long long *t_lo, unsigned long long *t_hi,
                                           const mc_cell *cells, unsigned long long *s_next,
                                           unsigned long long *s_tot, uint32_t max_ev,
                                           int refill_k, const mc_coef &K, batch_lane &bl) {
  const int lane = __lane_id();
  mc_state p;
  p.index = 0;
  int64_t idx = 0;
  bool has = false, done = false;
  uint32_t ev = 0, col = 0;
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
            p.seed = a.soa.seed[i];
            p.x = a.soa.x[i];
            p.mu = a.soa.mu[i];
            mc_refresh_rcp(p);
            p.wmc = a.soa.wmc[i];
            p.index = a.soa.index[i];
            idx = i;
            has = true;
            ev = 0;
            col = 0;
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
        const double dw = mc_event<true>(p, cells, c, L.dx, K, coll, di, nullptr, &col);
        ++ev;
        mc_tally<false, HI_COND>(t_lo, t_hi, a.flags + sl.cell_off, &a.ctr->status, c, dw);
      }
      const bool alive = mc_alive(p, L);
      if (!alive || ev >= max_ev) {
        int code = mc_exit_code(p, L);
        if (alive) {
          bl.capped = true;
          code = 2;
        }
        a.soa.seed[idx] = p.seed;
        a.soa.x[idx] = p.x;
        a.soa.mu[idx] = p.mu;
        a.soa.wmc[idx] = p.wmc;
        a.soa.index[idx] = p.index;
        if (a.result8) a.result8[idx] = (int8_t)code;
        bl.ev_pending += ev;
        bl.col_pending += col;
        atomicAdd(&s_tot[code == -1 ? 2 : (code == 1 ? 3 : 4)], code == 2 ? 0ull : 1ull);
        has = false;
        if (bl.ev_pending > 0x40000000u) { /* keep the 32-bit lane counters safe */
          atomicAdd(&s_tot[0], (unsigned long long)bl.ev_pending);
          atomicAdd(&s_tot[1], (unsigned long long)bl.col_pending);
          bl.ev_pending = bl.col_pending = 0;
        }
      }
    }
  }
}

/* HI_COND: every particle weight is below 2^-24 (see mc_fx_atomic_add_always) */
template <bool HI_COND>
__global__ __launch_bounds__(1024) void k_batched(batch_args a, int refill_k) {
  extern __shared__ __align__(16) unsigned char smem[];
  const mc_coef K = mc_coef_live();
  __shared__ int s_tile;
  __shared__ unsigned long long s_next;
  __shared__ unsigned long long s_tot[5];
  unsigned long long *t_lds = reinterpret_cast<unsigned long long *>(smem);
  const int lane = __lane_id();
  const int B = blockDim.x;
  const int tid = threadIdx.x;
  if (tid < 5) s_tot[tid] = 0;
  batch_lane bl;
  bl.capped = false;
  bl.ev_pending = bl.col_pending = 0;
  const uint32_t max_ev = a.max_events > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)a.max_events;

  for (;;) {
    if (tid == 0) s_tile = (int)atomicAdd(&a.ctr->next, 1ull);
    __syncthreads();
    const int t = s_tile;
    if (t >= a.n_tiles) break; /* block-uniform */
    const mc_tile_k tile = a.tiles[t];
    const mc_slab_k sl = a.slabs[tile.slab];
    const int m = sl.m;
    mc_layer_k L;
    L.index_start = sl.index_start;
    L.m = m;
    L.dx = sl.dx;
    L.min_weight = sl.min_weight;
    const bool lds_tally = m <= a.lds_tally_max_m; /* block-uniform */
    const bool lds_cells = m <= a.lds_cells_max_m; /* implies lds_tally */
    mc_cell *lc = reinterpret_cast<mc_cell *>(smem + (size_t)16 * m);
    if (lds_tally) {
      for (int c = tid; c < m; c += B) {
        t_lds[c] = 0;
        t_lds[m + c] = 0;
        if (lds_cells) lc[c] = a.cells[sl.cell_off + c];
      }
    }
    if (tid == 0) s_next = (unsigned long long)tile.begin;
    __syncthreads();

    if (lds_cells)
      batch_tile<true, true, HI_COND>(a, tile, sl, L, t_lds, t_lds + m, lc, &s_next, s_tot, max_ev,
                             refill_k, K, bl);
    else if (lds_tally)
      batch_tile<true, false, HI_COND>(a, tile, sl, L, t_lds, t_lds + m, a.cells + sl.cell_off, &s_next,
                              s_tot, max_ev, refill_k, K, bl);
    else
      batch_tile<false, false, HI_COND>(a, tile, sl, L, a.fx + sl.cell_off,
                               a.fx + a.total_cells + sl.cell_off, a.cells + sl.cell_off,
                               &s_next, s_tot, max_ev, refill_k, K, bl);

    __syncthreads();
    if (lds_tally) {
      for (int c = tid; c < m; c += B) {
        mc_fx v;
        v.lo = t_lds[c];
        v.hi = t_lds[m + c];
        if (v.lo | v.hi) /* cells the tile never reached add nothing */
          mc_fx_atomic_add(&a.fx[sl.cell_off + c], &a.fx[a.total_cells + sl.cell_off + c], v);
      }
    }
    __syncthreads();
  }

  {
    const unsigned long long e = mc_wave_sum(bl.ev_pending), cl = mc_wave_sum(bl.col_pending);
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

__global__ void k_batch_setup(int64_t total, const double *sigs, const double *ar, mc_cell *cells,
                              unsigned long long *fx, uint32_t *flags, mc_counters *ctr) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < total) {
    cells[c] = mc_make_cell(sigs[c], ar[c]);
    fx[c] = 0;
    fx[total + c] = 0;
    flags[c] = 0;
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

__global__ void k_batch_finalize(int64_t total, const unsigned long long *fx,
                                 const uint32_t *flags, double *weights) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < total) weights[c] = weights[c] + mc_fx_finalize(fx[c], fx[total + c], flags[c]);
}

} // namespace

int mc_batched_plan_shape(mcgpu_context *ctx, int max_m, int block, int grid, int refill,
                          mc_batched_shape *out) {
  if (block > 1024 || block < 64 || block % 64)
    return mc_fail(ctx, MCGPU_ERR_ARG, "block_size must be a multiple of 64 and <= 1024");
  mc_batched_shape sh;
  sh.block = block;
  sh.refill = refill > 0 ? refill : 8;
  /* LDS: the largest slab's tally plus its cross sections, as far as they fit
     (128 B are left for the kernel's static LDS); slabs whose tally does not
     fit add into the global tally directly */
  const size_t budget = ctx->lds_per_block - 128;
  sh.lds = std::min(budget, (size_t)32 * (size_t)std::max(max_m, 1));
  sh.lds_cells_max_m = (int)(sh.lds / 32);
  sh.lds_tally_max_m = (int)(sh.lds / 16);
  if (sh.lds > 65536)
  {
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_batched<false>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh.lds);
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_batched<true>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh.lds);
  }
  sh.grid = grid;
  if (sh.grid <= 0) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batched<false>, block, sh.lds) !=
            hipSuccess ||
        per_cu < 1)
      per_cu = 1;
    sh.grid = per_cu * ctx->num_cus;
  }
  *out = sh;
  return MCGPU_OK;
}

int mc_launch_batched(mcgpu_context *ctx, const mc_batched_shape &sh, const mc_slab_k *slabs,
                      const mc_tile_k *tiles, int32_t n_tiles, const mcgpu_bank &bank,
                      const mc_cell *cells, unsigned long long *fx, int64_t total_cells,
                      uint32_t *flags, int8_t *result8, mc_counters *ctr, uint64_t max_events,
                      bool small_weights, hipStream_t s) {
  batch_args a;
  a.slabs = slabs;
  a.tiles = tiles;
  a.n_tiles = n_tiles;
  a.lds_cells_max_m = sh.lds_cells_max_m;
  a.lds_tally_max_m = sh.lds_tally_max_m;
  a.soa = bank;
  a.cells = cells;
  a.fx = fx;
  a.total_cells = total_cells;
  a.flags = flags;
  a.result8 = result8;
  a.ctr = ctr;
  a.max_events = max_events;
  if (small_weights)
    hipLaunchKernelGGL(k_batched<true>, dim3(sh.grid), dim3(sh.block), sh.lds, s, a, sh.refill);
  else
    hipLaunchKernelGGL(k_batched<false>, dim3(sh.grid), dim3(sh.block), sh.lds, s, a, sh.refill);
  MC_HIP(ctx, hipGetLastError());
  return MCGPU_OK;
}

struct mcgpu_batch {
  mcgpu_context *ctx = nullptr;
  std::vector<mcgpu_slab> host_slabs;
  mcgpu_options opt{};
  /* inputs (caller-owned device memory) */
  const double *sigs = nullptr, *ar = nullptr;
  double *weights = nullptr;
  int64_t total_cells = 0, total_particles = 0;
  mcgpu_bank bank{};
  int8_t *result = nullptr;
  /* plan-owned device memory */
  mc_slab_k *d_slabs = nullptr;
  mc_tile_k *d_tiles = nullptr;
  int32_t n_tiles = 0;
  size_t tiles_cap = 0;
  mc_cell *d_cells = nullptr;
  unsigned long long *d_fx = nullptr;
  uint32_t *d_flags = nullptr;
  mc_counters *d_ctr = nullptr;
  mc_counters *h_ctr = nullptr;
  int max_m = 0;
  size_t lds = 0;
  int lds_cells_max_m = 0;
  int lds_tally_max_m = 0;
  int block = 1024, grid = 0, refill = 8;
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
};

namespace {

int batch_tiles(mcgpu_batch *b) {
  mcgpu_context *ctx = b->ctx;
  const int ppl = b->opt.particles_per_lane > 0 ? b->opt.particles_per_lane : 16;
  const int64_t T = (int64_t)ppl * b->block;
  std::vector<mc_tile_k> tiles;
  std::vector<double> work;
  for (size_t s = 0; s < b->host_slabs.size(); ++s) {
    const mcgpu_slab &h = b->host_slabs[s];
    const int m = h.max_index - h.min_index;
    for (int64_t p0 = 0; p0 < h.n; p0 += T) {
      mc_tile_k t;
      t.slab = (int32_t)s;
      t.pad_ = 0;
      t.begin = h.particle_offset + p0;
      t.end = h.particle_offset + std::min(h.n, p0 + T);
      tiles.push_back(t);
      work.push_back((double)(t.end - t.begin) * (double)m);
    }
  }
  std::vector<size_t> order(tiles.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return work[x] > work[y]; });
  std::vector<mc_tile_k> sorted(tiles.size());
  for (size_t i = 0; i < order.size(); ++i) sorted[i] = tiles[order[i]];
  b->n_tiles = (int32_t)sorted.size();
  if (sorted.size() > b->tiles_cap) {
    if (b->d_tiles) MC_HIP(ctx, hipFree(b->d_tiles));
    b->d_tiles = nullptr;
    MC_HIP(ctx, hipMalloc(&b->d_tiles, sizeof(mc_tile_k) * (sorted.size() + 1)));
    b->tiles_cap = sorted.size();
  }
  if (!sorted.empty())
    MC_HIP(ctx, hipMemcpy(b->d_tiles, sorted.data(), sizeof(mc_tile_k) * sorted.size(),
                          hipMemcpyHostToDevice));
  return MCGPU_OK;
}

int batch_capture(mcgpu_batch *b) {
  mcgpu_context *ctx = b->ctx;
  b->block = b->opt.block_size > 0 ? b->opt.block_size : 1024;
  if (b->block > 1024 || b->block % 64)
    return mc_fail(ctx, MCGPU_ERR_ARG, "block_size must be a multiple of 64 and <= 1024");
  b->refill = b->opt.events_per_refill > 0 ? b->opt.events_per_refill : 8;
  int rc = batch_tiles(b);
  if (rc) return rc;
  mc_batched_shape sh;
  if ((rc = mc_batched_plan_shape(ctx, b->max_m, b->block, b->opt.grid_blocks, b->refill, &sh)))
    return rc;
  b->lds = sh.lds;
  b->lds_cells_max_m = sh.lds_cells_max_m;
  b->lds_tally_max_m = sh.lds_tally_max_m;
  b->grid = sh.grid;
  if (b->exec) {
    (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
  }
  if (b->graph) {
    (void)hipGraphDestroy(b->graph);
    b->graph = nullptr;
  }
  batch_args a;
  a.slabs = b->d_slabs;
  a.tiles = b->d_tiles;
  a.n_tiles = b->n_tiles;
  a.lds_cells_max_m = b->lds_cells_max_m;
  a.lds_tally_max_m = b->lds_tally_max_m;
  a.soa = b->bank;
  a.cells = b->d_cells;
  a.fx = b->d_fx;
  a.total_cells = b->total_cells;
  a.flags = b->d_flags;
  a.result8 = b->result;
  a.ctr = b->d_ctr;
  a.max_events = b->opt.max_events_per_history ? b->opt.max_events_per_history : (1ULL << 27);
  const int64_t tc = b->total_cells;
  MC_HIP(ctx, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  hipLaunchKernelGGL(k_batch_setup, dim3((unsigned)((tc + 255) / 256)), dim3(256), 0, b->stream,
                     tc, b->sigs, b->ar, b->d_cells, b->d_fx, b->d_flags, b->d_ctr);
  hipLaunchKernelGGL(k_batched<false>, dim3(sh.grid), dim3(sh.block), sh.lds, b->stream, a,
                     sh.refill);
  hipLaunchKernelGGL(k_batch_finalize, dim3((unsigned)((tc + 255) / 256)), dim3(256), 0,
                     b->stream, tc, (const unsigned long long *)b->d_fx,
                     (const uint32_t *)b->d_flags, b->weights);
  hipError_t e = hipStreamEndCapture(b->stream, &b->graph);
  if (e != hipSuccess) return mc_fail(ctx, MCGPU_ERR_HIP, "hipStreamEndCapture", e);
  MC_HIP(ctx, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  b->opt.block_size = b->block;
  b->opt.grid_blocks = b->grid;
  b->opt.events_per_refill = b->refill;
  if (b->opt.particles_per_lane <= 0) b->opt.particles_per_lane = 16;
  return MCGPU_OK;
}

int batch_launch(mcgpu_batch *b, hipStream_t s, mcgpu_stats *stats) {
  mcgpu_context *ctx = b->ctx;
  MC_HIP(ctx, hipEventRecord(b->e0, s));
  MC_HIP(ctx, hipGraphLaunch(b->exec, s));
  MC_HIP(ctx, hipEventRecord(b->e1, s));
  MC_HIP(ctx, hipMemcpyAsync(b->h_ctr, b->d_ctr, sizeof(mc_counters), hipMemcpyDeviceToHost, s));
  MC_HIP(ctx, hipStreamSynchronize(s));
  if (stats) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, b->e0, b->e1);
    stats->events = b->h_ctr->events;
    stats->collisions = b->h_ctr->collisions;
    stats->n_left = (int64_t)b->h_ctr->n_left;
    stats->n_right = (int64_t)b->h_ctr->n_right;
    stats->n_killed = (int64_t)b->h_ctr->n_killed;
    stats->kernel_ms = ms;
  }
  if (b->h_ctr->status & 1u)
    return mc_fail(ctx, MCGPU_ERR_EVENT_CAP, "a history exceeded max_events_per_history");
  if (b->h_ctr->status & 2u) return mc_fail(ctx, MCGPU_ERR_RANGE, "a tally term was outside +-2^39");
  return MCGPU_OK;
}

} // namespace

extern "C" {

int mcgpu_batch_destroy(mcgpu_batch *b) {
  if (!b) return MCGPU_OK;
  (void)hipSetDevice(b->ctx->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  void *bufs[] = {b->d_slabs, b->d_tiles, b->d_cells, b->d_fx, b->d_flags, b->d_ctr};
  for (void *p : bufs)
    if (p) (void)hipFree(p);
  if (b->h_ctr) (void)hipHostFree(b->h_ctr);
  if (b->e0) (void)hipEventDestroy(b->e0);
  if (b->e1) (void)hipEventDestroy(b->e1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return MCGPU_OK;
}

int mcgpu_batch_create(mcgpu_context *ctx, int32_t n_slabs, const mcgpu_slab *slabs,
                       const double *sigs, const double *absorption_rates,
                       double *weights_absorbed, int64_t total_cells, const mcgpu_bank *bank,
                       int8_t *result, const mcgpu_options *opt, mcgpu_batch **out) {
  if (!ctx || !out) return MCGPU_ERR_ARG;
  *out = nullptr;
  ctx->err.clear();
  if (n_slabs <= 0 || !slabs || !sigs || !absorption_rates || !weights_absorbed || !bank ||
      total_cells <= 0)
    return mc_fail(ctx, MCGPU_ERR_ARG, "batch: null pointer or empty batch");
  MC_HIP(ctx, hipSetDevice(ctx->device));
  mcgpu_batch *b = new mcgpu_batch();
  b->ctx = ctx;
  b->host_slabs.assign(slabs, slabs + n_slabs);
  mcgpu_default_options(&b->opt);
  if (opt) b->opt = *opt;
  b->sigs = sigs;
  b->ar = absorption_rates;
  b->weights = weights_absorbed;
  b->total_cells = total_cells;
  b->bank = *bank;
  b->result = result;
  std::vector<mc_slab_k> ks(n_slabs);
  int rc = MCGPU_OK;
  for (int s = 0; s < n_slabs && !rc; ++s) {
    const mcgpu_slab &h = slabs[s];
    const int m = h.max_index - h.min_index;
    if (m <= 0 || h.n < 0 || h.cell_offset < 0 || h.cell_offset + m > total_cells ||
        h.particle_offset < 0)
      rc = mc_fail(ctx, MCGPU_ERR_ARG, "batch: invalid slab descriptor");
    ks[s].index_start = h.min_index;
    ks[s].m = m;
    ks[s].dx = h.dx;
    ks[s].min_weight = h.particle_min_weight;
    ks[s].cell_off = h.cell_offset;
    b->max_m = std::max(b->max_m, m);
    b->total_particles = std::max(b->total_particles, h.particle_offset + h.n);
  }
  if (!rc && b->total_particles > 0 &&
      (!bank->seed || !bank->x || !bank->mu || !bank->wmc || !bank->index))
    rc = mc_fail(ctx, MCGPU_ERR_ARG, "batch: bank has a null array");
#define B_HIP(call)                                                           \
  if (!rc) {                                                                  \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess) rc = mc_fail(ctx, MCGPU_ERR_HIP, #call, e_);        \
  }
  B_HIP(hipMalloc(&b->d_slabs, sizeof(mc_slab_k) * n_slabs));
  B_HIP(hipMemcpy(b->d_slabs, ks.data(), sizeof(mc_slab_k) * n_slabs, hipMemcpyHostToDevice));
  B_HIP(hipMalloc(&b->d_cells, sizeof(mc_cell) * total_cells));
  B_HIP(hipMalloc(&b->d_fx, 16 * (size_t)total_cells));
  B_HIP(hipMalloc(&b->d_flags, 4 * (size_t)total_cells));
  B_HIP(hipMalloc(&b->d_ctr, sizeof(mc_counters)));
  B_HIP(hipHostMalloc(&b->h_ctr, sizeof(mc_counters), hipHostMallocDefault));
  B_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  B_HIP(hipEventCreate(&b->e0));
  B_HIP(hipEventCreate(&b->e1));
#undef B_HIP
  if (!rc) rc = batch_capture(b);
  if (rc) {
    mcgpu_batch_destroy(b);
    return rc;
  }
  *out = b;
  return MCGPU_OK;
}

int mcgpu_batch_run(mcgpu_batch *b, void *stream, mcgpu_stats *stats) {
  if (!b) return MCGPU_ERR_ARG;
  b->ctx->err.clear();
  MC_HIP(b->ctx, hipSetDevice(b->ctx->device));
  return batch_launch(b, reinterpret_cast<hipStream_t>(stream), stats);
}

int mcgpu_batch_options(const mcgpu_batch *b, mcgpu_options *opt) {
  if (!b || !opt) return MCGPU_ERR_ARG;
  *opt = b->opt;
  return MCGPU_OK;
}

int mcgpu_batch_autotune(mcgpu_batch *b, int32_t n_candidates, double *times_ms,
                         mcgpu_options *chosen) {
  if (!b) return MCGPU_ERR_ARG;
  mcgpu_context *ctx = b->ctx;
  ctx->err.clear();
  MC_HIP(ctx, hipSetDevice(ctx->device));
  struct cand {
    int block, ppl, refill;
  };
  static const cand cands[] = {{1024, 16, 8}, {1024, 4, 8}, {1024, 64, 8}, {512, 16, 8},
                               {1024, 16, 32}, {512, 64, 8}, {256, 64, 8}, {1024, 1, 8}};
  const int n_all = (int)(sizeof(cands) / sizeof(cands[0]));
  const int nc = n_candidates > 0 && n_candidates < n_all ? n_candidates : n_all;
  /* device copies of everything a run overwrites */
  const int64_t np = b->total_particles;
  void *save = nullptr;
  const size_t bank_bytes = (size_t)np * (8 + 8 + 8 + 8 + 4);
  const size_t w_bytes = 8 * (size_t)b->total_cells;
  MC_HIP(ctx, hipMalloc(&save, bank_bytes + w_bytes + 64));
  unsigned char *sv = (unsigned char *)save;
  auto snapshot = [&](bool restore) -> int {
    struct {
      void *p;
      size_t bytes;
    } parts[] = {{b->bank.seed, 8 * (size_t)np}, {b->bank.x, 8 * (size_t)np},
                 {b->bank.mu, 8 * (size_t)np},   {b->bank.wmc, 8 * (size_t)np},
                 {b->bank.index, 4 * (size_t)np}, {b->weights, w_bytes}};
    size_t off = 0;
    for (auto &pt : parts) {
      if (pt.bytes) {
        hipError_t e = restore ? hipMemcpyAsync(pt.p, sv + off, pt.bytes, hipMemcpyDeviceToDevice, b->stream)
                               : hipMemcpyAsync(sv + off, pt.p, pt.bytes, hipMemcpyDeviceToDevice, b->stream);
        if (e != hipSuccess) return mc_fail(ctx, MCGPU_ERR_HIP, "autotune snapshot", e);
      }
      off += pt.bytes;
    }
    hipError_t e = hipStreamSynchronize(b->stream);
    return e == hipSuccess ? MCGPU_OK : mc_fail(ctx, MCGPU_ERR_HIP, "autotune sync", e);
  };
  int rc = snapshot(false);
  double best = 1e300;
  int best_i = 0;
  const mcgpu_options base = b->opt;
  for (int i = 0; i < nc && !rc; ++i) {
    b->opt = base;
    b->opt.block_size = cands[i].block;
    b->opt.particles_per_lane = cands[i].ppl;
    b->opt.events_per_refill = cands[i].refill;
    b->opt.grid_blocks = 0;
    rc = batch_capture(b);
    mcgpu_stats st{};
    if (!rc) rc = batch_launch(b, b->stream, &st);
    if (!rc) rc = snapshot(true);
    if (times_ms) times_ms[i] = rc ? -1.0 : st.kernel_ms;
    if (!rc && st.kernel_ms < best) {
      best = st.kernel_ms;
      best_i = i;
    }
  }
  b->opt = base;
  b->opt.block_size = cands[best_i].block;
  b->opt.particles_per_lane = cands[best_i].ppl;
  b->opt.events_per_refill = cands[best_i].refill;
  b->opt.grid_blocks = 0;
  if (!rc) rc = batch_capture(b);
  (void)hipFree(save);
  if (!rc && chosen) *chosen = b->opt;
  return rc;
}

} /* extern "C" */
