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
 *    slots in structure-of-arrays form (seed, x, mu, wmc, index, id, the
 *    pass its event count started from, collision counter), so every field
 *    is a coalesced 512-byte wave access;
 *  - a pass streams the live part of the current half: load, one event,
 *    tally, then the survivors are compacted into the other half -- a wave
 *    ballot + mbcnt prefix gives each survivor its rank, one LDS atomic per
 *    wave reserves the wave's range; finished histories are written to their
 *    original bank slot (the particle id), so compaction order is free;
 *  - with compaction interval c > 1 (default 16), c-1 passes out of c write
 *    survivors back in place -- only the fields their event changed -- and
 *    mark finished slots dead instead of compacting;
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
/* Bit 31 of a live slot's id: the particle's last event crossed a cell edge,
   so its x is the edge its cell index and direction imply,
   (mu < 0 ? index + 1 : index) * dx -- the very product the crossing stored
   (src/layer.cpp:141-143,168-171) -- and the slot's x field is neither
   written nor read.  Most events are crossings, so a pass moves 36 bytes per
   particle each way instead of 44.  Particle ids stay below 2^31 - 1. */
#define ON_EDGE 0x80000000u

/* A block's pool: two halves of P slots each, structure of arrays.  Field
   k of slot j in half h lives at base + h*HALF + OFF_k*P + j*size_k, with
   the 8-byte fields first.  Addresses are formed from one base pointer so
   the compiler keeps a single scalar base per half live. */
constexpr size_t SLOT_BYTES = 8 * 4 + 4 * 4; /* 48 bytes per slot */

struct pool_ref {
  unsigned char *b; /* base of one half */
  int P;
  __device__ __forceinline__ uint64_t *seed() const { return reinterpret_cast<uint64_t *>(b); }
  __device__ __forceinline__ double *x() const { return reinterpret_cast<double *>(b + 8 * (size_t)P); }
  __device__ __forceinline__ double *mu() const { return reinterpret_cast<double *>(b + 16 * (size_t)P); }
  __device__ __forceinline__ double *wmc() const { return reinterpret_cast<double *>(b + 24 * (size_t)P); }
  __device__ __forceinline__ int32_t *index() const { return reinterpret_cast<int32_t *>(b + 32 * (size_t)P); }
  __device__ __forceinline__ uint32_t *id() const { return reinterpret_cast<uint32_t *>(b + 36 * (size_t)P); }
  /* the pass at which the particle's event count was 0: every pass gives
     every live slot one event, so the count before pass k is k - t0 */
  __device__ __forceinline__ uint32_t *t0() const { return reinterpret_cast<uint32_t *>(b + 40 * (size_t)P); }
  __device__ __forceinline__ uint32_t *col() const { return reinterpret_cast<uint32_t *>(b + 44 * (size_t)P); }
};

/* Pool accesses use ordinary loads and stores by default: measured on
   MI355X (10,000-cell slab, 2e7 particles) non-temporal hints made the kernel 40 %
   slower, as the pools (about 180 MB at 8 slots per lane) then no longer stay
   in the Infinity Cache.  The hints remain selectable for experiments. */
#ifndef MC_POOL_NT_LOAD
#define MC_POOL_NT_LOAD 0
#endif
#ifndef MC_POOL_NT_STORE
#define MC_POOL_NT_STORE 0
#endif
template <typename T> __device__ __forceinline__ T ld_nt(const T *p) {
#if MC_POOL_NT_LOAD
  return __builtin_nontemporal_load(p);
#else
  return *p;
#endif
}
template <typename T> __device__ __forceinline__ void st_nt(T *p, T v) {
#if MC_POOL_NT_STORE
  __builtin_nontemporal_store(v, p);
#else
  *p = v;
#endif
}

template <bool COLL>
__device__ __forceinline__ void put(const pool_ref &h, int j, const mc_state &p, uint32_t id,
                                    uint32_t t0, uint32_t col, bool on_edge) {
  st_nt(h.seed() + j, p.seed);
  if (!on_edge) st_nt(h.x() + j, p.x);
  st_nt(h.mu() + j, p.mu);
  st_nt(h.wmc() + j, p.wmc);
  st_nt(h.index() + j, p.index);
  st_nt(h.id() + j, on_edge ? (id | ON_EDGE) : id);
  st_nt(h.t0() + j, t0);
  if (COLL) st_nt(h.col() + j, col);
}

/* An in-place pass (no compaction) writes back only what the event changed:
   seed and wmc always; the cell index after a crossing; x and mu after a
   collision; the id when the on-edge flag changed.  (t0 stays.) */
template <bool COLL>
__device__ __forceinline__ void put_changed(const pool_ref &h, int j, const mc_state &p,
                                            uint32_t id, uint32_t col, bool on_edge,
                                            bool was_on_edge) {
  st_nt(h.seed() + j, p.seed);
  st_nt(h.wmc() + j, p.wmc);
  if (COLL) st_nt(h.col() + j, col);
  if (on_edge) {
    st_nt(h.index() + j, p.index);
  } else {
    st_nt(h.x() + j, p.x);
    st_nt(h.mu() + j, p.mu);
  }
  if (on_edge != was_on_edge) st_nt(h.id() + j, on_edge ? (id | ON_EDGE) : id);
}

/* the registers of one pool slot */
struct slot_regs {
  mc_state p;
  uint32_t id, ev, col;
  bool on_edge; /* the slot's x was implicit */
  bool valid;
};

template <bool COLL>
__device__ __forceinline__ void load_slot(const pool_ref &h, int j, double dx, uint32_t pass,
                                          slot_regs &r) {
  r.id = ld_nt(h.id() + j);
  r.p.seed = ld_nt(h.seed() + j);
  r.p.mu = ld_nt(h.mu() + j);
  r.p.wmc = ld_nt(h.wmc() + j);
  r.p.index = ld_nt(h.index() + j);
  r.ev = pass - ld_nt(h.t0() + j);
  r.col = COLL ? ld_nt(h.col() + j) : 0u;
  r.on_edge = r.id != DEAD_ID && (r.id & ON_EDGE);
  if (r.on_edge) {
    r.id &= ~ON_EDGE;
    r.p.x = (r.p.mu < 0 ? r.p.index + 1 : r.p.index) * dx;
  } else if (r.id != DEAD_ID) {
    r.p.x = ld_nt(h.x() + j);
  }
}

template <bool LDS_TALLY, bool LDS_CELLS, bool COLL>
__global__ __launch_bounds__(1024) void k_event(mc_transport_args a, unsigned char *pools, int P,
                                                int compact_every) {
  extern __shared__ __align__(16) unsigned char smem[];
  const mc_coef K = mc_coef_live();
  /* block totals: events, collisions, left, right, killed */
  __shared__ unsigned long long s_tot[5];
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
  if (threadIdx.x < 5) s_tot[threadIdx.x] = 0;
  __syncthreads();

  /* this wave's pool: two halves of P slots */
  const int lane = __lane_id();
  /* wave index made explicitly uniform so the pool addresses stay scalar */
  const int wave = __builtin_amdgcn_readfirstlane(
      (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)));
  const size_t half_bytes = SLOT_BYTES * (size_t)P;
  unsigned char *mine = pools + (size_t)wave * 2 * half_bytes;
  const uint32_t max_ev = a.max_events > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)a.max_events;

  int cur = 0;
  int n_pool = 0; /* slots in use in the current half (wave-uniform) */
  bool queue_done = false;
  bool capped = false;

  for (unsigned pass = 0;; ++pass) {
    /* ---- refill from the work queue (wave-uniform condition) */
    if (!queue_done && n_pool <= P - 64) {
      const int want = P - n_pool;
      unsigned long long base = 0;
      if (lane == 0) base = atomicAdd(&a.ctr->next, (unsigned long long)want);
      base = mc_readlane64(base, 0);
      int take = 0;
      if (base < (unsigned long long)a.n) {
        const unsigned long long left = (unsigned long long)a.n - base;
        take = left < (unsigned long long)want ? (int)left : want;
      }
      if (take < want) queue_done = true;
      const pool_ref h{mine + cur * half_bytes, P};
      for (int k = lane; k < take; k += 64) {
        mc_state p;
        load_state(a, (int64_t)(base + k), p);
        put<COLL>(h, n_pool + k, p, (uint32_t)(base + k), (uint32_t)pass, 0u, false);
      }
      n_pool += take;
      /* slots written by one lane are read by another lane of this wave */
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    }
    if (n_pool == 0) {
      if (queue_done) break;
      continue;
    }

    /* ---- one event per live particle */
    const bool compact = (pass % (unsigned)compact_every) == 0;
    const pool_ref src{mine + cur * half_bytes, P};
    const pool_ref dst{mine + (cur ^ 1) * half_bytes, P};
    int wpos = 0; /* wave-uniform output position of the compaction */
    unsigned int pass_ev = 0, pass_col = 0; /* this lane's events of the pass */
    /* software pipeline: the next slot is loaded while this one's event is
       computed, so its memory latency overlaps the FP64 work */
    slot_regs nxt;
    nxt.valid = lane < n_pool;
    if (nxt.valid) load_slot<COLL>(src, lane, a.L.dx, pass, nxt);
    for (int j0 = 0; j0 < n_pool; j0 += 64) { /* wave-uniform trip count */
      const int j = j0 + lane;
      mc_state p = nxt.p;
      uint32_t id = nxt.valid ? nxt.id : DEAD_ID, ev = nxt.ev, col = nxt.col;
      const bool was_on_edge = nxt.valid && nxt.on_edge;
      if (j0 + 64 < n_pool) {
        nxt.valid = j + 64 < n_pool;
        if (nxt.valid) load_slot<COLL>(src, j + 64, a.L.dx, pass, nxt);
      }
      const bool live = id != DEAD_ID;
      bool keep = false;
      bool on_edge = false; /* set by a crossing */
      if (live) {
        if (mc_alive(p, a.L) && ev < max_ev) {
          const int c = p.index - a.L.index_start;
          bool coll;
          double di;
          const double dw = mc_event(p, cells, c, a.L.dx, K, coll, di);
          on_edge = !coll;
          ++ev;
          ++pass_ev;
          col += coll ? 1u : 0u;
          pass_col += coll ? 1u : 0u;
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
          atomicAdd(&s_tot[code == -1 ? 2 : (code == 1 ? 3 : 4)], code == 2 ? 0ull : 1ull);
        }
      }
      if (compact) {
        /* stream compaction: survivors keep their order, ranks from the ballot */
        const unsigned long long mask = __ballot(keep);
        if (keep) put<COLL>(dst, wpos + mc_mbcnt64(mask), p, id, pass + 1u - ev, col, on_edge);
        wpos += __popcll(mask);
      } else if (live) {
        if (keep)
          put_changed<COLL>(src, j, p, id, col, on_edge, was_on_edge);
        else
          st_nt(src.id() + j, DEAD_ID);
      }
    }
    {
      const unsigned long long e = mc_wave_sum(pass_ev), cl = mc_wave_sum(pass_col);
      if (lane == 0) {
        atomicAdd(&s_tot[0], e);
        atomicAdd(&s_tot[1], cl);
      }
    }
    if (compact) {
      n_pool = wpos;
      cur ^= 1;
      /* compacted slots move between lanes: complete the stores first */
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    }
  }

  if (capped) atomicOr(&a.ctr->status, 1u);

  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(&a.ctr->events, s_tot[0]);
    atomicAdd(&a.ctr->collisions, s_tot[1]);
    atomicAdd(&a.ctr->n_left, s_tot[2]);
    atomicAdd(&a.ctr->n_right, s_tot[3]);
    atomicAdd(&a.ctr->n_killed, s_tot[4]);
  }
  if (LDS_TALLY) {
    for (int c = threadIdx.x; c < m; c += blockDim.x) {
      mc_fx v;
      v.lo = t_lo[c];
      v.hi = t_hi[c];
      mc_fx_atomic_add(&a.fx[c], &a.fx[m + c], v);
    }
  }
}

template <bool T, bool C, bool COLL>
int launch_one(mcgpu_context *ctx, const mc_transport_args &a, const mcgpu_options &o,
               hipStream_t s, size_t lds) {
  auto kern = k_event<T, C, COLL>;
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
  /* particles per lane: each wave's pool holds 64 * ppl slots */
  const int ppl = o.particles_per_lane > 0 ? o.particles_per_lane : 8;
  const int P = ppl * 64;
  int64_t per_block = (a.n + grid - 1) / grid;
  if (grid > 1 && per_block < block) { /* small banks: fewer blocks */
    grid = (int)((a.n + block - 1) / block);
    if (grid < 1) grid = 1;
  }
  /* compaction every 16th pass: the passes between write back only what
     changed (measured on the 10,000-cell slab: every pass 10.13 s, every 4th 9.25 s,
     8th 9.11 s, 16th 9.04 s, 32nd 9.02 s) */
  const int c = o.compaction_interval > 0 ? o.compaction_interval : 16;
  const size_t waves = (size_t)grid * (size_t)(block / 64);
  int rc = mc_ensure(ctx, &ctx->work, &ctx->work_bytes, waves * 2 * SLOT_BYTES * (size_t)P);
  if (rc) return rc;
  hipLaunchKernelGGL(kern, dim3(grid), dim3(block), lds, s, a, (unsigned char *)ctx->work, P, c);
  MC_HIP(ctx, hipGetLastError());
  return MCGPU_OK;
}

} // namespace

int mc_launch_event(mcgpu_context *ctx, const mc_transport_args &a, const mcgpu_options &o,
                    hipStream_t s) {
  if (a.n >= (int64_t)(ON_EDGE - 1))
    return mc_fail(ctx, MCGPU_ERR_ARG, "event kernel: bank larger than 2^31-2 particles");
  const size_t m = (size_t)a.L.m;
  const size_t tally = 16 * m, cells = sizeof(mc_cell) * m;
  const bool lds_tally = tally <= ctx->lds_per_block;
  const bool lds_cells = lds_tally && tally + cells <= 65536;
  /* collision counters ride in the pool only when the caller asked for them
     per particle; the call's total is kept either way */
  if (a.collisions) {
    if (lds_tally && lds_cells) return launch_one<true, true, true>(ctx, a, o, s, tally + cells);
    if (lds_tally) return launch_one<true, false, true>(ctx, a, o, s, tally);
    return launch_one<false, false, true>(ctx, a, o, s, 0);
  }
  if (lds_tally && lds_cells) return launch_one<true, true, false>(ctx, a, o, s, tally + cells);
  if (lds_tally) return launch_one<true, false, false>(ctx, a, o, s, tally);
  return launch_one<false, false, false>(ctx, a, o, s, 0);
}
