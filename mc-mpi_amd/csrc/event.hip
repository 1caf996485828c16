/*
 * event.hip -- event-based transport kernel.
 *
 * The reference advances each particle to the end of its history in one go
 * (src/layer.cpp:195-200).  Here a pass applies exactly ONE event
 * (Layer::particle_step, src/layer.cpp:123-190) to every live particle of a
 * bank, and the bank is compacted between passes, so all lanes of a wave
 * always work on the same step of different particles.
 *
 * Layout and schedule (persistent blocks, no launch per pass):
 *  - each wave owns a particle pool in device memory: two halves (ping-pong)
 *    of P = 64 * particles_per_lane slots in structure-of-arrays form
 *    ({seed, wmc} side by side, x, mu, cell word, id, the pass its event
 *    count started from, collision counter), so every field is a coalesced
 *    wave access;
 *  - a pass gives every live slot one event, 64 slots (a chunk) at a time,
 *    through a four-stage software pipeline (below); the survivors of a
 *    compaction pass are packed into the other half -- a wave ballot +
 *    mbcnt prefix gives each survivor its rank; finished histories are
 *    written to their original bank slot (the particle id), so compaction
 *    order is free;
 *  - with compaction interval c > 1 (default 64), c-1 passes out of c write
 *    survivors back in place -- only the fields their event changed -- and
 *    mark finished slots dead instead of compacting;
 *  - when the pool has room for a chunk, it is refilled from the call's bank
 *    through the global work queue (one atomic per refill), so pools stay
 *    full until the bank is exhausted;
 *  - the exact fixed-point tally lives in LDS for the block's lifetime and is
 *    added to the global tally once (same arithmetic as the history kernel).
 * Results are bit-identical to the history kernel and the oracle: histories
 * are independent and the tally is an exact integer sum.
 */
#include "internal.h"

namespace {

/* A slot's cell word: the particle's cell as an offset into the layer
   (index - index_start, in [0, m)) in bits 0..30, and in bit 31 whether its
   last event crossed a cell edge.  On an edge, x is the edge its cell and
   direction imply, (mu < 0 ? index + 1 : index) * dx -- the very product the
   crossing stored (src/layer.cpp:141-143,168-171) -- and the slot's x field
   is neither written nor read.  All ones marks a dead slot.  A pass reads
   the cell word, seed, mu, wmc (and x off an edge) and writes back what the
   event changed; the particle id and the pass its event count started from
   are read only by compaction passes, when the event cap can bind, or when
   the history ends. */
#define ON_EDGE 0x80000000u
#define DEAD_CW 0xFFFFFFFFu

/* A wave's pool: two halves of P slots each, structure of arrays, except
   that seed and wmc, which every event reads and writes, sit side by side
   (one 16-byte load and one 16-byte store per slot instead of two of 8:
   the texture unit's cost goes by instructions).  Field k of slot j in
   half h lives at base + h*HALF + OFF_k*P + j*size_k: {seed, wmc} at 0,
   x at 16P, mu at 24P, then the 4-byte cell word, id, count origin and
   collision count.  pool_ref addresses the fields for the refill; the
   passes reach them through buffer resources (below). */
constexpr size_t SLOT_BYTES = 8 * 4 + 4 * 4; /* 48 bytes per slot */

struct pool_ref {
  unsigned char *b; /* base of one half */
  int P;
  /* seed and wmc side by side (both change at every event) */
  __device__ __forceinline__ uint64_t *sw() const { return reinterpret_cast<uint64_t *>(b); }
  __device__ __forceinline__ double *x() const { return reinterpret_cast<double *>(b + 16 * (size_t)P); }
  __device__ __forceinline__ double *mu() const { return reinterpret_cast<double *>(b + 24 * (size_t)P); }
  __device__ __forceinline__ uint32_t *cw() const { return reinterpret_cast<uint32_t *>(b + 32 * (size_t)P); }
  __device__ __forceinline__ uint32_t *id() const { return reinterpret_cast<uint32_t *>(b + 36 * (size_t)P); }
  /* the pass at which the particle's event count was 0: every pass gives
     every live slot one event, so the count before pass k is k - t0 */
  __device__ __forceinline__ uint32_t *t0() const { return reinterpret_cast<uint32_t *>(b + 40 * (size_t)P); }
  __device__ __forceinline__ uint32_t *col() const { return reinterpret_cast<uint32_t *>(b + 44 * (size_t)P); }
};

/* a new particle (cell word cw, not on an edge) into slot j.  Pool
   accesses are ordinary ones: with non-temporal hints (measured on 8-slot
   pools) the kernel was 40 % slower, the pools no longer staying in the
   caches. */
template <bool COLL>
__device__ __forceinline__ void put(const pool_ref &h, int j, const mc_state &p, uint32_t id,
                                    uint32_t t0, uint32_t cw) {
  h.sw()[2 * j] = p.seed;
  h.sw()[2 * j + 1] = __builtin_bit_cast(uint64_t, p.wmc);
  h.x()[j] = p.x;
  h.mu()[j] = p.mu;
  h.cw()[j] = cw;
  h.id()[j] = id;
  h.t0()[j] = t0;
  if (COLL) h.col()[j] = 0u;
}

/* ------------------------------------------------------------ the pipeline
   A pass gives every live slot of the wave's pool one event, 64 slots (a
   chunk) at a time.  Vector memory operations of a wave complete in the
   order they were issued, so a wait for one load is a wait for every
   earlier one: a chunk that loads its state, then its cell, then its exp
   table row, waiting for each, pays three memory latencies in a row.  Here
   a chunk's event runs in four stages, in four successive iterations of the
   wave's loop; each iteration advances four chunks by one stage, the oldest
   first:
     D  the exp table row has arrived: exp, dw, wmc, the tally (LDS), the
        exit of a finished history (to its bank slot), the write-back of the
        slot (in place, or compacted into the other half);
     C  the state and the cell have arrived: the draw, the edge distance,
        the collision decision, the new x and mu; load the exp table row;
     B  the cell word has arrived: load the rest of the state (x only where
        it is not implicit) and the cell;
     A  load the cell word of the next chunk.
   Every wait of an iteration is then for operations issued in the
   iteration before.  Pool accesses are raw buffer loads and stores whose
   masked lanes take an offset beyond the pool (no memory access; loads give
   0), and cell and table loads take a valid address on every lane, so each
   stage issues the same memory instructions on every common path: the
   compiler's wait counts stay exact.  A masked lane still costs the texture
   address unit its cycles, so the x load and the x and mu stores, which only
   a collision needs (about one event in a thousand on thin cells), are
   skipped by a wave-uniform branch when no lane of the chunk needs them; the
   waits then count for the common path, and the rare one waits a little
   longer.  The rare reads (a history's exit, a compaction pass's id and
   count origin, a binding cap) sit in branches of their own and are waited
   for inside them.  The passes up to the next
   compaction update the slots in place, so their chunks stream through the
   pipeline without a break (a slot is written in stage D of one pass before
   stage A of the next pass reads it, given at least four chunks); the
   pipeline drains after each compaction pass and before the pool is
   refilled. */

typedef __amdgpu_buffer_rsrc_t rsrc_t;
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
constexpr int OOB = 0x7FFFFFF0; /* a byte offset beyond every pool half */

__device__ __forceinline__ rsrc_t half_rsrc(unsigned char *half, int P) {
  return __builtin_amdgcn_make_buffer_rsrc(half, (short)0, (int)(SLOT_BYTES * (size_t)P),
                                           0x00020000);
}
__device__ __forceinline__ uint64_t ld64(rsrc_t r, int voff, int soff) {
  const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
  return ((uint64_t)v.y << 32) | v.x;
}
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
/* two 8-byte fields side by side */
__device__ __forceinline__ void ld128(rsrc_t r, int voff, int soff, uint64_t &a, uint64_t &b) {
  const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
  a = ((uint64_t)v.y << 32) | v.x;
  b = ((uint64_t)v.w << 32) | v.z;
}
__device__ __forceinline__ void st128(rsrc_t r, int voff, int soff, uint64_t a, uint64_t b) {
  u32x4 w;
  w.x = (uint32_t)a;
  w.y = (uint32_t)(a >> 32);
  w.z = (uint32_t)b;
  w.w = (uint32_t)(b >> 32);
  __builtin_amdgcn_raw_buffer_store_b128(w, r, voff, soff, 0);
}
__device__ __forceinline__ uint32_t ld32(rsrc_t r, int voff, int soff) {
  return __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
}
__device__ __forceinline__ void st64(rsrc_t r, int voff, int soff, uint64_t v) {
  u32x2 w;
  w.x = (uint32_t)v;
  w.y = (uint32_t)(v >> 32);
  __builtin_amdgcn_raw_buffer_store_b64(w, r, voff, soff, 0);
}
__device__ __forceinline__ void st32(rsrc_t r, int voff, int soff, uint32_t v) {
  __builtin_amdgcn_raw_buffer_store_b32(v, r, voff, soff, 0);
}
__device__ __forceinline__ double as_f64(uint64_t v) { return __builtin_bit_cast(double, v); }
__device__ __forceinline__ uint64_t as_u64(double v) { return __builtin_bit_cast(uint64_t, v); }

/* A copy the compiler cannot merge with its source: a stage copies the
   loads it consumes, so the next load into the same variable needs no copy
   (and no wait) at the loop's back edge. */
__device__ __forceinline__ uint64_t opq64(uint64_t v) {
  uint64_t r;
  asm volatile("v_mov_b64 %0, %1" : "=v"(r) : "v"(v));
  return r;
}
__device__ __forceinline__ uint32_t opq32(uint32_t v) {
  uint32_t r;
  asm volatile("v_mov_b32 %0, %1" : "=v"(r) : "v"(v));
  return r;
}

/* A kernel argument as a scalar register value of its own.  The argument
   block arrives in wide scalar loads, and a field the loop uses would
   otherwise keep the whole block live, or reload all of it from its spill
   lanes at every use. */
template <typename T> __device__ __forceinline__ T own_sreg(T v) {
  asm volatile("" : "+s"(v));
  return v;
}

/* The call's arguments as a block's LDS copy: the refill and the exit of a
   finished history (a few per ten thousand events) read the bank pointers
   from there when they run, behind a compiler barrier that keeps the loads
   in place (not hoisted out of the loop), so the pointers hold no registers
   during the passes. */
typedef mc_transport_args sargs;
__device__ __forceinline__ const sargs &fresh(const sargs &v) {
  asm volatile("" ::: "memory");
  return v;
}

/* A pointer the compiler can no longer trace to a kernel argument (read from
   LDS, or passed through own_sreg) would be accessed with FLAT instructions;
   a FLAT access still in flight makes every later wait a wait for all
   memory operations, which would undo the pipeline.  Pointers to device
   memory are therefore retyped as global ones before use. */
#define MC_GLOBAL __attribute__((address_space(1)))
template <typename T> __device__ __forceinline__ MC_GLOBAL T *gbl(T *p) {
  return (MC_GLOBAL T *)p;
}
__device__ __forceinline__ mc_cell ld_cell(MC_GLOBAL const mc_cell *p) {
  mc_cell c;
  c.sig_a = p->sig_a;
  c.sig_i = p->sig_i;
  return c;
}

__device__ __forceinline__ void refill_state(const sargs &v, int64_t i, mc_state &p) {
  const mcgpu_particle *aos = v.aos;
  if (aos) {
    MC_GLOBAL const mcgpu_particle *q = gbl(aos) + i;
    p.seed = q->seed;
    p.x = q->x;
    p.mu = q->mu;
    p.wmc = q->wmc;
    p.index = q->index;
  } else {
    p.seed = gbl(v.soa.seed)[i];
    p.x = gbl(v.soa.x)[i];
    p.mu = gbl(v.soa.mu)[i];
    p.wmc = gbl(v.soa.wmc)[i];
    p.index = gbl(v.soa.index)[i];
  }
}

/* the same stores as store_state, then the per-history counts asked for */
__device__ __forceinline__ void exit_state(const sargs &v, int64_t i, const mc_state &p, int code,
                                           uint32_t ev, uint32_t col) {
  mcgpu_particle *aos = v.aos;
  if (aos) {
    MC_GLOBAL mcgpu_particle *q = gbl(aos) + i;
    q->seed = p.seed;
    q->x = p.x;
    q->mu = p.mu;
    q->wmc = p.wmc;
    q->index = p.index;
    int32_t *r = v.result32;
    if (r) gbl(r)[i] = code;
  } else {
    gbl(v.soa.seed)[i] = p.seed;
    gbl(v.soa.x)[i] = p.x;
    gbl(v.soa.mu)[i] = p.mu;
    gbl(v.soa.wmc)[i] = p.wmc;
    gbl(v.soa.index)[i] = p.index;
    int8_t *r = v.result8;
    if (r) gbl(r)[i] = (int8_t)code;
  }
  uint64_t *e = v.events;
  if (e) gbl(e)[i] = ev;
  uint64_t *c = v.collisions;
  if (c) gbl(c)[i] = col;
}

/* Threads per block at most: 1024, 16 waves per CU when the block's tally
   fills the LDS.  The four stages keep about 60 values per lane live across
   the calls of the rare math paths, which leave only the callee-saved
   registers intact (8 of every 16 from v40 up); the kernel fits in 128
   registers without spills as long as every wave-uniform value (the pool
   size, the pass and chunk counters, the pool's buffer resource) stays
   scalar -- a pool size taken as divergent puts the buffer resource in
   vector registers and every pool access in a loop over its lanes. */
constexpr int EVENT_MAX_BLOCK = 1024;

/* the kernel's own LDS (block totals, the arguments), an upper bound: the
   tally and cells take what is left of the block's */
constexpr size_t EVENT_STATIC_LDS = 2048;
static_assert(5 * 8 + sizeof(mc_transport_args) + 64 <= EVENT_STATIC_LDS, "static LDS bound");

template <bool LDS_TALLY, bool LDS_CELLS, bool COLL>
__global__ __launch_bounds__(EVENT_MAX_BLOCK) void k_event(mc_transport_args a,
                                                           unsigned char *pools, int P,
                                                           int compact_every) {
  extern __shared__ __align__(16) unsigned char smem[];
  /* the log and exp coefficients in vector registers (as constants each use
     would take two scalar moves) */
  const mc_coef K = mc_coef_live();
  /* block totals: events, collisions, left, right, killed */
  __shared__ unsigned long long s_tot[5];
  __shared__ mc_transport_args s_a;
  const sargs &v = s_a;
  if (threadIdx.x == 0) s_a = a;
  mc_layer_k L;
  L.index_start = own_sreg(a.L.index_start);
  L.m = own_sreg(a.L.m);
  L.dx = own_sreg(a.L.dx);
  L.min_weight = own_sreg(a.L.min_weight);
  const int m = L.m;
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
  /* the cells in LDS when they fit beside the tally, else in device memory */
  MC_GLOBAL const mc_cell *cells_g = gbl(own_sreg(a.cells));
  mc_cell *cells_l = reinterpret_cast<mc_cell *>(smem + off);
  if (LDS_CELLS)
    for (int c = threadIdx.x; c < m; c += blockDim.x) cells_l[c] = ld_cell(cells_g + c);
  if (threadIdx.x < 5) s_tot[threadIdx.x] = 0;
  __syncthreads();

  /* this wave's pool: two halves of P slots */
  const int lane = __lane_id();
  /* wave index made explicitly uniform so the pool addresses stay scalar */
  const int wave = __builtin_amdgcn_readfirstlane(
      (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)));
  const size_t half_bytes = SLOT_BYTES * (size_t)P;
  unsigned char *mine = pools + (size_t)wave * 2 * half_bytes;
  const uint32_t max_ev =
      own_sreg(a.max_events > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)a.max_events);
  const double dx = L.dx;
  /* field offsets in a half (bytes) */
  const int F_X = 16 * P, F_MU = 24 * P, F_CW = 32 * P, F_ID = 36 * P,
            F_T0 = 40 * P, F_COL = 44 * P;
  const unsigned ce = (unsigned)compact_every;

  int cur = 0;
  int n_pool = 0; /* slots in use in the current half (wave-uniform) */
  bool queue_done = false;
  bool capped = false;
  /* a finished history (ev events, col collisions) to its bank slot */
  auto finish = [&](const mc_state &p, uint32_t id, uint32_t ev, uint32_t col) {
    int code = mc_exit_code(p, L);
    if (mc_alive(p, L)) { /* alive: stopped by the event cap */
      capped = true;
      code = 2;
    }
    exit_state(fresh(v), (int64_t)id, p, code, ev, col);
    atomicAdd(&s_tot[code == -1 ? 2 : (code == 1 ? 3 : 4)], code == 2 ? 0ull : 1ull);
  };

  for (unsigned pass = 0;;) {
    /* ---- refill from the work queue (wave-uniform condition) */
    if (!queue_done && n_pool <= P - 64) {
      const int want = P - n_pool;
      const sargs &w = fresh(v);
      unsigned long long base = 0;
      if (lane == 0)
        base = __hip_atomic_fetch_add(&gbl(w.ctr)->next, (unsigned long long)want,
                                      __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      base = mc_readlane64(base, 0);
      /* a load from LDS counts as divergent: made uniform again, so that
         the pool size and everything derived from it stay scalar */
      const unsigned long long n_bank = mc_readlane64((unsigned long long)(int64_t)w.n, 0);
      int take = 0;
      if (base < n_bank) {
        const unsigned long long left = n_bank - base;
        take = left < (unsigned long long)want ? (int)left : want;
      }
      if (take < want) queue_done = true;
      const pool_ref h{mine + cur * half_bytes, P};
      for (int k = lane; k < take; k += 64) {
        mc_state p;
        const uint32_t id = (uint32_t)(base + k);
        refill_state(fresh(v), (int64_t)id, p);
        if (mc_alive(p, L) && max_ev > 0) {
          put<COLL>(h, n_pool + k, p, id, (uint32_t)pass, (uint32_t)(p.index - L.index_start));
        } else { /* no event to apply: finished as it came, its slot dead */
          finish(p, id, 0u, 0u);
          h.cw()[n_pool + k] = DEAD_CW;
        }
      }
      n_pool += take;
      /* slots written by one lane are read by another lane of this wave */
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    }
    if (n_pool == 0) {
      if (queue_done) break;
      ++pass;
      continue;
    }

    /* ---- a segment: the passes up to the next compaction (one pass when
       the pool holds fewer than four chunks), streamed through the stages */
    const int nch = (n_pool + 63) >> 6;
    const unsigned last = nch >= 4 ? pass + (ce - pass % ce) % ce : pass;
    const int total = (int)(last - pass + 1) * nch;
    const bool ends_compacting = last % ce == 0;
    /* the event cap can bind in this segment (a count after pass q is at
       most q + 1): the passes then read every slot's count origin */
    const bool cap_live = (unsigned long long)last + 1 >= max_ev;
    const rsrc_t src = half_rsrc(mine + cur * half_bytes, P);
    const rsrc_t dst = half_rsrc(mine + (cur ^ 1) * half_bytes, P);
    int wpos = 0; /* wave-uniform output position of the compaction */
    int a_j = 0;           /* stage A's next chunk: index in the pool */
    unsigned a_pass = pass; /* and pass */
    unsigned int lane_ev = 0, lane_col = 0;

    /* stage registers -- A -> B */
    uint32_t a_cw = 0;
    bool a_valid = false;
    int a_slot = 0;
    unsigned a_q = 0;
    /* B -> C */
    uint64_t b_seed = 0, b_mu = 0, b_wmc = 0, b_x = 0;
    mc_cell b_cell;
    b_cell.sig_a = 0.0;
    b_cell.sig_i = 0.0;
    uint32_t b_col = 0;
    int32_t b_index = 0;
    bool b_live = false, b_on_edge = false;
    int b_slot = 0;
    unsigned b_q = 0;
    /* C -> D */
    mc_state c_p;
    c_p.seed = 0;
    c_p.x = 0.0;
    c_p.mu = 0.0;
    c_p.wmc = 0.0;
    c_p.index = 0;
    uint32_t c_col = 0;
    int32_t c_index_new = 0;
    int c_cell = 0;
    bool c_live = false, c_event = false, c_coll = false, c_was_on_edge = false;
    double c_arg = 0.0;
    mc_exp_row c_row;
    c_row.th = 0.0;
    c_row.tl = 0.0;
    int c_slot = 0;
    unsigned c_q = 0;

    for (int t = 0; t < total + 3; ++t) { /* wave-uniform trip count */
      /* ======== D: chunk t-3 */
      {
        /* the pass of this chunk compacts: the segment's last pass, when
           that is a compaction pass (passes before it are not) */
        const bool compact = ends_compacting && c_q == last;
        bool keep = false, on_edge = false;
        if (c_live) {
          if (c_event) {
            const double e = mc_exp_k_row(c_arg, &c_row, &K);
            const double dw = (1 - e) * c_p.wmc;
            c_p.wmc -= dw;
            c_p.index = c_index_new;
            on_edge = !c_coll;
            ++lane_ev;
            c_col += c_coll ? 1u : 0u;
            lane_col += c_coll ? 1u : 0u;
            mc_tally_with(t_lo, t_hi, c_cell, dw, [&](uint32_t g) {
              const sargs &w = fresh(v);
              __hip_atomic_fetch_or(gbl(w.flags) + c_cell, g, __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT);
              if (g & 8u)
                __hip_atomic_fetch_or(&gbl(w.ctr)->status, 2u, __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
            });
          }
          const uint32_t done = c_event ? 1u : 0u;
          bool at_cap = false;
          if (cap_live) /* wave-uniform: the count can reach the cap */
            at_cap = c_q - ld32(src, 4 * c_slot, F_T0) + done >= max_ev;
          keep = mc_alive(c_p, L) && !at_cap;
          if (!keep) { /* history finished: write it to its bank slot */
            const uint32_t id = ld32(src, 4 * c_slot, F_ID);
            const uint32_t t0 = ld32(src, 4 * c_slot, F_T0);
            /* both loads complete on this path (t0 is used only when the
               counts are asked for): no load still in flight leaves it */
            asm volatile("" ::"v"(id), "v"(t0));
            finish(c_p, id, c_q - t0 + done, c_col);
          }
        }
        /* the write-back: the offsets per path, then the same stores on both
           (a store on one path only would make the compiler assume fewer
           operations in flight after the branches and wait longer at the
           next stage); masked lanes aimed beyond the pool.  x and mu only
           when some lane needs them: every compaction pass, else after a
           collision */
        const uint32_t cw_val =
            keep ? ((uint32_t)(c_p.index - L.index_start) | (on_edge ? ON_EDGE : 0u)) : DEAD_CW;
        const bool w_x = keep && !on_edge; /* x is stored unless implicit */
        int o_seed, o_x, o_mu, o_cw, o_col;
        if (compact) { /* wave-uniform: survivors packed into the other half */
          const unsigned long long km = __ballot(keep);
          const int pos = wpos + mc_mbcnt64(km);
          wpos += __popcll(km);
          o_seed = keep ? 16 * pos : OOB;
          o_x = w_x ? 8 * pos : OOB;
          o_mu = keep ? 8 * pos : OOB;
          o_cw = o_col = keep ? 4 * pos : OOB;
        } else { /* in place: only what the event changed */
          o_seed = keep ? 16 * c_slot : OOB;
          o_x = o_mu = w_x ? 8 * c_slot : OOB; /* a collision's x and mu */
          /* a crossing's new cell, a collision's cleared edge bit, a death */
          o_cw = (c_live && (!keep || on_edge || c_was_on_edge)) ? 4 * c_slot : OOB;
          o_col = keep ? 4 * c_slot : OOB;
        }
        const rsrc_t r = compact ? dst : src;
        st128(r, o_seed, 0, c_p.seed, as_u64(c_p.wmc));
        if (compact || __ballot(w_x) != 0ull) {
          st64(r, o_x, F_X, as_u64(c_p.x));
          st64(r, o_mu, F_MU, as_u64(c_p.mu));
        }
        st32(r, o_cw, F_CW, cw_val);
        if (COLL) st32(r, o_col, F_COL, c_col);
        if (compact) { /* the id and the count origin move with the slot (the
                          wait for them is one pass in 64) */
          const int o = keep ? 4 * c_slot : OOB;
          const uint32_t id = ld32(src, o, F_ID), t0 = ld32(src, o, F_T0);
          st32(dst, o_cw, F_ID, id);
          st32(dst, o_cw, F_T0, t0);
        }
      }
      /* the stages in this order: each reads its inputs before the next one
         overwrites them, so a load's destination is never copied while the
         load is in flight (which would force a wait) */
      __builtin_amdgcn_sched_barrier(0);
      /* ======== C: chunk t-2 */
      {
        mc_state p;
        p.seed = opq64(b_seed);
        p.mu = as_f64(opq64(b_mu));
        p.wmc = as_f64(opq64(b_wmc));
        p.index = b_index;
        const double x_ld = as_f64(opq64(b_x));
        p.x = b_on_edge ? (p.mu < 0 ? b_index + 1 : b_index) * dx : x_ld;
        mc_cell cell;
        cell.sig_a = as_f64(opq64(as_u64(b_cell.sig_a)));
        cell.sig_i = as_f64(opq64(as_u64(b_cell.sig_i)));
        const uint32_t col0 = COLL ? opq32(b_col) : 0u;
        bool event = false, coll = false;
        int index_new = b_index;
        double arg = 0.0;
        const int c = b_index - L.index_start;
        bool below_cap = true;
        if (cap_live) { /* wave-uniform: the count origin, read only here */
          const uint32_t t0 = ld32(src, b_live ? 4 * b_slot : OOB, F_T0);
          below_cap = b_q - t0 < max_ev;
          asm volatile("" ::"v"(t0)); /* waited for on this path only */
        }
        if (b_live && mc_alive(p, L) && below_cap) {
          event = true;
          arg = mc_event_decide(p, cell, dx, K, coll, index_new);
        }
        c_row = mc_exp_lookup(arg); /* every lane: a row of the table */
        c_p = p;
        c_col = col0;
        c_index_new = index_new;
        c_cell = c;
        c_live = b_live;
        c_event = event;
        c_coll = coll;
        c_was_on_edge = b_on_edge;
        c_arg = arg;
        c_slot = b_slot;
        c_q = b_q;
      }
      __builtin_amdgcn_sched_barrier(0);
      /* ======== B: chunk t-1 */
      {
        const bool live = a_valid && a_cw != DEAD_CW;
        const bool on_edge = live && (a_cw & ON_EDGE) != 0;
        ld128(src, live ? 16 * a_slot : OOB, 0, b_seed, b_wmc);
        b_mu = ld64(src, live ? 8 * a_slot : OOB, F_MU);
        if (__ballot(live && !on_edge) != 0ull) /* else b_x is not read */
          b_x = ld64(src, (live && !on_edge) ? 8 * a_slot : OOB, F_X);
        if (COLL) b_col = ld32(src, live ? 4 * a_slot : OOB, F_COL);
        /* the cell, at an index inside the layer on every lane */
        const unsigned cu = a_cw & ~ON_EDGE;
        const unsigned ci = cu < (unsigned)m ? cu : 0u;
        if (LDS_CELLS)
          b_cell = cells_l[ci];
        else
          b_cell = ld_cell(cells_g + ci);
        b_index = (int32_t)cu + L.index_start;
        b_live = live;
        b_on_edge = on_edge;
        b_slot = a_slot;
        b_q = a_q;
      }
      __builtin_amdgcn_sched_barrier(0);
      /* ======== A: chunk t */
      {
        const bool active = t < total; /* wave-uniform */
        const int slot = a_j * 64 + lane;
        a_valid = active && slot < n_pool;
        a_cw = ld32(src, a_valid ? 4 * slot : OOB, F_CW);
        a_slot = slot;
        a_q = a_pass;
        if (++a_j == nch) { /* chunk index and pass of the next chunk */
          a_j = 0;
          ++a_pass;
        }
      }
    }
    {
      const unsigned long long e = mc_wave_sum(lane_ev), cl = mc_wave_sum(lane_col);
      if (lane == 0) {
        atomicAdd(&s_tot[0], e);
        atomicAdd(&s_tot[1], cl);
      }
    }
    if (ends_compacting) {
      n_pool = wpos;
      cur ^= 1;
    }
    /* compacted slots move between lanes, and the next refill writes slots
       another lane reads: complete the stores first */
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    pass = last + 1;
  }

  mc_counters *ctr = a.ctr;
  if (capped) atomicOr(&ctr->status, 1u);

  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(&ctr->events, s_tot[0]);
    atomicAdd(&ctr->collisions, s_tot[1]);
    atomicAdd(&ctr->n_left, s_tot[2]);
    atomicAdd(&ctr->n_right, s_tot[3]);
    atomicAdd(&ctr->n_killed, s_tot[4]);
  }
  if (LDS_TALLY) {
    unsigned long long *fx = a.fx;
    for (int c = threadIdx.x; c < m; c += blockDim.x) {
      mc_fx t;
      t.lo = t_lo[c];
      t.hi = t_hi[c];
      mc_fx_atomic_add(&fx[c], &fx[m + c], t);
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
  int block = o.block_size > 0 ? o.block_size : (lds > 81920 ? EVENT_MAX_BLOCK : 256);
  if (block > EVENT_MAX_BLOCK || block % 64)
    return mc_fail(ctx, MCGPU_ERR_ARG, "block_size must be a multiple of 64 and <= 1024");
  int grid = o.grid_blocks;
  if (grid <= 0) {
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, block, lds);
    if (e != hipSuccess || per_cu < 1) per_cu = 1;
    grid = per_cu * ctx->num_cus;
  }
  /* particles per lane: each wave's pool holds 64 * ppl slots.  Four, the
     fewest that keep four chunks streaming through the stages: the pools
     (about 100 MB on a full chip) then stay in the Infinity Cache (measured
     on the 10,000-cell slab at 1e8 particles: 4 slots 7.49 s, 5 7.59 s,
     6 7.78 s, 8 8.00 s, 12 8.08 s) */
  const int ppl = o.particles_per_lane > 0 ? o.particles_per_lane : 4;
  const int P = ppl * 64;
  int64_t per_block = (a.n + grid - 1) / grid;
  if (grid > 1 && per_block < block) { /* small banks: fewer blocks */
    grid = (int)((a.n + block - 1) / block);
    if (grid < 1) grid = 1;
  }
  /* compaction every 64th pass: the passes between write back only what
     changed, and a segment's chunks stream through the stages without a
     break (measured on the 10,000-cell slab at 1e8 particles, four slots per
     lane: every 8th pass 7.45 s, 16th 7.13 s, 32nd 6.98 s, 64th 6.91 s,
     128th 6.93 s, 256th 6.96 s, 1024th 7.10 s) */
  const int c = o.compaction_interval > 0 ? o.compaction_interval : 64;
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
  const bool lds_tally = tally + EVENT_STATIC_LDS <= ctx->lds_per_block;
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
