/*
 * cycle.hip -- mcgpu_sync_cycle: the synchronous loop of the reference's
 * workers (WorkerSync::spin, src/worker_sync.cpp:24-135) with every rank's
 * layer resident on one GPU.
 *
 * Device memory:
 *  - one particle pool (structure of arrays) in which rank r owns the slots
 *    [base_r, base_r + capacity_r); its bank (the reference's
 *    Layer::particles) is the prefix of `size_r` slots, in vector order;
 *  - the cells, exact fixed-point tallies and weights of all ranks,
 *    concatenated (rank r's cells at its cell_offset);
 *  - escape buffers: rank r's particles_left at slots [2rS, 2rS + S) and
 *    particles_right at [(2r+1)S, (2r+2)S) of a second pool, S =
 *    nb_particles_per_cycle.
 * Per cycle (one stream, one host synchronisation):
 *  1. lazy creation into the source rank's bank (host decides the count,
 *     src/layer.cpp:89-121 and :247-248);
 *  2. one batched transport launch whose tiles cover every rank's last
 *     n_r = min(size_r, S) particles (src/layer.cpp:250-330), tallies in the
 *     ranks' cells, then tally -> weights (one rounding per cell and cycle, as
 *     each Layer::simulate call adds its own tally);
 *  3. partition: iteration i of the reference transports particles[size-1-i]
 *     and appends escapes in ascending i (src/layer.cpp:332-346), so the
 *     escapes are compacted in reverse slot order, in chunks of 1,024
 *     iterations: per-chunk counts, a per-rank scan of the chunk counts, a
 *     plan of the exchange (one thread per rank), then the scatter into the
 *     escape buffers;
 *  4. exchange: each rank's bank is cut to size_r - n_r and receives, in
 *     the order of the two phases of src/worker_sync.cpp:55-109, the right
 *     escapes of rank r-1 and the left escapes of rank r+1 (odd ranks take
 *     their right neighbour's first, even ranks their left neighbour's);
 *     escapes through a global border are disabled (src/layer.cpp:348-357);
 *  5. the host reads back the per-rank summary (new sizes, disabled counts)
 *     and the device counters, and checks termination
 *     (src/worker_sync.cpp:111-119).
 */
#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "internal.h"

namespace {

constexpr int64_t MAX_PARTICLES_VECTOR = 1000000; /* src/layer.cpp:15 */
constexpr int CHUNK = 1024;                       /* partition chunk: one block */

struct pool_k {
  uint64_t *seed;
  double *x, *mu, *wmc;
  int32_t *index;
};

/* one partition chunk: iterations [i0, i1) of a rank's transport, which
   read slots seg_end - 1 - i */
struct chunk_k {
  int32_t rank;
  int32_t pad_;
  int64_t i0, i1;
  int64_t seg_end;
};

/* per rank and cycle, from the host */
struct rank_in {
  int64_t pool_base;
  int64_t size; /* bank size after creation */
  int64_t n;    /* transported this cycle */
  int32_t chunk_first, chunk_count;
  int32_t left_border, right_border;
};

struct chunk_cnt {
  int32_t l, r, k, pad_;
};

struct chunk_off {
  int64_t l, r; /* offsets of the chunk's escapes in the rank's buffers */
};

/* per rank and cycle: escapes and kills of its transport */
struct rank_tot {
  int64_t n_left, n_right, n_killed;
};

/* per rank and cycle, to the host */
struct rank_sum {
  int64_t n_left, n_right, n_killed;
  int64_t keep;          /* bank slots kept: size - n */
  int64_t src1, cnt1;    /* first received buffer (escape-pool slot, count) */
  int64_t src2, cnt2;    /* second received buffer */
  int64_t new_size;
  int64_t disabled_add;  /* kills + escapes through a global border */
  int64_t stranded;      /* escapes with neither a border nor a neighbour */
};

__device__ __forceinline__ void copy_particle(const pool_k &src, int64_t i, const pool_k &dst,
                                              int64_t j) {
  dst.seed[j] = src.seed[i];
  dst.x[j] = src.x[i];
  dst.mu[j] = src.mu[i];
  dst.wmc[j] = src.wmc[i];
  dst.index[j] = src.index[i];
}

__global__ void k_cyc_cells(int64_t total, const double *sigs, const double *ar, mc_cell *cells) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < total) cells[c] = mc_make_cell(sigs[c], ar[c]);
}

__global__ void k_cyc_clear(int64_t total, unsigned long long *fx, uint32_t *flags,
                            mc_counters *ctr) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < total) {
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

__global__ void k_cyc_finalize(int64_t total, const unsigned long long *fx, const uint32_t *flags,
                               double *weights) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < total) weights[c] = weights[c] + mc_fx_finalize(fx[c], fx[total + c], flags[c]);
}

/* exit codes of a chunk's iterations (3 past the chunk's end) */
__device__ __forceinline__ int chunk_code(const chunk_k &ch, const int8_t *result8, int64_t i) {
  return i < ch.i1 ? (int)result8[ch.seg_end - 1 - i] : 3;
}

__global__ __launch_bounds__(CHUNK) void k_cyc_count(const chunk_k *chunks, const int8_t *result8,
                                                    chunk_cnt *cnt) {
  __shared__ int s[3];
  const int t = threadIdx.x;
  if (t < 3) s[t] = 0;
  __syncthreads();
  const chunk_k ch = chunks[blockIdx.x];
  const int code = chunk_code(ch, result8, ch.i0 + t);
  const unsigned long long bl = __ballot(code == -1), br = __ballot(code == 1),
                           bk = __ballot(code == 0);
  if (__lane_id() == 0) {
    atomicAdd(&s[0], __popcll(bl));
    atomicAdd(&s[1], __popcll(br));
    atomicAdd(&s[2], __popcll(bk));
  }
  __syncthreads();
  if (t == 0) {
    chunk_cnt c;
    c.l = s[0];
    c.r = s[1];
    c.k = s[2];
    c.pad_ = 0;
    cnt[blockIdx.x] = c;
  }
}

/* inclusive scan over a wave */
__device__ __forceinline__ int64_t wave_incl_scan(int64_t v) {
  const int lane = __lane_id();
  for (int off = 1; off < 64; off <<= 1) {
    const int64_t u = __shfl_up(v, off);
    if (lane >= off) v += u;
  }
  return v;
}

/* One block per rank: exclusive scan of its chunks' escape counts (at most
   CHUNK chunks: nb_particles_per_cycle <= 1e6) and the rank's totals. */
__global__ __launch_bounds__(CHUNK) void k_cyc_scan(const rank_in *rin, const chunk_cnt *cnt,
                                                   chunk_off *off, rank_tot *tot) {
  __shared__ int64_t s_wl[CHUNK / 64], s_wr[CHUNK / 64], s_wk[CHUNK / 64];
  const int r = blockIdx.x, t = threadIdx.x, w = t / 64, lane = __lane_id();
  const rank_in ri = rin[r];
  chunk_cnt c{0, 0, 0, 0};
  if (t < ri.chunk_count) c = cnt[ri.chunk_first + t];
  const int64_t il = wave_incl_scan(c.l), ir = wave_incl_scan(c.r), ik = wave_incl_scan(c.k);
  if (lane == 63) {
    s_wl[w] = il;
    s_wr[w] = ir;
    s_wk[w] = ik;
  }
  __syncthreads();
  int64_t pl = 0, pr = 0;
  for (int v = 0; v < w; ++v) {
    pl += s_wl[v];
    pr += s_wr[v];
  }
  if (t < ri.chunk_count) {
    chunk_off o;
    o.l = pl + il - c.l;
    o.r = pr + ir - c.r;
    off[ri.chunk_first + t] = o;
  }
  if (t == 0) {
    int64_t L = 0, R = 0, K = 0;
    for (int v = 0; v < CHUNK / 64; ++v) {
      L += s_wl[v];
      R += s_wr[v];
      K += s_wk[v];
    }
    rank_tot s;
    s.n_left = L;
    s.n_right = R;
    s.n_killed = K;
    tot[r] = s;
  }
}

/* The exchange plan, one thread per rank: where each rank's escapes go and
   what each bank receives, in the reference's phase order. */
__global__ void k_cyc_plan(int R, int64_t S, const rank_in *rin, const rank_tot *tot,
                           rank_sum *sum) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= R) return;
  const rank_in ri = rin[q];
  rank_sum s;
  s.n_left = tot[q].n_left;
  s.n_right = tot[q].n_right;
  s.n_killed = tot[q].n_killed;
  s.keep = ri.size - ri.n;
  /* left escapes of r go to r-1 unless r has the global left border;
     right escapes of r go to r+1 unless r has the global right border */
  s.disabled_add = s.n_killed + (ri.left_border ? s.n_left : 0) +
                   (ri.right_border ? s.n_right : 0);
  s.stranded = ((!ri.left_border && q == 0) ? s.n_left : 0) +
               ((!ri.right_border && q == R - 1) ? s.n_right : 0);
  int64_t from_left = 0, from_right = 0; /* right escapes of q-1, left escapes of q+1 */
  if (q > 0 && !rin[q - 1].right_border) from_left = tot[q - 1].n_right;
  if (q + 1 < R && !rin[q + 1].left_border) from_right = tot[q + 1].n_left;
  const int64_t src_left = (2 * (int64_t)(q - 1) + 1) * S; /* particles_right of q-1 */
  const int64_t src_right = 2 * (int64_t)(q + 1) * S;      /* particles_left of q+1 */
  /* phase 1 pairs (odd, odd+1), phase 2 pairs (even, even+1): an even rank
     hears from its left neighbour first, an odd rank from its right one */
  if ((q & 1) == 0) {
    s.src1 = src_left;
    s.cnt1 = from_left;
    s.src2 = src_right;
    s.cnt2 = from_right;
  } else {
    s.src1 = src_right;
    s.cnt1 = from_right;
    s.src2 = src_left;
    s.cnt2 = from_left;
  }
  s.new_size = s.keep + from_left + from_right;
  sum[q] = s;
}

/* escapes of a chunk into the rank's buffers, in ascending iteration order */
__global__ __launch_bounds__(CHUNK) void k_cyc_scatter(const chunk_k *chunks, const int8_t *result8,
                                                      const chunk_off *off, int64_t S, pool_k pool,
                                                      pool_k esc) {
  __shared__ int s_l[CHUNK / 64], s_r[CHUNK / 64];
  const int t = threadIdx.x, w = t / 64;
  const chunk_k ch = chunks[blockIdx.x];
  const int64_t i = ch.i0 + t;
  const int code = chunk_code(ch, result8, i);
  const unsigned long long bl = __ballot(code == -1), br = __ballot(code == 1);
  if (__lane_id() == 0) {
    s_l[w] = __popcll(bl);
    s_r[w] = __popcll(br);
  }
  __syncthreads();
  if (code == -1 || code == 1) {
    int64_t before = 0;
    const int *sw = code == -1 ? s_l : s_r;
    for (int v = 0; v < w; ++v) before += sw[v];
    const chunk_off o = off[blockIdx.x];
    const int64_t j = code == -1
                          ? 2 * (int64_t)ch.rank * S + o.l + before + mc_mbcnt64(bl)
                          : (2 * (int64_t)ch.rank + 1) * S + o.r + before + mc_mbcnt64(br);
    copy_particle(pool, ch.seg_end - 1 - i, esc, j);
  }
}

/* received buffers into the banks after their kept slots */
__global__ void k_cyc_exchange(const rank_in *rin, const rank_sum *sum, pool_k pool, pool_k esc) {
  const int q = blockIdx.y;
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const rank_sum s = sum[q];
  int64_t src;
  if (k < s.cnt1)
    src = s.src1 + k;
  else if (k < s.cnt1 + s.cnt2)
    src = s.src2 + (k - s.cnt1);
  else
    return;
  copy_particle(esc, src, pool, rin[q].pool_base + s.keep + k);
}

inline unsigned blocks_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

/* device allocations of one run, freed on every exit path */
struct cycle_mem {
  std::vector<void *> dev;
  std::vector<void *> host;
  std::vector<hipEvent_t> events;
  ~cycle_mem() {
    for (void *p : dev)
      if (p) (void)hipFree(p);
    for (void *p : host)
      if (p) (void)hipHostFree(p);
    for (hipEvent_t e : events)
      if (e) (void)hipEventDestroy(e);
  }
};

/* wall clock in seconds since the epoch (the reference's timer reads
   MPI_Wtime plus the offset to CLOCK_REALTIME, src/timer.cpp:7-23) */
double wall_now() {
  return std::chrono::duration<double>(std::chrono::system_clock::now().time_since_epoch())
      .count();
}

/* The statistics windows of WorkerSync::spin (src/worker_sync.cpp:25-27,
   121-133): each cycle's phase times go to the open window; after a cycle
   that did not end the loop, the window closes if it is older than
   statistics_cycle_time, and the next one opens; the last window closes
   when the loop ends, without counting the cycle that ended it. */
struct window_log {
  double period = 0;
  mcgpu_cycle_window_fn fn = nullptr;
  void *user = nullptr;
  mcgpu_cycle_window w{};
  void open(double t) {
    w = mcgpu_cycle_window{};
    w.starttime = t;
  }
  void close(double t) {
    w.endtime = t;
    if (fn) fn(user, &w);
  }
};

/* ---------------------------------------------------------------------
   One rank: every cycle at once.

   With one rank nothing is exchanged, and when both ends of its layer are
   global borders every particle a cycle transports leaves the bank for good
   (an escape through a border or a kill, src/layer.cpp:332-360).  Which
   particles cycle k transports -- the last min(size, S) of the bank after
   that cycle's lazy creation (src/layer.cpp:89-121,247-250) -- then follows
   from the counts alone, never from a transport result.  So the cycles'
   transports are independent and run concurrently in one batched launch,
   each cycle a slab of its own with its own exact tally; the tallies are
   then added to weights_absorbed one cycle after the other, each rounded
   once (as each Layer::simulate call adds its own).  Histories, counts and
   weights are those of the cycle-by-cycle loop, bit for bit.
   The bank's logical order is kept as a stack of segments (parts of the
   initial bank, chunks of the lazy source); each cycle's particles are laid
   out contiguously, in bank order, in a work buffer. */

struct one_seg {
  int kind;          /* 0: initial bank, 1: lazy source */
  int64_t begin, n;  /* initial-bank index, or position in the source's seed chain */
};

struct one_cycle {
  int64_t first, n;              /* the cycle's particles in the work buffer */
  std::vector<one_seg> pieces;   /* where they come from, in bank order */
};

/* the cycles' cell tables: the rank's m cells once per cycle */
__global__ void k_one_cells(int64_t G, int m, const double *sigs, const double *ar,
                            mc_cell *cells) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < G * m) {
    const int c = (int)(i % m);
    cells[i] = mc_make_cell(sigs[c], ar[c]);
  }
}

/* weights[c] += the tally of cycle g, rounded once, for g = 0 .. G-1 in order */
__global__ void k_one_finalize(int m, int64_t G, const unsigned long long *fx,
                               const uint32_t *flags, double *weights) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m) return;
  const int64_t total = G * m;
  double w = weights[c];
  for (int64_t g = 0; g < G; ++g) {
    const int64_t i = g * m + c;
    w = w + mc_fx_finalize(fx[i], fx[total + i], flags[i]);
  }
  weights[c] = w;
}

int alloc_pool(mcgpu_context *ctx, int64_t slots, pool_k *p, int8_t **result8) {
  MC_HIP(ctx, hipMalloc(&p->seed, 8 * (size_t)slots));
  MC_HIP(ctx, hipMalloc(&p->x, 8 * (size_t)slots));
  MC_HIP(ctx, hipMalloc(&p->mu, 8 * (size_t)slots));
  MC_HIP(ctx, hipMalloc(&p->wmc, 8 * (size_t)slots));
  MC_HIP(ctx, hipMalloc(&p->index, 4 * (size_t)slots));
  if (result8) MC_HIP(ctx, hipMalloc(result8, (size_t)slots));
  return MCGPU_OK;
}

void free_pool(pool_k *p, int8_t **result8) {
  void *ptrs[] = {p->seed, p->x, p->mu, p->wmc, p->index};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  *p = pool_k{};
  if (result8 && *result8) {
    (void)hipFree(*result8);
    *result8 = nullptr;
  }
}

struct pool_guard {
  pool_k p{};
  int8_t *result8 = nullptr;
  ~pool_guard() { free_pool(&p, &result8); }
};

/* The single-rank path applies when the rank's layer has global borders at
   both ends and its counts close (bank + source + disabled = nb_particles):
   the loop then ends exactly when bank and source are empty. */
bool one_rank_eligible(int R, const mcgpu_cycle_rank *ranks, int64_t nb_particles) {
  if (R != 1) return false;
  const mcgpu_cycle_rank &k = ranks[0];
  return k.left_border && k.right_border &&
         k.n_particles + k.nb_particles_create + k.nb_disabled == nb_particles;
}

/* mcgpu_sync_cycle for one eligible rank (see k_one_cells above).  Cycles run
   in groups bounded by the work buffer (particles) and the per-cycle
   tallies (cells x cycles); each group is one batched launch. */
int sync_one_rank(mcgpu_context *ctx, mcgpu_cycle_rank &k, const double *sigs, const double *ar,
                  double *weights, int64_t S, int64_t max_cycles, const mcgpu_options &o,
                  window_log &wl, mcgpu_cycle_stats &st, int *status) {
  hipStream_t s = ctx->stream;
  const int m = k.max_index - k.min_index;
  const int block = o.block_size > 0 ? o.block_size : 256;
  const int ppl = o.particles_per_lane > 0 ? o.particles_per_lane : 16;
  const int64_t T = (int64_t)block * ppl;
  const uint64_t max_events = o.max_events_per_history ? o.max_events_per_history : (1ULL << 27);
  mc_batched_shape sh;
  int rc = mc_batched_plan_shape(ctx, m, block, o.grid_blocks, o.events_per_refill, &sh);
  if (rc) return rc;

  const int64_t n0 = k.n_particles;
  const int64_t all = n0 + k.nb_particles_create;
  /* work buffer and per-cycle tallies of one group */
  const int64_t pcap = std::max(S, std::min<int64_t>(all, (int64_t)1 << 26));
  const int64_t gcap = std::max<int64_t>(1, ((int64_t)1 << 23) / m);
  const int64_t max_tiles = pcap / T + gcap + 1;

  cycle_mem mem;
  auto dalloc = [&](void **p, size_t bytes) -> int {
    MC_HIP(ctx, hipMalloc(p, bytes));
    mem.dev.push_back(*p);
    return MCGPU_OK;
  };
  double *d_in, *d_w;
  mc_cell *d_cells;
  unsigned long long *d_fx;
  uint32_t *d_flags;
  mc_counters *d_ctr;
  mc_slab_k *d_slabs;
  mc_tile_k *d_tiles;
  void *h_ctr;
  if ((rc = dalloc((void **)&d_in, 16 * (size_t)m)) || (rc = dalloc((void **)&d_w, 8 * (size_t)m)) ||
      (rc = dalloc((void **)&d_cells, sizeof(mc_cell) * (size_t)(gcap * m))) ||
      (rc = dalloc((void **)&d_fx, 16 * (size_t)(gcap * m))) ||
      (rc = dalloc((void **)&d_flags, 4 * (size_t)(gcap * m))) ||
      (rc = dalloc((void **)&d_ctr, sizeof(mc_counters))) ||
      (rc = dalloc((void **)&d_slabs, sizeof(mc_slab_k) * (size_t)gcap)) ||
      (rc = dalloc((void **)&d_tiles, sizeof(mc_tile_k) * (size_t)max_tiles)))
    return rc;
  MC_HIP(ctx, hipHostMalloc(&h_ctr, sizeof(mc_counters), hipHostMallocDefault));
  mem.host.push_back(h_ctr);
  hipEvent_t ev_start, ev_comp;
  MC_HIP(ctx, hipEventCreate(&ev_start));
  mem.events.push_back(ev_start);
  MC_HIP(ctx, hipEventCreate(&ev_comp));
  mem.events.push_back(ev_comp);

  MC_HIP(ctx, hipMemcpyAsync(d_in, sigs + k.cell_offset, 8 * (size_t)m, hipMemcpyHostToDevice, s));
  MC_HIP(ctx, hipMemcpyAsync(d_in + m, ar + k.cell_offset, 8 * (size_t)m, hipMemcpyHostToDevice,
                             s));
  MC_HIP(ctx, hipMemcpyAsync(d_w, weights + k.cell_offset, 8 * (size_t)m, hipMemcpyHostToDevice,
                             s));
  hipLaunchKernelGGL(k_one_cells, dim3(blocks_for(gcap * m, 256)), dim3(256), 0, s, gcap, m, d_in,
                     d_in + m, d_cells);
  MC_HIP(ctx, hipGetLastError());

  /* the initial bank, host AoS -> device SoA */
  pool_guard init, work;
  double wmax = k.nb_particles_create > 0 ? std::fabs(k.wmc) : 0.0;
  if (n0 > 0) {
    if ((rc = alloc_pool(ctx, n0, &init.p, nullptr))) return rc;
    std::vector<uint64_t> hs(n0);
    std::vector<double> hx(n0), hm(n0), hw(n0);
    std::vector<int32_t> hi(n0);
    for (int64_t i = 0; i < n0; ++i) {
      const mcgpu_particle &q = k.particles[i];
      hs[i] = q.seed;
      hx[i] = q.x;
      hm[i] = q.mu;
      hw[i] = q.wmc;
      hi[i] = q.index;
      wmax = std::max(wmax, std::fabs(q.wmc));
    }
    MC_HIP(ctx, hipMemcpyAsync(init.p.seed, hs.data(), 8 * n0, hipMemcpyHostToDevice, s));
    MC_HIP(ctx, hipMemcpyAsync(init.p.x, hx.data(), 8 * n0, hipMemcpyHostToDevice, s));
    MC_HIP(ctx, hipMemcpyAsync(init.p.mu, hm.data(), 8 * n0, hipMemcpyHostToDevice, s));
    MC_HIP(ctx, hipMemcpyAsync(init.p.wmc, hw.data(), 8 * n0, hipMemcpyHostToDevice, s));
    MC_HIP(ctx, hipMemcpyAsync(init.p.index, hi.data(), 4 * n0, hipMemcpyHostToDevice, s));
    MC_HIP(ctx, hipStreamSynchronize(s)); /* the host vectors go out of scope */
  }
  if ((rc = alloc_pool(ctx, pcap, &work.p, nullptr))) return rc;
  const int32_t src_index = k.nb_particles_create > 0 ? (int32_t)(k.x_ini / k.dx) : 0;

  wl.open(wall_now()); /* the loop starts (the reference's spin, src/worker_sync.cpp:26) */
  /* the bank: segments in bank order (bottom first) */
  std::vector<one_seg> stack;
  if (n0 > 0) stack.push_back({0, 0, n0});
  int64_t size = n0, create = k.nb_particles_create, chain = 0;
  int64_t win_cycles = 0;
  bool first_group = true;
  std::vector<mc_slab_k> h_slabs;
  std::vector<mc_tile_k> h_tiles;
  while (size > 0 || create > 0) {
    if (max_cycles > 0 && st.cycles >= max_cycles) {
      *status = mc_fail(ctx, MCGPU_ERR_ARG, "sync_cycle: no termination within max_cycles");
      break;
    }
    const double t_group = wall_now();
    if (first_group) {
      wl.w.time_idle += t_group - wl.w.starttime;
      first_group = false;
    }
    /* ---- the group's cycles: creation and the last n of the bank each */
    std::vector<one_cycle> cyc;
    int64_t used = 0;
    while ((size > 0 || create > 0) && (int64_t)cyc.size() < gcap && used + S <= pcap &&
           (max_cycles <= 0 || st.cycles + (int64_t)cyc.size() < max_cycles)) {
      if (size < S && create > 0) { /* src/layer.cpp:89-121,247-248 */
        const int64_t n_new = std::min(create, std::max(MAX_PARTICLES_VECTOR - size, S));
        stack.push_back({1, chain, n_new});
        chain += n_new;
        create -= n_new;
        size += n_new;
      }
      one_cycle c;
      c.first = used;
      c.n = std::min(size, S);
      int64_t need = c.n;
      while (need > 0) {
        one_seg &t = stack.back();
        const int64_t take = std::min(need, t.n);
        c.pieces.push_back({t.kind, t.begin + t.n - take, take});
        t.n -= take;
        need -= take;
        if (t.n == 0) stack.pop_back();
      }
      std::reverse(c.pieces.begin(), c.pieces.end());
      size -= c.n;
      used += c.n;
      cyc.push_back(std::move(c));
    }
    const int64_t G = (int64_t)cyc.size();
    if (G == 0) break;
    /* ---- lay out the particles, tables, one launch, weights in cycle order */
    MC_HIP(ctx, hipEventRecord(ev_start, s));
    h_slabs.assign(G, mc_slab_k{});
    h_tiles.clear();
    for (int64_t g = 0; g < G; ++g) {
      int64_t pos = cyc[g].first;
      for (const one_seg &q : cyc[g].pieces) {
        if (q.kind == 0) {
          MC_HIP(ctx, hipMemcpyAsync(work.p.seed + pos, init.p.seed + q.begin, 8 * q.n,
                                     hipMemcpyDeviceToDevice, s));
          MC_HIP(ctx, hipMemcpyAsync(work.p.x + pos, init.p.x + q.begin, 8 * q.n,
                                     hipMemcpyDeviceToDevice, s));
          MC_HIP(ctx, hipMemcpyAsync(work.p.mu + pos, init.p.mu + q.begin, 8 * q.n,
                                     hipMemcpyDeviceToDevice, s));
          MC_HIP(ctx, hipMemcpyAsync(work.p.wmc + pos, init.p.wmc + q.begin, 8 * q.n,
                                     hipMemcpyDeviceToDevice, s));
          MC_HIP(ctx, hipMemcpyAsync(work.p.index + pos, init.p.index + q.begin, 4 * q.n,
                                     hipMemcpyDeviceToDevice, s));
        } else {
          const mcgpu_bank b{work.p.seed + pos, work.p.x + pos, work.p.mu + pos, work.p.wmc + pos,
                             work.p.index + pos};
          if ((rc = mc_launch_create(ctx, k.seed, q.begin, q.n, k.x_ini, k.wmc, src_index, b, s)))
            return rc;
        }
        pos += q.n;
      }
      mc_slab_k &sl = h_slabs[g];
      sl.index_start = k.min_index;
      sl.m = m;
      sl.dx = k.dx;
      sl.min_weight = k.particle_min_weight;
      sl.cell_off = g * m;
      const int64_t end = cyc[g].first + cyc[g].n;
      for (int64_t p0 = cyc[g].first; p0 < end; p0 += T) {
        mc_tile_k t;
        t.slab = (int32_t)g;
        t.pad_ = 0;
        t.begin = p0;
        t.end = std::min(end, p0 + T);
        h_tiles.push_back(t);
      }
    }
    MC_HIP(ctx, hipMemcpyAsync(d_slabs, h_slabs.data(), sizeof(mc_slab_k) * G,
                               hipMemcpyHostToDevice, s));
    MC_HIP(ctx, hipMemcpyAsync(d_tiles, h_tiles.data(), sizeof(mc_tile_k) * h_tiles.size(),
                               hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_cyc_clear, dim3(blocks_for(G * m, 256)), dim3(256), 0, s, G * m, d_fx,
                       d_flags, d_ctr);
    MC_HIP(ctx, hipGetLastError());
    const mcgpu_bank bank{work.p.seed, work.p.x, work.p.mu, work.p.wmc, work.p.index};
    MC_HIP(ctx, hipEventRecord(ctx->ev0, s));
    if ((rc = mc_launch_batched(ctx, sh, d_slabs, d_tiles, (int32_t)h_tiles.size(), bank, d_cells,
                                d_fx, G * m, d_flags, nullptr, d_ctr, max_events,
                                wmax < 0x1p-24, s)))
      return rc;
    MC_HIP(ctx, hipEventRecord(ctx->ev1, s));
    hipLaunchKernelGGL(k_one_finalize, dim3(blocks_for(m, 256)), dim3(256), 0, s, m, G,
                       (const unsigned long long *)d_fx, (const uint32_t *)d_flags, d_w);
    MC_HIP(ctx, hipGetLastError());
    MC_HIP(ctx, hipEventRecord(ev_comp, s));
    MC_HIP(ctx, hipMemcpyAsync(h_ctr, d_ctr, sizeof(mc_counters), hipMemcpyDeviceToHost, s));
    MC_HIP(ctx, hipStreamSynchronize(s)); /* also: h_slabs / h_tiles reusable */
    const mc_counters *hc = reinterpret_cast<const mc_counters *>(h_ctr);
    float ms = 0.f, comp_ms = 0.f;
    (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    (void)hipEventElapsedTime(&comp_ms, ev_start, ev_comp);
    st.kernel_ms += ms;
    st.events += hc->events;
    st.collisions += hc->collisions;
    st.cycles += G;
    k.n_left += (int64_t)hc->n_left;
    k.n_right += (int64_t)hc->n_right;
    k.n_killed += (int64_t)hc->n_killed;
    k.nb_disabled += used; /* every transported particle leaves through a border or is killed */
    {
      const double t_end = wall_now();
      wl.w.time_comp += comp_ms * 1e-3;
      wl.w.time_recv += std::max(0.0, (t_end - t_group) - comp_ms * 1e-3);
    }
    if (hc->status & 1u) {
      *status = mc_fail(ctx, MCGPU_ERR_EVENT_CAP, "a history exceeded max_events_per_history");
      break;
    }
    if (hc->status & 2u) {
      *status = mc_fail(ctx, MCGPU_ERR_RANGE, "a tally term was outside +-2^39");
      break;
    }
    /* statistics windows: the group's cycles as a unit; the cycle that ends
       the loop is not counted */
    const bool ended = size == 0 && create == 0;
    win_cycles += ended ? G - 1 : G;
    if (!ended) {
      const double t = wall_now();
      if (t > wl.w.starttime + wl.period) {
        wl.w.nb_cycles = win_cycles;
        wl.close(t);
        wl.open(t);
        win_cycles = 0;
      }
    }
  }
  wl.w.nb_cycles = win_cycles;
  wl.close(wall_now());
  MC_HIP(ctx, hipMemcpyAsync(weights + k.cell_offset, d_w, 8 * (size_t)m, hipMemcpyDeviceToHost,
                             s));
  MC_HIP(ctx, hipStreamSynchronize(s));
  k.nb_particles_create = create;
  k.seed = mc_seed_chain_at(k.seed, (uint64_t)chain);
  k.n_remaining = size;
  return MCGPU_OK;
}

} // namespace

extern "C" int mcgpu_sync_cycle(mcgpu_context *ctx, int32_t world_size, mcgpu_cycle_rank *ranks,
                                const double *sigs, const double *absorption_rates,
                                double *weights_absorbed, int64_t total_cells,
                                int32_t nb_particles_per_cycle, int64_t nb_particles,
                                int64_t max_cycles, const mcgpu_options *opt,
                                mcgpu_cycle_stats *stats) {
  if (!ctx) return MCGPU_ERR_ARG;
  ctx->err.clear();
  const auto t_start = std::chrono::steady_clock::now();
  window_log wl;
  int32_t cycle_by_cycle = 0;
  if (stats) {
    wl.period = stats->statistics_cycle_time;
    wl.fn = stats->on_window;
    wl.user = stats->window_user;
    cycle_by_cycle = stats->cycle_by_cycle;
  }
  const int R = world_size;
  const int64_t S = nb_particles_per_cycle;
  if (R < 1 || !ranks || !sigs || !absorption_rates || !weights_absorbed || total_cells <= 0 ||
      nb_particles < 0)
    return mc_fail(ctx, MCGPU_ERR_ARG, "sync_cycle: null pointer or empty input");
  if (S < 1 || S > MAX_PARTICLES_VECTOR)
    return mc_fail(ctx, MCGPU_ERR_ARG, "sync_cycle: nb_particles_per_cycle must be in [1, 1e6]");
  int max_m = 0;
  for (int r = 0; r < R; ++r) {
    const mcgpu_cycle_rank &k = ranks[r];
    const int64_t m = (int64_t)k.max_index - k.min_index;
    if (m <= 0 || k.cell_offset < 0 || k.cell_offset + m > total_cells || k.n_particles < 0 ||
        (k.n_particles > 0 && !k.particles) || k.nb_particles_create < 0 || !(k.dx > 0))
      return mc_fail(ctx, MCGPU_ERR_ARG, "sync_cycle: invalid rank descriptor");
    if (k.nb_particles_create > 0) {
      const double cell = k.x_ini / k.dx;
      if (!(cell >= 0) || !(cell < 2147483647.0))
        return mc_fail(ctx, MCGPU_ERR_ARG, "sync_cycle: x_ini / dx is not a cell index");
    }
    max_m = std::max(max_m, (int)m);
  }
  MC_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  mcgpu_options o;
  mcgpu_default_options(&o);
  if (opt) o = *opt;
  const int block = o.block_size > 0 ? o.block_size : 256;
  const int ppl = o.particles_per_lane > 0 ? o.particles_per_lane : 1;
  const int64_t T = (int64_t)block * ppl; /* particles per transport tile */
  const uint64_t max_events = o.max_events_per_history ? o.max_events_per_history : (1ULL << 27);
  if (one_rank_eligible(R, ranks, nb_particles) && !cycle_by_cycle) {
    mcgpu_cycle_rank &k = ranks[0];
    k.n_left = k.n_right = k.n_killed = 0;
    k.n_remaining = k.n_particles;
    mcgpu_cycle_stats st{};
    int status = MCGPU_OK;
    int rc = sync_one_rank(ctx, k, sigs, absorption_rates, weights_absorbed, S, max_cycles, o, wl,
                           st, &status);
    if (rc) return rc;
    st.wall_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start)
            .count();
    if (stats) {
      st.statistics_cycle_time = wl.period;
      st.on_window = wl.fn;
      st.window_user = wl.user;
      st.cycle_by_cycle = cycle_by_cycle;
      *stats = st;
    }
    return status;
  }
  mc_batched_shape sh;
  int rc = mc_batched_plan_shape(ctx, max_m, block, o.grid_blocks, o.events_per_refill, &sh);
  if (rc) return rc;

  /* ---- fixed device memory */
  cycle_mem mem;
  auto dalloc = [&](void **p, size_t bytes) -> int {
    MC_HIP(ctx, hipMalloc(p, bytes));
    mem.dev.push_back(*p);
    return MCGPU_OK;
  };
  auto halloc = [&](void **p, size_t bytes) -> int {
    MC_HIP(ctx, hipHostMalloc(p, bytes, hipHostMallocDefault));
    mem.host.push_back(*p);
    return MCGPU_OK;
  };
  const int64_t tc = total_cells;
  const int64_t max_tiles = (int64_t)R * ((S + T - 1) / T);
  const int64_t max_chunks = (int64_t)R * ((S + CHUNK - 1) / CHUNK);
  mc_cell *d_cells;
  unsigned long long *d_fx;
  uint32_t *d_flags;
  mc_counters *d_ctr;
  double *d_w, *d_in;
  mc_slab_k *d_slabs;
  mc_tile_k *d_tiles;
  chunk_k *d_chunks;
  rank_in *d_rin;
  chunk_cnt *d_cnt;
  chunk_off *d_off;
  rank_sum *d_sum;
  rank_tot *d_tot;
  const size_t table_bytes = sizeof(mc_tile_k) * max_tiles + sizeof(chunk_k) * max_chunks +
                             sizeof(rank_in) * R;
  void *h_table, *h_back;
  if ((rc = dalloc((void **)&d_cells, sizeof(mc_cell) * tc)) ||
      (rc = dalloc((void **)&d_fx, 16 * (size_t)tc)) ||
      (rc = dalloc((void **)&d_flags, 4 * (size_t)tc)) ||
      (rc = dalloc((void **)&d_ctr, sizeof(mc_counters))) ||
      (rc = dalloc((void **)&d_w, 8 * (size_t)tc)) ||
      (rc = dalloc((void **)&d_in, 16 * (size_t)tc)) ||
      (rc = dalloc((void **)&d_slabs, sizeof(mc_slab_k) * R)) ||
      (rc = dalloc((void **)&d_tiles, table_bytes)) ||
      (rc = dalloc((void **)&d_cnt, sizeof(chunk_cnt) * max_chunks)) ||
      (rc = dalloc((void **)&d_off, sizeof(chunk_off) * max_chunks)) ||
      (rc = dalloc((void **)&d_sum, sizeof(rank_sum) * R)) ||
      (rc = dalloc((void **)&d_tot, sizeof(rank_tot) * R)) ||
      (rc = halloc(&h_table, table_bytes)) ||
      (rc = halloc(&h_back, sizeof(rank_sum) * R + sizeof(mc_counters))))
    return rc;
  d_chunks = reinterpret_cast<chunk_k *>(d_tiles + max_tiles);
  d_rin = reinterpret_cast<rank_in *>(d_chunks + max_chunks);
  mc_tile_k *h_tiles = reinterpret_cast<mc_tile_k *>(h_table);
  chunk_k *h_chunks = reinterpret_cast<chunk_k *>(h_tiles + max_tiles);
  rank_in *h_rin = reinterpret_cast<rank_in *>(h_chunks + max_chunks);
  rank_sum *h_sum = reinterpret_cast<rank_sum *>(h_back);
  mc_counters *h_ctr = reinterpret_cast<mc_counters *>(h_sum + R);

  MC_HIP(ctx, hipMemcpyAsync(d_in, sigs, 8 * (size_t)tc, hipMemcpyHostToDevice, s));
  MC_HIP(ctx, hipMemcpyAsync(d_in + tc, absorption_rates, 8 * (size_t)tc, hipMemcpyHostToDevice, s));
  MC_HIP(ctx, hipMemcpyAsync(d_w, weights_absorbed, 8 * (size_t)tc, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_cyc_cells, dim3(blocks_for(tc, 256)), dim3(256), 0, s, tc, d_in, d_in + tc,
                     d_cells);
  MC_HIP(ctx, hipGetLastError());
  {
    std::vector<mc_slab_k> slabs(R);
    for (int r = 0; r < R; ++r) {
      slabs[r].index_start = ranks[r].min_index;
      slabs[r].m = ranks[r].max_index - ranks[r].min_index;
      slabs[r].dx = ranks[r].dx;
      slabs[r].min_weight = ranks[r].particle_min_weight;
      slabs[r].cell_off = ranks[r].cell_offset;
    }
    MC_HIP(ctx, hipMemcpyAsync(d_slabs, slabs.data(), sizeof(mc_slab_k) * R,
                               hipMemcpyHostToDevice, s));
    MC_HIP(ctx, hipStreamSynchronize(s)); /* slabs is a stack vector */
  }

  /* ---- escape buffers and the particle pool */
  pool_guard esc, pool;
  if ((rc = alloc_pool(ctx, 2 * (int64_t)R * S, &esc.p, nullptr))) return rc;
  std::vector<int64_t> base(R), cap(R), size(R), create(R), stranded(R, 0);
  std::vector<uint64_t> seed(R);
  for (int r = 0; r < R; ++r) {
    size[r] = ranks[r].n_particles;
    create[r] = ranks[r].nb_particles_create;
    seed[r] = ranks[r].seed;
    ranks[r].n_left = ranks[r].n_right = ranks[r].n_killed = 0;
  }
  /* room for the bank, one creation and two received buffers */
  auto need_of = [&](int r) {
    return size[r] + (create[r] > 0 ? std::max(MAX_PARTICLES_VECTOR, S) : 0) + 2 * S;
  };
  auto build_pool = [&](const std::vector<int64_t> &new_cap) -> int {
    pool_k np{};
    int8_t *nres = nullptr;
    int64_t total = 0;
    std::vector<int64_t> nbase(R);
    for (int r = 0; r < R; ++r) {
      nbase[r] = total;
      total += new_cap[r];
    }
    int e = alloc_pool(ctx, total, &np, &nres);
    if (e) {
      free_pool(&np, &nres);
      return e;
    }
    if (pool.p.seed) {
      for (int r = 0; r < R && !e; ++r) {
        const size_t n = (size_t)size[r];
        if (!n) continue;
        const int64_t a = base[r], b = nbase[r];
        hipError_t h = hipMemcpyAsync(np.seed + b, pool.p.seed + a, 8 * n, hipMemcpyDeviceToDevice, s);
        if (h == hipSuccess) h = hipMemcpyAsync(np.x + b, pool.p.x + a, 8 * n, hipMemcpyDeviceToDevice, s);
        if (h == hipSuccess) h = hipMemcpyAsync(np.mu + b, pool.p.mu + a, 8 * n, hipMemcpyDeviceToDevice, s);
        if (h == hipSuccess) h = hipMemcpyAsync(np.wmc + b, pool.p.wmc + a, 8 * n, hipMemcpyDeviceToDevice, s);
        if (h == hipSuccess) h = hipMemcpyAsync(np.index + b, pool.p.index + a, 4 * n, hipMemcpyDeviceToDevice, s);
        if (h != hipSuccess) e = mc_fail(ctx, MCGPU_ERR_HIP, "sync_cycle: pool copy", h);
      }
      const hipError_t h = hipStreamSynchronize(s);
      if (!e && h != hipSuccess) e = mc_fail(ctx, MCGPU_ERR_HIP, "sync_cycle: pool copy", h);
    }
    if (e) {
      free_pool(&np, &nres);
      return e;
    }
    free_pool(&pool.p, &pool.result8);
    pool.p = np;
    pool.result8 = nres;
    base = nbase;
    cap = new_cap;
    return MCGPU_OK;
  };
  /* weights only decrease; when every initial one is below 2^-24 the
     tally's high words only take carries (mc_launch_batched, small_weights) */
  double wmax = 0.0;
  for (int r = 0; r < R; ++r)
    if (ranks[r].nb_particles_create > 0) wmax = std::max(wmax, std::fabs(ranks[r].wmc));
  {
    std::vector<int64_t> c0(R);
    for (int r = 0; r < R; ++r) c0[r] = need_of(r);
    if ((rc = build_pool(c0))) return rc;
    /* initial banks: host AoS -> device SoA */
    for (int r = 0; r < R; ++r) {
      const int64_t n = size[r];
      if (!n) continue;
      std::vector<uint64_t> hs(n);
      std::vector<double> hx(n), hm(n), hw(n);
      std::vector<int32_t> hi(n);
      for (int64_t i = 0; i < n; ++i) {
        const mcgpu_particle &q = ranks[r].particles[i];
        hs[i] = q.seed;
        hx[i] = q.x;
        hm[i] = q.mu;
        hw[i] = q.wmc;
        hi[i] = q.index;
        wmax = std::max(wmax, std::fabs(q.wmc));
      }
      const int64_t b = base[r];
      MC_HIP(ctx, hipMemcpyAsync(pool.p.seed + b, hs.data(), 8 * n, hipMemcpyHostToDevice, s));
      MC_HIP(ctx, hipMemcpyAsync(pool.p.x + b, hx.data(), 8 * n, hipMemcpyHostToDevice, s));
      MC_HIP(ctx, hipMemcpyAsync(pool.p.mu + b, hm.data(), 8 * n, hipMemcpyHostToDevice, s));
      MC_HIP(ctx, hipMemcpyAsync(pool.p.wmc + b, hw.data(), 8 * n, hipMemcpyHostToDevice, s));
      MC_HIP(ctx, hipMemcpyAsync(pool.p.index + b, hi.data(), 4 * n, hipMemcpyHostToDevice, s));
      MC_HIP(ctx, hipStreamSynchronize(s)); /* the host vectors go out of scope */
    }
  }

  /* cycle phase boundaries: start, end of the tally (computation), end of
     the exchange (send) */
  hipEvent_t ev_start, ev_comp, ev_send;
  MC_HIP(ctx, hipEventCreate(&ev_start));
  mem.events.push_back(ev_start);
  MC_HIP(ctx, hipEventCreate(&ev_comp));
  mem.events.push_back(ev_comp);
  MC_HIP(ctx, hipEventCreate(&ev_send));
  mem.events.push_back(ev_send);

  int64_t disabled_total = 0;
  for (int r = 0; r < R; ++r) disabled_total += ranks[r].nb_disabled;
  mcgpu_cycle_stats st{};
  int status = MCGPU_OK;
  wl.open(wall_now());
  bool first_cycle = true;
  int64_t win_cycles = 0;
  while (disabled_total != nb_particles) {
    if (max_cycles > 0 && st.cycles >= max_cycles) {
      status = mc_fail(ctx, MCGPU_ERR_ARG, "sync_cycle: no termination within max_cycles");
      break;
    }
    const double t_cycle = wall_now();
    if (first_cycle) {
      wl.w.time_idle += t_cycle - wl.w.starttime;
      first_cycle = false;
    }
    /* room for this cycle */
    bool grow = false;
    std::vector<int64_t> want(cap);
    for (int r = 0; r < R; ++r)
      if (need_of(r) > cap[r]) {
        want[r] = 2 * need_of(r);
        grow = true;
      }
    if (grow && (rc = build_pool(want))) return rc;
    const double t_grown = wall_now();
    MC_HIP(ctx, hipEventRecord(ev_start, s));

    /* 1. lazy creation (src/layer.cpp:247-248, :89-121) */
    int64_t n_tiles = 0, n_chunks = 0;
    for (int r = 0; r < R; ++r) {
      mcgpu_cycle_rank &k = ranks[r];
      if (size[r] < S && create[r] > 0) {
        int64_t n = std::max(MAX_PARTICLES_VECTOR - size[r], S);
        n = std::min(create[r], n);
        const int64_t at = base[r] + size[r];
        mcgpu_bank b{pool.p.seed + at, pool.p.x + at, pool.p.mu + at, pool.p.wmc + at,
                     pool.p.index + at};
        if ((rc = mc_launch_create(ctx, seed[r], 0, n, k.x_ini, k.wmc, (int32_t)(k.x_ini / k.dx),
                                   b, s)))
          return rc;
        seed[r] = mc_seed_chain_at(seed[r], (uint64_t)n);
        create[r] -= n;
        size[r] += n;
      }
      /* 2. tiles of the last n particles, 3. partition chunks */
      const int64_t n = std::min(size[r], S);
      const int64_t end = base[r] + size[r];
      for (int64_t p0 = end - n; p0 < end; p0 += T) {
        mc_tile_k t;
        t.slab = r;
        t.pad_ = 0;
        t.begin = p0;
        t.end = std::min(end, p0 + T);
        h_tiles[n_tiles++] = t;
      }
      rank_in &ri = h_rin[r];
      ri.pool_base = base[r];
      ri.size = size[r];
      ri.n = n;
      ri.chunk_first = (int32_t)n_chunks;
      ri.left_border = k.left_border != 0;
      ri.right_border = k.right_border != 0;
      for (int64_t i0 = 0; i0 < n; i0 += CHUNK) {
        chunk_k c;
        c.rank = r;
        c.pad_ = 0;
        c.i0 = i0;
        c.i1 = std::min(n, i0 + CHUNK);
        c.seg_end = end;
        h_chunks[n_chunks++] = c;
      }
      ri.chunk_count = (int32_t)(n_chunks - ri.chunk_first);
    }
    MC_HIP(ctx, hipMemcpyAsync(d_tiles, h_table, table_bytes, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_cyc_clear, dim3(blocks_for(tc, 256)), dim3(256), 0, s, tc, d_fx, d_flags,
                       d_ctr);
    MC_HIP(ctx, hipGetLastError());
    const mcgpu_bank bank{pool.p.seed, pool.p.x, pool.p.mu, pool.p.wmc, pool.p.index};
    MC_HIP(ctx, hipEventRecord(ctx->ev0, s));
    if (n_tiles > 0 &&
        (rc = mc_launch_batched(ctx, sh, d_slabs, d_tiles, (int32_t)n_tiles, bank, d_cells, d_fx,
                                tc, d_flags, pool.result8, d_ctr, max_events,
                                wmax < 0x1p-24, s)))
      return rc;
    MC_HIP(ctx, hipEventRecord(ctx->ev1, s));
    hipLaunchKernelGGL(k_cyc_finalize, dim3(blocks_for(tc, 256)), dim3(256), 0, s, tc,
                       (const unsigned long long *)d_fx, (const uint32_t *)d_flags, d_w);
    MC_HIP(ctx, hipGetLastError());
    MC_HIP(ctx, hipEventRecord(ev_comp, s));
    if (n_chunks > 0) {
      hipLaunchKernelGGL(k_cyc_count, dim3((unsigned)n_chunks), dim3(CHUNK), 0, s,
                         (const chunk_k *)d_chunks, (const int8_t *)pool.result8, d_cnt);
      MC_HIP(ctx, hipGetLastError());
    }
    hipLaunchKernelGGL(k_cyc_scan, dim3(R), dim3(CHUNK), 0, s, (const rank_in *)d_rin,
                       (const chunk_cnt *)d_cnt, d_off, d_tot);
    MC_HIP(ctx, hipGetLastError());
    hipLaunchKernelGGL(k_cyc_plan, dim3(blocks_for(R, 256)), dim3(256), 0, s, R, S,
                       (const rank_in *)d_rin, (const rank_tot *)d_tot, d_sum);
    MC_HIP(ctx, hipGetLastError());
    if (n_chunks > 0) {
      hipLaunchKernelGGL(k_cyc_scatter, dim3((unsigned)n_chunks), dim3(CHUNK), 0, s,
                         (const chunk_k *)d_chunks, (const int8_t *)pool.result8,
                         (const chunk_off *)d_off, S, pool.p, esc.p);
      MC_HIP(ctx, hipGetLastError());
    }
    hipLaunchKernelGGL(k_cyc_exchange, dim3(blocks_for(2 * S, 256), R), dim3(256), 0, s,
                       (const rank_in *)d_rin, (const rank_sum *)d_sum, pool.p, esc.p);
    MC_HIP(ctx, hipGetLastError());
    MC_HIP(ctx, hipEventRecord(ev_send, s));
    MC_HIP(ctx, hipMemcpyAsync(h_sum, d_sum, sizeof(rank_sum) * R, hipMemcpyDeviceToHost, s));
    MC_HIP(ctx, hipMemcpyAsync(h_ctr, d_ctr, sizeof(mc_counters), hipMemcpyDeviceToHost, s));
    MC_HIP(ctx, hipStreamSynchronize(s));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    st.kernel_ms += ms;
    {
      float comp_ms = 0.f, send_ms = 0.f;
      (void)hipEventElapsedTime(&comp_ms, ev_start, ev_comp);
      (void)hipEventElapsedTime(&send_ms, ev_comp, ev_send);
      const double comp = (t_grown - t_cycle) + comp_ms * 1e-3, send = send_ms * 1e-3;
      const double t_end = wall_now();
      wl.w.time_comp += comp;
      wl.w.time_send += send;
      wl.w.time_recv += std::max(0.0, (t_end - t_cycle) - comp - send);
    }
    st.events += h_ctr->events;
    st.collisions += h_ctr->collisions;
    ++st.cycles;
    for (int r = 0; r < R; ++r) {
      const rank_sum &q = h_sum[r];
      size[r] = q.new_size;
      ranks[r].nb_disabled += q.disabled_add;
      ranks[r].n_left += q.n_left;
      ranks[r].n_right += q.n_right;
      ranks[r].n_killed += q.n_killed;
      stranded[r] += q.stranded;
      disabled_total += q.disabled_add;
    }
    if (h_ctr->status & 1u) {
      status = mc_fail(ctx, MCGPU_ERR_EVENT_CAP, "a history exceeded max_events_per_history");
      break;
    }
    if (h_ctr->status & 2u) {
      status = mc_fail(ctx, MCGPU_ERR_RANGE, "a tally term was outside +-2^39");
      break;
    }
    if (disabled_total == nb_particles) break; /* the ending cycle is not counted */
    const double t = wall_now();
    if (t > wl.w.starttime + wl.period) {
      wl.w.nb_cycles = win_cycles;
      wl.close(t);
      wl.open(t);
      win_cycles = 0;
    }
    ++win_cycles;
  }
  wl.w.nb_cycles = win_cycles;
  wl.close(wall_now());

  MC_HIP(ctx, hipMemcpyAsync(weights_absorbed, d_w, 8 * (size_t)tc, hipMemcpyDeviceToHost, s));
  MC_HIP(ctx, hipStreamSynchronize(s));
  for (int r = 0; r < R; ++r) {
    ranks[r].nb_particles_create = create[r];
    ranks[r].seed = seed[r];
    ranks[r].n_remaining = size[r] + stranded[r];
  }
  st.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start)
                   .count();
  if (stats) {
    st.statistics_cycle_time = wl.period;
    st.on_window = wl.fn;
    st.window_user = wl.user;
    st.cycle_by_cycle = cycle_by_cycle;
    *stats = st;
  }
  return status;
}
