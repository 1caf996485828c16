/*
 * internal.h -- shared definitions of the library's translation units:
 * the context, device counter block and kernel launchers.
 */
#ifndef MCGPU_INTERNAL_H
#define MCGPU_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/mcgpu.h"
#include "mc_device.h"

/* Device-side counters of one transport call. */
struct mc_counters {
  unsigned long long next;       /* work-queue head (history kernel) */
  unsigned long long events;
  unsigned long long collisions;
  unsigned long long n_left;
  unsigned long long n_right;
  unsigned long long n_killed;
  unsigned int status;           /* bit 0: event cap hit, bit 1: tally range */
  unsigned int pad_;
  unsigned long long aux[8];     /* kernel-specific (event kernel pass data) */
};

struct mcgpu_context {
  int device = 0;
  int num_cus = 0;
  size_t lds_per_block = 0;
  bool quiet = false; /* no error text (a stack context of the no-allocation entry point) */
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string err;

  /* grow-only scratch (device) */
  void *cells = nullptr;      size_t cells_bytes = 0;     /* mc_cell[m] */
  void *fx = nullptr;         size_t fx_bytes = 0;        /* u64[2m] tally */
  void *flags = nullptr;      size_t flags_bytes = 0;     /* u32[m] */
  void *counters = nullptr;                               /* mc_counters */
  void *host_counters = nullptr;                          /* pinned mc_counters */
  void *aos = nullptr;        size_t aos_bytes = 0;       /* staging mcgpu_particle[n] */
  void *res32 = nullptr;      size_t res32_bytes = 0;     /* int32[n] */
  void *sig_in = nullptr;     size_t sig_in_bytes = 0;    /* double[2m] sigs, ar */
  void *w_in = nullptr;       size_t w_in_bytes = 0;      /* double[m] weights */
  void *work = nullptr;       size_t work_bytes = 0;      /* event kernel banks */
  void *tables = nullptr;     size_t tables_bytes = 0;    /* batched calls: slab + tile tables */
};

/* grow-only device buffer */
int mc_ensure(mcgpu_context *ctx, void **buf, size_t *cap, size_t need);
int mc_fail(mcgpu_context *ctx, int status, const char *what, hipError_t e = hipSuccess);
#define MC_HIP(ctx, call)                                          \
  do {                                                             \
    hipError_t e_ = (call);                                        \
    if (e_ != hipSuccess) return mc_fail((ctx), MCGPU_ERR_HIP, #call, e_); \
  } while (0)

/* Arguments common to the transport kernels. */
struct mc_transport_args {
  mc_layer_k L;
  int64_t n;
  mcgpu_particle *aos; /* exactly one of aos / soa is used */
  mcgpu_bank soa;
  const mc_cell *cells;
  unsigned long long *fx; /* [2m]: lo words then hi words */
  uint32_t *flags;
  int8_t *result8;
  int32_t *result32;
  uint64_t *events;
  uint64_t *collisions;
  mc_counters *ctr;
  uint64_t max_events;
  const double *g; /* optional per-cell target weights (history kernel) */
  double *score;   /* optional per-history sum of g[c] * dw, in event order */
};

/* particle i of the call's bank (AoS staging or SoA device bank) */
static __device__ __forceinline__ void load_state(const mc_transport_args &a, int64_t i, mc_state &p) {
  if (a.aos) {
    const mcgpu_particle *q = a.aos + i;
    p.seed = q->seed;
    p.x = q->x;
    p.mu = q->mu;
    p.wmc = q->wmc;
    p.index = q->index;
  } else {
    p.seed = a.soa.seed[i];
    p.x = a.soa.x[i];
    p.mu = a.soa.mu[i];
    p.wmc = a.soa.wmc[i];
    p.index = a.soa.index[i];
  }
}

static __device__ __forceinline__ void store_state(const mc_transport_args &a, int64_t i,
                                            const mc_state &p, int code) {
  if (a.aos) {
    mcgpu_particle *q = a.aos + i;
    q->seed = p.seed;
    q->x = p.x;
    q->mu = p.mu;
    q->wmc = p.wmc;
    q->index = p.index;
    if (a.result32) a.result32[i] = code;
  } else {
    a.soa.seed[i] = p.seed;
    a.soa.x[i] = p.x;
    a.soa.mu[i] = p.mu;
    a.soa.wmc[i] = p.wmc;
    a.soa.index[i] = p.index;
    if (a.result8) a.result8[i] = (int8_t)code;
  }
}

/* Threads per block of a persistent kernel with `lds` bytes of dynamic LDS
   per block, when the caller gave none: of 256, 512 and 1024, the one with
   the most resident waves per CU by the occupancy calculator (which counts
   registers and LDS); ties go to the smaller block.  With LDS-resident
   tables a small block can leave most of the CU idle: 56 KB per block fits
   two blocks in 160 KB, 8 waves at 256 threads but 16 at 512.
   *per_cu receives the resident blocks per CU of the choice (at least 1). */
template <typename K>
static inline int mc_auto_block(K kern, size_t lds, int *per_cu) {
  int best = 256, best_waves = -1, best_per_cu = 1;
  for (int b = 256; b <= 1024; b *= 2) {
    int pc = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&pc, kern, b, lds) != hipSuccess) {
      (void)hipGetLastError();
      pc = 0;
    }
    const int waves = pc * (b / 64);
    if (waves > best_waves) {
      best = b;
      best_waves = waves;
      best_per_cu = pc > 0 ? pc : 1;
    }
  }
  *per_cu = best_per_cu;
  return best;
}

int mc_launch_history(mcgpu_context *ctx, const mc_transport_args &a, const mcgpu_options &o,
                      hipStream_t s);

/* The steps of one transport call around the kernel (capi.hip), on buffers
   the caller provides:
   mc_launch_setup:    cells[c] = (sigs*ar, sigs*(1-ar)), tally words and
                       flags of m cells and the counter block zeroed (a
                       kernel, no memset);
   mc_launch_finalize: weights[c] += the tally of cell c, rounded once;
   mc_status_from_counters: stats from the counter block read back (with
                       the kernel time given) and the call's status. */
int mc_launch_setup(mcgpu_context *ctx, int m, const double *sigs, const double *ar,
                    mc_cell *cells, unsigned long long *fx, uint32_t *flags, mc_counters *ctr,
                    hipStream_t s);
int mc_launch_finalize(mcgpu_context *ctx, int m, const unsigned long long *fx,
                       const uint32_t *flags, double *weights, hipStream_t s);
int mc_status_from_counters(mcgpu_context *ctx, const mc_counters &hc, float kernel_ms,
                            mcgpu_stats *stats);

/* Device properties and the address of the no-allocation entry point's
   scratch array on the current device, set up once per device (noalloc.hip;
   also called by mcgpu_context_create for its device). */
int mc_noalloc_device_setup(int device);
int mc_launch_event(mcgpu_context *ctx, const mc_transport_args &a, const mcgpu_options &o,
                    hipStream_t s);

/* Particles first .. first+n-1 of a lazy source (Layer::create_particles,
   src/layer.cpp:89-121) into bank[0..n), on stream s (capi.hip). */
int mc_launch_create(mcgpu_context *ctx, uint64_t seed, int64_t first, int64_t n, double x_ini,
                     double wmc, int32_t index, const mcgpu_bank &bank, hipStream_t s);

/* The batched transport (batched.hip): a persistent launch over a work
   queue of tiles, each a particle range [begin, end) of the bank belonging
   to one slab (a layer: global cell range, dx, min weight, offset of its
   cells in the concatenated cell / tally arrays). */
struct mc_slab_k {
  int32_t index_start;
  int32_t m;
  double dx;
  double min_weight;
  int64_t cell_off;
};

struct mc_tile_k {
  int32_t slab;
  int32_t pad_;
  int64_t begin; /* absolute particle indices [begin, end) */
  int64_t end;
};

struct mc_batched_shape {
  int block, grid, refill;
  size_t lds;          /* dynamic LDS bytes per block */
  int lds_cells_max_m; /* slabs with m <= this stage their cells in LDS */
  int lds_tally_max_m; /* slabs with m <= this keep their tally in LDS */
};

/* Launch shape for slabs of up to max_m cells; grid <= 0 takes resident
   blocks per CU x CUs. */
int mc_batched_plan_shape(mcgpu_context *ctx, int max_m, int block, int grid, int refill,
                          mc_batched_shape *out);
/* One launch over n_tiles tiles; fx / flags / ctr must be zeroed first
   (ctr->next is the tile queue head).  small_weights: every particle weight
   is below 2^-24 (a faster tally add; results are the same either way). */
int mc_launch_batched(mcgpu_context *ctx, const mc_batched_shape &sh, const mc_slab_k *slabs,
                      const mc_tile_k *tiles, int32_t n_tiles, const mcgpu_bank &bank,
                      const mc_cell *cells, unsigned long long *fx, int64_t total_cells,
                      uint32_t *flags, int8_t *result8, mc_counters *ctr, uint64_t max_events,
                      bool small_weights, hipStream_t s);

#endif
