/*
 * mcgpu.h -- C ABI of the MI355X (gfx950) particle-transport library that
 * replaces the hot path of mc-mpi's Layer::simulate.
 *
 * Reference interface replaced: the GPU branch of Layer::simulate
 * (src/layer.cpp:254-300) calls
 *
 *   void cusimulate(int n, Particle *particles, float const *sigs,
 *                   float const *absorption_rates, float *weights_absorbed,
 *                   int min_index, int max_index, float dx);
 *                                       (include/culayer/culayer.hpp:6-13)
 *
 * and the CPU branch runs the OpenMP loop of src/layer.cpp:303-330 over the
 * same last-n particles.  mcgpu_simulate() is the drop-in for both: same
 * argument meaning and order, FP64 instead of FP32, plus the minimum weight
 * the CPU branch applies (src/layer.cpp:195) and an exit code per particle so
 * the caller can fill particles_left / particles_right exactly as
 * src/layer.cpp:332-346 does.
 *
 * Conventions
 *  - Every entry point returns an int status (MCGPU_OK or a negative code) and
 *    never terminates the process (the reference's gpu_errcheck calls exit(),
 *    include/gpu_errcheck/gpu_errcheck.hpp:10-18).
 *  - Per-particle outputs are indexed by array position p.  The reference
 *    transports particles[size-1-i] in iteration i; its result[i] is our
 *    result[n-1-i].
 *  - weights_absorbed is accumulated (+=), like the reference.  Each call's
 *    tally is an exact sum (fixed point, 2^-88 resolution), rounded once to
 *    double and then added, so it is independent of thread scheduling.
 *  - All arithmetic is IEEE binary64; log and exp are correctly rounded.
 *  - Functions taking a `void *stream` accept a hipStream_t; NULL is HIP's
 *    default (null) stream.  mcgpu_simulate (host buffers) uses the
 *    context's own stream.  No torch or HIP types appear in signatures.
 *  - Every call returns after its work has finished.  A context (and a
 *    batch plan) holds scratch buffers reused by its calls: use it from one
 *    thread at a time; separate contexts are independent.
 */
#ifndef MCGPU_H
#define MCGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------- status */
enum {
  MCGPU_OK = 0,
  MCGPU_ERR_ARG = -1,       /* invalid argument (null pointer, bad size, ...) */
  MCGPU_ERR_HIP = -2,       /* a HIP runtime call failed; see mcgpu_context_error */
  MCGPU_ERR_NOMEM = -3,     /* device allocation failed */
  MCGPU_ERR_NO_DEVICE = -4, /* no HIP device / bad device ordinal */
  MCGPU_ERR_EVENT_CAP = -5, /* a history exceeded options.max_events_per_history */
  MCGPU_ERR_RANGE = -6      /* a tally term outside +-2^39 (cell tally set to NaN) */
};

/* ------------------------------------------------------------- types */

/* include/types/particle.hpp:7-17 with real_t = double.  40 bytes. */
typedef struct mcgpu_particle {
  uint64_t seed;  /* live RNG state of the history */
  double x;
  double mu;
  double wmc;
  int32_t index;  /* global cell index */
  int32_t pad_;
} mcgpu_particle;

/* A particle bank in device memory, structure of arrays (n entries each). */
typedef struct mcgpu_bank {
  uint64_t *seed;
  double *x;
  double *mu;
  double *wmc;
  int32_t *index;
} mcgpu_bank;

enum {
  MCGPU_KERNEL_HISTORY = 0, /* one history per lane, lanes refilled from a work queue */
  MCGPU_KERNEL_EVENT = 1    /* one event per live particle per pass, SoA banks, compaction */
};

typedef struct mcgpu_options {
  int32_t kernel;                  /* MCGPU_KERNEL_* */
  int32_t block_size;              /* threads per block, 0 = automatic */
  int32_t grid_blocks;             /* blocks, 0 = automatic */
  int32_t events_per_refill;       /* history kernel: events between refills, 0 = automatic */
  int32_t compaction_interval;     /* event kernel: passes between compactions, 0 = automatic (64) */
  int32_t particles_per_lane;      /* event kernel: pool slots per thread, 0 = automatic (4) */
  uint64_t max_events_per_history; /* guard against runaway histories, 0 = 2^27 */
} mcgpu_options;

/* Device-counted totals of one call. */
typedef struct mcgpu_stats {
  uint64_t events;     /* particle_step calls (src/layer.cpp:123) */
  uint64_t collisions; /* events that took the collision branch */
  int64_t n_left;      /* histories that left through the low face */
  int64_t n_right;     /* ... through the high face */
  int64_t n_killed;    /* histories ended by the minimum weight */
  double kernel_ms;    /* device time of the transport kernel(s) */
} mcgpu_stats;

typedef struct mcgpu_context mcgpu_context;

/* ------------------------------------------------------------- context */

/* Create a context on HIP device `device`: a stream of default priority and
   grow-only device scratch buffers reused by later calls. */
int mcgpu_context_create(int device, mcgpu_context **out);
int mcgpu_context_destroy(mcgpu_context *ctx);
const char *mcgpu_status_string(int status);
/* Text of the last failure on this context ("" if none). */
const char *mcgpu_context_error(const mcgpu_context *ctx);
void mcgpu_default_options(mcgpu_options *opt);
/* Library build identification, e.g. "mcgpu gfx950 ...". */
const char *mcgpu_version(void);

/* ------------------------------------------------------------- transport */

/*
 * Drop-in for cusimulate (src/culayer.cu:41-92) and the CPU loop of
 * Layer::simulate (src/layer.cpp:303-330).
 *   n                 number of particles; particles[0..n) are transported
 *                     in place (host memory, in/out)
 *   sigs, absorption_rates  host arrays of max_index - min_index cells
 *   weights_absorbed  host array of max_index - min_index cells, +=
 *   min_index, max_index  the layer's global cell range [min, max)
 *   dx                the layer's cell width (Layer::dx)
 *   particle_min_weight   Layer::particle_min_weight
 *   result            host, n entries (may be NULL): exit code of
 *                     particles[p]: -1 left face, +1 right face, 0 killed
 *   opt, stats        may be NULL
 */
int mcgpu_simulate(mcgpu_context *ctx, int32_t n, mcgpu_particle *particles,
                   const double *sigs, const double *absorption_rates,
                   double *weights_absorbed, int32_t min_index, int32_t max_index, double dx,
                   double particle_min_weight, int32_t *result, const mcgpu_options *opt,
                   mcgpu_stats *stats);

/*
 * Same transport on a device-resident bank (for banks that live in HBM).
 * All pointers are device pointers; result/events/collisions may be NULL.
 * The call is stream-ordered on `stream` and returns after the transport
 * finished (it reads back the device counters).
 */
int mcgpu_simulate_device(mcgpu_context *ctx, int64_t n, const mcgpu_bank *bank,
                          const double *sigs, const double *absorption_rates,
                          double *weights_absorbed, int32_t min_index, int32_t max_index,
                          double dx, double particle_min_weight, int8_t *result,
                          uint64_t *events, uint64_t *collisions, const mcgpu_options *opt,
                          mcgpu_stats *stats, void *stream);

/*
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code that must not call hipMalloc, such as tight serving loops
 * or harnesses that forbid allocation).  mcgpu_simulate_device without a
 * context: the per-call scratch (cell table, exact tally, flags, counters)
 * lives in a fixed device array of the library, one per GPU, held until the
 * process ends; the call allocates nothing (no hipMalloc / hipFree, no
 * events; stats->kernel_ms is 0).  It runs on the current HIP device, on
 * `stream` (which must belong to that device), with the history kernel only
 * (options.kernel must be MCGPU_KERNEL_HISTORY), for layers of at most
 * MCGPU_NOALLOC_MAX_CELLS cells; larger layers, the event kernel or bad
 * arguments return MCGPU_ERR_ARG.  Results are those of
 * mcgpu_simulate_device.  All calls on one GPU share the array and each
 * call clears it first: calls that overlap on one GPU (other threads, or
 * streams not ordered with each other) race and must not be made.
 */
#define MCGPU_NOALLOC_MAX_CELLS 65536
int mcgpu_simulate_device_noalloc(int64_t n, const mcgpu_bank *bank, const double *sigs,
                                  const double *absorption_rates, double *weights_absorbed,
                                  int32_t min_index, int32_t max_index, double dx,
                                  double particle_min_weight, int8_t *result, uint64_t *events,
                                  uint64_t *collisions, const mcgpu_options *opt,
                                  mcgpu_stats *stats, void *stream);

/* ------------------------------------------------------------- batched
 * Many independent slab problems in one persistent launch (a work queue of
 * (slab, particle range) tiles), replayed from a captured hipGraph.  Each
 * slab is what one mcgpu_simulate_device call would transport; results are
 * bit-identical to running the slabs one call at a time.
 *
 * Inputs are concatenations in device memory: slab s owns cells
 * [cell_offset, cell_offset + max_index - min_index) of sigs, absorption_rates
 * and weights_absorbed, and particles [particle_offset, particle_offset + n)
 * of the bank (and of result, if given).
 */
typedef struct mcgpu_slab {
  int32_t min_index;       /* the slab's global cell range [min_index, max_index) */
  int32_t max_index;
  double dx;
  double particle_min_weight;
  int64_t cell_offset;
  int64_t particle_offset;
  int64_t n;
} mcgpu_slab;

typedef struct mcgpu_batch mcgpu_batch;

/* Build a plan: copies the slab table to the device, orders the tiles
   (largest work first) and captures setup + transport + finalize into a
   hipGraph.  opt: block_size, grid_blocks, events_per_refill and
   particles_per_lane (tile = particles_per_lane * block_size particles).
   Slabs of any cell count: a tile keeps its slab's tally (and cross
   sections) in LDS when they fit and adds to the global tally otherwise. */
int mcgpu_batch_create(mcgpu_context *ctx, int32_t n_slabs, const mcgpu_slab *slabs,
                       const double *sigs, const double *absorption_rates,
                       double *weights_absorbed, int64_t total_cells, const mcgpu_bank *bank,
                       int8_t *result, const mcgpu_options *opt, mcgpu_batch **out);
/* Launch the captured graph on `stream` and wait; stats may be NULL. */
int mcgpu_batch_run(mcgpu_batch *batch, void *stream, mcgpu_stats *stats);
int mcgpu_batch_destroy(mcgpu_batch *batch);
/* The options a plan was built with (after autotuning, the chosen ones). */
int mcgpu_batch_options(const mcgpu_batch *batch, mcgpu_options *opt);
/*
 * Autotuner: times the plan's transport for each candidate launch shape
 * (block size x particles per lane x events per refill, up to n_candidates
 * of the built-in list) on the plan's own inputs, restoring the bank and
 * weights from a device copy between trials, then re-captures the graph with
 * the fastest.  On return the bank and weights hold their original values.
 * times_ms (may be NULL) receives each candidate's time.
 */
int mcgpu_batch_autotune(mcgpu_batch *batch, int32_t n_candidates, double *times_ms,
                         mcgpu_options *chosen);

/* ------------------------------------------------------------- adjoint
 * Gradient of a weighted tally T = sum_c g[c] * tally[c] with respect to the
 * per-cell cross sections the transport step uses, sig_a = sigs * ar and
 * sig_i = sigs * (1.0 - ar) (src/layer.cpp:131-133).  These entry points take
 * the two cell arrays directly (instead of sigs / absorption_rates), so a
 * gradient and a finite difference refer to exactly the same parameters.
 * No counterpart exists in the reference.
 *
 * mcgpu_forward_cells: mcgpu_simulate_device with (sig_a, sig_i) given per
 *   cell; with g and history_score (device, n doubles) non-NULL it also
 *   stores each history's T_h = sum over its events of g[c] * dw (history
 *   kernel only).
 * mcgpu_adjoint_cells: replays every history of `initial` (the bank as it was
 *   before the forward call; not modified) and adds dT/dsig_a and dT/dsig_i
 *   to grad_sig_a / grad_sig_i (device, m doubles, +=): the pathwise
 *   derivative for sig_a and the score-function derivative for sig_i (each
 *   event's term weighted by the target tally from that event on).  With
 *   history_score NULL each history is replayed twice (score, then gradient).
 *   The weight kill at particle_min_weight is not differentiated.
 */
int mcgpu_forward_cells(mcgpu_context *ctx, int64_t n, const mcgpu_bank *bank,
                        const double *sig_a, const double *sig_i, double *weights_absorbed,
                        int32_t min_index, int32_t max_index, double dx,
                        double particle_min_weight, const double *g, double *history_score,
                        int8_t *result, const mcgpu_options *opt, mcgpu_stats *stats,
                        void *stream);
int mcgpu_adjoint_cells(mcgpu_context *ctx, int64_t n, const mcgpu_bank *initial,
                        const double *sig_a, const double *sig_i, int32_t min_index,
                        int32_t max_index, double dx, double particle_min_weight,
                        const double *g, const double *history_score, double *grad_sig_a,
                        double *grad_sig_i, const mcgpu_options *opt, mcgpu_stats *stats,
                        void *stream);

/*
 * Batched forms of the two calls above: many independent slabs (the
 * mcgpu_slab layout of mcgpu_batch_create) in one persistent launch each.
 * sig_a, sig_i, g, weights_absorbed, grad_sig_a and grad_sig_i are
 * concatenations over the slabs' cells (slab s owns [cell_offset, +m)), and
 * bank / initial, history_score and result over their particles (slab s owns
 * [particle_offset, +n)).  Every slab's outputs are bit-identical to the
 * single-slab call on that slab.  options: block_size, grid_blocks,
 * events_per_refill, particles_per_lane (tile = particles_per_lane x
 * block_size histories).
 */
int mcgpu_forward_cells_batch(mcgpu_context *ctx, int32_t n_slabs, const mcgpu_slab *slabs,
                              int64_t total_cells, const mcgpu_bank *bank, const double *sig_a,
                              const double *sig_i, double *weights_absorbed, const double *g,
                              double *history_score, int8_t *result, const mcgpu_options *opt,
                              mcgpu_stats *stats, void *stream);
int mcgpu_adjoint_cells_batch(mcgpu_context *ctx, int32_t n_slabs, const mcgpu_slab *slabs,
                              int64_t total_cells, const mcgpu_bank *initial,
                              const double *sig_a, const double *sig_i, const double *g,
                              const double *history_score, double *grad_sig_a,
                              double *grad_sig_i, const mcgpu_options *opt, mcgpu_stats *stats,
                              void *stream);

/* ------------------------------------------------------------- RNG helpers
 * Known-answer helpers for the reference RNG (src/random.cpp), evaluated on
 * the device (all pointers are device pointers):
 *   mcgpu_rng_real_step: seeds[i] <- one rnd_real step, reals[i] <- its value
 *                        (the kernel of src/curandom.cu:7-14, in FP64);
 *   mcgpu_rng_seed_chain: out[i] = the (i+1)-th rnd_seed output from `seed`,
 *                        computed with an affine jump-ahead (no serial loop).
 */
int mcgpu_rng_real_step(mcgpu_context *ctx, int64_t n, uint64_t *seeds, double *reals,
                        void *stream);
int mcgpu_rng_seed_chain(mcgpu_context *ctx, uint64_t seed, int64_t n, uint64_t *out,
                         void *stream);

/* ------------------------------------------------------------- sources
 * Device-side Layer::create_particles (src/layer.cpp:89-121): writes
 * particles first .. first+n-1 of the source that create_particles(x_ini,
 * wmc, ..., seed) (src/layer.cpp:71-82) would push, into bank[0..n) (device
 * SoA).  Particle i gets seed = the (first+i+1)-th rnd_seed output from
 * `seed`, then mu = 2 * rnd_real(&seed) - 1, x = x_ini, wmc = wmc and
 * index = (int)(x_ini / dx).  Chain positions are reached by jump-ahead, so
 * lazy creation in chunks passes first = particles already created.
 */
int mcgpu_create_particles(mcgpu_context *ctx, uint64_t seed, int64_t first, int64_t n,
                           double x_ini, double wmc, double dx, const mcgpu_bank *bank,
                           void *stream);

/* ------------------------------------------------------------- cycle loop
 * The synchronous loop of the reference's workers (WorkerSync::spin,
 * src/worker_sync.cpp:24-135, all ranks of `mpirun -n world_size ./main
 * config.yaml sync`) with every rank's layer resident on this GPU.  Per
 * cycle, each rank r does Layer::simulate(nb_particles_per_cycle)
 * (src/layer.cpp:239-361): lazy creation from its source when its bank holds
 * fewer than nb_particles_per_cycle particles (src/layer.cpp:89-121), then
 * transport of the last min(size, nb_particles_per_cycle) particles, whose
 * escapes form particles_left / particles_right in the reference's order
 * (iteration i transports particles[size-1-i]); escapes through a global
 * border (left_border / right_border) and kills count as disabled.  Then the
 * two exchange phases (odd ranks with their right neighbour, then even
 * ranks, src/worker_sync.cpp:55-109) append each escape buffer to the
 * neighbour's bank in sender order.  The loop ends when the disabled counts
 * add up to nb_particles (src/worker_sync.cpp:111-119).
 * All ranks' transports of one cycle share one launch (the batched
 * kernel's work queue), and banks, escape buffers and the exchange stay in
 * device memory: per cycle the host reads back one small summary.
 * Per history and per cycle the arithmetic is that of mcgpu_simulate, so
 * the tallies (one rounding into weights_absorbed per rank and cycle, as
 * Layer::simulate adds its call's tally) and the counts are bit-identical to
 * calling mcgpu_simulate per rank and cycle and exchanging on the host.
 * No counterpart exists in the reference's API (its ranks are processes).
 */
typedef struct mcgpu_cycle_rank {
  /* in: the rank's Layer */
  int32_t min_index;   /* global cell range [min_index, max_index) */
  int32_t max_index;
  double dx;
  double particle_min_weight;
  int32_t left_border; /* Layer::left_border / right_border */
  int32_t right_border;
  int64_t cell_offset; /* the rank's cells in the concatenated sigs / ar / weights */
  const mcgpu_particle *particles; /* host: the initial bank, particles[0..n_particles) */
  int64_t n_particles;
  /* in / out: the lazy source (Layer::create_particles(x_ini, wmc, n, seed)) */
  int64_t nb_particles_create;
  uint64_t seed;       /* rnd_seed chain state */
  double x_ini, wmc;
  /* in / out */
  int64_t nb_disabled;
  /* out: totals over the run */
  int64_t n_left, n_right, n_killed; /* escapes (either border or exchanged) and kills */
  int64_t n_remaining;               /* particles left in the bank */
} mcgpu_cycle_rank;

/* One statistics window of the loop: the Timer::State of the reference's
   workers (include/timer/timer.hpp:19-29, src/timer.cpp:48-58) and the
   cycles counted in it (src/worker_sync.cpp:121-133).  All ranks share one
   launch per cycle, so a window holds for every rank. */
typedef struct mcgpu_cycle_window {
  double starttime, endtime; /* wall clock, seconds since the epoch */
  double time_comp;          /* creation + transport + tally (device time) */
  double time_send;          /* partition + exchange (device time) */
  double time_recv;          /* the rest of each cycle: launch overhead, read-back
                                of the summary, termination test */
  double time_idle;          /* from the start of the loop to its first cycle */
  int64_t nb_cycles;
} mcgpu_cycle_window;

typedef void (*mcgpu_cycle_window_fn)(void *user, const mcgpu_cycle_window *window);

typedef struct mcgpu_cycle_stats {
  /* out */
  int64_t cycles;
  uint64_t events;
  uint64_t collisions;
  double kernel_ms;    /* the transport launches */
  double wall_ms;      /* the whole loop */
  /* in: statistics windows, kept as WorkerSync::spin keeps them
     (src/worker_sync.cpp:25-27,121-133): a window closes after a cycle that
     ends more than statistics_cycle_time seconds after the window opened,
     and when the loop ends (the cycle that ends the loop is not counted);
     on_window (may be NULL) is called with each closed window, in order */
  double statistics_cycle_time;
  mcgpu_cycle_window_fn on_window;
  void *window_user;
  /* in: 1 = always run the cycles one after the other (see below) */
  int32_t cycle_by_cycle;
  int32_t pad_;
} mcgpu_cycle_stats;

/* Runs cycles until the ranks' nb_disabled add up to nb_particles.
   sigs, absorption_rates, weights_absorbed: host arrays of total_cells (the
   ranks' cells at their cell_offset); weights_absorbed is added to.
   nb_particles_per_cycle must be in [1, 1000000] (MAX_PARTICLES_VECTOR).
   opt: block_size (default 256), particles_per_lane (default 1),
   events_per_refill, grid_blocks and max_events_per_history of the
   transport launch.  Returns MCGPU_ERR_ARG (and leaves the ranks' outputs
   as of the last cycle) if max_cycles > 0 cycles pass without termination.
   stats may be NULL; its statistics_cycle_time / on_window / window_user
   and cycle_by_cycle fields are read as inputs (zeroed: no windows
   reported).
   One rank whose layer has global borders at both ends (and whose bank,
   source and nb_disabled add up to nb_particles) exchanges nothing, and
   every particle a cycle transports leaves the bank: which particles each
   cycle transports then follows from the counts alone.  Unless
   cycle_by_cycle is set, all its cycles are transported concurrently
   (batched launches, each cycle with its own exact tally) and the tallies
   are added to weights_absorbed cycle after cycle, each rounded once:
   results bit-identical to the cycle-by-cycle loop.  Its statistics
   windows then cover whole launches (time_send 0). */
int mcgpu_sync_cycle(mcgpu_context *ctx, int32_t world_size, mcgpu_cycle_rank *ranks,
                     const double *sigs, const double *absorption_rates,
                     double *weights_absorbed, int64_t total_cells,
                     int32_t nb_particles_per_cycle, int64_t nb_particles, int64_t max_cycles,
                     const mcgpu_options *opt, mcgpu_cycle_stats *stats);

/* Evaluate the device log (which = 0) or exp (which = 1) used by the
   transport kernels on n inputs (device pointers). */
int mcgpu_math_eval(mcgpu_context *ctx, int which, int64_t n, const double *x, double *out,
                    void *stream);

#ifdef __cplusplus
}
#endif

#endif /* MCGPU_H */
