/*
 * capi.hip -- the extern "C" entry points of include/mcgpu.h.
 *
 * Per call: cross sections -> per-cell (sig_a, sig_i) table and zeroed exact
 * tally (k_setup); transport (history or event kernel); tally -> double,
 * added to weights_absorbed (k_finalize); device counters read back.  Device
 * scratch lives in the context and only grows, so repeated calls (the
 * reference calls Layer::simulate once per cycle, src/worker_sync.cpp:38)
 * allocate nothing after the first.
 */
#include <stdio.h>
#include <string.h>

#include "internal.h"

#ifndef MCGPU_BUILD_ID
#define MCGPU_BUILD_ID "dev"
#endif

/* ---------------------------------------------------------------- helpers */

int mc_fail(mcgpu_context *ctx, int status, const char *what, hipError_t e) {
  if (ctx && !ctx->quiet) {
    ctx->err = what ? what : "";
    if (e != hipSuccess) {
      ctx->err += ": ";
      ctx->err += hipGetErrorString(e);
    }
  }
  return status;
}

int mc_ensure(mcgpu_context *ctx, void **buf, size_t *cap, size_t need) {
  if (need == 0) need = 256;
  if (*buf && *cap >= need) return MCGPU_OK;
  if (*buf) {
    hipError_t e = hipFree(*buf);
    *buf = nullptr;
    *cap = 0;
    if (e != hipSuccess) return mc_fail(ctx, MCGPU_ERR_HIP, "hipFree", e);
  }
  size_t alloc = need + need / 4; /* headroom for slowly growing sizes */
  hipError_t e = hipMalloc(buf, alloc);
  if (e != hipSuccess) {
    *buf = nullptr;
    (void)hipGetLastError();
    return mc_fail(ctx, MCGPU_ERR_NOMEM, "hipMalloc", e);
  }
  *cap = alloc;
  return MCGPU_OK;
}

namespace {

__global__ void k_setup(int m, const double *sigs, const double *ar, mc_cell *cells,
                        unsigned long long *fx, uint32_t *flags, mc_counters *ctr) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < m) {
    cells[c] = mc_make_cell(sigs[c], ar[c]);
    fx[c] = 0;
    fx[m + c] = 0;
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
    for (int i = 0; i < 8; ++i) ctr->aux[i] = 0;
  }
}

__global__ void k_finalize(int m, const unsigned long long *fx, const uint32_t *flags,
                           double *weights) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < m) weights[c] = weights[c] + mc_fx_finalize(fx[c], fx[m + c], flags[c]);
}

__global__ void k_rng_real_step(int64_t n, uint64_t *seeds, double *reals) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    uint64_t s = seeds[i];
    reals[i] = mc_rnd_real(s);
    seeds[i] = s;
  }
}

/* out[i] = (i+1)-th output of rnd_seed from `seed`. */
__global__ void k_seed_chain(uint64_t seed, int64_t n, uint64_t *out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = mc_seed_chain_at(seed, (uint64_t)i + 1);
}

/* Layer::create_particles(int) (src/layer.cpp:89-121) for particles
   first .. first+n-1 of a source: particle i takes the (first+i+1)-th chain
   output as its seed, then draws mu = 2 rnd_real - 1 from it. */
__global__ void k_create_particles(uint64_t seed, int64_t first, int64_t n, double x_ini,
                                   double wmc, int32_t index, mcgpu_bank b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t s = mc_seed_chain_at(seed, (uint64_t)(first + i) + 1);
  const double mu = 2 * mc_rnd_real(s) - 1;
  b.seed[i] = s;
  b.x[i] = x_ini;
  b.mu[i] = mu;
  b.wmc[i] = wmc;
  b.index[i] = index;
}

__global__ void k_math_eval(int which, int64_t n, const double *x, double *out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = which ? mc_exp(x[i]) : mc_log(x[i]);
}

/* Device entry points follow HIP's convention: NULL is the default (null)
   stream, which orders the call after work its caller put there. */
inline hipStream_t pick(mcgpu_context *ctx, void *stream) {
  (void)ctx;
  return reinterpret_cast<hipStream_t>(stream);
}

inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

/* Prepare cells + zeroed tally + counters for an m-cell layer whose sigs/ar
   are device pointers. */
int prepare(mcgpu_context *ctx, int m, const double *d_sigs, const double *d_ar, hipStream_t s) {
  int rc;
  if ((rc = mc_ensure(ctx, &ctx->cells, &ctx->cells_bytes, sizeof(mc_cell) * (size_t)m))) return rc;
  if ((rc = mc_ensure(ctx, &ctx->fx, &ctx->fx_bytes, 16 * (size_t)m))) return rc;
  if ((rc = mc_ensure(ctx, &ctx->flags, &ctx->flags_bytes, 4 * (size_t)m))) return rc;
  return mc_launch_setup(ctx, m, d_sigs, d_ar, (mc_cell *)ctx->cells,
                         (unsigned long long *)ctx->fx, (uint32_t *)ctx->flags,
                         (mc_counters *)ctx->counters, s);
}

int transport(mcgpu_context *ctx, mc_transport_args &a, const mcgpu_options *opt,
              hipStream_t s, double *d_weights, mcgpu_stats *stats) {
  mcgpu_options o;
  mcgpu_default_options(&o);
  if (opt) o = *opt;
  a.cells = (const mc_cell *)ctx->cells;
  a.fx = (unsigned long long *)ctx->fx;
  a.flags = (uint32_t *)ctx->flags;
  a.ctr = (mc_counters *)ctx->counters;
  a.max_events = o.max_events_per_history ? o.max_events_per_history : (1ULL << 27);
  MC_HIP(ctx, hipEventRecord(ctx->ev0, s));
  int rc = MCGPU_OK;
  if (a.n > 0) {
    if (o.kernel == MCGPU_KERNEL_HISTORY)
      rc = mc_launch_history(ctx, a, o, s);
    else if (o.kernel == MCGPU_KERNEL_EVENT)
      rc = mc_launch_event(ctx, a, o, s);
    else
      rc = mc_fail(ctx, MCGPU_ERR_ARG, "unknown options.kernel");
  }
  if (rc) return rc;
  MC_HIP(ctx, hipEventRecord(ctx->ev1, s));
  if ((rc = mc_launch_finalize(ctx, a.L.m, a.fx, a.flags, d_weights, s))) return rc;
  MC_HIP(ctx, hipMemcpyAsync(ctx->host_counters, ctx->counters, sizeof(mc_counters),
                             hipMemcpyDeviceToHost, s));
  MC_HIP(ctx, hipStreamSynchronize(s));
  float ms = 0.f;
  if (stats) (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
  return mc_status_from_counters(ctx, *(const mc_counters *)ctx->host_counters, ms, stats);
}

int check_layer(mcgpu_context *ctx, int32_t min_index, int32_t max_index, double dx) {
  if (max_index <= min_index) return mc_fail(ctx, MCGPU_ERR_ARG, "max_index must exceed min_index");
  (void)dx;
  return MCGPU_OK;
}

} // namespace

int mc_launch_setup(mcgpu_context *ctx, int m, const double *sigs, const double *ar,
                    mc_cell *cells, unsigned long long *fx, uint32_t *flags, mc_counters *ctr,
                    hipStream_t s) {
  hipLaunchKernelGGL(k_setup, dim3(grid_for(m > 0 ? m : 1, 256)), dim3(256), 0, s, m, sigs, ar,
                     cells, fx, flags, ctr);
  MC_HIP(ctx, hipGetLastError());
  return MCGPU_OK;
}

int mc_launch_finalize(mcgpu_context *ctx, int m, const unsigned long long *fx,
                       const uint32_t *flags, double *weights, hipStream_t s) {
  hipLaunchKernelGGL(k_finalize, dim3(grid_for(m, 256)), dim3(256), 0, s, m, fx, flags, weights);
  MC_HIP(ctx, hipGetLastError());
  return MCGPU_OK;
}

int mc_status_from_counters(mcgpu_context *ctx, const mc_counters &hc, float kernel_ms,
                            mcgpu_stats *stats) {
  if (stats) {
    stats->events = hc.events;
    stats->collisions = hc.collisions;
    stats->n_left = (int64_t)hc.n_left;
    stats->n_right = (int64_t)hc.n_right;
    stats->n_killed = (int64_t)hc.n_killed;
    stats->kernel_ms = kernel_ms;
  }
  if (hc.status & 1u)
    return mc_fail(ctx, MCGPU_ERR_EVENT_CAP, "a history exceeded max_events_per_history");
  if (hc.status & 2u)
    return mc_fail(ctx, MCGPU_ERR_RANGE, "a tally term was outside +-2^39");
  return MCGPU_OK;
}

int mc_launch_create(mcgpu_context *ctx, uint64_t seed, int64_t first, int64_t n, double x_ini,
                     double wmc, int32_t index, const mcgpu_bank &bank, hipStream_t s) {
  if (n <= 0) return MCGPU_OK;
  hipLaunchKernelGGL(k_create_particles, dim3(grid_for(n, 256)), dim3(256), 0, s, seed, first, n,
                     x_ini, wmc, index, bank);
  MC_HIP(ctx, hipGetLastError());
  return MCGPU_OK;
}

/* =========================================================== C ABI */
extern "C" {

const char *mcgpu_version(void) { return "mcgpu gfx950 " MCGPU_BUILD_ID; }

const char *mcgpu_status_string(int status) {
  switch (status) {
  case MCGPU_OK: return "ok";
  case MCGPU_ERR_ARG: return "invalid argument";
  case MCGPU_ERR_HIP: return "HIP runtime error";
  case MCGPU_ERR_NOMEM: return "device out of memory";
  case MCGPU_ERR_NO_DEVICE: return "no HIP device";
  case MCGPU_ERR_EVENT_CAP: return "history exceeded the event cap";
  case MCGPU_ERR_RANGE: return "tally term out of range";
  }
  return "unknown status";
}

void mcgpu_default_options(mcgpu_options *opt) {
  if (!opt) return;
  memset(opt, 0, sizeof(*opt));
  opt->kernel = MCGPU_KERNEL_HISTORY;
}

int mcgpu_context_create(int device, mcgpu_context **out) {
  if (!out) return MCGPU_ERR_ARG;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
    (void)hipGetLastError();
    return MCGPU_ERR_NO_DEVICE;
  }
  if (device < 0 || device >= count) return MCGPU_ERR_NO_DEVICE;
  mcgpu_context *ctx = new mcgpu_context();
  ctx->device = device;
  int rc = MCGPU_OK;
  hipDeviceProp_t prop;
  if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess) {
    delete ctx;
    return MCGPU_ERR_HIP;
  }
  ctx->num_cus = prop.multiProcessorCount;
  ctx->lds_per_block = prop.sharedMemPerBlock;
  /* library set-up of this device for the no-allocation entry point */
  if (mc_noalloc_device_setup(device) != MCGPU_OK) {
    delete ctx;
    return MCGPU_ERR_HIP;
  }
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess ||
      hipMalloc(&ctx->counters, sizeof(mc_counters)) != hipSuccess ||
      hipHostMalloc(&ctx->host_counters, sizeof(mc_counters), hipHostMallocDefault) != hipSuccess) {
    rc = MCGPU_ERR_HIP;
  }
  if (rc) {
    mcgpu_context_destroy(ctx);
    return rc;
  }
  *out = ctx;
  return MCGPU_OK;
}

int mcgpu_context_destroy(mcgpu_context *ctx) {
  if (!ctx) return MCGPU_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  void *bufs[] = {ctx->cells, ctx->fx, ctx->flags, ctx->counters, ctx->aos,
                  ctx->res32, ctx->sig_in, ctx->w_in, ctx->work, ctx->tables};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  if (ctx->host_counters) (void)hipHostFree(ctx->host_counters);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return MCGPU_OK;
}

const char *mcgpu_context_error(const mcgpu_context *ctx) { return ctx ? ctx->err.c_str() : ""; }

int mcgpu_simulate(mcgpu_context *ctx, int32_t n, mcgpu_particle *particles, const double *sigs,
                   const double *absorption_rates, double *weights_absorbed, int32_t min_index,
                   int32_t max_index, double dx, double particle_min_weight, int32_t *result,
                   const mcgpu_options *opt, mcgpu_stats *stats) {
  if (!ctx) return MCGPU_ERR_ARG;
  ctx->err.clear();
  if (n < 0 || (n > 0 && !particles) || !sigs || !absorption_rates || !weights_absorbed)
    return mc_fail(ctx, MCGPU_ERR_ARG, "null pointer or negative n");
  int rc = check_layer(ctx, min_index, max_index, dx);
  if (rc) return rc;
  MC_HIP(ctx, hipSetDevice(ctx->device));
  const int m = max_index - min_index;
  hipStream_t s = ctx->stream;
  if ((rc = mc_ensure(ctx, &ctx->sig_in, &ctx->sig_in_bytes, 16 * (size_t)m))) return rc;
  if ((rc = mc_ensure(ctx, &ctx->w_in, &ctx->w_in_bytes, 8 * (size_t)m))) return rc;
  if ((rc = mc_ensure(ctx, &ctx->aos, &ctx->aos_bytes, sizeof(mcgpu_particle) * (size_t)n))) return rc;
  if ((rc = mc_ensure(ctx, &ctx->res32, &ctx->res32_bytes, 4 * (size_t)n))) return rc;
  double *d_sigs = (double *)ctx->sig_in, *d_ar = d_sigs + m, *d_w = (double *)ctx->w_in;
  MC_HIP(ctx, hipMemcpyAsync(d_sigs, sigs, 8 * (size_t)m, hipMemcpyHostToDevice, s));
  MC_HIP(ctx, hipMemcpyAsync(d_ar, absorption_rates, 8 * (size_t)m, hipMemcpyHostToDevice, s));
  MC_HIP(ctx, hipMemcpyAsync(d_w, weights_absorbed, 8 * (size_t)m, hipMemcpyHostToDevice, s));
  if (n > 0)
    MC_HIP(ctx, hipMemcpyAsync(ctx->aos, particles, sizeof(mcgpu_particle) * (size_t)n,
                               hipMemcpyHostToDevice, s));
  if ((rc = prepare(ctx, m, d_sigs, d_ar, s))) return rc;
  mc_transport_args a;
  memset(&a, 0, sizeof a);
  a.L.index_start = min_index;
  a.L.m = m;
  a.L.dx = dx;
  a.L.min_weight = particle_min_weight;
  a.n = n;
  a.aos = (mcgpu_particle *)ctx->aos;
  a.result32 = (int32_t *)ctx->res32;
  rc = transport(ctx, a, opt, s, d_w, stats);
  if (rc != MCGPU_OK && rc != MCGPU_ERR_EVENT_CAP && rc != MCGPU_ERR_RANGE) return rc;
  /* copy results back (also after a capped / range status, for inspection) */
  MC_HIP(ctx, hipMemcpyAsync(weights_absorbed, d_w, 8 * (size_t)m, hipMemcpyDeviceToHost, s));
  if (n > 0) {
    MC_HIP(ctx, hipMemcpyAsync(particles, ctx->aos, sizeof(mcgpu_particle) * (size_t)n,
                               hipMemcpyDeviceToHost, s));
    if (result)
      MC_HIP(ctx, hipMemcpyAsync(result, ctx->res32, 4 * (size_t)n, hipMemcpyDeviceToHost, s));
  }
  MC_HIP(ctx, hipStreamSynchronize(s));
  return rc;
}

int mcgpu_simulate_device(mcgpu_context *ctx, int64_t n, const mcgpu_bank *bank,
                          const double *sigs, const double *absorption_rates,
                          double *weights_absorbed, int32_t min_index, int32_t max_index,
                          double dx, double particle_min_weight, int8_t *result,
                          uint64_t *events, uint64_t *collisions, const mcgpu_options *opt,
                          mcgpu_stats *stats, void *stream) {
  if (!ctx) return MCGPU_ERR_ARG;
  ctx->err.clear();
  if (n < 0 || !bank || !sigs || !absorption_rates || !weights_absorbed)
    return mc_fail(ctx, MCGPU_ERR_ARG, "null pointer or negative n");
  if (n > 0 && (!bank->seed || !bank->x || !bank->mu || !bank->wmc || !bank->index))
    return mc_fail(ctx, MCGPU_ERR_ARG, "bank has a null array");
  int rc = check_layer(ctx, min_index, max_index, dx);
  if (rc) return rc;
  MC_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t s = pick(ctx, stream);
  const int m = max_index - min_index;
  if ((rc = prepare(ctx, m, sigs, absorption_rates, s))) return rc;
  mc_transport_args a;
  memset(&a, 0, sizeof a);
  a.L.index_start = min_index;
  a.L.m = m;
  a.L.dx = dx;
  a.L.min_weight = particle_min_weight;
  a.n = n;
  a.soa = *bank;
  a.result8 = result;
  a.events = events;
  a.collisions = collisions;
  return transport(ctx, a, opt, s, weights_absorbed, stats);
}

int mcgpu_rng_real_step(mcgpu_context *ctx, int64_t n, uint64_t *seeds, double *reals,
                        void *stream) {
  if (!ctx || n < 0 || (n > 0 && (!seeds || !reals))) return MCGPU_ERR_ARG;
  if (n == 0) return MCGPU_OK;
  MC_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t s = pick(ctx, stream);
  hipLaunchKernelGGL(k_rng_real_step, dim3(grid_for(n, 256)), dim3(256), 0, s, n, seeds, reals);
  MC_HIP(ctx, hipGetLastError());
  MC_HIP(ctx, hipStreamSynchronize(s));
  return MCGPU_OK;
}

int mcgpu_rng_seed_chain(mcgpu_context *ctx, uint64_t seed, int64_t n, uint64_t *out,
                         void *stream) {
  if (!ctx || n < 0 || (n > 0 && !out)) return MCGPU_ERR_ARG;
  if (n == 0) return MCGPU_OK;
  MC_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t s = pick(ctx, stream);
  hipLaunchKernelGGL(k_seed_chain, dim3(grid_for(n, 256)), dim3(256), 0, s, seed, n, out);
  MC_HIP(ctx, hipGetLastError());
  MC_HIP(ctx, hipStreamSynchronize(s));
  return MCGPU_OK;
}

int mcgpu_create_particles(mcgpu_context *ctx, uint64_t seed, int64_t first, int64_t n,
                           double x_ini, double wmc, double dx, const mcgpu_bank *bank,
                           void *stream) {
  if (!ctx || n < 0 || first < 0) return MCGPU_ERR_ARG;
  if (n == 0) return MCGPU_OK;
  if (!bank || !bank->seed || !bank->x || !bank->mu || !bank->wmc || !bank->index)
    return mc_fail(ctx, MCGPU_ERR_ARG, "bank has a null array");
  const double cell = x_ini / dx;
  if (!(dx > 0) || !(cell >= 0) || !(cell < 2147483647.0))
    return mc_fail(ctx, MCGPU_ERR_ARG, "x_ini / dx is not a cell index");
  MC_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t s = pick(ctx, stream);
  hipLaunchKernelGGL(k_create_particles, dim3(grid_for(n, 256)), dim3(256), 0, s, seed, first,
                     n, x_ini, wmc, (int32_t)cell, *bank);
  MC_HIP(ctx, hipGetLastError());
  MC_HIP(ctx, hipStreamSynchronize(s));
  return MCGPU_OK;
}

int mcgpu_math_eval(mcgpu_context *ctx, int which, int64_t n, const double *x, double *out,
                    void *stream) {
  if (!ctx || n < 0 || (n > 0 && (!x || !out)) || (which != 0 && which != 1)) return MCGPU_ERR_ARG;
  if (n == 0) return MCGPU_OK;
  MC_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t s = pick(ctx, stream);
  hipLaunchKernelGGL(k_math_eval, dim3(grid_for(n, 256)), dim3(256), 0, s, which, n, x, out);
  MC_HIP(ctx, hipGetLastError());
  MC_HIP(ctx, hipStreamSynchronize(s));
  return MCGPU_OK;
}

} /* extern "C" */
