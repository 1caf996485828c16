/*
 * noalloc.hip -- mcgpu_simulate_device_noalloc: the device-resident
 * transport for callers that cannot allocate device memory.
 *
 * mcgpu_simulate_device keeps its scratch (per-cell cross-section pairs,
 * the exact fixed-point tally, its flag words and the counter block) in
 * grow-only buffers of a context, allocated on first use.  This entry point
 * takes no context and calls no allocator: the same scratch lives in one
 * file-scope device array, `scratch`, of fixed size, which every GPU that
 * loads the library holds until the process ends.  Host code never uses
 * the array's name as an address (on the host it is only a stand-in
 * symbol): the device address is looked up once per device with
 * hipGetSymbolAddress, with that device current, and kept.
 *
 * Limits: layers of at most MCGPU_NOALLOC_MAX_CELLS cells, the history
 * kernel only (the event kernel's particle pools are hundreds of MB).
 * Every call on one GPU uses the same array: calls that overlap on one GPU
 * (two host threads, or two streams without ordering) race.  Each call
 * clears what must start at zero (tally words, flags, counters) with a
 * kernel on the caller's stream before the transport.
 */
#include <string.h>

#include <mutex>

#include "internal.h"

namespace {

constexpr size_t kCounterBytes = 256; /* the mc_counters block, padded */
static_assert(sizeof(mc_counters) <= kCounterBytes, "counter block exceeds its slot");
constexpr size_t kMaxCells = MCGPU_NOALLOC_MAX_CELLS;
/* counters | cells (16 B) | tally lo, hi (8 B each) | flags (4 B), per cell */
constexpr size_t kCellsOff = kCounterBytes;
constexpr size_t kFxOff = kCellsOff + sizeof(mc_cell) * kMaxCells;
constexpr size_t kFlagsOff = kFxOff + 16 * kMaxCells;
constexpr size_t kScratchBytes = kFlagsOff + 4 * kMaxCells;

} // namespace

/* The scratch array: 256 + 36 * MCGPU_NOALLOC_MAX_CELLS bytes (2.25 MiB) per
   device.  Only kernels touch it, through the address below. */
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

/* Per-device set-up: device address of `scratch` and the properties the
   launchers need. */
constexpr int kMaxDevices = 64;
struct dev_setup {
  bool ready = false;
  unsigned char *base = nullptr; /* hipGetSymbolAddress of scratch on this device */
  int num_cus = 0;
  size_t lds_per_block = 0;
};
dev_setup g_dev[kMaxDevices];
std::mutex g_dev_mu;

} // namespace

/* Once per device, with that device current. */
int mc_noalloc_device_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return MCGPU_ERR_NO_DEVICE;
  std::lock_guard<std::mutex> lock(g_dev_mu);
  dev_setup &d = g_dev[device];
  if (d.ready) return MCGPU_OK;
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess || cur != device) {
    (void)hipGetLastError();
    return MCGPU_ERR_HIP;
  }
  void *ptr = nullptr;
  if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
    (void)hipGetLastError();
    return MCGPU_ERR_HIP;
  }
  int cus = 0, lds = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
      hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) !=
          hipSuccess) {
    (void)hipGetLastError();
    return MCGPU_ERR_HIP;
  }
  d.base = static_cast<unsigned char *>(ptr);
  d.num_cus = cus;
  d.lds_per_block = (size_t)lds;
  d.ready = true;
  return MCGPU_OK;
}

extern "C" int mcgpu_simulate_device_noalloc(int64_t n, const mcgpu_bank *bank,
                                             const double *sigs, const double *absorption_rates,
                                             double *weights_absorbed, int32_t min_index,
                                             int32_t max_index, double dx,
                                             double particle_min_weight, int8_t *result,
                                             uint64_t *events, uint64_t *collisions,
                                             const mcgpu_options *opt, mcgpu_stats *stats,
                                             void *stream) {
  if (n < 0 || !bank || !sigs || !absorption_rates || !weights_absorbed) return MCGPU_ERR_ARG;
  if (n > 0 && (!bank->seed || !bank->x || !bank->mu || !bank->wmc || !bank->index))
    return MCGPU_ERR_ARG;
  if (max_index <= min_index) return MCGPU_ERR_ARG;
  const int64_t m = (int64_t)max_index - (int64_t)min_index;
  if (m > (int64_t)kMaxCells) return MCGPU_ERR_ARG; /* beyond the scratch array */
  mcgpu_options o;
  mcgpu_default_options(&o);
  if (opt) o = *opt;
  if (o.kernel != MCGPU_KERNEL_HISTORY) return MCGPU_ERR_ARG;
  if (o.block_size != 0 && (o.block_size < 0 || o.block_size > 1024 || o.block_size % 64))
    return MCGPU_ERR_ARG;

  int device = -1;
  if (hipGetDevice(&device) != hipSuccess) {
    (void)hipGetLastError();
    return MCGPU_ERR_NO_DEVICE;
  }
  int rc = mc_noalloc_device_setup(device);
  if (rc) return rc;
  const dev_setup &d = g_dev[device];

  /* a context on the stack: launch parameters only, no buffers, no text */
  mcgpu_context tmp;
  tmp.device = device;
  tmp.num_cus = d.num_cus;
  tmp.lds_per_block = d.lds_per_block;
  tmp.quiet = true;

  mc_counters *ctr = reinterpret_cast<mc_counters *>(d.base);
  mc_cell *cells = reinterpret_cast<mc_cell *>(d.base + kCellsOff);
  unsigned long long *fx = reinterpret_cast<unsigned long long *>(d.base + kFxOff);
  uint32_t *flags = reinterpret_cast<uint32_t *>(d.base + kFlagsOff);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);

  if ((rc = mc_launch_setup(&tmp, (int)m, sigs, absorption_rates, cells, fx, flags, ctr, s)))
    return rc;
  mc_transport_args a;
  memset(&a, 0, sizeof a);
  a.L.index_start = min_index;
  a.L.m = (int32_t)m;
  a.L.dx = dx;
  a.L.min_weight = particle_min_weight;
  a.n = n;
  a.soa = *bank;
  a.result8 = result;
  a.events = events;
  a.collisions = collisions;
  a.cells = cells;
  a.fx = fx;
  a.flags = flags;
  a.ctr = ctr;
  a.max_events = o.max_events_per_history ? o.max_events_per_history : (1ULL << 27);
  if (n > 0 && (rc = mc_launch_history(&tmp, a, o, s))) return rc;
  if ((rc = mc_launch_finalize(&tmp, (int)m, fx, flags, weights_absorbed, s))) return rc;
  mc_counters hc;
  if (hipMemcpyAsync(&hc, ctr, sizeof hc, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    (void)hipGetLastError();
    return MCGPU_ERR_HIP;
  }
  return mc_status_from_counters(&tmp, hc, 0.f, stats);
}
