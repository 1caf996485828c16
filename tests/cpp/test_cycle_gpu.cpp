// test_cycle_gpu.cpp -- mcgpu::SyncCycle (every rank of a decomposition on
// one GPU) against the oracle's restatement of the synchronous worker loop
// (orc64_sync_run, src/worker_sync.cpp:24-135).
//
// For each case both run the whole loop, SyncCycle once with its
// device-resident loop (mcgpu_sync_cycle) and once with the host loop over
// Layer::simulate; then the number of cycles, the device-counted events, and
// for every rank the particles, escape buffers (bytes), nb_disabled and
// weights_absorbed (bits; both add each call's exact tally) must be
// identical.
#include <stdio.h>
#include <string.h>

#include <vector>

#include "cycle.hpp"
#include "oracle.h"

static int failures = 0;
#define CHECK(cond, ...)                                                   \
  do {                                                                     \
    if (!(cond)) {                                                         \
      fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__);                 \
      fprintf(stderr, __VA_ARGS__);                                        \
      fprintf(stderr, "\n");                                               \
      ++failures;                                                          \
    }                                                                      \
  } while (0)

template <class T>
static bool same_bytes(const std::vector<T> &a, const void *b, int64_t nb) {
  if ((int64_t)a.size() != nb) return false;
  return nb == 0 || memcmp(a.data(), b, sizeof(T) * nb) == 0;
}

struct Case {
  const char *name;
  int world_size, nb_cells, nb_particles, per_cycle;
  double x_ini, min_w;
  int kernel;
  double wscale = 1.0; /* initial weights and the minimum weight scaled by this */
};

/* every particle's weight times s (all particles are created up front when
   nb_particles <= MAX_PARTICLES_VECTOR) */
static void scale_weights(mcgpu_particle *p, int64_t n, double s) {
  for (int64_t i = 0; i < n; ++i) p[i].wmc *= s;
}

static void run(const Case &c, bool device_resident, bool cycle_by_cycle = false) {
  mcgpu::MCMPIOptions opt{};
  opt.world_size = c.world_size;
  opt.nb_cells = c.nb_cells;
  opt.x_min = 0.0;
  opt.x_max = 1.0;
  opt.x_ini = c.x_ini;
  opt.particle_min_weight = c.min_w * c.wscale;
  opt.nb_particles = c.nb_particles;
  opt.cycle_time = 1e-5;
  opt.nb_particles_per_cycle = c.per_cycle;
  opt.nthread = 1;
  opt.statistics_cycle_time = 0.1;
  mcgpu::SyncCycle g(opt);
  for (auto &L : g.layers) {
    L.options.kernel = c.kernel;
    scale_weights(L.particles.data(), (int64_t)L.particles.size(), c.wscale);
  }
  g.device_resident = device_resident;
  g.cycle_by_cycle = cycle_by_cycle;
  g.spin(10000000);
  const char *loop = !device_resident ? "host loop"
                     : cycle_by_cycle ? "device loop, cycle by cycle"
                                      : "device loop";

  std::vector<orc64_Layer *> O(c.world_size);
  for (int r = 0; r < c.world_size; ++r) {
    O[r] = orc64_decompose_domain(0.0, 1.0, c.x_ini, c.world_size, r, c.nb_cells, c.nb_particles,
                                  c.min_w * c.wscale);
    orc64_layer_set_tally_exact(O[r], 1);
    void *bank;
    const int64_t nb = orc64_layer_array(O[r], 3, &bank);
    scale_weights(static_cast<mcgpu_particle *>(bank), nb, c.wscale);
  }
  const int64_t cycles = orc64_sync_run(O.data(), c.world_size, c.nb_particles, c.per_cycle, 8,
                                        10000000);
  CHECK(cycles == g.nb_cycles, "%s (%s) cycles %lld vs %lld", c.name, loop,
        (long long)g.nb_cycles, (long long)cycles);
  unsigned long long events = 0;
  long long disabled = 0;
  for (int r = 0; r < c.world_size; ++r) {
    mcgpu::Layer &G = g.layers[r];
    void *p;
    int64_t n;
    n = orc64_layer_array(O[r], 3, &p);
    CHECK(same_bytes(G.particles, p, n), "%s (%s) rank %d particles", c.name, loop, r);
    n = orc64_layer_array(O[r], 4, &p);
    CHECK(same_bytes(G.particles_left, p, n), "%s (%s) rank %d particles_left", c.name, loop, r);
    n = orc64_layer_array(O[r], 5, &p);
    CHECK(same_bytes(G.particles_right, p, n), "%s (%s) rank %d particles_right", c.name, loop, r);
    n = orc64_layer_array(O[r], 2, &p);
    CHECK(same_bytes(G.weights_absorbed, p, n), "%s (%s) rank %d weights_absorbed", c.name, loop, r);
    CHECK(G.nb_disabled == orc64_layer_int(O[r], 0), "%s (%s) rank %d nb_disabled %d vs %lld",
          c.name, loop, r, G.nb_disabled, (long long)orc64_layer_int(O[r], 0));
    events += (unsigned long long)orc64_layer_int(O[r], 7);
    disabled += G.nb_disabled;
    orc64_layer_free(O[r]);
  }
  CHECK(events == g.events, "%s (%s) events %llu vs %llu", c.name, loop, g.events, events);
  CHECK(disabled == c.nb_particles, "%s (%s) disabled %lld", c.name, loop, disabled);
  /* statistics windows: every rank counts all cycles but the ending one */
  for (int r = 0; r < c.world_size; ++r) {
    long long counted = 0;
    for (const mcgpu::TimerWindow &w : g.windows[r]) counted += w.nb_cycles;
    CHECK(!g.windows[r].empty() && counted == g.nb_cycles - 1,
          "%s (%s) rank %d windows count %lld cycles of %lld", c.name, loop, r, counted,
          g.nb_cycles);
  }
  printf("ok %s (%s): %lld cycles, %llu events\n", c.name, loop, g.nb_cycles, g.events);
  fflush(stdout);
}

/* A device-resident spin() stopped by max_cycles leaves the layers without
   their banks: the SyncCycle then refuses further spin() and dump(). */
static void failed_spin_is_not_reused() {
  mcgpu::MCMPIOptions opt{};
  opt.world_size = 3;
  opt.nb_cells = 300;
  opt.x_min = 0.0;
  opt.x_max = 1.0;
  opt.x_ini = 0.5;
  opt.particle_min_weight = 1e-9;
  opt.nb_particles = 3000;
  opt.cycle_time = 1e-5;
  opt.nb_particles_per_cycle = 10;
  opt.nthread = 1;
  opt.statistics_cycle_time = 0.1;
  mcgpu::SyncCycle g(opt);
  int thrown = 0;
  try {
    g.spin(2);
  } catch (const mcgpu::Error &e) {
    thrown += e.status == MCGPU_ERR_ARG;
  }
  try {
    g.spin();
  } catch (const mcgpu::Error &) {
    ++thrown;
  }
  try {
    g.dump("out_failed", "WA_failed.out");
  } catch (const mcgpu::Error &) {
    ++thrown;
  }
  CHECK(thrown == 3, "a failed device-resident spin: %d of 3 calls threw", thrown);
  printf("ok failed device-resident spin is not reused\n");
  fflush(stdout);
}

/* One rank, with part of the source already in the bank when the loop
   starts (Layer::create_particles(int) before spin): the concurrent cycles
   of the device loop, its cycle-by-cycle form and the host loop agree bit
   for bit (the source's chunks and the initial bank both feed cycles). */
static void one_rank_with_initial_bank() {
  mcgpu::MCMPIOptions opt{};
  opt.world_size = 1;
  opt.nb_cells = 250;
  opt.x_min = 0.0;
  opt.x_max = 1.0;
  opt.x_ini = 0.37;
  opt.particle_min_weight = 1e-8;
  opt.nb_particles = 1300000;
  opt.cycle_time = 1e-5;
  opt.nb_particles_per_cycle = 90000;
  opt.nthread = 1;
  opt.statistics_cycle_time = 0.1;
  std::vector<mcgpu::real_t> w[3];
  long long cycles[3];
  unsigned long long ev[3];
  for (int mode = 0; mode < 3; ++mode) {
    mcgpu::SyncCycle g(opt);
    g.layers[0].create_particles(123457); /* part of the source into the bank */
    g.device_resident = mode != 2;
    g.cycle_by_cycle = mode == 1;
    g.spin(10000000);
    w[mode] = g.layers[0].weights_absorbed;
    cycles[mode] = g.nb_cycles;
    ev[mode] = g.events;
    CHECK(g.layers[0].nb_disabled == 1300000 && g.layers[0].particles.empty(),
          "one rank with an initial bank, mode %d: disabled %d", mode, g.layers[0].nb_disabled);
  }
  for (int mode = 1; mode < 3; ++mode) {
    CHECK(cycles[mode] == cycles[0] && ev[mode] == ev[0], "one rank: mode %d cycles %lld vs %lld",
          mode, cycles[mode], cycles[0]);
    CHECK(w[mode].size() == w[0].size() &&
              memcmp(w[mode].data(), w[0].data(), sizeof(mcgpu::real_t) * w[0].size()) == 0,
          "one rank: mode %d weights differ", mode);
  }
  printf("ok one rank with an initial bank: %lld cycles, %llu events\n", cycles[0], ev[0]);
  fflush(stdout);
}

int main() {
  const double x_ini = 7.071067690849304199e-01, min_w = 9.999999960041972003e-13;
  const std::vector<Case> cases = {
      {"config.yaml source, 1 rank", 1, 1000, 20000, 500, x_ini, min_w, MCGPU_KERNEL_HISTORY},
      /* one rank, lazy creation in chunks of MAX_PARTICLES_VECTOR between
         cycles (the concurrent path's source segments) */
      {"lazy creation, 1 rank, 2.5e6 particles", 1, 77, 2500000, 300000, 0.45, 1e-9,
       MCGPU_KERNEL_HISTORY},
      {"config.yaml source, 2 ranks", 2, 1000, 20000, 500, x_ini, min_w, MCGPU_KERNEL_HISTORY},
      {"config.yaml source, 5 ranks", 5, 1000, 20000, 500, x_ini, min_w, MCGPU_KERNEL_HISTORY},
      {"uneven split, 7 ranks, source in rank 0", 7, 1000, 10000, 300, 0.05, 1e-6,
       MCGPU_KERNEL_HISTORY},
      {"event kernel, 4 ranks", 4, 1024, 10000, 700, 0.3, 1e-6, MCGPU_KERNEL_EVENT},
      /* more particles than MAX_PARTICLES_VECTOR: lazy creation in the source
         rank during the loop */
      {"lazy creation, 3 ranks, 1.2e6 particles", 3, 99, 1200000, 100000, 0.5, 1e-9,
       MCGPU_KERNEL_HISTORY},
      /* many ranks, tiny cycles: 16 layers of 62-63 cells, 7 particles per
         cycle and rank */
      {"16 ranks, 7 per cycle", 16, 1000, 2000, 7, 0.41, 1e-7, MCGPU_KERNEL_HISTORY},
      /* the source in the last rank (the right border) */
      {"source in the right-border rank, 6 ranks", 6, 600, 8000, 250, 0.97, 1e-7,
       MCGPU_KERNEL_HISTORY},
      /* kills dominate: weights 1e-4 against a minimum of 5e-5 */
      {"kills dominate, 4 ranks", 4, 400, 10000, 333, 0.6, 5e-5, MCGPU_KERNEL_HISTORY},
      /* weights below 2^-24 (the device loop's tally then adds the high
         words only for carries) */
      {"weights of 5e-9, 5 ranks", 5, 1000, 20000, 500, 7.071067690849304199e-01,
       9.999999960041972003e-13, MCGPU_KERNEL_HISTORY, 1e-4},
  };
  try {
    failed_spin_is_not_reused();
    one_rank_with_initial_bank();
    for (const Case &c : cases) {
      printf("case %s\n", c.name);
      fflush(stdout);
      run(c, true);
      if (c.world_size == 1) run(c, true, true);
      run(c, false);
    }
  } catch (const std::exception &e) {
    fprintf(stderr, "exception: %s\n", e.what());
    return 1;
  }
  if (failures) {
    fprintf(stderr, "%d failures\n", failures);
    return 1;
  }
  printf("all ok\n");
  return 0;
}
