/*
 * cycle.cpp -- mcgpu::SyncCycle (src/worker_sync.cpp:24-135 on one GPU).
 */
#include "cycle.hpp"

#include <stdio.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>

namespace mcgpu {

namespace {

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

/* One MPI_Sendrecv pair of src/worker_sync.cpp:55-109 between ranks r and
   r+1: r's right escapes are appended to r+1's particles and r+1's left
   escapes to r's, each in the sender's order; the sent buffers are cleared. */
void sync_pair(Layer &a, Layer &b) {
  b.particles.insert(b.particles.end(), a.particles_right.begin(), a.particles_right.end());
  a.particles.insert(a.particles.end(), b.particles_left.begin(), b.particles_left.end());
  a.particles_right.clear();
  b.particles_left.clear();
}

} // namespace

SyncCycle::SyncCycle(const MCMPIOptions &opt) : options(opt) {
  if (opt.world_size < 1) throw Error(MCGPU_ERR_ARG, "SyncCycle: world_size must be >= 1");
  if (opt.nb_particles_per_cycle < 1)
    throw Error(MCGPU_ERR_ARG, "SyncCycle: nb_particles_per_cycle must be >= 1");
  layers.reserve(opt.world_size);
  for (int r = 0; r < opt.world_size; ++r)
    layers.push_back(decompose_domain(opt.x_min, opt.x_max, opt.x_ini, opt.world_size, r,
                                      opt.nb_cells, (int)opt.nb_particles,
                                      opt.particle_min_weight));
  times.assign(opt.world_size, RankTimes{});
}

long long SyncCycle::spin(long long max_cycles) {
  if (device_resident && options.nb_particles_per_cycle <= MAX_PARTICLES_VECTOR)
    return spin_device(max_cycles);
  return spin_host(max_cycles);
}

long long SyncCycle::spin_device(long long max_cycles) {
  const int R = options.world_size;
  std::vector<mcgpu_cycle_rank> rk(R);
  std::vector<int64_t> off(R + 1, 0);
  for (int r = 0; r < R; ++r) off[r + 1] = off[r] + layers[r].get_m();
  const int64_t total = off[R];
  std::vector<real_t> sigs(total), ar(total), w(total);
  for (int r = 0; r < R; ++r) {
    Layer &L = layers[r];
    std::copy(L.sigs.begin(), L.sigs.end(), sigs.begin() + off[r]);
    std::copy(L.absorption_rates.begin(), L.absorption_rates.end(), ar.begin() + off[r]);
    std::copy(L.weights_absorbed.begin(), L.weights_absorbed.end(), w.begin() + off[r]);
    if (!L.particles_left.empty() || !L.particles_right.empty())
      throw Error(MCGPU_ERR_ARG, "SyncCycle: escape buffers must be empty before spin()");
    mcgpu_cycle_rank &k = rk[r];
    k = mcgpu_cycle_rank{};
    k.min_index = L.get_index_start();
    k.max_index = L.get_index_start() + L.get_m();
    k.dx = L.dx;
    k.particle_min_weight = L.particle_min_weight;
    k.left_border = L.left_border;
    k.right_border = L.right_border;
    k.cell_offset = off[r];
    k.particles = L.particles.data();
    k.n_particles = (int64_t)L.particles.size();
    k.nb_particles_create = L.nb_particles_create;
    k.seed = L.seed;
    k.x_ini = L.x_ini;
    k.wmc = L.wmc;
    k.nb_disabled = L.nb_disabled;
  }
  mcgpu_cycle_stats st{};
  mcgpu_context *ctx = default_context();
  const int rc = mcgpu_sync_cycle(ctx, R, rk.data(), sigs.data(), ar.data(), w.data(), total,
                                  options.nb_particles_per_cycle, options.nb_particles, max_cycles,
                                  &device_options, &st);
  for (int r = 0; r < R; ++r) {
    Layer &L = layers[r];
    const mcgpu_cycle_rank &k = rk[r];
    std::copy(w.begin() + off[r], w.begin() + off[r + 1], L.weights_absorbed.begin());
    L.nb_disabled = (int)k.nb_disabled;
    L.nb_particles_create = (int)k.nb_particles_create;
    L.seed = k.seed;
    if (k.n_remaining == 0) L.particles.clear();
    L.total_stats.n_left += k.n_left;
    L.total_stats.n_right += k.n_right;
    L.total_stats.n_killed += k.n_killed;
    times[r].comp += st.kernel_ms * 1e-3;
    times[r].send += (st.wall_ms - st.kernel_ms) * 1e-3;
  }
  events += st.events;
  nb_cycles += st.cycles;
  elapsed = st.wall_ms * 1e-3;
  if (rc != MCGPU_OK)
    throw Error(rc, std::string("mcgpu_sync_cycle: ") + mcgpu_status_string(rc) + ": " +
                        mcgpu_context_error(ctx));
  return st.cycles;
}

long long SyncCycle::spin_host(long long max_cycles) {
  const int R = options.world_size;
  const double t_start = now();
  long long cycles = 0;
  while (true) {
    if (max_cycles > 0 && cycles >= max_cycles)
      throw Error(MCGPU_ERR_ARG, "SyncCycle::spin: no termination within max_cycles");
    for (int r = 0; r < R; ++r) {
      const double t0 = now();
      const unsigned long long before = layers[r].total_stats.events;
      layers[r].simulate(options.nb_particles_per_cycle, options.nthread);
      events += layers[r].total_stats.events - before;
      times[r].comp += now() - t0;
    }
    const double t1 = now();
    for (int r = 1; r + 1 < R; r += 2) sync_pair(layers[r], layers[r + 1]);
    for (int r = 0; r + 1 < R; r += 2) sync_pair(layers[r], layers[r + 1]);
    const double t2 = now();
    ++cycles;
    long long disabled = 0;
    for (int r = 0; r < R; ++r) disabled += layers[r].nb_disabled;
    const double t3 = now();
    for (int r = 0; r < R; ++r) {
      times[r].send += t2 - t1;
      times[r].recv += t3 - t2;
    }
    if (disabled == options.nb_particles) break;
  }
  nb_cycles += cycles;
  elapsed = now() - t_start;
  return cycles;
}

std::vector<real_t> SyncCycle::gather_weights_absorbed(std::vector<int> *displs) const {
  std::vector<real_t> all;
  if (displs) displs->clear();
  for (const Layer &L : layers) {
    if (displs) displs->push_back((int)all.size());
    all.insert(all.end(), L.weights_absorbed.begin(), L.weights_absorbed.end());
  }
  return all;
}

void SyncCycle::dump(const std::string &folder, const std::string &wa_path) const {
  mkdir(folder.c_str(), 0777);
  const std::string cfg = folder + "/config.yaml";
  FILE *f = fopen(cfg.c_str(), "w");
  if (!f) throw Error(MCGPU_ERR_ARG, "cannot write " + cfg);
  char host[1000] = "";
  gethostname(host, sizeof host);
  fprintf(f, "# %s\n", "Read from config");
  fprintf(f, "%s: %d\n", "nb_cells", options.nb_cells);
  fprintf(f, "%s: %.18e\n", "x_min", options.x_min);
  fprintf(f, "%s: %.18e\n", "x_max", options.x_max);
  fprintf(f, "%s: %.18e\n", "x_ini", options.x_ini);
  fprintf(f, "%s: %.18e\n", "particle_min_weight", options.particle_min_weight);
  fprintf(f, "%s: %d\n", "nb_particles", (int)options.nb_particles);
  fprintf(f, "%s: %d\n", "buffer_size", 0);
  fprintf(f, "%s: %.18e\n", "cycle_time", options.cycle_time);
  fprintf(f, "%s: %d\n", "nb_particles_per_cycle", options.nb_particles_per_cycle);
  fprintf(f, "%s: %d\n", "nthread", options.nthread);
  fprintf(f, "%s: %.18e\n", "statistics_cycle_time", options.statistics_cycle_time);
  fprintf(f, "\n");
  fprintf(f, "# %s\n", "Other values");
  fprintf(f, "%s: %d\n", "world_size", options.world_size);
  fprintf(f, "%s: %s\n", "hostname", host);
  fclose(f);

  std::vector<int> displs;
  const std::vector<real_t> w = gather_weights_absorbed(&displs);
  const std::string wpath = folder + "/weights.csv";
  f = fopen(wpath.c_str(), "w");
  if (!f) throw Error(MCGPU_ERR_ARG, "cannot write " + wpath);
  fprintf(f, "proc, x, weight\n");
  const real_t dx = layers[0].dx;
  int proc = 0;
  for (int i = 0; i < (int)w.size(); ++i) {
    if (proc < options.world_size - 1 && displs[proc + 1] == i) proc++;
    fprintf(f, "%d, %.18e, %.18e\n", proc, dx * (i + 0.5), w[i] / dx);
  }
  fclose(f);

  f = fopen(wa_path.c_str(), "w");
  if (!f) throw Error(MCGPU_ERR_ARG, "cannot write " + wa_path);
  const Layer &L0 = layers[0];
  for (int i = 0; i < (int)L0.weights_absorbed.size(); ++i)
    fprintf(f, "%.4e %.3e\n", L0.x_min + (i * L0.dx) + 0.5 * L0.dx, L0.weights_absorbed[i] / L0.dx);
  fclose(f);

  const std::string spath = folder + "/stats.csv";
  f = fopen(spath.c_str(), "w");
  if (!f) throw Error(MCGPU_ERR_ARG, "cannot write " + spath);
  fprintf(f, "rank, %s, %s, %s, %s, %s, %s, nb_cycles, \n", "starttime", "endtime", "time_comp",
          "time_send", "time_recv", "time_idle");
  for (int r = 0; r < options.world_size; ++r)
    fprintf(f, "%d, %.18e, %.18e, %.18e, %.18e, %.18e, %.18e, %lld, \n", r, 0.0, elapsed,
            times[r].comp, times[r].send, times[r].recv, 0.0, nb_cycles);
  fclose(f);
}

} // namespace mcgpu
