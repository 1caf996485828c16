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

/* wall clock, seconds since the epoch (the reference's timer: MPI_Wtime
   plus its offset to CLOCK_REALTIME, src/timer.cpp:7-23) */
double wall_now() {
  return std::chrono::duration<double>(std::chrono::system_clock::now().time_since_epoch())
      .count();
}

/* mcgpu_cycle_stats::on_window: every rank gets the window */
void on_window(void *user, const mcgpu_cycle_window *w) {
  auto *all = static_cast<std::vector<std::vector<TimerWindow>> *>(user);
  TimerWindow t;
  t.starttime = w->starttime;
  t.endtime = w->endtime;
  t.time_comp = w->time_comp;
  t.time_send = w->time_send;
  t.time_recv = w->time_recv;
  t.time_idle = w->time_idle;
  t.nb_cycles = w->nb_cycles;
  for (auto &v : *all) v.push_back(t);
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
  windows.assign(opt.world_size, {});
}

void SyncCycle::check_usable() const {
  if (!usable_)
    throw Error(MCGPU_ERR_ARG, "SyncCycle: a device-resident spin() failed; the layers' banks "
                               "were not brought back, so this run's state is undefined");
}

long long SyncCycle::spin(long long max_cycles) {
  check_usable();
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
  st.statistics_cycle_time = options.statistics_cycle_time;
  st.on_window = on_window;
  st.window_user = &windows;
  st.cycle_by_cycle = cycle_by_cycle ? 1 : 0;
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
  if (rc != MCGPU_OK) {
    usable_ = false;
    throw Error(rc, std::string("mcgpu_sync_cycle: ") + mcgpu_status_string(rc) + ": " +
                        mcgpu_context_error(ctx));
  }
  return st.cycles;
}

long long SyncCycle::spin_host(long long max_cycles) {
  const int R = options.world_size;
  const double t_start = now();
  /* statistics windows (src/worker_sync.cpp:25-27,121-133), per rank */
  std::vector<TimerWindow> win(R);
  for (TimerWindow &w : win) w.starttime = wall_now();
  long long win_cycles = 0;
  bool first_cycle = true;
  auto close_windows = [&](double t) {
    for (int r = 0; r < R; ++r) {
      win[r].endtime = t;
      win[r].nb_cycles = win_cycles;
      windows[r].push_back(win[r]);
      win[r] = TimerWindow{};
      win[r].starttime = t;
    }
    win_cycles = 0;
  };
  long long cycles = 0;
  try {
    while (true) {
      if (max_cycles > 0 && cycles >= max_cycles)
        throw Error(MCGPU_ERR_ARG, "SyncCycle::spin: no termination within max_cycles");
      const double tw0 = wall_now();
      if (first_cycle) {
        for (TimerWindow &w : win) w.time_idle += tw0 - w.starttime;
        first_cycle = false;
      }
      std::vector<double> comp(R);
      for (int r = 0; r < R; ++r) {
        const double t0 = now();
        const unsigned long long before = layers[r].total_stats.events;
        layers[r].simulate(options.nb_particles_per_cycle, options.nthread);
        events += layers[r].total_stats.events - before;
        comp[r] = now() - t0;
        times[r].comp += comp[r];
      }
      const double t1 = now();
      for (int r = 1; r + 1 < R; r += 2) sync_pair(layers[r], layers[r + 1]);
      for (int r = 0; r + 1 < R; r += 2) sync_pair(layers[r], layers[r + 1]);
      const double t2 = now();
      ++cycles;
      long long disabled = 0;
      for (int r = 0; r < R; ++r) disabled += layers[r].nb_disabled;
      const double t3 = now();
      const double tw1 = wall_now();
      for (int r = 0; r < R; ++r) {
        times[r].send += t2 - t1;
        times[r].recv += t3 - t2;
        win[r].time_comp += comp[r];
        win[r].time_recv += t3 - t2;
        win[r].time_send += std::max(0.0, (tw1 - tw0) - comp[r] - (t3 - t2));
      }
      if (disabled == options.nb_particles) break; /* the ending cycle is not counted */
      const double t = wall_now();
      if (t > win[0].starttime + options.statistics_cycle_time) close_windows(t);
      ++win_cycles;
    }
  } catch (...) {
    close_windows(wall_now());
    nb_cycles += cycles;
    elapsed = now() - t_start;
    throw;
  }
  close_windows(wall_now());
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
  check_usable();
  mkdir(folder.c_str(), 0777);
  const std::string cfg = folder + "/config.yaml";
  FILE *f = fopen(cfg.c_str(), "w");
  if (!f) throw Error(MCGPU_ERR_ARG, "cannot write " + cfg);
  char host[1000] = "";
  gethostname(host, sizeof host);
  write_config_yaml<real_t>(f, options, host);
  fclose(f);

  std::vector<int> displs;
  const std::vector<real_t> w = gather_weights_absorbed(&displs);
  const std::string wpath = folder + "/weights.csv";
  f = fopen(wpath.c_str(), "w");
  if (!f) throw Error(MCGPU_ERR_ARG, "cannot write " + wpath);
  write_weights_csv<real_t>(f, w.data(), (int)w.size(), displs.data(), options.world_size,
                            layers[0].dx);
  fclose(f);

  f = fopen(wa_path.c_str(), "w");
  if (!f) throw Error(MCGPU_ERR_ARG, "cannot write " + wa_path);
  const Layer &L0 = layers[0];
  write_wa<real_t>(f, L0.weights_absorbed.data(), (int)L0.weights_absorbed.size(), L0.x_min,
                   L0.dx);
  fclose(f);

  const std::string spath = folder + "/stats.csv";
  f = fopen(spath.c_str(), "w");
  if (!f) throw Error(MCGPU_ERR_ARG, "cannot write " + spath);
  write_stats_csv(f, windows);
  fclose(f);
}

} // namespace mcgpu
