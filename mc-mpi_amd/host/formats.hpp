/*
 * formats.hpp -- the reference's output files, written byte for byte as the
 * reference writes them:
 *   WA.out           Layer::dump_WA              (src/layer.cpp:363-380)
 *   out/weights.csv  Worker::dump_weights_absorbed (src/worker.cpp:287-313)
 *   out/config.yaml  Worker::dump_config         (src/worker.cpp:218-241,
 *                                                 src/yaml_dumper.cpp:13-30)
 *   out/stats.csv    Worker::write_file          (src/worker.cpp:63-181,
 *                                                 src/timer.cpp:48-58)
 * The number writers are templates on the real type: with Real = float the
 * expressions round exactly as in the reference's FP32 build (its mixed
 * float / double promotions included), with Real = double they are the
 * FP64 build this library runs.
 */
#ifndef MCGPU_FORMATS_HPP
#define MCGPU_FORMATS_HPP

#include <stdio.h>

#include <vector>

#include "yaml_loader.hpp"

namespace mcgpu {

/* src/layer.cpp:373-377: x_min + (i * dx) + 0.5 * dx and W / dx */
template <class Real>
void write_wa(FILE *f, const Real *weights, int m, Real x_min, Real dx) {
  for (int i = 0; i < m; ++i)
    fprintf(f, "%.4e %.3e\n", x_min + (i * dx) + 0.5 * dx, weights[i] / dx);
}

/* src/worker.cpp:300-310: every rank's tallies in rank order (displs[r] =
   rank r's first cell), x and W / dx of the cell with rank 0's dx */
template <class Real>
void write_weights_csv(FILE *f, const Real *weights, int total_len, const int *displs,
                       int world_size, Real dx) {
  fprintf(f, "proc, x, weight\n");
  int proc = 0;
  for (int i = 0; i < total_len; ++i) {
    if (proc < world_size - 1 && displs[proc + 1] == i) proc++;
    fprintf(f, "%d, %.18e, %.18e\n", proc, dx * (i + 0.5), weights[i] / dx);
  }
}

/* src/worker.cpp:218-241.  Real: the type the options hold (the reference
   keeps x_min, x_max, x_ini and particle_min_weight as real_t and
   cycle_time, statistics_cycle_time as double).  buffer_size is written as
   0: the reference never sets it (src/worker.cpp:315-333 leaves it
   uninitialised). */
template <class Real>
void write_config_yaml(FILE *f, const MCMPIOptions &o, const char *hostname) {
  auto dump_int = [f](const char *k, int v) { fprintf(f, "%s: %d\n", k, v); };
  auto dump_double = [f](const char *k, double v) { fprintf(f, "%s: %.18e\n", k, v); };
  fprintf(f, "# %s\n", "Read from config");
  dump_int("nb_cells", o.nb_cells);
  dump_double("x_min", (Real)o.x_min);
  dump_double("x_max", (Real)o.x_max);
  dump_double("x_ini", (Real)o.x_ini);
  dump_double("particle_min_weight", (Real)o.particle_min_weight);
  dump_int("nb_particles", (int)o.nb_particles);
  dump_int("buffer_size", 0);
  dump_double("cycle_time", o.cycle_time);
  dump_int("nb_particles_per_cycle", o.nb_particles_per_cycle);
  dump_int("nthread", o.nthread);
  dump_double("statistics_cycle_time", o.statistics_cycle_time);
  fprintf(f, "\n");
  fprintf(f, "# %s\n", "Other values");
  dump_int("world_size", o.world_size);
  fprintf(f, "%s: %s\n", "hostname", hostname);
}

/* One statistics window of a rank: Timer::State (include/timer/timer.hpp:
   19-29, seconds since the epoch for start and end) and the cycles counted
   in it (src/worker_sync.cpp:121-133). */
struct TimerWindow {
  double starttime = 0, endtime = 0;
  double time_comp = 0, time_send = 0, time_recv = 0, time_idle = 0;
  long long nb_cycles = 0;
};

/* src/worker.cpp:63-99 gathered in rank order (src/worker.cpp:101-178):
   the header, then every rank's windows.  Without any window the header
   has no timer columns, as in the reference. */
void write_stats_csv(FILE *f, const std::vector<std::vector<TimerWindow>> &rank_windows);

} // namespace mcgpu

#endif
