/*
 * cycle.hpp -- mcgpu::SyncCycle: mc-mpi's synchronous worker
 * (src/worker_sync.cpp:24-135) with every rank of a domain decomposition held
 * in one process and transported on one GPU.
 *
 * The reference runs one MPI rank per layer.  Each cycle every rank calls
 * Layer::simulate(nb_particles_per_cycle); then escape buffers go to the
 * neighbours in two phases (odd ranks with their right neighbour, then even
 * ranks with theirs, src/worker_sync.cpp:55-109), each received buffer being
 * appended to the receiver's particles in sender order; the loop ends when
 * the ranks' nb_disabled add up to nb_particles (src/worker_sync.cpp:111-119).
 * Here the layers are mcgpu::Layer objects (GPU transport) and the exchange
 * moves vectors in host memory, in the same order, so the run reproduces
 * `mpirun -n world_size ./main config.yaml sync` without MPI.
 *
 * dump() writes the reference's output files (src/worker.cpp:37-61):
 * folder/config.yaml (src/worker.cpp:218-241), folder/weights.csv
 * (src/worker.cpp:287-313), WA.out of rank 0's layer (src/layer.cpp:363-380)
 * and folder/stats.csv (src/worker.cpp:63-181): one row per rank and
 * statistics window, the windows kept as src/worker_sync.cpp:25-27,121-133
 * keeps them (a window closes after a cycle ending more than
 * statistics_cycle_time seconds after it opened, and at the end, the ending
 * cycle not counted).
 */
#ifndef MCGPU_CYCLE_HPP
#define MCGPU_CYCLE_HPP

#include <string>
#include <vector>

#include "formats.hpp"
#include "layer.hpp"
#include "yaml_loader.hpp"

namespace mcgpu {

class SyncCycle {
public:
  /* decompose_domain(..., options.world_size, r, ...) for every rank r */
  explicit SyncCycle(const MCMPIOptions &options);

  /* WorkerSync::spin; returns the number of cycles.  Throws mcgpu::Error if
     max_cycles (> 0) pass without every particle being disabled.
     With device_resident (the default) the whole loop runs on the GPU
     (mcgpu_sync_cycle: banks, escape buffers and the exchange in device
     memory, all ranks' transports of a cycle in one launch); otherwise each
     rank's Layer::simulate is called in turn and the escape buffers move
     between the Layer objects in host memory.  Both give the same cycles,
     events, nb_disabled and weights_absorbed, bit for bit.  On return from
     the device-resident loop every Layer holds its final state (tallies,
     nb_disabled, source, empty banks); its per-call stats (last_stats /
     total_stats) count escapes and kills per rank, and events only in
     SyncCycle::events.  nb_particles_per_cycle above MAX_PARTICLES_VECTOR
     takes the host loop (Layer::simulate then chunks each call).
     After the device-resident loop fails (max_cycles, event cap, tally
     range) the Layers' banks are not brought back from the device, so
     their state is not a consistent one: spin() and dump() then throw. */
  long long spin(long long max_cycles = -1);

  bool device_resident = true;
  /* the device-resident loop runs a single rank's cycles concurrently when
     that is exact (see mcgpu_sync_cycle); true: always one after the other */
  bool cycle_by_cycle = false;
  /* launch shape of the device-resident loop's transport (0 = defaults) */
  mcgpu_options device_options{};

  /* Worker::gather_weights_absorbed (src/worker.cpp:183-216): the ranks'
     tallies concatenated in rank order; displs receives each rank's offset. */
  std::vector<real_t> gather_weights_absorbed(std::vector<int> *displs = nullptr) const;

  /* Worker::dump (src/worker.cpp:37-61) */
  void dump(const std::string &folder, const std::string &wa_path = "WA.out") const;

  MCMPIOptions options;
  std::vector<Layer> layers;
  long long nb_cycles = 0;
  unsigned long long events = 0; /* device-counted over all ranks and cycles */

  struct RankTimes {
    double comp = 0, send = 0, recv = 0; /* seconds: transport, exchange, termination sum */
  };
  std::vector<RankTimes> times;
  double elapsed = 0; /* seconds of the last spin() */
  /* per rank, the statistics windows of every spin() so far (stats.csv).
     The device-resident loop runs all ranks in one launch per cycle, so
     every rank gets the same windows; the host loop times each rank's
     Layer::simulate on its own and counts the rest of the cycle (the other
     ranks' turns and the exchange, where an MPI rank would wait in
     MPI_Sendrecv) as its send time. */
  std::vector<std::vector<TimerWindow>> windows;

private:
  bool usable_ = true;
  void check_usable() const;
  long long spin_host(long long max_cycles);
  long long spin_device(long long max_cycles);
};

} // namespace mcgpu

#endif
