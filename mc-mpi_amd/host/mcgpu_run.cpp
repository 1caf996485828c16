/*
 * mcgpu_run -- the run of `mpirun -n R ./main config.yaml sync`
 * (src/main.cpp:51-96, src/worker_sync.cpp:24-135) with all R layers on one
 * GPU: decompose_domain for every rank, synchronous cycles until every
 * particle is disabled, then the reference's outputs (src/worker.cpp:37-61):
 * out/config.yaml, out/weights.csv, out/stats.csv and WA.out of rank 0.
 * Prints the elapsed seconds on stdout, like src/main.cpp:90-91.
 *
 *   mcgpu_run config.yaml [history|event] [device] [--ranks R] [--out DIR] [--host-loop]
 *             [--cycle-by-cycle]
 *
 * By default the loop runs device-resident (mcgpu_sync_cycle); --host-loop
 * calls each rank's Layer::simulate in turn and exchanges on the host (the
 * kernel argument applies to that loop).  --cycle-by-cycle keeps a single
 * rank's cycles one after the other (otherwise they run concurrently).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>

#include "cycle.hpp"

int main(int argc, char **argv) {
  if (argc < 2) {
    fprintf(stderr,
            "usage: %s config.yaml [history|event] [device] [--ranks R] [--out DIR] "
            "[--host-loop] [--cycle-by-cycle]\n",
            argv[0]);
    return 2;
  }
  int ranks = 1, kernel = MCGPU_KERNEL_HISTORY, device = -1, positional = 0;
  bool host_loop = false, cycle_by_cycle = false;
  std::string out = "out";
  for (int i = 2; i < argc; ++i) {
    if (!strcmp(argv[i], "--ranks") && i + 1 < argc) {
      ranks = atoi(argv[++i]);
    } else if (!strcmp(argv[i], "--host-loop")) {
      host_loop = true;
    } else if (!strcmp(argv[i], "--cycle-by-cycle")) {
      cycle_by_cycle = true;
    } else if (!strcmp(argv[i], "--out") && i + 1 < argc) {
      out = argv[++i];
    } else if (positional == 0) {
      kernel = strcmp(argv[i], "event") == 0 ? MCGPU_KERNEL_EVENT : MCGPU_KERNEL_HISTORY;
      ++positional;
    } else {
      device = atoi(argv[i]);
    }
  }
  if (ranks < 1) {
    fprintf(stderr, "mcgpu_run: --ranks must be >= 1\n");
    return 2;
  }
  try {
    if (device >= 0) mcgpu::set_device(device);
    const mcgpu::MCMPIOptions opt = mcgpu::options_from_config(argv[1], ranks);
    mcgpu::SyncCycle run(opt);
    for (mcgpu::Layer &L : run.layers) L.options.kernel = kernel;
    run.device_resident = !host_loop;
    run.cycle_by_cycle = cycle_by_cycle;
    run.spin();
    printf("%lf\n", run.elapsed);
    run.dump(out);
    long long disabled = 0;
    double kernel_ms = 0;
    for (const mcgpu::Layer &L : run.layers) {
      disabled += L.nb_disabled;
      kernel_ms += L.total_stats.kernel_ms;
    }
    if (!host_loop) kernel_ms = run.times.empty() ? 0.0 : run.times[0].comp * 1e3;
    fprintf(stderr, "events %llu, disabled %lld, cycles %lld, ranks %d, kernel_ms %.3f, loop %s\n",
            run.events, disabled, run.nb_cycles, ranks, kernel_ms, host_loop ? "host" : "device");
  } catch (const std::exception &e) {
    fprintf(stderr, "mcgpu_run: %s\n", e.what());
    return 1;
  }
  return 0;
}
