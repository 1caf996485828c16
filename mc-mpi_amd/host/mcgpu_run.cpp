/*
 * mcgpu_run -- single-GPU run of a config.yaml, the one-rank case of
 * `mpirun -n 1 ./main config.yaml sync` (src/main.cpp:51-96,
 * src/worker_sync.cpp:24-135): decompose_domain for rank 0 of 1, then
 * Layer::simulate(nb_particles_per_cycle) cycles until every particle is
 * disabled; writes WA.out (src/layer.cpp:363-380) and weights.csv
 * (src/worker.cpp:287-313) and prints the elapsed seconds.
 *
 *   mcgpu_run config.yaml [history|event] [device]
 */
#include <stdio.h>
#include <string.h>

#include <chrono>

#include "layer.hpp"
#include "yaml_loader.hpp"

int main(int argc, char **argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: %s config.yaml [history|event] [device]\n", argv[0]);
    return 2;
  }
  try {
    const mcgpu::MCMPIOptions opt = mcgpu::options_from_config(argv[1], 1);
    if (argc > 3) mcgpu::set_device(atoi(argv[3]));
    auto t0 = std::chrono::steady_clock::now();
    mcgpu::Layer layer = mcgpu::decompose_domain(opt.x_min, opt.x_max, opt.x_ini, 1, 0,
                                                 opt.nb_cells, (int)opt.nb_particles,
                                                 opt.particle_min_weight);
    if (argc > 2 && strcmp(argv[2], "event") == 0) layer.options.kernel = MCGPU_KERNEL_EVENT;
    unsigned long long events = 0;
    while (layer.nb_disabled < opt.nb_particles) {
      layer.simulate(opt.nb_particles_per_cycle, opt.nthread, true);
      events += layer.last_stats.events;
      if (layer.nb_active() == 0 && layer.particles_left.empty() && layer.particles_right.empty())
        break;
    }
    const double secs =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    layer.dump_WA();
    FILE *f = fopen("weights.csv", "w");
    if (f) {
      fprintf(f, "proc, x, weight\n");
      for (size_t i = 0; i < layer.weights_absorbed.size(); ++i)
        fprintf(f, "%d, %.18e, %.18e\n", 0, layer.dx * (i + 0.5), layer.weights_absorbed[i] / layer.dx);
      fclose(f);
    }
    printf("%lf\n", secs);
    fprintf(stderr, "events %llu, disabled %d\n", events, layer.nb_disabled);
  } catch (const std::exception &e) {
    fprintf(stderr, "mcgpu_run: %s\n", e.what());
    return 1;
  }
  return 0;
}
