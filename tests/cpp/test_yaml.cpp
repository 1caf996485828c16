// test_yaml.cpp -- the config reader (src/yaml_loader.cpp:4-40 semantics),
// CPU only: parses the reference's config.yaml format, throws on a missing key
// (std::out_of_range from map::at) and on an unreadable file (mcgpu::Error).
#include <stdio.h>

#include <stdexcept>

#include "layer.hpp"
#include "yaml_loader.hpp"

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  const char *path = argv[1];
  FILE *f = fopen(path, "w");
  fprintf(f, "x_min: 0.000000000000000000e+00\nx_max: 1.000000000000000000e+00\n"
             "x_ini: 7.071067690849304199e-01\nparticle_min_weight: 9.999999960041972003e-13\n"
             "nb_cells: 1000\nnb_particles: 100000\ncycle_time: 1.e-05\n"
             "nb_particles_per_cycle: 500\nnthread: 1\nstatistics_cycle_time: 1.e-1\n");
  fclose(f);
  mcgpu::MCMPIOptions o = mcgpu::options_from_config(path, 1);
  int bad = 0;
  bad += o.nb_cells != 1000;
  bad += o.nb_particles != 100000;
  bad += o.x_ini != 7.071067690849304199e-01;
  bad += o.nb_particles_per_cycle != 500;
  bad += o.particle_min_weight != 9.999999960041972003e-13;
  bad += o.statistics_cycle_time != 0.1;
  mcgpu::YamlLoader y(path);
  try {
    (void)y.load_int("buffer_size");
    ++bad;
  } catch (const std::out_of_range &) {
  }
  try {
    mcgpu::YamlLoader missing("/nonexistent/config.yaml");
    ++bad;
  } catch (const mcgpu::Error &) {
  }
  f = fopen(path, "w");
  fprintf(f, "no colon here\n");
  fclose(f);
  try {
    mcgpu::YamlLoader broken(path);
    ++bad;
  } catch (const mcgpu::Error &) {
  }
  remove(path);
  printf(bad ? "FAIL\n" : "all ok\n");
  return bad ? 1 : 0;
}
