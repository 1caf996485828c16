/* yaml_loader.cpp -- see yaml_loader.hpp (src/yaml_loader.cpp:4-40). */
#include "yaml_loader.hpp"

#include <fstream>

#include "layer.hpp"

namespace mcgpu {

YamlLoader::YamlLoader(const std::string &filepath) {
  std::ifstream in(filepath);
  if (!in) throw Error(MCGPU_ERR_ARG, "Couldn't open Yaml File " + filepath + ".");
  std::string s;
  while (std::getline(in, s)) {
    const size_t n = s.find(':');
    if (n == std::string::npos)
      throw Error(MCGPU_ERR_ARG, "Yaml File " + filepath + " was not correctly formatted.");
    variables.emplace(s.substr(0, n), s.substr(n + 1));
  }
}

int YamlLoader::load_int(const std::string &variable) const {
  return std::stoi(variables.at(variable));
}

double YamlLoader::load_double(const std::string &variable) const {
  return std::stod(variables.at(variable));
}

MCMPIOptions options_from_config(const std::string &filepath, int world_size) {
  YamlLoader y(filepath);
  MCMPIOptions opt;
  opt.world_size = world_size;
  opt.nb_cells = y.load_int("nb_cells");
  opt.x_min = y.load_double("x_min");
  opt.x_max = y.load_double("x_max");
  opt.x_ini = y.load_double("x_ini");
  opt.particle_min_weight = y.load_double("particle_min_weight");
  opt.nb_particles = y.load_int("nb_particles");
  opt.cycle_time = y.load_double("cycle_time");
  opt.nb_particles_per_cycle = y.load_int("nb_particles_per_cycle");
  opt.nthread = y.load_int("nthread");
  opt.statistics_cycle_time = y.load_double("statistics_cycle_time");
  return opt;
}

} // namespace mcgpu
