/*
 * yaml_loader.hpp -- the reference's flat `key: value` config reader
 * (include/yaml/yaml_loader.hpp, src/yaml_loader.cpp:4-40) and its options
 * (include/mcmpi/mcmpi_options.hpp:12-45, src/worker.cpp:315-333).
 * A missing key throws std::out_of_range (map::at), as in the reference; an
 * unreadable or malformed file throws mcgpu::Error instead of exit(1).
 */
#ifndef MCGPU_YAML_LOADER_HPP
#define MCGPU_YAML_LOADER_HPP

#include <map>
#include <string>

namespace mcgpu {

class YamlLoader {
public:
  explicit YamlLoader(const std::string &filepath);
  int load_int(const std::string &variable) const;
  double load_double(const std::string &variable) const;
  bool has(const std::string &variable) const { return variables.count(variable) != 0; }

private:
  std::map<std::string, std::string> variables;
};

/* include/mcmpi/mcmpi_options.hpp:12-45 (MPI tags and buffer size omitted) */
struct MCMPIOptions {
  int world_size;
  int nb_cells;
  double x_min, x_max, x_ini;
  double particle_min_weight;
  long long nb_particles;
  double cycle_time;
  int nb_particles_per_cycle;
  int nthread;
  double statistics_cycle_time;
};

/* src/worker.cpp:315-333 */
MCMPIOptions options_from_config(const std::string &filepath, int world_size);

} // namespace mcgpu
#endif
