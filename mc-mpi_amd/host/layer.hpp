/*
 * layer.hpp -- mcgpu::Layer: the reference's Layer (include/layer/layer.hpp
 * :14-114) with real_t = double, whose simulate() transports on the GPU
 * through the C ABI of include/mcgpu.h.
 *
 * Same members, methods, argument meaning and post-conditions as the
 * reference (src/layer.cpp:17-380): lazy particle creation in chunks of
 * MAX_PARTICLES_VECTOR, the last-n particles transported per call, escape
 * buffers filled in iteration order, nb_disabled, clearing of escape buffers
 * at the global borders, WA.out format.  Differences:
 *  - real_t is double (the reference only builds with float);
 *  - simulate() always runs on the GPU; `nthread` and `use_gpu` are accepted
 *    for source compatibility and ignored (there is no CPU transport here);
 *  - failures throw mcgpu::Error instead of calling exit();
 *  - particle_step / simulate_particle (per-particle CPU steps) are not
 *    provided: the GPU kernel replaces the whole loop they form.
 */
#ifndef MCGPU_LAYER_HPP
#define MCGPU_LAYER_HPP

#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mcgpu.h"

namespace mcgpu {

typedef unsigned long long seed_t;  /* include/types/types.hpp:6 */
typedef double real_t;              /* FP64 restatement of include/types/types.hpp:14 */
typedef mcgpu_particle Particle;    /* include/types/particle.hpp:7-17 */

constexpr int MAX_PARTICLES_VECTOR = 1000000; /* src/layer.cpp:15 */
constexpr real_t EPS_PRECISION = 1e-4;        /* include/types/types.hpp:18 */

class Error : public std::runtime_error {
public:
  Error(int status, const std::string &what) : std::runtime_error(what), status(status) {}
  int status;
};

/* The process-wide GPU context Layer::simulate uses (device 0 unless
   set_device is called before the first simulate). */
mcgpu_context *default_context();
void set_device(int device);

class Layer {
public:
  Layer(real_t x_min, real_t x_max, int index_start, int m, real_t particle_min_weight);

  void create_particles(real_t x_ini, real_t wmc, int n, seed_t seed);
  void simulate(int nb_particles, int nthread = -1, bool use_gpu = true);
  void dump_WA();
  int nb_active() const;

  void create_particles(int n);
  void simulate_helper(int nb_particles, int nthread, bool use_gpu);

  /* options of the GPU transport (kernel choice, launch shape) */
  mcgpu_options options;
  /* device-counted totals of the last transport (mcgpu_simulate call) made
     by simulate(); zero when the last simulate() transported nothing */
  mcgpu_stats last_stats;
  /* the same, summed over every transport of this Layer */
  mcgpu_stats total_stats;

  // -- Data --
  const real_t x_min, x_max;

private:
  const int m;
  const int index_start;

public:
  std::vector<real_t> weights_absorbed;
  std::vector<Particle> particles;
  std::vector<Particle> particles_left;
  std::vector<Particle> particles_right;
  const real_t dx;
  int nb_disabled = 0;

  const bool left_border, right_border;

public:
  std::vector<real_t> sigs;
  std::vector<real_t> absorption_rates;
  const real_t particle_min_weight;

  seed_t seed = 0;
  real_t x_ini = 0, wmc = 0;
  int nb_particles_create = 0;

  int get_m() const { return m; }
  int get_index_start() const { return index_start; }
};

Layer decompose_domain(real_t x_min, real_t x_max, real_t x_ini, int world_size, int world_rank,
                       int nb_cells, int nb_particles, real_t particle_min_weight);

/* The reference's Particle as it is built (real_t = float,
   include/types/particle.hpp:7-17, include/types/types.hpp:12-17): 24 bytes,
   the layout its MPI datatype describes (src/worker_sync.cpp:7-21) and its
   communicators copy byte for byte.  For exchanging particles with ranks
   that run the reference. */
struct WireParticle {
  seed_t seed;
  float x;
  float mu;
  float wmc;
  int index;
};
static_assert(sizeof(WireParticle) == 24, "reference Particle is 24 bytes");
static_assert(offsetof(WireParticle, seed) == 0 && offsetof(WireParticle, x) == 8 &&
                  offsetof(WireParticle, mu) == 12 && offsetof(WireParticle, wmc) == 16 &&
                  offsetof(WireParticle, index) == 20,
              "field offsets of the reference's MPI particle datatype");

/* FP64 -> FP32 (x, mu, wmc rounded to nearest), seed and index unchanged */
WireParticle to_wire(const Particle &p);
/* FP32 -> FP64, exact */
Particle from_wire(const WireParticle &w);
std::vector<WireParticle> to_wire(const std::vector<Particle> &ps);
std::vector<Particle> from_wire(const std::vector<WireParticle> &ws);

/* src/random.cpp:24-29 */
seed_t rnd_seed(seed_t *seed);

} // namespace mcgpu

#endif
