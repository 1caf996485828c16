/*
 * layer.cpp -- mcgpu::Layer, the reference's Layer (src/layer.cpp) with the
 * transport loop replaced by the GPU kernels behind include/mcgpu.h.
 */
#include "layer.hpp"
#include "formats.hpp"

#include <math.h>
#include <stdio.h>

#include <algorithm>
#include <mutex>

#include "mc_math.h" /* correctly rounded exp for the default cross sections */

namespace mcgpu {

namespace {
std::mutex ctx_mutex;
mcgpu_context *ctx_global = nullptr;
int ctx_device = 0;

/* src/random.cpp:12-16 in FP64 */
real_t rnd_real(seed_t *seed) {
  *seed = (6364136223846793005ull * *seed + 1442695040888963407ull) % (1ull << 63);
  return (real_t)(*seed) * (1.0 / (real_t)(1ull << 63));
}
} // namespace

seed_t rnd_seed(seed_t *seed) {
  *seed = (5177284530976225183ull * *seed + 2096348467109453893ull) % (1ull << 63);
  return *seed;
}

mcgpu_context *default_context() {
  std::lock_guard<std::mutex> lock(ctx_mutex);
  if (!ctx_global) {
    int rc = mcgpu_context_create(ctx_device, &ctx_global);
    if (rc != MCGPU_OK)
      throw Error(rc, std::string("mcgpu_context_create: ") + mcgpu_status_string(rc));
  }
  return ctx_global;
}

void set_device(int device) {
  std::lock_guard<std::mutex> lock(ctx_mutex);
  ctx_device = device;
}

/* src/layer.cpp:17-42 */
Layer decompose_domain(real_t x_min, real_t x_max, real_t x_ini, int world_size, int world_rank,
                       int nb_cells, int nb_particles, real_t particle_min_weight) {
  if (!(x_max > x_min)) throw Error(MCGPU_ERR_ARG, "decompose_domain: x_max must exceed x_min");
  const int cells_per_layer = nb_cells / world_size;
  const int num_with_extra = nb_cells % world_size;
  const int nb_my_cells = cells_per_layer + (world_rank < num_with_extra);
  const int start_index = world_rank * cells_per_layer + std::min(world_rank, num_with_extra);

  const real_t dx = (x_max - x_min) / ((float)nb_cells);
  const int cell_ini = (int)((x_ini - x_min) / dx);

  Layer layer(x_min + start_index * dx, x_min + (start_index + nb_my_cells) * dx, start_index,
              nb_my_cells, particle_min_weight);
  if ((cell_ini >= start_index) && (cell_ini < start_index + nb_my_cells)) {
    seed_t seed = 5127801;
    layer.create_particles(x_ini, 1.0 / nb_particles, nb_particles, seed);
  }
  return layer;
}

/* src/layer.cpp:44-69 */
Layer::Layer(real_t x_min, real_t x_max, int index_start, int m, real_t particle_min_weight)
    : x_min(x_min), x_max(x_max), m(m), index_start(index_start), dx((x_max - x_min) / m),
      left_border(fabs(x_min) < EPS_PRECISION), right_border(fabs(x_max - 1.0) < EPS_PRECISION),
      particle_min_weight(particle_min_weight) {
  mcgpu_default_options(&options);
  last_stats = mcgpu_stats{};
  total_stats = mcgpu_stats{};
  sigs.reserve(m);
  real_t x_mid;
  for (int i = 0; i < m; ++i) {
    x_mid = x_min + (i * dx) + 0.5 * dx;
    sigs.push_back(mc_exp(-x_mid));
  }
  absorption_rates = std::vector<real_t>(m, 0.5);
  weights_absorbed = std::vector<real_t>(m, 0.0);
  particles.reserve(10000);
}

/* src/layer.cpp:71-82 */
void Layer::create_particles(real_t x_ini, real_t wmc, int n, seed_t seed) {
  if (x_ini > x_min && x_ini < x_max) {
    this->x_ini = x_ini;
    this->wmc = wmc;
    this->nb_particles_create = n;
    this->seed = seed;
    this->create_particles(MAX_PARTICLES_VECTOR);
  }
}

/* src/layer.cpp:84-87 */
int Layer::nb_active() const { return (int)particles.size() + nb_particles_create; }

/* src/layer.cpp:89-121 */
void Layer::create_particles(int n) {
  if (nb_particles_create <= 0) return;
  const int best_num_particles = MAX_PARTICLES_VECTOR - (int)particles.size();
  n = std::max(best_num_particles, n);
  n = std::min(nb_particles_create, n);
  nb_particles_create -= n;

  Particle particle = Particle();
  particle.x = x_ini;
  particle.wmc = wmc;
  particle.index = static_cast<int>(x_ini / dx);

  particles.reserve(particles.size() + n);
  for (int i = 0; i < n; ++i) {
    seed_t s = rnd_seed(&seed);
    particle.mu = 2 * rnd_real(&s) - 1;
    particle.seed = s;
    particles.push_back(particle);
  }
}

/* src/layer.cpp:220-237 */
void Layer::simulate_helper(int nb_particles, int nthread, bool use_gpu) {
  if (nb_particles == -1) {
    while ((particles.size() > 0) || (nb_particles_create > 0))
      simulate(MAX_PARTICLES_VECTOR, nthread, use_gpu);
  }
  while ((nb_particles > 0) && (particles.size() > 0 || nb_particles_create > 0)) {
    const int nb_particles_this_call = std::min(nb_particles, MAX_PARTICLES_VECTOR);
    simulate(nb_particles_this_call, nthread, use_gpu);
    nb_particles -= nb_particles_this_call;
  }
}

/* src/layer.cpp:239-361, transport through mcgpu_simulate */
void Layer::simulate(int nb_particles, int nthread, bool use_gpu) {
  if ((nb_particles == -1) || (nb_particles > MAX_PARTICLES_VECTOR))
    simulate_helper(nb_particles, nthread, use_gpu);

  if (((int)particles.size() < nb_particles) && nb_particles_create > 0)
    create_particles(nb_particles);

  nb_particles = std::min((int)particles.size(), nb_particles);
  if (nb_particles <= 0) {
    last_stats = mcgpu_stats{};
    return;
  }

  const int particles_size = (int)particles.size();
  Particle *bank = particles.data() + (particles_size - nb_particles);
  std::vector<int32_t> result(nb_particles);
  const int rc = mcgpu_simulate(default_context(), nb_particles, bank, sigs.data(),
                                absorption_rates.data(), weights_absorbed.data(), index_start,
                                index_start + m, dx, particle_min_weight, result.data(), &options,
                                &last_stats);
  if (rc != MCGPU_OK)
    throw Error(rc, std::string("mcgpu_simulate: ") + mcgpu_status_string(rc) + ": " +
                        mcgpu_context_error(default_context()));
  total_stats.events += last_stats.events;
  total_stats.collisions += last_stats.collisions;
  total_stats.n_left += last_stats.n_left;
  total_stats.n_right += last_stats.n_right;
  total_stats.n_killed += last_stats.n_killed;
  total_stats.kernel_ms += last_stats.kernel_ms;

  /* iteration i transported particles[size-1-i] (src/layer.cpp:332-346) */
  for (int i = 0; i < nb_particles; i++) {
    const int p = nb_particles - 1 - i;
    switch (result[p]) {
    case -1: particles_left.push_back(bank[p]); break;
    case 1: particles_right.push_back(bank[p]); break;
    case 0: nb_disabled++; break;
    }
  }
  particles.resize(particles_size - nb_particles);

  if (left_border) {
    nb_disabled += (int)particles_left.size();
    particles_left.clear();
  }
  if (right_border) {
    nb_disabled += (int)particles_right.size();
    particles_right.clear();
  }
}

WireParticle to_wire(const Particle &p) {
  WireParticle w;
  w.seed = p.seed;
  w.x = (float)p.x;
  w.mu = (float)p.mu;
  w.wmc = (float)p.wmc;
  w.index = p.index;
  return w;
}

Particle from_wire(const WireParticle &w) {
  Particle p{};
  p.seed = w.seed;
  p.x = w.x;
  p.mu = w.mu;
  p.wmc = w.wmc;
  p.index = w.index;
  return p;
}

std::vector<WireParticle> to_wire(const std::vector<Particle> &ps) {
  std::vector<WireParticle> out(ps.size());
  for (size_t i = 0; i < ps.size(); ++i) out[i] = to_wire(ps[i]);
  return out;
}

std::vector<Particle> from_wire(const std::vector<WireParticle> &ws) {
  std::vector<Particle> out(ws.size());
  for (size_t i = 0; i < ws.size(); ++i) out[i] = from_wire(ws[i]);
  return out;
}

/* src/layer.cpp:363-380 */
void Layer::dump_WA() {
  FILE *file = fopen("WA.out", "w");
  if (!file) throw Error(MCGPU_ERR_ARG, "Couldn't open file WA.out for writing.");
  write_wa<real_t>(file, weights_absorbed.data(), (int)weights_absorbed.size(), x_min, dx);
  fclose(file);
}

} // namespace mcgpu
