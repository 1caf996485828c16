// ref_shim.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A C interface (for ctypes) over the reference's own FP32 Layer and RNG,
// compiled together with the reference sources src/layer.cpp and
// src/random.cpp where they lie (see Makefile target `ref`).  Used only to pin
// the oracle's FP32 instance against the reference itself.  This file is ours;
// it contains no reference code, only calls into it.
#include <stddef.h>
#include <stdint.h>

#include "layer.hpp"
#include "random.hpp"

extern "C" {

void *ref_layer_decompose(float x_min, float x_max, float x_ini, int world_size, int world_rank,
                          int nb_cells, int nb_particles, float particle_min_weight) {
  return new Layer(decompose_domain(x_min, x_max, x_ini, world_size, world_rank, nb_cells,
                                    nb_particles, particle_min_weight));
}

void *ref_layer_new(float x_min, float x_max, int index_start, int m, float particle_min_weight) {
  return new Layer(x_min, x_max, index_start, m, particle_min_weight);
}

void ref_layer_free(void *p) { delete static_cast<Layer *>(p); }

void ref_layer_simulate(void *p, int nb_particles, int nthread) {
  static_cast<Layer *>(p)->simulate(nb_particles, nthread);
}

void ref_layer_dump_WA(void *p) { static_cast<Layer *>(p)->dump_WA(); }

// 0 sigs, 1 absorption_rates, 2 weights_absorbed, 3 particles, 4 left, 5 right
int64_t ref_layer_array(void *p, int which, void **ptr) {
  Layer *L = static_cast<Layer *>(p);
  switch (which) {
  case 0: *ptr = L->sigs.data(); return (int64_t)L->sigs.size();
  case 1: *ptr = L->absorption_rates.data(); return (int64_t)L->absorption_rates.size();
  case 2: *ptr = L->weights_absorbed.data(); return (int64_t)L->weights_absorbed.size();
  case 3: *ptr = L->particles.data(); return (int64_t)L->particles.size();
  case 4: *ptr = L->particles_left.data(); return (int64_t)L->particles_left.size();
  case 5: *ptr = L->particles_right.data(); return (int64_t)L->particles_right.size();
  }
  *ptr = nullptr;
  return -1;
}

// 0 nb_disabled, 1 nb_active, 2 left_border, 3 right_border
int64_t ref_layer_int(void *p, int which) {
  Layer *L = static_cast<Layer *>(p);
  switch (which) {
  case 0: return L->nb_disabled;
  case 1: return L->nb_active();
  case 2: return L->left_border;
  case 3: return L->right_border;
  }
  return -1;
}

double ref_layer_dx(void *p) { return static_cast<Layer *>(p)->dx; }

void ref_layer_push(void *p, const Particle *src, int64_t n) {
  Layer *L = static_cast<Layer *>(p);
  for (int64_t i = 0; i < n; ++i) L->particles.push_back(src[i]);
}

// Transport every particle of `bank` with the reference's Layer::particle_step,
// looping on the condition of Layer::simulate_particle, and count its events;
// result[i] is the reference's simulate_particle return value for bank[i].
void ref_transport_bank(void *p, Particle *bank, int64_t n, int index_start, int m,
                        int32_t *result, uint64_t *events, float *tally) {
  Layer *L = static_cast<Layer *>(p);
  std::vector<real_t> local(L->weights_absorbed.size(), 0.0f);
  for (int64_t i = n - 1; i >= 0; --i) {
    Particle &q = bank[i];
    uint64_t ev = 0;
    while ((q.wmc >= L->particle_min_weight) && (q.index < index_start + m) &&
           (q.index >= index_start)) {
      L->particle_step(q, local);
      ++ev;
    }
    events[i] = ev;
    Particle copy = bank[i];
    copy.index = q.index;
    result[i] = L->simulate_particle(copy, local); // loop body does not run again
  }
  for (std::size_t j = 0; j < local.size(); ++j) tally[j] += local[j];
}

float ref_rnd_real(unsigned long long *seed) { return rnd_real(seed); }
unsigned long long ref_rnd_seed(unsigned long long *seed) { return rnd_seed(seed); }

} // extern "C"
