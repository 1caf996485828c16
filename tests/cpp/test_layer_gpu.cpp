// test_layer_gpu.cpp -- mcgpu::Layer (GPU transport behind the C ABI) against
// the oracle's FP64 restatement of the reference Layer (src/layer.cpp),
// following the recipe of the reference's src/test_layer.cpp:39-95.
//
// For each scenario both Layers run the same sequence of simulate(n) calls;
// after every call particles, particles_left, particles_right (bytes),
// nb_disabled, nb_active and weights_absorbed (bits; the oracle adds each
// call's exact tally, as the GPU does) must be identical.  The FP64 WA.out of
// the golden recipe must match the oracle's text.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "layer.hpp"
#include "oracle.h"

static int failures = 0;
#define CHECK(cond, ...)                                                   \
  do {                                                                     \
    if (!(cond)) {                                                         \
      fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__);                 \
      fprintf(stderr, __VA_ARGS__);                                        \
      fprintf(stderr, "\n");                                               \
      ++failures;                                                          \
    }                                                                      \
  } while (0)

template <class T>
static bool same_bytes(const std::vector<T> &a, const void *b, int64_t nb) {
  if ((int64_t)a.size() != nb) return false;
  return nb == 0 || memcmp(a.data(), b, sizeof(T) * nb) == 0;
}

static void compare(const char *tag, mcgpu::Layer &G, orc64_Layer *O) {
  void *p;
  int64_t n;
  n = orc64_layer_array(O, 3, &p);
  CHECK(same_bytes(G.particles, p, n), "%s particles (%zu vs %lld)", tag, G.particles.size(), (long long)n);
  n = orc64_layer_array(O, 4, &p);
  CHECK(same_bytes(G.particles_left, p, n), "%s particles_left (%zu vs %lld)", tag, G.particles_left.size(), (long long)n);
  n = orc64_layer_array(O, 5, &p);
  CHECK(same_bytes(G.particles_right, p, n), "%s particles_right (%zu vs %lld)", tag, G.particles_right.size(), (long long)n);
  n = orc64_layer_array(O, 2, &p);
  CHECK(same_bytes(G.weights_absorbed, p, n), "%s weights_absorbed", tag);
  n = orc64_layer_array(O, 0, &p);
  CHECK(same_bytes(G.sigs, p, n), "%s sigs", tag);
  CHECK(G.nb_disabled == orc64_layer_int(O, 0), "%s nb_disabled %d vs %lld", tag, G.nb_disabled, (long long)orc64_layer_int(O, 0));
  CHECK(G.nb_active() == orc64_layer_int(O, 1), "%s nb_active", tag);
  CHECK(G.left_border == (bool)orc64_layer_int(O, 2) && G.right_border == (bool)orc64_layer_int(O, 3), "%s borders", tag);
}

struct Scenario {
  const char *name;
  double x_min, x_max, x_ini;
  int world_size, world_rank, nb_cells, nb_particles;
  double min_w;
  std::vector<int> calls;
  int kernel;
};

static void run(const Scenario &s) {
  mcgpu::Layer G = mcgpu::decompose_domain(s.x_min, s.x_max, s.x_ini, s.world_size, s.world_rank,
                                           s.nb_cells, s.nb_particles, s.min_w);
  G.options.kernel = s.kernel;
  orc64_Layer *O = orc64_decompose_domain(s.x_min, s.x_max, s.x_ini, s.world_size, s.world_rank,
                                          s.nb_cells, s.nb_particles, s.min_w);
  orc64_layer_set_tally_exact(O, 1);
  compare(s.name, G, O);
  for (int n : s.calls) {
    G.simulate(n);
    orc64_simulate(O, n, 8);
    char tag[128];
    snprintf(tag, sizeof tag, "%s after simulate(%d)", s.name, n);
    compare(tag, G, O);
  }
  orc64_layer_free(O);
  printf("ok %s (kernel %d)\n", s.name, s.kernel);
  fflush(stdout);
}

static std::string slurp(const char *path) {
  std::ifstream f(path);
  std::stringstream b;
  b << f.rdbuf();
  return b.str();
}

int main() {
  const double x_ini = sqrt(2.0) / 2.0;
  std::vector<Scenario> scen = {
      {"golden recipe", 0.0, 1.0, x_ini, 1, 0, 100, 100, 0.0, {-1}, 0},
      {"single layer, partial calls", 0.0, 1.0, 0.3, 1, 0, 200, 5000, 1e-3, {37, 1000, 1, -1}, 0},
      {"interior rank keeps escape buffers", 0.0, 1.0, 0.55, 3, 1, 90, 3000, 1e-3, {700, 1, 5000}, 0},
      {"rank 0 of 4 (left border)", 0.0, 1.0, 0.1, 4, 0, 400, 4000, 1e-2, {-1}, 0},
      {"more than MAX_PARTICLES_VECTOR", 0.0, 1.0, 0.9, 1, 0, 50, 1200000, 1e-1, {1100000, -1}, 0},
      {"event kernel, interior rank", 0.0, 1.0, 0.55, 3, 1, 90, 3000, 1e-3, {700, 1, 5000}, 1},
      {"event kernel, golden recipe", 0.0, 1.0, x_ini, 1, 0, 100, 100, 0.0, {-1}, 1},
  };
  try {
    for (const auto &s : scen) run(s);
    // FP64 WA.out of the golden recipe (src/test_layer.cpp:41-57) vs the oracle's
    mcgpu::Layer G = mcgpu::decompose_domain(0.0, 1.0, x_ini, 1, 0, 100, 100, 0.0);
    G.simulate(-1);
    G.dump_WA();
    const std::string mine = slurp("WA.out");
    orc64_Layer *O = orc64_decompose_domain(0.0, 1.0, x_ini, 1, 0, 100, 100, 0.0);
    orc64_layer_set_tally_exact(O, 1);
    orc64_simulate(O, -1, 1);
    orc64_dump_WA(O, "WA_oracle.out");
    orc64_layer_free(O);
    CHECK(mine == slurp("WA_oracle.out") && !mine.empty(), "WA.out differs from the oracle");
    remove("WA.out");
    remove("WA_oracle.out");
  } catch (const std::exception &e) {
    fprintf(stderr, "exception: %s\n", e.what());
    return 1;
  }
  if (failures) {
    fprintf(stderr, "%d failures\n", failures);
    return 1;
  }
  printf("all ok\n");
  return 0;
}
