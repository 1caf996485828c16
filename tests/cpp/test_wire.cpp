// test_wire.cpp -- the reference's 24-byte FP32 Particle (wire layout of its
// MPI datatype) and the FP64 <-> FP32 conversions of the C++ host side.
// Writes the wire bytes of the reference's serialization literal
// (src/test_async_serialization.cpp:29-31) to the file named by argv[1], for
// the Python side to compare with numpy's packing of the same record.
#include <stdio.h>
#include <string.h>

#include "layer.hpp"

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  int failures = 0;
  mcgpu::WireParticle w{5127801, 0.172634512365276f, 123.12342351f, -1231.123486123f, 2134512};
  const mcgpu::Particle p = mcgpu::from_wire(w);
  const mcgpu::WireParticle back = mcgpu::to_wire(p);
  if (memcmp(&w, &back, sizeof w) != 0) {
    fprintf(stderr, "FAIL: FP32 -> FP64 -> FP32 is not the identity\n");
    ++failures;
  }
  if (p.x != (double)w.x || p.seed != w.seed || p.index != w.index) {
    fprintf(stderr, "FAIL: widening is not exact\n");
    ++failures;
  }
  mcgpu::Particle q{};
  q.seed = 42;
  q.x = 0.1;
  q.mu = -1.0 / 3.0;
  q.wmc = 1e-3;
  q.index = 7;
  const mcgpu::WireParticle wq = mcgpu::to_wire(q);
  if (wq.x != 0.1f || wq.mu != (float)(-1.0 / 3.0) || wq.wmc != 1e-3f || wq.index != 7) {
    fprintf(stderr, "FAIL: narrowing is not round to nearest\n");
    ++failures;
  }
  std::vector<mcgpu::Particle> v(3, q);
  if (mcgpu::from_wire(mcgpu::to_wire(v)).size() != 3) ++failures;
  FILE *f = fopen(argv[1], "wb");
  if (!f || fwrite(&w, sizeof w, 1, f) != 1) return 1;
  fclose(f);
  if (failures) return 1;
  printf("all ok\n");
  return 0;
}
