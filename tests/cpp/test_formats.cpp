// test_formats.cpp -- the output-file writers of the host library
// (host/formats.hpp), CPU only.  Writes, into <outdir>, the files the
// reference's program writes, from the inputs the Python test gives it:
//   weights.csv    write_weights_csv<float> of weights.bin (float32, every
//                  rank's tallies in rank order) with rank 0's dx
//   WA.out         write_wa<float> of rank 0's cells
//   config.yaml    write_config_yaml<float>  (the reference's FP32 options)
//   config64.yaml  write_config_yaml<double> (this library's FP64 options)
//   stats.csv      write_stats_csv of one synthetic window per rank
// The Python side compares them with the reference program's own files.
//
//   test_formats config.yaml world_size weights.bin displs(comma list)
//                dx(hex float) x_min(hex float) m0 outdir
#include <stdio.h>
#include <stdlib.h>

#include <string>
#include <vector>

#include "formats.hpp"
#include "yaml_loader.hpp"

static FILE *open_out(const std::string &dir, const char *name) {
  const std::string p = dir + "/" + name;
  FILE *f = fopen(p.c_str(), "w");
  if (!f) {
    fprintf(stderr, "cannot write %s\n", p.c_str());
    exit(1);
  }
  return f;
}

int main(int argc, char **argv) {
  if (argc != 9) {
    fprintf(stderr, "usage: %s config.yaml world_size weights.bin displs dx x_min m0 outdir\n",
            argv[0]);
    return 2;
  }
  const int world_size = atoi(argv[2]);
  const mcgpu::MCMPIOptions opt = mcgpu::options_from_config(argv[1], world_size);

  std::vector<float> w;
  FILE *in = fopen(argv[3], "rb");
  if (!in) return 1;
  float v;
  while (fread(&v, sizeof v, 1, in) == 1) w.push_back(v);
  fclose(in);

  std::vector<int> displs;
  for (const char *p = argv[4]; *p;) {
    char *end;
    displs.push_back((int)strtol(p, &end, 10));
    p = *end == ',' ? end + 1 : end;
  }
  if ((int)displs.size() != world_size) return 1;
  const float dx = strtof(argv[5], nullptr);
  const float x_min = strtof(argv[6], nullptr);
  const int m0 = atoi(argv[7]);
  const std::string out = argv[8];

  FILE *f = open_out(out, "weights.csv");
  mcgpu::write_weights_csv<float>(f, w.data(), (int)w.size(), displs.data(), world_size, dx);
  fclose(f);
  f = open_out(out, "WA.out");
  mcgpu::write_wa<float>(f, w.data(), m0, x_min, dx);
  fclose(f);
  f = open_out(out, "config.yaml");
  mcgpu::write_config_yaml<float>(f, opt, "HOSTNAME");
  fclose(f);
  f = open_out(out, "config64.yaml");
  mcgpu::write_config_yaml<double>(f, opt, "HOSTNAME");
  fclose(f);

  std::vector<std::vector<mcgpu::TimerWindow>> win(world_size);
  for (int r = 0; r < world_size; ++r) {
    mcgpu::TimerWindow t;
    t.starttime = 1.5e9 + r;
    t.endtime = t.starttime + 0.25;
    t.time_comp = 0.125;
    t.time_send = 0.0625;
    t.time_recv = 0.0625;
    t.time_idle = 0.0;
    t.nb_cycles = 10 + r;
    win[r].push_back(t);
  }
  f = open_out(out, "stats.csv");
  mcgpu::write_stats_csv(f, win);
  fclose(f);
  std::vector<std::vector<mcgpu::TimerWindow>> none(world_size);
  f = open_out(out, "stats_empty.csv");
  mcgpu::write_stats_csv(f, none);
  fclose(f);
  printf("all ok\n");
  return 0;
}
