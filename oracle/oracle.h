/*
 * oracle.h -- TEST INFRASTRUCTURE ONLY: C declarations of the FP64 oracle's
 * Layer restatement (liboracle.so), for the C++ tests under tests/cpp.
 */
#ifndef ORACLE_H
#define ORACLE_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif
typedef struct orc64_Layer_s orc64_Layer;
orc64_Layer *orc64_layer_new(double x_min, double x_max, int index_start, int m, double min_w);
orc64_Layer *orc64_decompose_domain(double x_min, double x_max, double x_ini, int world_size,
                                    int world_rank, int nb_cells, int nb_particles, double min_w);
void orc64_layer_free(orc64_Layer *L);
void orc64_create_particles(orc64_Layer *L, double x_ini, double wmc, int n, uint64_t seed);
void orc64_simulate(orc64_Layer *L, int nb_particles, int nthread);
int orc64_dump_WA(const orc64_Layer *L, const char *path);
/* 0 sigs, 1 absorption_rates, 2 weights_absorbed, 3 particles, 4 left, 5 right */
int64_t orc64_layer_array(orc64_Layer *L, int which, void **ptr);
/* 0 nb_disabled, 1 nb_active, 2 left_border, 3 right_border, 7 events, 8 collisions */
int64_t orc64_layer_int(const orc64_Layer *L, int which);
void orc64_layer_set_tally_exact(orc64_Layer *L, int on);
int orc64_layer_push(orc64_Layer *L, const void *p, int64_t n);
/* src/worker_sync.cpp:24-135 over layers[0..world_size) (see orc_impl.inc) */
int64_t orc64_sync_run(orc64_Layer **layers, int world_size, int nb_particles,
                       int nb_particles_per_cycle, int nthread, int64_t max_cycles);
void orc_set_math_mode(int mode);
#ifdef __cplusplus
}
#endif
#endif
