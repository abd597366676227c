/*
 * axx_oracle.h — CPU oracle (test infrastructure only; see axx_oracle.c).
 * Uses the descriptor structs of include/axonexa.h so that a test hands the
 * same inputs to the oracle and to the product library.
 */
#ifndef AXX_ORACLE_H
#define AXX_ORACLE_H
#include <stdint.h>
#include "../include/axonexa.h"
#ifdef __cplusplus
extern "C" {
#endif
void  orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
float orc_unif(uint32_t w);
/* -2 ln u(w[k]) of the Gaussian step generator, k < n */
void orc_neg2log_unif(const uint32_t* w, float* out, int64_t n);
void  orc_sincos_turn(uint32_t w, float* s, float* c);
void  orc_sincos_turn_n(const uint32_t* w, float* s, float* c, int64_t n);
void  orc_sincos_u24(uint32_t v, float* s, float* c);
void  orc_azimuth_table(float* c, float* s);
int   orc_validate(const axx_substrate* s, const axx_acq* q);
int   orc_axes_mask(const axx_acq* q);
/* bin_width <= 0: reference-faithful linear scan over all primitives
 * (lattice.cuh:184-199); > 0: the same walk through cell bins. */
int   orc_walk(const axx_substrate* s, const axx_acq* q, const axx_run* run, int step_begin,
               int step_end, uint32_t* state, int nthreads, double bin_width);
int   orc_reduce(const axx_substrate* s, const axx_acq* q, const axx_run* run,
                 const uint32_t* state, axx_result* out);
/* orc_walk that also counts boundary events at circles and spheres:
 * ev += {from inside, transmitted of those, from outside, transmitted}. */
int   orc_walk_events(const axx_substrate* s, const axx_acq* q, const axx_run* run, int step_begin,
                      int step_end, uint32_t* state, int nthreads, double bin_width, int64_t ev[4]);
/* Per-walker T2 exponent (the T2factor of kernelLattice.cuh:101-103, negated)
 * of a finished walk. */
int   orc_exponents(const axx_substrate* s, const axx_acq* q, const axx_run* run, const uint32_t* state,
                    double* out);
/* Magnetization walk of one group of fields (state: AXX_F_MOM + 6 F words
 * per walker) and its sums; the double-precision sine/cosine and square
 * root it uses, elementwise. */
int   orc_walk_mag(const axx_substrate* s, const axx_field* f, const axx_run* run, int group,
                   int step_begin, int step_end, uint32_t* state, int nthreads, double bin_width);
void  orc_reduce_mag(const uint32_t* state, int64_t n_walkers, int F, double* out);
void  orc_sincos64(const double* x, double* s, double* c, int64_t n);
void  orc_sqrt_cr(const double* x, double* y, int64_t n);
int   orc_image_count(const axx_substrate* s, const axx_acq* q);
/* 1 (default): a substrate whose primitives repeat along x and y is walked in
 * its unit cell, as the product does; 0: always in the whole cell. */
void  orc_set_unit_cells(int on);
int   orc_unit_cells(const axx_substrate* s, const axx_acq* q, const axx_run* run);
int   orc_hit_test(int type, float cx, float cy, float cz, float R, int inside, const float s3[3],
                   const float u3[3], float ell, float* t);
#ifdef __cplusplus
}
#endif
#endif
