/*
 * gnx_rng.h -- host random-number stream of the reference (random.c over
 * dSFMT-132049), for callers that draw fields, momenta and noise in the
 * reference's order without linking the reference itself.
 *
 * The reference keeps the RNG on the host (random.c:9, one global dSFMT
 * state); so does this library: these are host functions over host memory.
 * A stream created with seed s yields the values the reference yields after
 * Initrand(s) (random.c:52-66, dsfmt_init_gen_rand): the same uniforms bit
 * for bit, and the same Gaussians where libm's log and sqrt are the same.
 */
#ifndef GNX_RNG_H
#define GNX_RNG_H

#include "gnx.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gnx_rng gnx_rng;

/* Initrand(seed) (random.c:52-66); seed 0 is, as there, replaced by a
 * clock-based one */
gnx_status gnx_rng_create(unsigned int seed, gnx_rng **out);
gnx_status gnx_rng_destroy(gnx_rng *r);

/* Casuale: uniform in (0, 1), dsfmt_genrand_open_open (random.c:28-35) */
double gnx_rng_casuale(gnx_rng *r);
/* Gauss1: polar method, first variate only (random.c:88-102) */
double gnx_rng_gauss1(gnx_rng *r);
/* Gauss2MeanVariance (random.c:129-143) */
void gnx_rng_gauss2mv(gnx_rng *r, double mean, double variance, double *r1, double *r2);

/* Bulk draws into host memory, in the reference's order:
 *   GNX_DRAW_CASUALE     out[n]: n Casuale
 *   GNX_DRAW_GAUSS1      out[n]: n Gauss1 (momenta, moleculardynamics.c:781-786;
 *                        real pseudo-fermion sources, fermion.c:337-345)
 *   GNX_DRAW_SIGMA       out[n], n even: Gauss2MeanVariance(0, var) pairs into
 *                        (x, x + n/2) for x < n/2 (sigma.c:29-30)
 *   GNX_DRAW_Z2_COMPLEX  out[2n]: per site re then im, each Casuale() <= 0.5 ? 1 : -1
 *   GNX_DRAW_Z2_REAL     out[n]: per site Casuale() <= 0.5 ? 1 : -1 (condensate.c:52-60)
 *   GNX_DRAW_GAUSS2_PAIRS out[n], n even: Gauss2MeanVariance(0, var) pairs into
 *                        (2r, 2r + 1) (link momenta, moleculardynamics.c:44)
 */
typedef enum gnx_draw {
    GNX_DRAW_CASUALE = 0,
    GNX_DRAW_GAUSS1 = 1,
    GNX_DRAW_SIGMA = 2,
    GNX_DRAW_Z2_COMPLEX = 3,
    GNX_DRAW_Z2_REAL = 4,
    GNX_DRAW_GAUSS2_PAIRS = 5
} gnx_draw;
gnx_status gnx_rng_fill(gnx_rng *r, gnx_draw kind, long n, double var, double *out);

#ifdef __cplusplus
}
#endif
#endif /* GNX_RNG_H */
