/*
 * bksp_gen.h -- seeded generators of the benchmark systems (host, C).
 *
 * No counterpart in the reference (its only inputs are the Matrix Market files
 * read by src/read_matrix.c). These build the synthetic systems the benchmarks
 * and parity tests use, identically for the GPU path and the CPU oracle.
 *
 * RNG: splitmix64(seed) seeds xoshiro256**; U[0,1) = (next() >> 11) * 2^-53;
 * N(0,1) by Box-Muller on (1 - U1, U2), both outputs used.
 * All matrices are 0-based CSR, int32 indices, ascending columns per row.
 */
#ifndef BKSP_GEN_H
#define BKSP_GEN_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int n;
    int nnz;
    int *row_ptr;
    int *col;
    double *val;
} bkh_csr_host;

void bkh_csr_host_free(bkh_csr_host *m);

/* U[lo, hi) and N(0,1) streams. */
void bkh_gen_uniform(uint64_t seed, long long n, double lo, double hi, double *out);
void bkh_gen_normal(uint64_t seed, long long n, double *out);

/* -Lap(u) + (vx, vy).grad(u) on the unit square, m x m interior grid, h = 1/(m+1),
 * 5-point centred differences, Dirichlet, lexicographic order (x fastest). */
int bkh_gen_convdiff2d(int m, double vx, double vy, bkh_csr_host *out);

/* 3D 7-point Poisson on m^3 (stencil 6, -1), Dirichlet. lower = 1 stores only
 * the lower triangle (the bksp "symmetric" form), lower = 0 the full matrix. */
int bkh_gen_poisson3d(int m, int lower, bkh_csr_host *out);

/* 3D 7-point convection-diffusion on m^3, h^2-scaled: diagonal 6, neighbour
 * -1 - p_i below and -1 + p_i above along axis i (cell Peclet numbers p). */
int bkh_gen_convdiff3d(int m, double px, double py, double pz, bkh_csr_host *out);

/* One random nonsymmetric system: n rows, each with U{5..27} distinct random
 * off-diagonal columns of N(0,1) values; diagonal = sum|a_ij| * (1 + U[0.05,0.5]). */
int bkh_gen_random_rows(uint64_t seed, int n, bkh_csr_host *out);
/* Sizes of a batch: n_i = round(2^(lo2 + (hi2 - lo2) U)). */
void bkh_gen_batch_sizes(uint64_t seed, int count, double lo2, double hi2, int *n_out);

/* kappa = exp(g) on m^3, g a Gaussian random field (squared-exponential
 * covariance, correlation length `corr` cells, standard deviation `sigma`) made by
 * separable Gaussian filtering of white noise (periodic), normalised by the
 * filter's exact variance. */
int bkh_gen_grf(uint64_t seed, int m, double sigma, double corr, double *kappa);
/* 7-point finite-volume diffusion with harmonic-mean face coefficients and
 * Dirichlet boundaries (boundary face coefficient 2*kappa), lower triangle if
 * lower = 1. */
int bkh_gen_fv_diffusion(int m, const double *kappa, int lower, bkh_csr_host *out);

#ifdef __cplusplus
}
#endif
#endif
