/*
 * bksp_kernels.h -- the reference library's exported kernel symbols, for callers
 * that use them directly (the reference's tests/test.c calls sp_norm_f and, through
 * include/bksp_internal.h's inline wrappers, the CSR kernels).
 *
 *   dcsrgemm0 / dcsrsymm0 / dcsrgemm1 / dcsrsymm1   reference include/bksp_internal.h:19-26
 *                                                    (defined in src/matvec.c:11-246)
 *   sp_norm_f                                        reference src/norm.c:37-47 (undeclared there)
 *   dhorner / comb                                   reference include/bksp_internal.h:132-135
 *   zcsrgemm0 / zcsrgemm1 / zhorner                  complex counterparts (src/matvec.c:55-98,
 *                                                    174-217; src/util_shift.c:29-53)
 *
 * The CSR kernels take host arrays, run on the GPU (upload, CSR-stream SpMV per
 * column, download) and reproduce the reference's per-entry summation order.
 * B and C are column-major with leading dimension m. 0-based (…0) or 1-based (…1)
 * indices as in the reference. They return 0, or -1 with "invalid parameter 0."
 * printed for an unknown `type`, like the reference; -5 on a device error.
 */
#ifndef BKSP_KERNELS_H
#define BKSP_KERNELS_H
#include "bksp.h"

#ifdef __cplusplus
extern "C" {
#endif

int dcsrgemm0(char type, int m, int n, double *A, int *Ai, int *Aj, int nnz, double *B, double *C);
int dcsrsymm0(char type, int m, int n, double *A, int *Ai, int *Aj, int nnz, double *B, double *C);
int dcsrgemm1(char type, double alpha, int m, int n, double *A, int *Ai, int *Aj, int nnz,
              double *B, double beta, double *C);
int dcsrsymm1(char type, double alpha, int m, int n, double *A, int *Ai, int *Aj, int nnz,
              double *B, double beta, double *C);
double sp_norm_f(dspmat *A);
int dhorner(double *gamma, int l, double sigma, double *out);
int comb(int n, int r);
#ifndef __cplusplus
/* complex kernels, reference include/bksp_internal.h:28-33 and :133-134 */
int zcsrgemm0(char type, int m, int n, double _Complex *A, int *Ai, int *Aj, int nnz,
              double _Complex *B, double _Complex *C);
int zcsrgemm1(char type, double _Complex *alpha, int m, int n, double _Complex *A, int *Ai, int *Aj,
              int nnz, double _Complex *B, double _Complex *beta, double _Complex *C);
int zhorner(double _Complex *gamma, int l, double _Complex sigma, double _Complex *out);
#endif

#ifdef __cplusplus
}
#endif
#endif
