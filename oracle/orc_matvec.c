/*
 * orc_matvec.c -- test infrastructure only (see oracle.h).
 *
 * Restates the reference's CSR kernels (src/matvec.c) loop for loop, including the
 * summation order (sequential per output entry, multiply then add, no FMA), and the
 * MKL-named dispatch wrappers of include/bksp_internal.h:35-79.
 * B and C are column-major with leading dimension m (matvec.c:21-24).
 */
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* matvec.c:11-53 -- 0-based general CSR: 'n' gathers C = A*B, 't' scatters C = A^T*B. */
int orc_dcsrgemm0(char type, int m, int n, const double *A, const int *Ai,
                  const int *Aj, const double *B, double *C) {
    if (type == 'n') {
        for (int i = 0; i < n; i++)
            for (int j = 0; j < m; j++) {
                double *c = &C[(size_t)i * m + j];
                *c = 0.0;
                for (int k = Ai[j]; k < Ai[j + 1]; k++)
                    *c = *c + A[k] * B[(size_t)i * m + Aj[k]];
            }
    } else if (type == 't') {
        for (int i = 0; i < n; i++)
            for (int j = 0; j < m; j++) C[(size_t)i * m + j] = 0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < m; j++)
                for (int k = Ai[j]; k < Ai[j + 1]; k++)
                    C[(size_t)i * m + Aj[k]] = C[(size_t)i * m + Aj[k]] + A[k] * B[(size_t)i * m + j];
    } else {
        return -1;
    }
    return 0;
}

/* matvec.c:100-127 -- one stored triangle: gather the stored row, scatter the mirror.
 * C is accumulated into (the caller zeroes it, bksp_internal.h:39). */
int orc_dcsrsymm0(char type, int m, int n, const double *A, const int *Ai,
                  const int *Aj, const double *B, double *C) {
    if (type != 'l') return -1;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < m; j++)
            for (int k = Ai[j]; k < Ai[j + 1]; k++) {
                if (Aj[k] != j)
                    C[(size_t)i * m + Aj[k]] = C[(size_t)i * m + Aj[k]] + A[k] * B[(size_t)i * m + j];
                C[(size_t)i * m + j] = C[(size_t)i * m + j] + A[k] * B[(size_t)i * m + Aj[k]];
            }
    return 0;
}

/* matvec.c:130-172 -- 1-based general CSR, C = beta*C + alpha*A*B (or A^T*B). */
int orc_dcsrgemm1(char type, double alpha, int m, int n, const double *A,
                  const int *Ai, const int *Aj, const double *B, double beta,
                  double *C) {
    if (type == 'n') {
        for (int i = 0; i < n; i++)
            for (int j = 0; j < m; j++) {
                double *c = &C[(size_t)i * m + j];
                *c = beta * *c;
                for (int k = Ai[j] - 1; k < Ai[j + 1] - 1; k++)
                    *c = *c + alpha * A[k] * B[(size_t)i * m + Aj[k] - 1];
            }
    } else if (type == 't') {
        for (int i = 0; i < n; i++)
            for (int j = 0; j < m; j++) C[(size_t)i * m + j] = beta * C[(size_t)i * m + j];
        for (int i = 0; i < n; i++)
            for (int j = 0; j < m; j++)
                for (int k = Ai[j] - 1; k < Ai[j + 1] - 1; k++)
                    C[(size_t)i * m + Aj[k] - 1] = C[(size_t)i * m + Aj[k] - 1] + alpha * A[k] * B[(size_t)i * m + j];
    } else {
        return -1;
    }
    return 0;
}

/* matvec.c:219-246 -- 1-based stored triangle, accumulates alpha*A*B into C
 * (the beta scaling is done by the caller, bksp_internal.h:73). */
int orc_dcsrsymm1(char type, double alpha, int m, int n, const double *A,
                  const int *Ai, const int *Aj, const double *B, double beta,
                  double *C) {
    (void)beta;
    if (type != 'n') return -1;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < m; j++)
            for (int k = Ai[j] - 1; k < Ai[j + 1] - 1; k++) {
                if (Aj[k] - 1 != j)
                    C[(size_t)i * m + Aj[k] - 1] = C[(size_t)i * m + Aj[k] - 1] + alpha * A[k] * B[(size_t)i * m + j];
                C[(size_t)i * m + j] = C[(size_t)i * m + j] + alpha * A[k] * B[(size_t)i * m + Aj[k] - 1];
            }
    return 0;
}

/* bksp_internal.h:35-43: y = 0, then the triangle kernel. */
void orc_csrsymv(int m, const double *a, const int *ia, const int *ja,
                 const double *x, double *y) {
    orc_dscal(m, 0.0, y);
    orc_dcsrsymm0('l', m, 1, a, ia, ja, x, y);
}

/* bksp_internal.h:45-52. */
void orc_csrgemv(char trans, int m, const double *a, const int *ia,
                 const int *ja, const double *x, double *y) {
    orc_dcsrgemm0(trans, m, 1, a, ia, ja, x, y);
}

/* bksp_internal.h:63-79, called with the 1-based copies the solvers build
 * (blcg.c:42-49, blbicgstabl.c:51-58). kind 'G' or 'S'. */
void orc_csrmm(char kind, int m, int n, double alpha, const double *val,
               const int *row_ptr0, const int *col0, const double *b,
               double beta, double *c) {
    int nnz = row_ptr0[m];
    int *rp = (int *)malloc((size_t)(m + 1) * sizeof(int));
    int *ci = (int *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int));
    for (int i = 0; i <= m; i++) rp[i] = row_ptr0[i] + 1;
    for (int i = 0; i < nnz; i++) ci[i] = col0[i] + 1;
    if (kind == 'G') {
        orc_dcsrgemm1('n', alpha, m, n, val, rp, ci, b, beta, c);
    } else {
        orc_dscal(m * n, beta, c);
        orc_dcsrsymm1('n', alpha, m, n, val, rp, ci, b, beta, c);
    }
    free(ci);
    free(rp);
}
