/*
 * kernels_api.c -- the reference's exported kernel symbols (bksp_kernels.h) and the
 * remaining bksp.h helpers, on top of the device layer (bksp_hip.h).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "../../../include/bksp.h"
#include "../../../include/bksp_hip.h"
#include "../../../include/bksp_kernels.h"

/* One SpMV per column: C[:, i] = alpha * op(A) B[:, i] + beta * C[:, i], with the
 * reference's summation order (bkh_matrix_create orders the rows that way). */
static int csr_apply(char op, int triangle, double alpha, int m, int n, const double *A,
                     const int *rp0, const int *ci0, const double *B, double beta, double *C) {
    bkh_matrix *M = bkh_matrix_create(m, rp0, ci0, A, triangle, op == 't', 0);
    if (!M) return BKH_DEVICE_ERROR;
    double *dx = (double *)bkh_malloc(sizeof(double) * (size_t)m);
    double *dy = (double *)bkh_malloc(sizeof(double) * (size_t)m);
    int rc = (dx && dy) ? 0 : BKH_DEVICE_ERROR;
    for (int i = 0; i < n && rc == 0; i++) {
        const double *b = B + (size_t)i * m;
        double *c = C + (size_t)i * m;
        if (bkh_h2d(dx, b, sizeof(double) * m) || bkh_h2d(dy, c, sizeof(double) * m) ||
            bkh_spmv(M, op, alpha, dx, beta, dy) || bkh_d2h(c, dy, sizeof(double) * m))
            rc = BKH_DEVICE_ERROR;
    }
    bkh_free(dx);
    bkh_free(dy);
    bkh_matrix_destroy(M);
    return rc;
}

static int to_zero_based(int m, const int *Ai, const int *Aj, int **rp, int **ci) {
    int nnz = Ai[m] - 1;
    *rp = (int *)malloc(sizeof(int) * ((size_t)m + 1));
    *ci = (int *)malloc(sizeof(int) * (size_t)(nnz > 0 ? nnz : 1));
    if (!*rp || !*ci) return -1;
    for (int i = 0; i <= m; i++) (*rp)[i] = Ai[i] - 1;
    for (int k = 0; k < nnz; k++) (*ci)[k] = Aj[k] - 1;
    return 0;
}

/* reference src/matvec.c:11-53 */
int dcsrgemm0(char type, int m, int n, double *A, int *Ai, int *Aj, int nnz, double *B, double *C) {
    (void)nnz;
    if (type != 'n' && type != 't') { printf("invalid parameter 0.\n"); return -1; }
    return csr_apply(type, 0, 1.0, m, n, A, Ai, Aj, B, 0.0, C);
}

/* reference src/matvec.c:100-127: accumulates into C */
int dcsrsymm0(char type, int m, int n, double *A, int *Ai, int *Aj, int nnz, double *B, double *C) {
    (void)nnz;
    if (type != 'l') { printf("invalid parameter 0.\n"); return -1; }
    return csr_apply('n', 1, 1.0, m, n, A, Ai, Aj, B, 1.0, C);
}

/* reference src/matvec.c:130-172 (1-based) */
int dcsrgemm1(char type, double alpha, int m, int n, double *A, int *Ai, int *Aj, int nnz,
              double *B, double beta, double *C) {
    (void)nnz;
    if (type != 'n' && type != 't') { printf("invalid parameter 0.\n"); return -1; }
    int *rp, *ci;
    if (to_zero_based(m, Ai, Aj, &rp, &ci)) return BKH_DEVICE_ERROR;
    int rc = csr_apply(type, 0, alpha, m, n, A, rp, ci, B, beta, C);
    free(rp);
    free(ci);
    return rc;
}

/* reference src/matvec.c:219-246 (1-based, accumulates alpha*A*B into C) */
int dcsrsymm1(char type, double alpha, int m, int n, double *A, int *Ai, int *Aj, int nnz,
              double *B, double beta, double *C) {
    (void)nnz;
    (void)beta;
    if (type != 'n') { printf("invalid parameter 0.\n"); return -1; }
    int *rp, *ci;
    if (to_zero_based(m, Ai, Aj, &rp, &ci)) return BKH_DEVICE_ERROR;
    int rc = csr_apply('n', 1, alpha, m, n, A, rp, ci, B, 1.0, C);
    free(rp);
    free(ci);
    return rc;
}

/* reference src/norm.c:37-47: -1.0 for an unknown type */
double sp_norm_f(dspmat *A) {
    int sym;
    if (strcmp(A->type, "symmetric") == 0) sym = 1;
    else if (strcmp(A->type, "unsymmetric") == 0) sym = 0;
    else return -1.0;
    double out = 0.0;
    if (bkh_sp_norm_f(*A->row_size, A->row, A->col, A->value, sym, &out)) return -1.0;
    return out;
}

/* reference src/myqr.c:4-39 on the GPU: Householder QR in LAPACK's convention
 * (dgeqr2 + dorg2r, what dgeqrf + dorgqr run below 32 columns), so Q and R carry
 * LAPACK's signs; returns dorgqr's info (0, or -1/-2 for bad m/n). */
int myqr(int m, int n, double *mat, double *r) {
    if (m < 0) return -1;
    if (n < 0 || n > m) return -2;
    if (n == 0) return 0;
    double *d = (double *)bkh_malloc(sizeof(double) * (size_t)m * n);
    if (!d) return BKH_DEVICE_ERROR;
    int rc = bkh_h2d(d, mat, sizeof(double) * (size_t)m * n);
    if (!rc) rc = bkh_qr_householder(m, n, d, r);
    if (!rc) rc = bkh_d2h(mat, d, sizeof(double) * (size_t)m * n);
    bkh_free(d);
    return rc;
}

/* reference src/util_shift.c:5-27 (scalar helper; the solver runs its own copy on
 * the device inside the shifted MR step). */
int dhorner(double *gamma, int l, double sigma, double *out) {
    out[l] = -gamma[l - 1];
    for (int i = l - 1; i > 0; i--) out[i] = -sigma * out[i + 1] - gamma[i - 1];
    out[0] = -sigma * out[1] + 1.0;
    for (int i = 1; i < l; i++)
        for (int j = l - 1; j >= i; j--) out[j] += -sigma * out[j + 1];
    for (int i = 1; i <= l; i++) out[i] = -out[i] / out[0];
    return 0;
}

/* reference src/util_shift.c:55-63 */
int comb(int n, int r) {
    if (n < 0 || r < 0 || n < r) return 0;
    if (n == r || !r) return 1;
    long long c = 1;
    if (r > n - r) r = n - r;
    for (int i = 1; i <= r; i++) c = c * (n - r + i) / i;
    return (int)c;
}
