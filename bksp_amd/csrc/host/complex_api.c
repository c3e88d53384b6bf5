/*
 * complex_api.c -- the complex entry points of bksp.h and the complex kernel symbols
 * (C host code over the device layer's bkh_z* C-ABI).
 *
 *   zbicg          reference src/bicg.c:91-177
 *   zbicgstabl     reference src/bicgstabl.c:141-276
 *   zshbicgstabl   reference src/shbicgstabl.c:288-592
 *   zblbicgstabl   reference src/blbicgstabl.c:255-486 (prints its summary, :461-469)
 *   zqreco         reference src/myqr.c:41-76 (Householder QR on the GPU, zgeqr2 + zung2r)
 *   zmat_create / zmat_free   src/util.c:72-88
 *   zhorner        src/util_shift.c:29-53 (host, scalar)
 *   zcsrgemm0 / zcsrgemm1     src/matvec.c:55-98, 174-217 (on the GPU)
 *
 * Like the real entry points: x holds the initial guess and receives the solution,
 * the return value is the 0-based break index or a bksp.h code, and a device
 * failure returns BKH_DEVICE_ERROR (-5). The reference never checks the matrix
 * type of these solvers; neither does this layer.
 */
#include <complex.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "../../../include/bksp.h"
#include "../../../include/bksp_hip.h"
#include "../../../include/bksp_kernels.h"

typedef struct {
    bkh_zmatrix *M;
    double *d_b, *d_x;
} zupload_t;

static int zupload(zspmat *A, const double _Complex *b, size_t nb, const double _Complex *x, size_t nx,
                   zupload_t *u) {
    memset(u, 0, sizeof *u);
    u->M = bkh_zmatrix_create(*A->row_size, A->row, A->col, (const double *)A->value);
    if (!u->M) return BKH_DEVICE_ERROR;
    u->d_b = (double *)bkh_malloc(sizeof(double _Complex) * nb);
    u->d_x = (double *)bkh_malloc(sizeof(double _Complex) * nx);
    if (!u->d_b || !u->d_x) return BKH_DEVICE_ERROR;
    if (bkh_h2d(u->d_b, b, sizeof(double _Complex) * nb) || bkh_h2d(u->d_x, x, sizeof(double _Complex) * nx))
        return BKH_DEVICE_ERROR;
    return 0;
}

static void zrelease(zupload_t *u) {
    bkh_free(u->d_b);
    bkh_free(u->d_x);
    if (u->M) bkh_zmatrix_destroy(u->M);
}

int zbicg(zspmat *A, double _Complex *b, double tol, int max_iter, double _Complex *x) {
    if (max_iter <= 0) return 0;                /* no iteration: the error test sees 0 */
    int n = *A->row_size;
    zupload_t u;
    int ret = zupload(A, b, n, x, n, &u);
    if (ret == 0) {
        ret = bkh_zbicg(u.M, u.d_b, u.d_x, tol, max_iter, NULL);
        if (ret > BKH_DEVICE_ERROR && bkh_d2h(x, u.d_x, sizeof(double _Complex) * n)) ret = BKH_DEVICE_ERROR;
    }
    zrelease(&u);
    return ret;
}

int zbicgstabl(zspmat *A, double _Complex *b, int ell, double tol, int max_iter, double _Complex *x) {
    if (max_iter <= 0) return 0;                /* bicgstabl.c:145, 271-275 */
    int n = *A->row_size;
    zupload_t u;
    int ret = zupload(A, b, n, x, n, &u);
    if (ret == 0) {
        ret = bkh_zbicgstabl(u.M, u.d_b, ell, u.d_x, tol, max_iter, NULL);
        if (ret > BKH_DEVICE_ERROR && bkh_d2h(x, u.d_x, sizeof(double _Complex) * n)) ret = BKH_DEVICE_ERROR;
    }
    zrelease(&u);
    return ret;
}

int zshbicgstabl(zspmat *A, double _Complex *b, int l, double _Complex *sigma, int sigma_size, double tol,
                 int max_iter, double _Complex *x) {
    if (max_iter <= 0) return 0;                /* shbicgstabl.c:295, 586-591 */
    int n = *A->row_size;
    size_t nx = (size_t)n * ((size_t)sigma_size + 1);
    zupload_t u;
    int ret = zupload(A, b, n, x, nx, &u);
    if (ret == 0) {
        ret = bkh_zshbicgstabl(u.M, u.d_b, l, (const double *)sigma, sigma_size, u.d_x, tol, max_iter, NULL,
                               NULL);
        if (ret > BKH_DEVICE_ERROR && bkh_d2h(x, u.d_x, sizeof(double _Complex) * nx)) ret = BKH_DEVICE_ERROR;
    }
    zrelease(&u);
    return ret;
}

int zblbicgstabl(zspmat *A, zmat *B, int l, double tol, int max_iter, zmat *X) {
    int m = *A->row_size, s = B->col_size;
    if (s < 1 || s > 16 || l < 1 || l > 8) return INVALID_PARAMETERS;
    size_t ms = (size_t)m * s;
    bkh_info info;
    memset(&info, 0, sizeof info);
    zupload_t u;
    int ret = zupload(A, B->value, ms, X->value, ms, &u);
    if (ret == 0) {
        ret = bkh_zblbicgstabl(u.M, s, u.d_b, l, u.d_x, tol, max_iter, &info);
        if (ret > BKH_DEVICE_ERROR && bkh_d2h(X->value, u.d_x, sizeof(double _Complex) * ms)) ret = BKH_DEVICE_ERROR;
    }
    zrelease(&u);
    if (ret > BKH_DEVICE_ERROR)
        printf("Block BiCGSTAB(%d) has converged at:\n"
               "        The number of iteration ... %d\n"
               "        Frobenius norm of the relative residual ... %e\n"
               "        Frobenius norm of the true residual ... %e\n",
               l, max_iter <= 0 ? 0 : info.iters, info.rec_res, info.true_res);
    fflush(stdout);
    return ret;
}

int zqreco(int m, int n, double _Complex *mat, double _Complex *r) {
    if (m < 0) return -1;                      /* zungqr's info */
    if (n < 0 || n > m) return -2;
    if (n == 0) return 0;
    size_t bytes = sizeof(double _Complex) * (size_t)m * n;
    double *d = (double *)bkh_malloc(bytes);
    if (!d) return BKH_DEVICE_ERROR;
    int rc = bkh_h2d(d, mat, bytes);
    if (!rc) rc = bkh_zqr_householder(m, n, d, (double *)r);
    if (!rc) rc = bkh_d2h(mat, d, bytes);
    bkh_free(d);
    return rc;
}

zmat *zmat_create(int m, int n) {
    zmat *mat = (zmat *)malloc(sizeof(zmat));
    mat->value = (double _Complex *)calloc(sizeof(double _Complex), (size_t)m * n);
    mat->row_size = m;
    mat->col_size = n;
    return mat;
}

int zmat_free(zmat *mat) {
    free(mat->value);
    free(mat);
    return 0;
}

int zhorner(double _Complex *gamma, int l, double _Complex sigma, double _Complex *out) {
    out[l] = -gamma[l - 1];
    for (int i = l - 1; i > 0; i--) out[i] = -sigma * out[i + 1] - gamma[i - 1];
    out[0] = -sigma * out[1] + 1.0;
    for (int i = 1; i < l; i++)
        for (int j = l - 1; j >= i; j--) out[j] += -sigma * out[j + 1];
    for (int i = 1; i <= l; i++) out[i] = -out[i] / out[0];
    return 0;
}

/* C[:, i] = beta C[:, i] + alpha op(A) B[:, i] on the GPU; 0-based CSR. */
static int zcsr_apply(char op, const double _Complex *alpha, int m, int n, const double _Complex *A,
                      const int *rp0, const int *ci0, const double _Complex *B, const double _Complex *beta,
                      double _Complex *C) {
    bkh_zmatrix *M = bkh_zmatrix_create(m, rp0, ci0, (const double *)A);
    if (!M) return BKH_DEVICE_ERROR;
    size_t bytes = sizeof(double _Complex) * (size_t)m * n;
    double *dx = (double *)bkh_malloc(bytes), *dy = (double *)bkh_malloc(bytes);
    int rc = (dx && dy) ? 0 : BKH_DEVICE_ERROR;
    if (rc == 0 && (bkh_h2d(dx, B, bytes) || bkh_h2d(dy, C, bytes) ||
                    bkh_zspmm1(M, op, (const double *)alpha, n, dx, (const double *)beta, dy) ||
                    bkh_d2h(C, dy, bytes)))
        rc = BKH_DEVICE_ERROR;
    bkh_free(dx);
    bkh_free(dy);
    bkh_zmatrix_destroy(M);
    return rc;
}

/* reference src/matvec.c:55-98 (0-based): C = op(A) B */
int zcsrgemm0(char type, int m, int n, double _Complex *A, int *Ai, int *Aj, int nnz, double _Complex *B,
              double _Complex *C) {
    (void)nnz;
    if (type != 'n' && type != 't') { printf("invalid parameter 0.\n"); return -1; }
    const double _Complex one = 1.0, zero = 0.0;
    memset(C, 0, sizeof(double _Complex) * (size_t)m * n);
    return zcsr_apply(type, &one, m, n, A, Ai, Aj, B, &zero, C);
}

/* reference src/matvec.c:174-217 (1-based): C = beta C + alpha op(A) B */
int zcsrgemm1(char type, double _Complex *alpha, int m, int n, double _Complex *A, int *Ai, int *Aj, int nnz,
              double _Complex *B, double _Complex *beta, double _Complex *C) {
    (void)nnz;
    if (type != 'n' && type != 't') { printf("invalid parameter 0.\n"); return -1; }
    int nz = Ai[m] - 1;
    int *rp = (int *)malloc(sizeof(int) * ((size_t)m + 1));
    int *ci = (int *)malloc(sizeof(int) * (size_t)(nz > 0 ? nz : 1));
    if (!rp || !ci) { free(rp); free(ci); return BKH_DEVICE_ERROR; }
    for (int i = 0; i <= m; i++) rp[i] = Ai[i] - 1;
    for (int k = 0; k < nz; k++) ci[k] = Aj[k] - 1;
    int rc = zcsr_apply(type, alpha, m, n, A, rp, ci, B, beta, C);
    free(rp);
    free(ci);
    return rc;
}
