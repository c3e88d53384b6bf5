/*
 * orc_block.c -- test infrastructure only (see oracle.h).
 *
 * Block solvers, restated call for call from the reference:
 *   orc_blcg        src/blcg.c:15-119
 *   orc_blbicgstabl src/blbicgstabl.c:16-252
 *   orc_blbicgstab  src/blbicgstab.c:13-123
 * Dense blocks are m x s column-major (bksp.h:23-27, dmat).
 *
 * Documented readings of reference quirks:
 *  - blcg.c:32-34,61-63 leaves beta's off-diagonal uninitialised (malloc); the
 *    first direction is P = V + P*beta^T. p0_mode 0 uses the calloc reading
 *    (beta = I, so P0 = 2V); p0_mode 1 is the corrected P0 = V.
 *  - blbicgstabl.c:83-86 forms B + A*X0 (sign slip); for X0 = 0, which every
 *    reference caller passes, this equals B - A*X0. This restatement forms
 *    B - A*X0.
 *  - blbicgstabl.c:210 breaks on NaN but then returns the index as success;
 *    kept as is here (the oracle mirrors the reference).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define NOT_CONVERGED_CODE (-1)

static double *dalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

/* 1-based copies built once per call (blcg.c:42-49). */
static void one_based(int m, const int *rp, const int *ci, int **rp1, int **ci1) {
    int nnz = rp[m];
    *rp1 = (int *)malloc((size_t)(m + 1) * sizeof(int));
    *ci1 = (int *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int));
    for (int i = 0; i <= m; i++) (*rp1)[i] = rp[i] + 1;
    for (int i = 0; i < nnz; i++) (*ci1)[i] = ci[i] + 1;
}

/* mkl_dcsrmm 'S' (bksp_internal.h:72-76): C = beta*C, then C += alpha*A*B. */
static void symm(int m, int n, double alpha, const double *v, const int *rp1,
                 const int *ci1, const double *B, double beta, double *C) {
    orc_dscal(m * n, beta, C);
    orc_dcsrsymm1('n', alpha, m, n, v, rp1, ci1, B, beta, C);
}

int orc_blcg(int m, const int *rp, const int *ci, const double *v,
             int n, const double *B, double tol, int max_iter, double *X,
             int p0_mode) {
    int i;
    double fnb, error = 0;
    double *base0 = dalloc((size_t)n * n * 3);
    double *alpha = base0, *beta = base0 + n * n, *C = base0 + 2 * n * n;
    double *base1 = dalloc((size_t)m * n * 4);
    double *P = base1, *V = base1 + (size_t)m * n, *ap = base1 + (size_t)2 * m * n;
    double *tmp = base1 + (size_t)3 * m * n;
    int *ipiv = (int *)malloc((size_t)n * sizeof(int));
    int *rp1, *ci1;
    one_based(m, rp, ci, &rp1, &ci1);

    symm(m, n, -1, v, rp1, ci1, X, 0, tmp);                  /* blcg.c:54-63 */
    orc_daxpy(m * n, 1, B, tmp);
    orc_dcopy(m * n, tmp, V);
    orc_myqr(m, n, V, C);
    orc_dcopy(m * n, V, P);
    fnb = orc_dnrm2(m * n, B);
    for (i = 0; i < n; i++) beta[i * n + i] = 1.0;
    if (p0_mode == 1) for (i = 0; i < n * n; i++) beta[i] = 0.0;

    for (i = 0; i < max_iter; i++) {                         /* blcg.c:66-106 */
        orc_dgemm('N', 'T', m, n, n, 1, P, m, beta, n, 0, tmp, m);
        orc_daxpy(m * n, 1, V, tmp);
        orc_dcopy(m * n, tmp, P);
        symm(m, n, 1, v, rp1, ci1, P, 0, ap);
        orc_dgemm('T', 'N', n, n, m, 1, P, m, ap, m, 0, alpha, n);
        orc_dgetrf(n, alpha, n, ipiv);
        orc_dgetri(n, alpha, n, ipiv);
        orc_dgemm('N', 'N', n, n, n, 1, alpha, n, C, n, 0, tmp, n);
        orc_dgemm('N', 'N', m, n, n, 1, P, m, tmp, n, 1, X, m);
        orc_dgemm('N', 'N', m, n, n, -1, ap, m, alpha, n, 1, V, m);
        orc_myqr(m, n, V, beta);
        orc_dgemm('N', 'N', n, n, n, 1, beta, n, C, n, 0, tmp, n);
        orc_dcopy(n * n, tmp, C);
        error = orc_dnrm2(n * n, C) / fnb;
        if (error < tol) break;
    }
    free(ci1); free(rp1); free(ipiv); free(base1); free(base0);
    if (error >= tol) return NOT_CONVERGED_CODE;
    return i;
}

/* mkl_dcsrmm 'G' (bksp_internal.h:68-71): C = beta*C + alpha*A*B. */
static void gemm1(int m, int n, double alpha, const double *v, const int *rp1,
                  const int *ci1, const double *B, double beta, double *C) {
    orc_dcsrgemm1('n', alpha, m, n, v, rp1, ci1, B, beta, C);
}

int orc_blbicgstabl(int m, const int *rp, const int *ci, const double *v,
                    int n, const double *B, int l, double tol, int max_iter,
                    double *X, double *true_res) {
    int iter, i, j;
    double fnb, error = 0, t_local;
    size_t mn = (size_t)m * n;
    double *alpha = dalloc(n * n), *beta = dalloc(n * n), *t = dalloc(n * n);
    double *xi = dalloc(n * n), *gamma0 = dalloc(n * n), *gamma1 = dalloc(n * n);
    double *omega = dalloc(l + 1), *omega_d = dalloc(l + 1), *sigma = dalloc(l + 1);
    double *tau = dalloc((size_t)(l + 1) * (l + 1));
    double *Rs = dalloc(mn), *R_base = dalloc((l + 1) * mn), *PV_base = dalloc(2 * (l + 1) * mn);
    double *tmp = dalloc(mn);
    double **P = (double **)malloc((l + 1) * sizeof(double *));
    double **V = (double **)malloc((l + 1) * sizeof(double *));
    double **R = (double **)malloc((l + 1) * sizeof(double *));
    int *ipiv = (int *)malloc((size_t)n * sizeof(int));
    int *rp1, *ci1;
    one_based(m, rp, ci, &rp1, &ci1);
    omega[l] = 1;
    for (i = 0; i <= l; i++) {                               /* blbicgstabl.c:59-66 */
        P[i] = PV_base + (size_t)(2 * i) * mn;
        V[i] = PV_base + (size_t)(2 * i + 1) * mn;
        R[i] = R_base + (size_t)i * mn;
    }
    for (i = 0; i < n; i++)
        for (j = 0; j < n; j++) gamma0[i * n + j] = (i == j) ? 1 : 0;

    gemm1(m, n, -1, v, rp1, ci1, X, 0, tmp);                /* :83-93 (B - A X0) */
    orc_daxpy((int)mn, 1, B, tmp);
    orc_dcopy((int)mn, tmp, V[0]);
    orc_myqr(m, n, V[0], xi);
    orc_dcopy((int)mn, V[0], P[0]);
    orc_dcopy((int)mn, V[0], Rs);
    fnb = orc_dnrm2((int)mn, B);

    for (iter = 0; iter < max_iter; iter++) {
        orc_dscal(n * n, -omega[l], gamma0);                 /* :98 */
        for (j = 0; j < l; j++) {
            orc_dgemm('T', 'N', n, n, m, 1, Rs, m, V[j], m, 0, gamma1, n);   /* :102-108 */
            orc_dcopy(n * n, gamma1, tmp);
            orc_dgesv(n, n, gamma0, n, ipiv, tmp, n);
            orc_dgemm('N', 'N', n, n, n, -1, alpha, n, tmp, n, 0, beta, n);
            orc_dcopy(n * n, gamma1, gamma0);
            for (i = 0; i <= j; i++) {                       /* :111-116 */
                orc_dgemm('N', 'N', m, n, n, 1, P[i], m, beta, n, 0, tmp, m);
                orc_daxpy((int)mn, 1, V[i], tmp);
                orc_dcopy((int)mn, tmp, P[i]);
            }
            gemm1(m, n, 1, v, rp1, ci1, P[j], 0, P[j + 1]);
            orc_dcopy(n * n, gamma1, alpha);                 /* :121-124 */
            orc_dgemm('T', 'N', n, n, m, 1, Rs, m, P[j + 1], m, 0, tmp, n);
            orc_dgesv(n, n, tmp, n, ipiv, alpha, n);
            for (i = 0; i <= j; i++) {                       /* :127-157 */
                if (i == 0) {
                    orc_dgemm('N', 'N', m, n, n, 1, P[0], m, alpha, n, 0, tmp, m);
                    gemm1(m, n, -1, v, rp1, ci1, tmp, 1, V[0]);
                    orc_myqr(m, n, V[0], t);
                    orc_dgemm('N', 'N', m, n, n, 1, tmp, m, xi, n, 1, X, m);
                    orc_dgemm('N', 'N', n, n, n, 1, t, n, xi, n, 0, tmp, n);
                    orc_dcopy(n * n, tmp, xi);
                    orc_dgetrf(n, t, n, ipiv);
                    orc_dgetri(n, t, n, ipiv);
                } else {
                    orc_dgemm('N', 'N', m, n, n, -1, P[i + 1], m, alpha, n, 1, V[i], m);
                    orc_dgemm('N', 'N', m, n, n, 1, V[i], m, t, n, 0, tmp, m);
                    orc_dcopy((int)mn, tmp, V[i]);
                }
            }
            gemm1(m, n, 1, v, rp1, ci1, V[j], 0, V[j + 1]);  /* :159-160 */
        }
        for (i = 0; i <= l; i++)                              /* :163-166 */
            orc_dgemm('N', 'N', m, n, n, 1, V[i], m, xi, n, 0, R[i], m);
        for (j = 1; j <= l; j++) {                           /* MR :170-178 */
            for (i = 1; i < j; i++) {
                tau[i * l + j] = orc_ddot((int)mn, R[j], R[i]) / sigma[i];
                orc_daxpy((int)mn, -tau[i * l + j], R[i], R[j]);
            }
            sigma[j] = orc_ddot((int)mn, R[j], R[j]);
            omega_d[j] = orc_ddot((int)mn, R[0], R[j]) / sigma[j];
        }
        omega[l] = omega_d[l];                               /* :179-187 */
        for (j = l - 1; j > 0; j--) {
            t_local = 0;
            for (i = j + 1; i <= l; i++) t_local += tau[j * l + i] * omega[i];
            omega[j] = omega_d[j] - t_local;
        }
        orc_dscal((int)mn, 0, tmp);                          /* :198-207 */
        for (i = 1; i <= l; i++) {
            orc_daxpy((int)mn, omega[i], V[i - 1], tmp);
            orc_daxpy((int)mn, -omega[i], P[i], P[0]);
        }
        gemm1(m, n, -1, v, rp1, ci1, tmp, 1, V[0]);
        orc_dgemm('N', 'N', m, n, n, 1, tmp, m, xi, n, 1, X, m);
        error = orc_dnrm2(n * n, xi) / fnb;                  /* :209-212 */
        if (error < tol || isnan(error)) break;
    }
    gemm1(m, n, -1, v, rp1, ci1, X, 0, tmp);                /* :216-224 */
    orc_daxpy((int)mn, 1, B, tmp);
    if (true_res) *true_res = orc_dnrm2((int)mn, tmp) / fnb;

    free(ci1); free(rp1); free(ipiv); free(R); free(V); free(P); free(tmp);
    free(PV_base); free(R_base); free(Rs); free(tau); free(sigma); free(omega_d);
    free(omega); free(gamma1); free(gamma0); free(xi); free(t); free(beta); free(alpha);
    if (error >= tol) return NOT_CONVERGED_CODE;
    return iter;
}

/* src/blbicgstab.c:13-123: block BiCGSTAB with an orthonormalised direction block.
 * R0 = B - A X0 (:49-54); per iteration V = A P, alpha = (Rs^T V)^-1 Rs^T R by dgesv
 * (whose LU of Rs^T V stays in tmp for beta's dgetrs, :65-69, :92-94), R -= V alpha,
 * T = A R, omega = <T,R>/<T,T> over all m*n entries, X += P alpha + omega R,
 * R -= omega T, beta = -(Rs^T V)^-1 Rs^T T, P = (P - omega V) beta + R, then
 * myqr(P) (:104) keeps only P's orthonormal basis. The stop test is ||R||_F/||B||_F. */
int orc_blbicgstab(int m, const int *rp, const int *ci, const double *v,
                   int n, const double *B, double tol, int max_iter, double *X) {
    int i;
    double omega, fnb, error = 0;
    size_t mn = (size_t)m * n;
    double *base0 = dalloc((size_t)n * n * 2);
    double *alpha = base0, *beta = base0 + (size_t)n * n;
    double *base1 = dalloc(mn * 6);
    double *P = base1, *R = base1 + mn, *Rs = base1 + 2 * mn, *V = base1 + 3 * mn, *T = base1 + 4 * mn;
    double *tmp = base1 + 5 * mn;
    int *ipiv = (int *)malloc((size_t)n * n * sizeof(int));
    int *rp1, *ci1;
    one_based(m, rp, ci, &rp1, &ci1);

    gemm1(m, n, -1, v, rp1, ci1, X, 0, tmp);                /* blbicgstab.c:49-55 */
    orc_daxpy((int)mn, 1, B, tmp);
    orc_dcopy((int)mn, tmp, R);
    orc_dcopy((int)mn, R, P);
    orc_dcopy((int)mn, R, Rs);
    fnb = orc_dnrm2((int)mn, B);
    for (i = 0; i < max_iter; i++) {                         /* :58-110 */
        gemm1(m, n, 1, v, rp1, ci1, P, 0, V);
        orc_dgemm('T', 'N', n, n, m, 1, Rs, m, V, m, 0, tmp, n);
        orc_dgemm('T', 'N', n, n, m, 1, Rs, m, R, m, 0, alpha, n);
        orc_dgesv(n, n, tmp, n, ipiv, alpha, n);
        orc_dgemm('N', 'N', m, n, n, -1, V, m, alpha, n, 1, R, m);
        gemm1(m, n, 1, v, rp1, ci1, R, 0, T);
        omega = orc_ddot((int)mn, T, R) / orc_ddot((int)mn, T, T);
        orc_dgemm('N', 'N', m, n, n, 1, P, m, alpha, n, 1, X, m);
        orc_daxpy((int)mn, omega, R, X);
        orc_daxpy((int)mn, -omega, T, R);
        orc_dgemm('T', 'N', n, n, m, -1, Rs, m, T, m, 0, beta, n);
        orc_dgetrs(n, n, tmp, n, ipiv, beta, n);
        orc_daxpy((int)mn, -omega, V, P);
        orc_dgemm('N', 'N', m, n, n, 1, P, m, beta, n, 0, tmp, m);
        orc_daxpy((int)mn, 1, R, tmp);
        orc_dcopy((int)mn, tmp, P);
        orc_myqr(m, n, P, tmp);
        error = orc_dnrm2((int)mn, R) / fnb;
        if (error < tol) break;
    }
    free(ci1); free(rp1); free(ipiv); free(base1); free(base0);
    if (error >= tol) return NOT_CONVERGED_CODE;
    return i;
}
