/*
 * orc_complex.c -- test infrastructure only (see oracle.h).
 *
 * The complex FP64 family, restated call for call from the reference:
 *   orc_zcsrgemm0   src/matvec.c:55-98     (mkl_cspblas_zcsrgemv, include/bksp_internal.h:54-59)
 *   orc_zcsrgemm1   src/matvec.c:174-217   (mkl_zcsrmm 'G', include/bksp_internal.h:81-95)
 *   orc_zhorner     src/util_shift.c:29-53
 *   orc_zbicg       src/bicg.c:91-177
 *   orc_zbicgstabl  src/bicgstabl.c:141-276
 *   orc_zshbicgstabl src/shbicgstabl.c:288-592
 *   orc_zqreco      src/myqr.c:41-76        (zgeqrf + zungqr restated below)
 *   orc_zblbicgstabl src/blbicgstabl.c:255-486
 * and the CBLAS level-1 routines they call, from the Netlib reference algorithms
 * (zdotc: sum of conj(x_i) y_i in index order; dznrm2: the scaled sum of squares
 * over real and imaginary parts; zaxpy: quick return when alpha == 0).
 *
 * Complex products and sums use C99 `double complex` arithmetic as the reference
 * does (gcc, no FMA: re = ac - bd, im = ad + bc for finite operands).
 *
 * Reference quirk: zbicgstabl passes `&one` with `double one = 1` to cblas_zaxpy
 * (bicgstabl.c:147, :177, :194), which reads a complex alpha whose imaginary part is
 * the next double in the caller's frame. The restatement reads the intended 1 + 0i.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#define ORC_COMPLEX_IMPL
#include "oracle.h"

#define NOT_CONVERGED_CODE (-1)

static double complex *zalloc(size_t n) { return (double complex *)calloc(n ? n : 1, sizeof(double complex)); }

/* ---------------------------------------------------------------- level 1 */

double complex orc_zdotc(int n, const double complex *x, const double complex *y) {
    double complex s = 0.0;
    for (int i = 0; i < n; i++) s = s + conj(x[i]) * y[i];
    return s;
}

void orc_zdotc_sub(int n, const double complex *x, const double complex *y, double complex *out) {
    *out = orc_zdotc(n, x, y);
}

double orc_dznrm2(int n, const double complex *x) {
    if (n < 1) return 0.0;
    double scale = 0.0, ssq = 1.0;
    for (int i = 0; i < n; i++) {
        const double part[2] = {creal(x[i]), cimag(x[i])};
        for (int h = 0; h < 2; h++) {
            if (part[h] != 0.0) {
                const double a = fabs(part[h]);
                if (scale < a) { const double q = scale / a; ssq = 1.0 + ssq * q * q; scale = a; }
                else           { const double q = a / scale; ssq += q * q; }
            }
        }
    }
    return scale * sqrt(ssq);
}

void orc_zaxpy(int n, double complex a, const double complex *x, double complex *y) {
    if (fabs(creal(a)) + fabs(cimag(a)) == 0.0) return;
    for (int i = 0; i < n; i++) y[i] = y[i] + a * x[i];
}

void orc_zscal(int n, double complex a, double complex *x) {
    for (int i = 0; i < n; i++) x[i] = a * x[i];
}

void orc_zdscal(int n, double a, double complex *x) {
    for (int i = 0; i < n; i++) x[i] = CMPLX(a * creal(x[i]), a * cimag(x[i]));
}

void orc_zcopy(int n, const double complex *x, double complex *y) {
    memmove(y, x, (size_t)n * sizeof(double complex));
}

/* ------------------------------------------------------------ CSR kernels */

int orc_zcsrgemm0(char type, int m, int n, const double complex *A, const int *Ai, const int *Aj,
                  const double complex *B, double complex *C) {
    if (type == 'n') {                                   /* matvec.c:60-73 */
        for (int i = 0; i < n; i++)
            for (int j = 0; j < m; j++) {
                double complex c = 0.0;
                for (int k = Ai[j]; k < Ai[j + 1]; k++) c = c + A[k] * B[i * m + Aj[k]];
                C[i * m + j] = c;
            }
    } else if (type == 't') {                            /* :74-90: conj(A)^T by scatter */
        for (int i = 0; i < n; i++)
            for (int j = 0; j < m; j++) C[i * m + j] = 0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < m; j++)
                for (int k = Ai[j]; k < Ai[j + 1]; k++)
                    C[i * m + Aj[k]] = C[i * m + Aj[k]] + conj(A[k]) * B[i * m + j];
    } else {
        return -1;
    }
    return 0;
}

int orc_zcsrgemm1(char type, const double complex *alpha_p, int m, int n, const double complex *A,
                  const int *Ai, const int *Aj, const double complex *B, const double complex *beta_p,
                  double complex *C) {
    const double complex alpha = *alpha_p, beta = *beta_p;
    if (type == 'n') {                                   /* matvec.c:179-191, 1-based */
        for (int i = 0; i < n; i++)
            for (int j = 0; j < m; j++) {
                double complex c = beta * C[i * m + j];
                for (int k = Ai[j] - 1; k < Ai[j + 1] - 1; k++) c = c + alpha * A[k] * B[i * m + Aj[k] - 1];
                C[i * m + j] = c;
            }
    } else if (type == 't') {                            /* :192-209 */
        for (int i = 0; i < n; i++)
            for (int j = 0; j < m; j++) C[i * m + j] = beta * C[i * m + j];
        for (int i = 0; i < n; i++)
            for (int j = 0; j < m; j++)
                for (int k = Ai[j] - 1; k < Ai[j + 1] - 1; k++)
                    C[i * m + Aj[k] - 1] = C[i * m + Aj[k] - 1] + alpha * conj(A[k]) * B[i * m + j];
    } else {
        return -1;
    }
    return 0;
}

/* y = op(A) x through the zcsrgemv dispatch (bksp_internal.h:54-59, 0-based). */
static void zgemv(char t, int n, const int *rp, const int *ci, const double complex *v,
                  const double complex *x, double complex *y) {
    orc_zcsrgemm0(t, n, 1, v, rp, ci, x, y);
}

int orc_zhorner(const double complex *gamma, int l, const double complex *sigma_p, double complex *out) {
    const double complex sigma = *sigma_p;
    out[l] = -gamma[l - 1];                              /* util_shift.c:32-51 */
    for (int i = l - 1; i > 0; i--) out[i] = -sigma * out[i + 1] - gamma[i - 1];
    out[0] = -sigma * out[1] + 1.0;
    for (int i = 1; i < l; i++)
        for (int j = l - 1; j >= i; j--) out[j] += -sigma * out[j + 1];
    for (int i = 1; i <= l; i++) out[i] = -out[i] / out[0];
    return 0;
}

/* ------------------------------------------------------------------ solvers */

int orc_zbicg(int n, const int *rp, const int *ci, const double complex *v, const double complex *b,
              double tol, int max_iter, double complex *x) {
    int i;
    double error = 0, norm_b;
    double complex alpha, beta, r_dot, r_dot_old;
    double complex *r = zalloc(n), *sr = zalloc(n), *p = zalloc(n), *sp = zalloc(n);
    double complex *ap = zalloc(n), *atp = zalloc(n), *tmp = zalloc(n);
    zgemv('n', n, rp, ci, v, x, tmp);                    /* bicg.c:106-115 */
    orc_zdscal(n, -1, tmp);
    orc_zaxpy(n, 1, b, tmp);
    orc_zcopy(n, tmp, r);
    orc_zcopy(n, tmp, sr);
    orc_zcopy(n, r, p);
    orc_zcopy(n, sr, sp);
    r_dot_old = orc_zdotc(n, sr, r);
    norm_b = orc_dznrm2(n, b);
    for (i = 0; i < max_iter; i++) {                     /* :118-160 */
        zgemv('n', n, rp, ci, v, p, ap);
        zgemv('t', n, rp, ci, v, sp, atp);
        alpha = r_dot_old / orc_zdotc(n, sp, ap);
        orc_zaxpy(n, alpha, p, x);
        orc_zaxpy(n, -alpha, ap, r);
        orc_zaxpy(n, -conj(alpha), atp, sr);
        r_dot = orc_zdotc(n, sr, r);
        beta = r_dot / r_dot_old;
        r_dot_old = r_dot;
        orc_zscal(n, beta, p);
        orc_zaxpy(n, 1, r, p);
        orc_zscal(n, conj(beta), sp);
        orc_zaxpy(n, 1, sr, sp);
        error = orc_dznrm2(n, r) / norm_b;
        if (error < tol) break;
    }
    free(tmp); free(atp); free(ap); free(sp); free(p); free(sr); free(r);
    if (error >= tol) return NOT_CONVERGED_CODE;
    return i;
}

int orc_zbicgstabl(int n, const int *rp, const int *ci, const double complex *v, const double complex *b,
                   int l, double tol, int max_iter, double complex *x) {
    int i, j, m, l2 = l + 1;
    double error = 0, fnb;
    double complex beta, gamma1, alpha = 0, gamma0 = 1, t_local;
    double complex *base = zalloc((size_t)n * (l2 * 2 + 2));
    double complex **r = (double complex **)malloc(sizeof(double complex *) * l2 * 2), **p = r + l2;
    for (i = 0; i < l2; i++) { r[i] = base + (size_t)n * i; p[i] = base + (size_t)n * (l2 + i); }
    double complex *rs = base + (size_t)n * l2 * 2, *tmp = base + (size_t)n * (l2 * 2 + 1);
    double complex *wb = zalloc((size_t)l2 * 4);
    double complex *omega = wb, *omega_d = wb + l2, *omega_dd = wb + 2 * l2, *sigma = wb + 3 * l2;
    double complex *tau = zalloc((size_t)l2 * l2);
    omega[l] = 1;                                        /* bicgstabl.c:168 */

    zgemv('n', n, rp, ci, v, x, tmp);                    /* :175-181 */
    orc_zdscal(n, -1, tmp);
    orc_zaxpy(n, 1, b, tmp);
    orc_zcopy(n, tmp, r[0]);
    orc_zcopy(n, r[0], p[0]);
    orc_zcopy(n, r[0], rs);
    fnb = orc_dznrm2(n, b);

    for (m = 0; m < max_iter; m++) {
        gamma0 = -omega[l] * gamma0;                     /* :186 */
        for (j = 0; j < l; j++) {                        /* BiCG part :188-210 */
            gamma1 = orc_zdotc(n, rs, r[j]);
            beta = -alpha * gamma1 / gamma0;
            gamma0 = gamma1;
            for (i = 0; i <= j; i++) {
                orc_zscal(n, beta, p[i]);
                orc_zaxpy(n, 1, r[i], p[i]);
            }
            zgemv('n', n, rp, ci, v, p[j], p[j + 1]);
            alpha = gamma1 / orc_zdotc(n, rs, p[j + 1]);
            for (i = 0; i <= j; i++) orc_zaxpy(n, -alpha, p[i + 1], r[i]);
            zgemv('n', n, rp, ci, v, r[j], r[j + 1]);
            orc_zaxpy(n, alpha, p[0], x);
        }
        for (j = 1; j <= l; j++) {                       /* MR part :214-224 */
            for (i = 1; i < j; i++) {
                tau[j * l + i] = orc_zdotc(n, r[i], r[j]) / sigma[i];
                orc_zaxpy(n, -tau[j * l + i], r[i], r[j]);
            }
            sigma[j] = orc_zdotc(n, r[j], r[j]);
            omega_d[j] = orc_zdotc(n, r[j], r[0]) / sigma[j];
        }
        omega[l] = omega_d[l];                           /* :226-242 */
        for (j = l - 1; j > 0; j--) {
            t_local = 0;
            for (i = j + 1; i <= l; i++) t_local += tau[i * l + j] * omega[i];
            omega[j] = omega_d[j] - t_local;
        }
        for (j = 1; j < l; j++) {
            t_local = 0;
            for (i = j + 1; i < l; i++) t_local += tau[i * l + j] * omega[i + 1];
            omega_dd[j] = omega[j + 1] + t_local;
        }
        orc_zaxpy(n, omega[1], r[0], x);                 /* update :245-256 */
        orc_zaxpy(n, -omega_d[l], r[l], r[0]);
        orc_zaxpy(n, -omega[l], p[l], p[0]);
        for (i = 1; i < l; i++) {
            orc_zaxpy(n, -omega[i], p[i], p[0]);
            orc_zaxpy(n, -omega_d[i], r[i], r[0]);
            orc_zaxpy(n, omega_dd[i], r[i], x);
        }
        error = orc_dznrm2(n, r[0]) / fnb;               /* :258-261 */
        if (error < tol) break;
    }
    free(tau); free(wb); free(r); free(base);
    if (error >= tol) return NOT_CONVERGED_CODE;
    return m;
}

int orc_zshbicgstabl(int n, const int *rp, const int *ci, const double complex *v, const double complex *b,
                     int l, const double complex *sigma, int ns, double tol, int max_iter, double complex *x) {
    int iter, i, j, k, l2 = l + 1;
    double error = 0, norm_b;
    double complex alpha, beta, alpha_old, rho, rho_old, t_local, psi, e;
    double complex *r[32], *p[32];
    if (l2 > 32) return -3;
    for (i = 0; i < l2; i++) { r[i] = zalloc(n); p[i] = zalloc(n); }
    double complex *rs = zalloc(n), *tmp = zalloc(n > l2 ? n : l2);
    double complex *omega = zalloc(l), *omega_d = zalloc(l), *omega_dd = zalloc(l);
    double complex *s = zalloc(l), *tau = zalloc((size_t)l * l);
    omega[l - 1] = 1;                                    /* shbicgstabl.c:315 */

    zgemv('n', n, rp, ci, v, x, tmp);                    /* :322-329 */
    orc_zdscal(n, -1, tmp);
    orc_zaxpy(n, 1, b, tmp);
    orc_zcopy(n, tmp, r[0]);
    orc_zcopy(n, r[0], p[0]);
    orc_zcopy(n, r[0], rs);
    norm_b = orc_dznrm2(n, b);
    alpha = 0.0; alpha_old = 1.0; rho_old = 1.0;

    double complex *alpha_s = zalloc(ns), *beta_s = zalloc(ns), *pi = zalloc(ns);
    double complex *pi_new = zalloc(ns), *pi_old = zalloc(ns), *theta = zalloc(ns);
    double complex *pbase = zalloc((size_t)n * l2 * ns);
    double complex *om_s = zalloc(l), *om_sd = zalloc(l), *om_sdd = zalloc(l);
    double complex *M = zalloc((size_t)ns * l * l);
#define PS(k_, i_) (pbase + ((size_t)(k_) * l2 + (i_)) * (size_t)n)
#define XS(k_) (x + (size_t)((k_) + 1) * n)
    for (i = 0; i < ns; i++) { pi_old[i] = 1; pi[i] = 1; theta[i] = 1; }
    for (k = 0; k < ns; k++)                             /* :374-380 */
        for (j = 0; j < l; j++)
            for (i = 0; i <= j; i++)
                M[(size_t)k * l * l + j * l + i] = cpow(sigma[k], j - i) * orc_comb(j, i);

    for (iter = 0; iter < max_iter; iter++) {
        rho_old = -omega[l - 1] * rho_old;               /* :387 */
        for (j = 0; j < l; j++) {
            rho = orc_zdotc(n, rs, r[j]);                /* :390-410 */
            beta = alpha * rho / rho_old;
            for (i = 0; i <= j; i++) {
                orc_zscal(n, -beta, p[i]);
                orc_zaxpy(n, 1, r[i], p[i]);
            }
            zgemv('n', n, rp, ci, v, p[j], p[j + 1]);
            alpha = rho / orc_zdotc(n, rs, p[j + 1]);
            orc_zaxpy(n, alpha, p[0], x);
            for (k = 0; k < ns; k++) {                   /* :413-440 */
                pi_new[k] = (1.0 + alpha * sigma[k]) * pi[k]
                    + alpha * beta / alpha_old * (pi_old[k] - pi[k]);
                beta_s[k] = (pi_old[k] / pi[k]) * (pi_old[k] / pi[k]) * beta;
                alpha_s[k] = pi[k] / pi_new[k] * alpha;
                for (i = 0; i <= j; i++) {
                    orc_zscal(n, -beta_s[k], PS(k, i));
                    orc_zaxpy(n, 1.0 / (pi[k] * theta[k]), r[i], PS(k, i));
                }
                orc_zcopy(n, r[j], PS(k, j + 1));
                orc_zscal(n, 1.0 / pi[k], PS(k, j + 1));
                orc_zaxpy(n, alpha_s[k], PS(k, 0), XS(k));
            }
            for (i = 0; i <= j; i++) orc_zaxpy(n, -alpha, p[i + 1], r[i]);   /* :443-446 */
            zgemv('n', n, rp, ci, v, r[j], r[j + 1]);
            for (k = 0; k < ns; k++) {                   /* :452-461 */
                orc_zaxpy(n, -1.0 / pi_new[k], r[j], PS(k, j + 1));
                orc_zscal(n, 1.0 / (theta[k] * alpha_s[k]), PS(k, j + 1));
                orc_zaxpy(n, -sigma[k], PS(k, j), PS(k, j + 1));
            }
            alpha_old = alpha;                           /* :464-469 */
            rho_old = rho;
            for (k = 0; k < ns; k++) { pi_old[k] = pi[k]; pi[k] = pi_new[k]; }
        }
        for (j = 0; j < l; j++) {                        /* MR :475-484 */
            for (i = 0; i < j; i++) {
                tau[j * l + i] = orc_zdotc(n, r[i + 1], r[j + 1]) / s[i];
                orc_zaxpy(n, -tau[j * l + i], r[i + 1], r[j + 1]);
            }
            s[j] = orc_zdotc(n, r[j + 1], r[j + 1]);
            omega_d[j] = orc_zdotc(n, r[j + 1], r[0]) / s[j];
        }
        omega[l - 1] = omega_d[l - 1];                   /* :486-502 */
        for (j = l - 2; j >= 0; j--) {
            t_local = 0;
            for (i = j + 1; i < l; i++) t_local += tau[i * l + j] * omega[i];
            omega[j] = omega_d[j] - t_local;
        }
        for (j = 0; j < l - 1; j++) {
            t_local = 0;
            for (i = j + 1; i < l - 1; i++) t_local += tau[i * l + j] * omega[i + 1];
            omega_dd[j] = omega[j + 1] + t_local;
        }
        orc_zaxpy(n, omega[0], r[0], x);                 /* :505-511 */
        orc_zaxpy(n, -omega[l - 1], p[l], p[0]);
        for (i = 1; i < l; i++) {
            orc_zaxpy(n, -omega[i - 1], p[i], p[0]);
            orc_zaxpy(n, omega_dd[i - 1], r[i], x);
        }
        for (k = 0; k < ns; k++) {                       /* :514-558 */
            orc_zhorner(omega, l, &sigma[k], tmp);
            psi = tmp[0];
            for (i = 0; i < l; i++) om_s[i] = tmp[i + 1];
            e = theta[k] * pi[k];
            theta[k] = theta[k] * psi;
            for (j = 0; j < l; j++) {
                double complex acc = 0;
                for (i = j; i < l; i++) acc += M[(size_t)k * l * l + i * l + j] * om_s[i];
                om_sd[j] = acc;
            }
            for (j = 0; j < l - 1; j++) {
                double complex acc = 0;
                for (i = j + 1; i < l - 1; i++) acc += tau[i * l + j] * om_sd[i + 1];
                om_sdd[j] = om_sd[j + 1] + acc;
            }
            orc_zaxpy(n, om_sd[0] / e, r[0], XS(k));
            orc_zaxpy(n, -omega[l - 1], PS(k, l), PS(k, 0));
            for (i = 1; i < l; i++) {
                orc_zaxpy(n, -omega[i - 1], PS(k, i), PS(k, 0));
                orc_zaxpy(n, om_sdd[i - 1] / e, r[i], XS(k));
            }
            orc_zscal(n, 1.0 / psi, PS(k, 0));
        }
        for (i = 1; i <= l; i++) orc_zaxpy(n, -omega_d[i - 1], r[i], r[0]);   /* :561-564 */
        error = orc_dznrm2(n, r[0]) / norm_b;
        if (error < tol) break;
    }
#undef PS
#undef XS
    free(M); free(om_sdd); free(om_sd); free(om_s); free(pbase);
    free(theta); free(pi_old); free(pi_new); free(pi); free(beta_s); free(alpha_s);
    free(tau); free(s); free(omega_dd); free(omega_d); free(omega);
    free(tmp); free(rs);
    for (i = 0; i < l2; i++) { free(r[i]); free(p[i]); }
    if (error >= tol) return NOT_CONVERGED_CODE;
    return iter;
}

/* ------------------------------------------------ LAPACK-level restatements
 * From the Netlib reference algorithms (zgemm, zgetf2/zgetrs/zgetri/zgesv,
 * zgeqr2 + zlarfg + zlarf, zung2r); the safe-minimum rescaling branches of
 * zlarfg/zgetf2 are omitted (they only act on values near underflow) and ZLADIV
 * is C99 complex division. Pinned by invariants (Q^H Q = I, QR = V, LU solves). */

/* C = alpha op(A) op(B) + beta C, op 'N' or 'T' (transpose without conjugation). */
void orc_zgemm(char ta, char tb, int m, int n, int k, double complex alpha, const double complex *A, int lda,
               const double complex *B, int ldb, double complex beta, double complex *C, int ldc) {
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) {
            double complex s = 0.0;
            for (int l = 0; l < k; l++) {
                const double complex a = (ta == 'N') ? A[i + (size_t)l * lda] : A[l + (size_t)i * lda];
                const double complex b = (tb == 'N') ? B[l + (size_t)j * ldb] : B[j + (size_t)l * ldb];
                s += a * b;
            }
            double complex *c = &C[i + (size_t)j * ldc];
            *c = (beta == 0.0) ? alpha * s : alpha * s + beta * *c;
        }
}

static double cabs1(double complex z) { return fabs(creal(z)) + fabs(cimag(z)); }

int orc_zgetrf(int n, double complex *A, int lda, int *ipiv) {
    int info = 0;
    for (int j = 0; j < n; j++) {
        int p = j;
        double best = cabs1(A[j + (size_t)j * lda]);
        for (int i = j + 1; i < n; i++)
            if (cabs1(A[i + (size_t)j * lda]) > best) { best = cabs1(A[i + (size_t)j * lda]); p = i; }
        ipiv[j] = p + 1;
        if (A[p + (size_t)j * lda] != 0.0) {
            if (p != j)
                for (int c = 0; c < n; c++) {
                    double complex t = A[j + (size_t)c * lda];
                    A[j + (size_t)c * lda] = A[p + (size_t)c * lda];
                    A[p + (size_t)c * lda] = t;
                }
            const double complex r = 1.0 / A[j + (size_t)j * lda];
            for (int i = j + 1; i < n; i++) A[i + (size_t)j * lda] *= r;
        } else if (info == 0) {
            info = j + 1;
        }
        for (int c = j + 1; c < n; c++) {
            const double complex u = A[j + (size_t)c * lda];
            for (int i = j + 1; i < n; i++) A[i + (size_t)c * lda] -= A[i + (size_t)j * lda] * u;
        }
    }
    return info;
}

int orc_zgetrs(int n, int nrhs, const double complex *LU, int lda, const int *ipiv, double complex *B, int ldb) {
    for (int c = 0; c < nrhs; c++) {
        double complex *b = B + (size_t)c * ldb;
        for (int i = 0; i < n; i++) {
            const int p = ipiv[i] - 1;
            if (p != i) { double complex t = b[i]; b[i] = b[p]; b[p] = t; }
        }
        for (int j = 0; j < n; j++)
            for (int i = j + 1; i < n; i++) b[i] -= LU[i + (size_t)j * lda] * b[j];
        for (int j = n - 1; j >= 0; j--) {
            b[j] /= LU[j + (size_t)j * lda];
            for (int i = 0; i < j; i++) b[i] -= LU[i + (size_t)j * lda] * b[j];
        }
    }
    return 0;
}

int orc_zgetri(int n, double complex *A, int lda, const int *ipiv) {
    double complex *lu = zalloc((size_t)n * n);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) lu[i + (size_t)j * n] = A[i + (size_t)j * lda];
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) A[i + (size_t)j * lda] = (i == j) ? 1.0 : 0.0;
    orc_zgetrs(n, n, lu, n, ipiv, A, lda);
    free(lu);
    return 0;
}

int orc_zgesv(int n, int nrhs, double complex *A, int lda, int *ipiv, double complex *B, int ldb) {
    const int info = orc_zgetrf(n, A, lda, ipiv);
    if (info) return info;
    return orc_zgetrs(n, nrhs, A, lda, ipiv, B, ldb);
}

static double dlapy3(double x, double y, double z) {
    const double ax = fabs(x), ay = fabs(y), az = fabs(z);
    double w = ax > ay ? ax : ay;
    w = w > az ? w : az;
    if (w == 0.0) return ax + ay + az;
    return w * sqrt((ax / w) * (ax / w) + (ay / w) * (ay / w) + (az / w) * (az / w));
}

/* zlarfg: H^H (alpha; x) = (beta; 0), beta real, H = I - tau v v^H, v = (1; x scaled). */
static void zlarfg(int n, double complex *alpha, double complex *x, double complex *tau) {
    if (n <= 0) { *tau = 0.0; return; }
    const double xnorm = orc_dznrm2(n - 1, x);
    const double ar = creal(*alpha), ai = cimag(*alpha);
    if (xnorm == 0.0 && ai == 0.0) { *tau = 0.0; return; }
    const double beta = -copysign(dlapy3(ar, ai, xnorm), ar);
    *tau = CMPLX((beta - ar) / beta, -ai / beta);
    orc_zscal(n - 1, 1.0 / (*alpha - beta), x);
    *alpha = beta;
}

/* zlarf('L'): C (m x n, ldc) := (I - tau v v^H) C. */
static void zlarf_left(int m, int n, const double complex *v, double complex tau, double complex *C, int ldc) {
    if (tau == 0.0) return;
    for (int j = 0; j < n; j++) {
        double complex *c = C + (size_t)j * ldc;
        double complex w = 0.0;
        for (int i = 0; i < m; i++) w += conj(c[i]) * v[i];
        for (int i = 0; i < m; i++) c[i] -= tau * v[i] * conj(w);
    }
}

int orc_zgeqrf(int m, int n, double complex *A, int lda, double complex *tau) {
    const int k = m < n ? m : n;
    for (int i = 0; i < k; i++) {
        double complex *aii = &A[i + (size_t)i * lda];
        zlarfg(m - i, aii, &A[(i + 1 < m ? i + 1 : m - 1) + (size_t)i * lda], &tau[i]);
        if (i < n - 1) {
            const double complex keep = *aii;
            *aii = 1.0;
            zlarf_left(m - i, n - i - 1, aii, conj(tau[i]), &A[i + (size_t)(i + 1) * lda], lda);
            *aii = keep;
        }
    }
    return 0;
}

int orc_zungqr(int m, int n, int k, double complex *A, int lda, const double complex *tau) {
    for (int j = k; j < n; j++) {
        for (int l = 0; l < m; l++) A[l + (size_t)j * lda] = 0.0;
        A[j + (size_t)j * lda] = 1.0;
    }
    for (int i = k - 1; i >= 0; i--) {
        double complex *aii = &A[i + (size_t)i * lda];
        if (i < n - 1) {
            *aii = 1.0;
            zlarf_left(m - i, n - i - 1, aii, tau[i], &A[i + (size_t)(i + 1) * lda], lda);
        }
        if (i < m - 1) orc_zscal(m - i - 1, -tau[i], &A[i + 1 + (size_t)i * lda]);
        *aii = 1.0 - tau[i];
        for (int l = 0; l < i; l++) A[l + (size_t)i * lda] = 0.0;
    }
    return 0;
}

/* zqreco (myqr.c:41-76): thin QR of an m x n column-major block, Q in place, R (n x n). */
int orc_zqreco(int m, int n, double complex *mat, double complex *r) {
    double complex *tau = zalloc((size_t)n * n);
    orc_zgeqrf(m, n, mat, m, tau);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) r[i * n + j] = (j > i) ? 0.0 : mat[i * m + j];
    orc_zungqr(m, n, n, mat, m, tau);
    free(tau);
    return 0;
}

/* mkl_zcsrmm 'G' (bksp_internal.h:81-95) on 1-based CSR: C = beta C + alpha A B. */
static void zgemm1(int m, int n, double complex alpha, const double complex *v, const int *rp1, const int *ci1,
                   const double complex *B, double complex beta, double complex *C) {
    orc_zcsrgemm1('n', &alpha, m, n, v, rp1, ci1, B, &beta, C);
}

int orc_zblbicgstabl(int m, const int *rp, const int *ci, const double complex *v, int n,
                     const double complex *B, int l, double tol, int max_iter, double complex *X,
                     double *true_res) {
    int iter, i, j;
    double fnb, error = 0;
    const size_t mn = (size_t)m * n;
    double complex *alpha = zalloc((size_t)n * n), *beta = zalloc((size_t)n * n), *t = zalloc((size_t)n * n);
    double complex *xi = zalloc((size_t)n * n), *gamma0 = zalloc((size_t)n * n), *gamma1 = zalloc((size_t)n * n);
    double complex *omega = zalloc(l), *omega_d = zalloc(l), *sigma = zalloc(l), *tau = zalloc((size_t)l * l);
    double complex *Rs = zalloc(mn), *R_base = zalloc((l + 1) * mn), *PV_base = zalloc(2 * (l + 1) * mn);
    double complex *tmp = zalloc(mn), t_local;
    double complex **P = (double complex **)malloc((l + 1) * sizeof(double complex *));
    double complex **V = (double complex **)malloc((l + 1) * sizeof(double complex *));
    double complex **R = (double complex **)malloc((l + 1) * sizeof(double complex *));
    int *ipiv = (int *)malloc((size_t)n * n * sizeof(int));
    int *rp1 = (int *)malloc(sizeof(int) * (m + 1)), *ci1 = (int *)malloc(sizeof(int) * (rp[m] > 0 ? rp[m] : 1));
    for (i = 0; i <= m; i++) rp1[i] = rp[i] + 1;               /* blbicgstabl.c:299-304 */
    for (i = 0; i < rp[m]; i++) ci1[i] = ci[i] + 1;
    omega[l - 1] = 1;                                          /* :283 */
    for (i = 0; i <= l; i++) {                                 /* :305-312 */
        P[i] = PV_base + (size_t)(2 * i) * mn;
        V[i] = PV_base + (size_t)(2 * i + 1) * mn;
        R[i] = R_base + (size_t)i * mn;
    }
    for (i = 0; i < n; i++) gamma0[i * n + i] = 1.0;

    zgemm1(m, n, -1, v, rp1, ci1, X, 0, tmp);                  /* :322-330 */
    orc_zdscal((int)mn, -1, tmp);
    orc_zaxpy((int)mn, 1, B, tmp);
    orc_zcopy((int)mn, tmp, V[0]);
    orc_zqreco(m, n, V[0], xi);
    orc_zcopy((int)mn, V[0], P[0]);
    orc_zcopy((int)mn, V[0], Rs);
    fnb = orc_dznrm2((int)mn, B);

    for (iter = 0; iter < max_iter; iter++) {
        orc_zscal(n * n, -omega[l - 1], gamma0);               /* :339-340 */
        for (j = 0; j < l; j++) {
            orc_zgemm('T', 'N', n, n, m, 1, Rs, m, V[j], m, 0, gamma1, n);   /* :344-350 */
            orc_zcopy(n * n, gamma1, tmp);
            orc_zgesv(n, n, gamma0, n, ipiv, tmp, n);
            orc_zgemm('N', 'N', n, n, n, 1, alpha, n, tmp, n, 0, beta, n);
            orc_zcopy(n * n, gamma1, gamma0);
            for (i = 0; i <= j; i++) {                         /* :353-358 */
                orc_zgemm('N', 'N', m, n, n, -1, P[i], m, beta, n, 0, tmp, m);
                orc_zaxpy((int)mn, 1, V[i], tmp);
                orc_zcopy((int)mn, tmp, P[i]);
            }
            zgemm1(m, n, 1, v, rp1, ci1, P[j], 0, P[j + 1]);
            orc_zcopy(n * n, gamma1, alpha);                   /* :363-366 */
            orc_zgemm('T', 'N', n, n, m, 1, Rs, m, P[j + 1], m, 0, tmp, n);
            orc_zgesv(n, n, tmp, n, ipiv, alpha, n);
            for (i = 0; i <= j; i++) {                         /* :369-400 */
                if (i == 0) {
                    orc_zgemm('N', 'N', m, n, n, 1, P[0], m, alpha, n, 0, tmp, m);
                    zgemm1(m, n, -1, v, rp1, ci1, tmp, 1, V[0]);
                    orc_zqreco(m, n, V[0], t);
                    orc_zgemm('N', 'N', m, n, n, 1, tmp, m, xi, n, 1, X, m);
                    orc_zgemm('N', 'N', n, n, n, 1, t, n, xi, n, 0, tmp, n);
                    orc_zcopy(n * n, tmp, xi);
                    orc_zgetrf(n, t, n, ipiv);
                    orc_zgetri(n, t, n, ipiv);
                } else {
                    orc_zgemm('N', 'N', m, n, n, -1, P[i + 1], m, alpha, n, 1, V[i], m);
                    orc_zgemm('N', 'N', m, n, n, 1, V[i], m, t, n, 0, tmp, m);
                    orc_zcopy((int)mn, tmp, V[i]);
                }
            }
            zgemm1(m, n, 1, v, rp1, ci1, V[j], 0, V[j + 1]);   /* :402-403 */
        }
        for (i = 0; i <= l; i++)                                /* :406-409 */
            orc_zgemm('N', 'N', m, n, n, 1, V[i], m, xi, n, 0, R[i], m);
        for (j = 0; j < l; j++) {                              /* MR :413-423 */
            for (i = 0; i < j; i++) {
                tau[j * l + i] = orc_zdotc((int)mn, R[i + 1], R[j + 1]) / sigma[i];
                orc_zaxpy((int)mn, -tau[j * l + i], R[i + 1], R[j + 1]);
            }
            sigma[j] = orc_zdotc((int)mn, R[j + 1], R[j + 1]);
            omega_d[j] = orc_zdotc((int)mn, R[j + 1], R[0]) / sigma[j];
        }
        omega[l - 1] = omega_d[l - 1];                         /* :424-432 */
        for (j = l - 2; j >= 0; j--) {
            t_local = 0;
            for (i = j + 1; i < l; i++) t_local += tau[i * l + j] * omega[i];
            omega[j] = omega_d[j] - t_local;
        }
        orc_zdscal((int)mn, 0, tmp);                           /* :442-452 */
        for (i = 0; i < l; i++) {
            orc_zaxpy((int)mn, omega[i], V[i], tmp);
            orc_zaxpy((int)mn, -omega[i], P[i + 1], P[0]);
        }
        zgemm1(m, n, -1, v, rp1, ci1, tmp, 1, V[0]);
        orc_zgemm('N', 'N', m, n, n, 1, tmp, m, xi, n, 1, X, m);
        error = orc_dznrm2(n * n, xi) / fnb;                   /* :454-457 */
        if (error < tol || isnan(error)) break;
    }
    zgemm1(m, n, -1, v, rp1, ci1, X, 0, tmp);                  /* :461-469 */
    orc_zaxpy((int)mn, 1, B, tmp);
    if (true_res) *true_res = orc_dznrm2((int)mn, tmp) / fnb;

    free(ci1); free(rp1); free(ipiv); free(R); free(V); free(P); free(tmp);
    free(PV_base); free(R_base); free(Rs); free(tau); free(sigma); free(omega_d); free(omega);
    free(gamma1); free(gamma0); free(xi); free(t); free(beta); free(alpha);
    if (error >= tol || isnan(error)) return NOT_CONVERGED_CODE;
    return iter;
}
