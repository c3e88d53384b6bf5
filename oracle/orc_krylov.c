/*
 * orc_krylov.c -- test infrastructure only (see oracle.h).
 *
 * Single right-hand-side solvers, restated call for call from the reference:
 *   orc_cg          src/cg.c:6-71
 *   orc_cr          src/cr.c:6-80
 *   orc_bicg        src/bicg.c:7-88
 *   orc_bicgstabl   src/bicgstabl.c:13-138
 *   orc_shbicgstabl src/shbicgstabl.c:7-285
 *   orc_bicgstab    src/bicgstab.c:5-72
 *   orc_shbicg      src/shbicg.c:5-129
 * Each returns the 0-based loop index at the break, or -1 (NOT_CONVERGED), with
 * the reference's final test `error >= tol` (so a NaN error returns max_iter).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define NOT_CONVERGED_CODE (-1)

static double *dalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

int orc_cg(int n, const int *rp, const int *ci, const double *v,
           const double *b, double tol, int max_iter, double *x) {
    int i;
    double error = 0, alpha, beta, fnb, r_dot, r_dot_old;
    double *r = dalloc(n), *p = dalloc(n), *tmp = dalloc(n);
    orc_csrsymv(n, v, rp, ci, x, tmp);                      /* cg.c:23-24 */
    orc_dscal(n, -1, tmp);
    orc_daxpy(n, 1, b, tmp);
    orc_dcopy(n, tmp, r);
    orc_dcopy(n, r, p);
    r_dot_old = orc_ddot(n, r, r);
    fnb = orc_dnrm2(n, b);
    for (i = 0; i < max_iter; i++) {                         /* cg.c:33-60 */
        orc_csrsymv(n, v, rp, ci, p, tmp);
        alpha = r_dot_old / orc_ddot(n, tmp, p);
        orc_daxpy(n, alpha, p, x);
        orc_daxpy(n, -alpha, tmp, r);
        r_dot = orc_ddot(n, r, r);
        beta = r_dot / r_dot_old;
        r_dot_old = r_dot;
        orc_dscal(n, beta, p);
        orc_daxpy(n, 1, r, p);
        error = orc_dnrm2(n, r) / fnb;
        if (error < tol) break;
    }
    free(tmp); free(p); free(r);
    if (error >= tol) return NOT_CONVERGED_CODE;
    return i;
}

int orc_cr(int n, const int *rp, const int *ci, const double *v,
           const double *b, double tol, int max_iter, double *x) {
    int i;
    double error = 0, alpha, beta, norm_b, r_dot, r_dot_old;
    double *r = dalloc(n), *p = dalloc(n), *ap = dalloc(n), *tmp = dalloc(n);
    orc_csrsymv(n, v, rp, ci, x, tmp);                      /* cr.c:23-34 */
    orc_dscal(n, -1, tmp);
    orc_daxpy(n, 1, b, tmp);
    orc_dcopy(n, tmp, r);
    orc_dcopy(n, r, p);
    orc_csrsymv(n, v, rp, ci, r, tmp);
    orc_dcopy(n, tmp, ap);
    r_dot_old = orc_ddot(n, r, tmp);
    norm_b = orc_dnrm2(n, b);
    for (i = 1; i < max_iter; i++) {                         /* cr.c:37-68 */
        alpha = r_dot_old / orc_ddot(n, ap, ap);
        orc_daxpy(n, alpha, p, x);
        orc_daxpy(n, -alpha, ap, r);
        orc_csrsymv(n, v, rp, ci, r, tmp);
        r_dot = orc_ddot(n, r, tmp);
        beta = r_dot / r_dot_old;
        r_dot_old = r_dot;
        orc_dscal(n, beta, p);
        orc_daxpy(n, 1, r, p);
        orc_dscal(n, beta, ap);
        orc_daxpy(n, 1, tmp, ap);
        error = orc_dnrm2(n, r) / norm_b;
        if (error < tol) break;
    }
    free(tmp); free(ap); free(p); free(r);
    if (error >= tol) return NOT_CONVERGED_CODE;
    return i;
}

int orc_bicg(int n, const int *rp, const int *ci, const double *v,
             const double *b, double tol, int max_iter, double *x) {
    int i;
    double error = 0, alpha, beta, norm_b, r_dot, r_dot_old;
    double *r = dalloc(n), *sr = dalloc(n), *p = dalloc(n), *sp = dalloc(n);
    double *ap = dalloc(n), *atp = dalloc(n), *tmp = dalloc(n);
    orc_csrgemv('n', n, v, rp, ci, x, tmp);                 /* bicg.c:23-32 */
    orc_dscal(n, -1, tmp);
    orc_daxpy(n, 1, b, tmp);
    orc_dcopy(n, tmp, r);
    orc_dcopy(n, tmp, sr);
    orc_dcopy(n, r, p);
    orc_dcopy(n, sr, sp);
    r_dot_old = orc_ddot(n, r, sr);
    norm_b = orc_dnrm2(n, b);
    for (i = 0; i < max_iter; i++) {                         /* bicg.c:35-73 */
        orc_csrgemv('n', n, v, rp, ci, p, ap);
        orc_csrgemv('t', n, v, rp, ci, sp, atp);
        alpha = r_dot_old / orc_ddot(n, ap, sp);
        orc_daxpy(n, alpha, p, x);
        orc_daxpy(n, -alpha, ap, r);
        orc_daxpy(n, -alpha, atp, sr);
        r_dot = orc_ddot(n, r, sr);
        beta = r_dot / r_dot_old;
        r_dot_old = r_dot;
        orc_dscal(n, beta, p);
        orc_daxpy(n, 1, r, p);
        orc_dscal(n, beta, sp);
        orc_daxpy(n, 1, sr, sp);
        error = orc_dnrm2(n, r) / norm_b;
        if (error < tol) break;
    }
    free(tmp); free(atp); free(ap); free(sp); free(p); free(sr); free(r);
    if (error >= tol) return NOT_CONVERGED_CODE;
    return i;
}

int orc_bicgstabl(int n, const int *rp, const int *ci, const double *v,
                  const double *b, int l, double tol, int max_iter, double *x) {
    int i, j, m, l2 = l + 1;
    double error = 0, beta, gamma1, fnb, alpha = 0, gamma0 = 1, t_local;
    double *base = dalloc((size_t)n * (l2 * 2 + 2));
    double **r = (double **)malloc(sizeof(double *) * l2 * 2), **p = r + l2;
    for (i = 0; i < l2; i++) { r[i] = base + (size_t)n * i; p[i] = base + (size_t)n * (l2 + i); }
    double *rs = base + (size_t)n * l2 * 2, *tmp = base + (size_t)n * (l2 * 2 + 1);
    double *wb = dalloc((size_t)l2 * 4);
    double *omega = wb, *omega_d = wb + l2, *omega_dd = wb + 2 * l2, *sigma = wb + 3 * l2;
    double *tau = dalloc((size_t)l2 * l2);
    omega[l] = 1;                                            /* bicgstabl.c:38 */

    orc_csrgemv('n', n, v, rp, ci, x, tmp);                 /* :44-51 */
    orc_dscal(n, -1, tmp);
    orc_daxpy(n, 1, b, tmp);
    orc_dcopy(n, tmp, r[0]);
    orc_dcopy(n, r[0], p[0]);
    orc_dcopy(n, r[0], rs);
    fnb = orc_dnrm2(n, b);

    for (m = 0; m < max_iter; m++) {
        gamma0 = -omega[l] * gamma0;                         /* :57 */
        for (j = 0; j < l; j++) {                            /* BiCG part :59-80 */
            gamma1 = orc_ddot(n, rs, r[j]);
            beta = -alpha * gamma1 / gamma0;
            gamma0 = gamma1;
            for (i = 0; i <= j; i++) {
                orc_dscal(n, beta, p[i]);
                orc_daxpy(n, 1, r[i], p[i]);
            }
            orc_csrgemv('n', n, v, rp, ci, p[j], p[j + 1]);
            alpha = gamma1 / orc_ddot(n, rs, p[j + 1]);
            for (i = 0; i <= j; i++) orc_daxpy(n, -alpha, p[i + 1], r[i]);
            orc_csrgemv('n', n, v, rp, ci, r[j], r[j + 1]);
            orc_daxpy(n, alpha, p[0], x);
        }
        for (j = 1; j <= l; j++) {                           /* MR part :84-91 */
            for (i = 1; i < j; i++) {
                tau[i * l + j] = orc_ddot(n, r[j], r[i]) / sigma[i];
                orc_daxpy(n, -tau[i * l + j], r[i], r[j]);
            }
            sigma[j] = orc_ddot(n, r[j], r[j]);
            omega_d[j] = orc_ddot(n, r[0], r[j]) / sigma[j];
        }
        omega[l] = omega_d[l];                               /* :93-109 */
        for (j = l - 1; j > 0; j--) {
            t_local = 0;
            for (i = j + 1; i <= l; i++) t_local += tau[j * l + i] * omega[i];
            omega[j] = omega_d[j] - t_local;
        }
        for (j = 1; j < l; j++) {
            t_local = 0;
            for (i = j + 1; i < l; i++) t_local += tau[j * l + i] * omega[i + 1];
            omega_dd[j] = omega[j + 1] + t_local;
        }
        orc_daxpy(n, omega[1], r[0], x);                     /* update :113-120 */
        orc_daxpy(n, -omega_d[l], r[l], r[0]);
        orc_daxpy(n, -omega[l], p[l], p[0]);
        for (i = 1; i < l; i++) {
            orc_daxpy(n, -omega[i], p[i], p[0]);
            orc_daxpy(n, -omega_d[i], r[i], r[0]);
            orc_daxpy(n, omega_dd[i], r[i], x);
        }
        error = orc_dnrm2(n, r[0]) / fnb;                    /* :122-125 */
        if (error < tol) break;
    }
    free(tau); free(wb); free(r); free(base);
    if (error >= tol) return NOT_CONVERGED_CODE;
    return m;
}

int orc_shbicgstabl(int n, const int *rp, const int *ci, const double *v,
                    const double *b, int l, const double *sigma, int ns,
                    double tol, int max_iter, double *x) {
    int iter, i, j, k, l2 = l + 1;
    double error = 0, alpha, beta, alpha_old, rho, rho_old, norm_b, t_local;
    double *r[32], *p[32];
    if (l2 > 32) return -3;
    for (i = 0; i < l2; i++) { r[i] = dalloc(n); p[i] = dalloc(n); }
    double *rs = dalloc(n), *tmp = dalloc(n > l2 ? n : l2);
    double *omega = dalloc(l), *omega_d = dalloc(l), *omega_dd = dalloc(l);
    double *s = dalloc(l), *tau = dalloc((size_t)l * l);
    omega[l - 1] = 1;                                        /* shbicgstabl.c:24-25 */

    orc_csrgemv('n', n, v, rp, ci, x, tmp);                 /* :32-42 */
    orc_dscal(n, -1, tmp);
    orc_daxpy(n, 1, b, tmp);
    orc_dcopy(n, tmp, r[0]);
    orc_dcopy(n, r[0], p[0]);
    orc_dcopy(n, r[0], rs);
    norm_b = orc_dnrm2(n, b);
    alpha = 0.0; alpha_old = 1.0; rho_old = 1.0;

    double *alpha_s = dalloc(ns), *beta_s = dalloc(ns), *pi = dalloc(ns);
    double *pi_new = dalloc(ns), *pi_old = dalloc(ns), *theta = dalloc(ns);
    double *pbase = dalloc((size_t)n * l2 * ns);             /* :53 calloc */
    double *om_s = dalloc(l), *om_sd = dalloc(l), *om_sdd = dalloc(l);
    double *M = dalloc((size_t)ns * l * l);
    double psi, e;
#define PS(k_, i_) (pbase + ((size_t)(k_) * l2 + (i_)) * (size_t)n)
#define XS(k_) (x + (size_t)((k_) + 1) * n)
    for (i = 0; i < ns; i++) { pi_old[i] = 1; pi[i] = 1; theta[i] = 1; }
    for (k = 0; k < ns; k++)                                 /* :75-81 */
        for (j = 0; j < l; j++)
            for (i = 0; i <= j; i++)
                M[(size_t)k * l * l + j * l + i] = pow(sigma[k], j - i) * orc_comb(j, i);

    for (iter = 0; iter < max_iter; iter++) {
        rho_old = -omega[l - 1] * rho_old;                   /* :88 */
        for (j = 0; j < l; j++) {
            rho = orc_ddot(n, r[j], rs);                     /* :91-108 */
            beta = alpha * rho / rho_old;
            for (i = 0; i <= j; i++) {
                orc_dscal(n, -beta, p[i]);
                orc_daxpy(n, 1, r[i], p[i]);
            }
            orc_csrgemv('n', n, v, rp, ci, p[j], p[j + 1]);
            alpha = rho / orc_ddot(n, rs, p[j + 1]);
            orc_daxpy(n, alpha, p[0], x);
            for (k = 0; k < ns; k++) {                       /* :111-134 */
                pi_new[k] = (1.0 + alpha * sigma[k]) * pi[k]
                    + alpha * beta / alpha_old * (pi_old[k] - pi[k]);
                beta_s[k] = (pi_old[k] / pi[k]) * (pi_old[k] / pi[k]) * beta;
                alpha_s[k] = pi[k] / pi_new[k] * alpha;
                for (i = 0; i <= j; i++) {
                    orc_dscal(n, -beta_s[k], PS(k, i));
                    orc_daxpy(n, 1.0 / (pi[k] * theta[k]), r[i], PS(k, i));
                }
                orc_dcopy(n, r[j], PS(k, j + 1));
                orc_dscal(n, 1.0 / pi[k], PS(k, j + 1));
                orc_daxpy(n, alpha_s[k], PS(k, 0), XS(k));
            }
            for (i = 0; i <= j; i++) orc_daxpy(n, -alpha, p[i + 1], r[i]);   /* :137-139 */
            orc_csrgemv('n', n, v, rp, ci, r[j], r[j + 1]);
            for (k = 0; k < ns; k++) {                       /* :146-153 */
                orc_daxpy(n, -1.0 / pi_new[k], r[j], PS(k, j + 1));
                orc_dscal(n, 1.0 / (theta[k] * alpha_s[k]), PS(k, j + 1));
                orc_daxpy(n, -sigma[k], PS(k, j), PS(k, j + 1));
            }
            alpha_old = alpha;                               /* :156-161 */
            rho_old = rho;
            for (k = 0; k < ns; k++) { pi_old[k] = pi[k]; pi[k] = pi_new[k]; }
        }
        for (j = 0; j < l; j++) {                            /* MR :166-173 */
            for (i = 0; i < j; i++) {
                tau[i * l + j] = orc_ddot(n, r[j + 1], r[i + 1]) / s[i];
                orc_daxpy(n, -tau[i * l + j], r[i + 1], r[j + 1]);
            }
            s[j] = orc_ddot(n, r[j + 1], r[j + 1]);
            omega_d[j] = orc_ddot(n, r[0], r[j + 1]) / s[j];
        }
        omega[l - 1] = omega_d[l - 1];                       /* :175-191 */
        for (j = l - 2; j >= 0; j--) {
            t_local = 0;
            for (i = j + 1; i < l; i++) t_local += tau[j * l + i] * omega[i];
            omega[j] = omega_d[j] - t_local;
        }
        for (j = 0; j < l - 1; j++) {
            t_local = 0;
            for (i = j + 1; i < l - 1; i++) t_local += tau[j * l + i] * omega[i + 1];
            omega_dd[j] = omega[j + 1] + t_local;
        }
        orc_daxpy(n, omega[0], r[0], x);                     /* :194-199 */
        orc_daxpy(n, -omega[l - 1], p[l], p[0]);
        for (i = 1; i < l; i++) {
            orc_daxpy(n, -omega[i - 1], p[i], p[0]);
            orc_daxpy(n, omega_dd[i - 1], r[i], x);
        }
        for (k = 0; k < ns; k++) {                           /* :202-237 */
            orc_dhorner(omega, l, sigma[k], tmp);
            psi = tmp[0];
            for (i = 0; i < l; i++) om_s[i] = tmp[i + 1];
            e = theta[k] * pi[k];
            theta[k] = theta[k] * psi;
            for (j = 0; j < l; j++) {
                double acc = 0;
                for (i = j; i < l; i++) acc += M[(size_t)k * l * l + i * l + j] * om_s[i];
                om_sd[j] = acc;
            }
            for (j = 0; j < l - 1; j++) {
                double acc = 0;
                for (i = j + 1; i < l - 1; i++) acc += tau[j * l + i] * om_sd[i + 1];
                om_sdd[j] = om_sd[j + 1] + acc;
            }
            orc_daxpy(n, om_sd[0] / e, r[0], XS(k));
            orc_daxpy(n, -omega[l - 1], PS(k, l), PS(k, 0));
            for (i = 1; i < l; i++) {
                orc_daxpy(n, -omega[i - 1], PS(k, i), PS(k, 0));
                orc_daxpy(n, om_sdd[i - 1] / e, r[i], XS(k));
            }
            orc_dscal(n, 1.0 / psi, PS(k, 0));
        }
        for (i = 1; i <= l; i++) orc_daxpy(n, -omega_d[i - 1], r[i], r[0]);   /* :240-242 */
        error = orc_dnrm2(n, r[0]) / norm_b;
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

/* src/bicgstab.c:5-72: BiCGSTAB. beta = alpha / omega * rho_new / rho evaluated
 * left to right (:56); x gets alpha p and omega s as two separate updates (:49-50). */
int orc_bicgstab(int n, const int *rp, const int *ci, const double *v,
                 const double *b, double tol, int max_iter, double *x) {
    int i;
    double error = 0, alpha, beta, omega, rho, rho_new, norm_b;
    double *base = dalloc((size_t)n * 7);
    double *r = base, *rs = base + n, *s = base + 2 * (size_t)n, *p = base + 3 * (size_t)n;
    double *ap = base + 4 * (size_t)n, *as = base + 5 * (size_t)n, *tmp = base + 6 * (size_t)n;
    orc_csrgemv('n', n, v, rp, ci, x, tmp);                  /* bicgstab.c:22-30 */
    orc_dscal(n, -1, tmp);
    orc_daxpy(n, 1, b, tmp);
    orc_dcopy(n, tmp, r);
    orc_dcopy(n, r, p);
    orc_dcopy(n, r, rs);
    norm_b = orc_dnrm2(n, b);
    rho = orc_ddot(n, rs, r);
    for (i = 0; i < max_iter; i++) {                         /* bicgstab.c:33-65 */
        orc_csrgemv('n', n, v, rp, ci, p, ap);
        alpha = rho / orc_ddot(n, rs, ap);
        orc_dcopy(n, r, s);
        orc_daxpy(n, -alpha, ap, s);
        orc_csrgemv('n', n, v, rp, ci, s, as);
        omega = orc_ddot(n, as, s) / orc_ddot(n, as, as);
        orc_daxpy(n, alpha, p, x);
        orc_daxpy(n, omega, s, x);
        orc_dcopy(n, s, r);
        orc_daxpy(n, -omega, as, r);
        rho_new = orc_ddot(n, rs, r);
        beta = alpha / omega * rho_new / rho;
        rho = rho_new;
        orc_daxpy(n, -omega, ap, p);
        orc_dscal(n, beta, p);
        orc_daxpy(n, 1, r, p);
        error = orc_dnrm2(n, r) / norm_b;
        if (error < tol) break;
    }
    free(base);
    if (error >= tol) return NOT_CONVERGED_CODE;
    return i;
}

/* src/shbicg.c:5-129: shifted BiCG. x holds n*(ns+1) entries, the seed first and
 * shift k at x + (k+1) n; the seed loop is BiCG with beta = -rho/rho_old (:53) and
 * p = -beta p + r (:56-57); each shift k carries pi_k, p_k and x_k (:78-96). */
int orc_shbicg(int n, const int *rp, const int *ci, const double *v, const double *b,
               const double *sigma, int ns, double tol, int max_iter, double *x) {
    int i, k;
    double error = 0, alpha, beta, alpha_old, rho, rho_old, fnb;
    double *base = dalloc((size_t)n * 7);
    double *r = base, *sr = base + n, *p = base + 2 * (size_t)n, *sp = base + 3 * (size_t)n;
    double *ap = base + 4 * (size_t)n, *atp = base + 5 * (size_t)n, *tmp = base + 6 * (size_t)n;
    double *shb = dalloc((size_t)ns * 5);
    double *alpha_shift = shb, *beta_shift = shb + ns, *pi = shb + 2 * ns, *pi_new = shb + 3 * ns,
           *pi_old = shb + 4 * ns;
    double *ps = dalloc((size_t)n * ns);
    orc_csrgemv('n', n, v, rp, ci, x, tmp);                  /* shbicg.c:21-29 */
    orc_dscal(n, -1, tmp);
    orc_daxpy(n, 1, b, tmp);
    orc_dcopy(n, tmp, r);
    orc_dcopy(n, tmp, sr);
    fnb = orc_dnrm2(n, b);
    alpha_old = 1;
    rho_old = 1;
    for (k = 0; k < ns; k++) { pi_old[k] = 1; pi[k] = 1; }  /* :41-46 */
    for (i = 0; i < max_iter; i++) {                         /* :49-116 */
        rho = orc_ddot(n, r, sr);
        beta = -rho / rho_old;
        orc_dscal(n, -beta, p);
        orc_daxpy(n, 1, r, p);
        orc_dscal(n, -beta, sp);
        orc_daxpy(n, 1, sr, sp);
        orc_csrgemv('n', n, v, rp, ci, p, ap);
        orc_csrgemv('t', n, v, rp, ci, sp, atp);
        alpha = rho / orc_ddot(n, ap, sr);
        orc_daxpy(n, alpha, p, x);
        for (k = 0; k < ns; k++) {
            double *pk = ps + (size_t)k * n, *xk = x + (size_t)(k + 1) * n;
            pi_new[k] = (1 + alpha * sigma[k]) * pi[k] + alpha * beta / alpha_old * (pi_old[k] - pi[k]);
            beta_shift[k] = (pi_old[k] / pi[k]) * (pi_old[k] / pi[k]) * beta;
            alpha_shift[k] = pi[k] / pi_new[k] * alpha;
            orc_dscal(n, -beta_shift[k], pk);
            orc_daxpy(n, 1.0 / pi[k], r, pk);
            orc_daxpy(n, alpha_shift[k], pk, xk);
        }
        orc_daxpy(n, -alpha, ap, r);
        orc_daxpy(n, -alpha, atp, sr);
        alpha_old = alpha;
        rho_old = rho;
        for (k = 0; k < ns; k++) { pi_old[k] = pi[k]; pi[k] = pi_new[k]; }
        error = orc_dnrm2(n, r) / fnb;
        if (error < tol) break;
    }
    free(ps); free(shb); free(base);
    if (error >= tol) return NOT_CONVERGED_CODE;
    return i;
}
