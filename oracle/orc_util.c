/*
 * orc_util.c -- test infrastructure only (see oracle.h).
 * Restates src/util_shift.c (Horner re-expansion, binomial), src/norm.c and
 * src/myqr.c of the reference.
 */
#include <math.h>
#include <stdlib.h>
#include "oracle.h"

/* util_shift.c:5-27. gamma[0..l-1] are the MR coefficients; out[0..l] receives
 * psi = out[0] and the re-expanded coefficients out[1..l] for the shift sigma. */
int orc_dhorner(const double *gamma, int l, double sigma, double *out) {
    out[l] = -gamma[l - 1];
    for (int i = l - 1; i > 0; i--) out[i] = -sigma * out[i + 1] - gamma[i - 1];
    out[0] = -sigma * out[1] + 1.0;
    for (int i = 1; i < l; i++)
        for (int j = l - 1; j >= i; j--) out[j] += -sigma * out[j + 1];
    for (int i = 1; i <= l; i++) out[i] = -out[i] / out[0];
    return 0;
}

/* util_shift.c:55-63 (Pascal's rule; iterative here, same values). */
int orc_comb(int n, int r) {
    if (n < 0 || r < 0 || n < r) return 0;
    if (r > n - r) r = n - r;
    long long c = 1;
    for (int i = 1; i <= r; i++) c = c * (n - r + i) / i;
    return (int)c;
}

/* norm.c:5-47. The symmetric branch counts each stored off-diagonal twice. */
double orc_sp_norm_f(int n, const int *rp, const int *ci, const double *v, int symmetric) {
    double s = 0;
    if (!symmetric) {
        for (int i = 0; i < rp[n]; i++) s += v[i] * v[i];
        return sqrt(s);
    }
    for (int i = 0; i < n; i++)
        for (int j = rp[i]; j < rp[i + 1]; j++) {
            double t = v[j] * v[j];
            if (ci[j] == i) s += t; else s += 2 * t;
        }
    return sqrt(s);
}

/* myqr.c:4-39: thin Householder QR. Q overwrites mat (m x n, ld m); r gets the
 * n x n upper-triangular R (column-major), LAPACK sign convention. */
int orc_myqr(int m, int n, double *mat, double *r) {
    double *tau = (double *)malloc((size_t)n * sizeof(double));
    orc_dgeqrf(m, n, mat, m, tau);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            r[i * n + j] = (j <= i) ? mat[(size_t)i * m + j] : 0.0;
    orc_dorgqr(m, n, n, mat, m, tau);
    free(tau);
    return 0;
}
