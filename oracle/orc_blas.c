/*
 * orc_blas.c -- test infrastructure only (see oracle.h).
 *
 * The reference links CBLAS and LAPACKE (not vendored, no pinned version:
 * Makefile:8-17). The routines below restate the published Netlib reference
 * algorithms of the calls the solvers make, in their documented operation order:
 *   ddot/daxpy/dscal/dcopy  -- straight loops (Netlib reference BLAS);
 *   dnrm2                   -- the scale/ssq recurrence of Netlib dnrm2 (pre-3.10);
 *   dgemm                   -- textbook triple loop, column-major;
 *   dgeqrf/dorgqr           -- unblocked Householder dgeqr2/dorg2r with dlarfg's sign;
 *   dgetrf/dgetrs/dgesv     -- unblocked partial-pivoting dgetf2 + dgetrs;
 *   dgetri                  -- inverse by solving A X = I with the LU factors
 *                              (LAPACK's dtrtri route differs only in rounding).
 * One deliberate convention: dscal with a == 0 writes exact zeros (the reference
 * relies on it to clear malloc'd buffers, bksp_internal.h:39 and :73).
 *
 * Built with -DORC_THREADS -fopenmp (liboracle_mt.so) the vector routines run on all
 * host cores, as a multithreaded BLAS runs the reference's BLAS-1 calls, while the
 * CSR kernels stay serial as the reference's matvec.c does (its Makefile's -openmp is
 * parsed as an output-file option); the reduction order then depends on the thread
 * count. That build is only a timing baseline; the checker is the serial build.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "oracle.h"

#ifdef ORC_THREADS
#define ORC_PAR _Pragma("omp parallel for schedule(static) if (n > 32768)")
#define ORC_PAR_SUM _Pragma("omp parallel for schedule(static) reduction(+ : s) if (n > 32768)")
#else
#define ORC_PAR
#define ORC_PAR_SUM
#endif

double orc_ddot(int n, const double *x, const double *y) {
    double s = 0.0;
    ORC_PAR_SUM
    for (int i = 0; i < n; i++) s += x[i] * y[i];
    return s;
}

void orc_daxpy(int n, double a, const double *x, double *y) {
    if (a == 0.0) return;                 /* Netlib daxpy quick return */
    ORC_PAR
    for (int i = 0; i < n; i++) y[i] += a * x[i];
}

void orc_dscal(int n, double a, double *x) {
    if (a == 0.0) {
        ORC_PAR
        for (int i = 0; i < n; i++) x[i] = 0.0;
        return;
    }
    ORC_PAR
    for (int i = 0; i < n; i++) x[i] = a * x[i];
}

void orc_dcopy(int n, const double *x, double *y) {
#ifdef ORC_THREADS
    if (x + n <= y || y + n <= x) {
        ORC_PAR
        for (int i = 0; i < n; i++) y[i] = x[i];
        return;
    }
#endif
    memmove(y, x, (size_t)n * sizeof(double));
}

double orc_dnrm2(int n, const double *x) {
#ifdef ORC_THREADS
    if (n > 32768) {                      /* max-scaled two-pass norm */
        double mx = 0.0;
        _Pragma("omp parallel for schedule(static) reduction(max : mx)")
        for (int i = 0; i < n; i++) mx = fmax(mx, fabs(x[i]));
        if (mx == 0.0) return 0.0;
        double s = 0.0;
        const double inv = 1.0 / mx;
        _Pragma("omp parallel for schedule(static) reduction(+ : s)")
        for (int i = 0; i < n; i++) s += (x[i] * inv) * (x[i] * inv);
        return mx * sqrt(s);
    }
#endif
    if (n < 1) return 0.0;
    if (n == 1) return fabs(x[0]);
    double scale = 0.0, ssq = 1.0;
    for (int i = 0; i < n; i++) {
        if (x[i] != 0.0) {
            double a = fabs(x[i]);
            if (scale < a) { double q = scale / a; ssq = 1.0 + ssq * q * q; scale = a; }
            else           { double q = a / scale; ssq += q * q; }
        }
    }
    return scale * sqrt(ssq);
}

void orc_dgemm(char ta, char tb, int m, int n, int k, double alpha,
               const double *A, int lda, const double *B, int ldb,
               double beta, double *C, int ldc) {
    for (int j = 0; j < n; j++) {
        for (int i = 0; i < m; i++) {
            double s = 0.0;
            for (int l = 0; l < k; l++) {
                double a = (ta == 'N') ? A[i + (size_t)l * lda] : A[l + (size_t)i * lda];
                double b = (tb == 'N') ? B[l + (size_t)j * ldb] : B[j + (size_t)l * ldb];
                s += a * b;
            }
            double *c = &C[i + (size_t)j * ldc];
            *c = (beta == 0.0) ? alpha * s : alpha * s + beta * *c;
        }
    }
}

static double lapy2(double a, double b) {
    double x = fabs(a), y = fabs(b);
    double w = x > y ? x : y, z = x > y ? y : x;
    if (z == 0.0) return w;
    double q = z / w;
    return w * sqrt(1.0 + q * q);
}

/* dlarfg: H*(alpha;x) = (beta;0), H = I - tau*v*v^T, v = (1; x_scaled). */
static void larfg(int n, double *alpha, double *x, double *tau) {
    if (n <= 1) { *tau = 0.0; return; }
    double xnorm = orc_dnrm2(n - 1, x);
    if (xnorm == 0.0) { *tau = 0.0; return; }
    double beta = -copysign(lapy2(*alpha, xnorm), *alpha);
    *tau = (beta - *alpha) / beta;
    orc_dscal(n - 1, 1.0 / (*alpha - beta), x);
    *alpha = beta;
}

/* dlarf('L'): C(m x n, ldc) := (I - tau v v^T) C. */
static void larf_left(int m, int n, const double *v, double tau, double *C, int ldc) {
    if (tau == 0.0) return;
    for (int j = 0; j < n; j++) {
        double *c = C + (size_t)j * ldc;
        double w = 0.0;
        for (int i = 0; i < m; i++) w += c[i] * v[i];
        for (int i = 0; i < m; i++) c[i] -= tau * w * v[i];
    }
}

int orc_dgeqrf(int m, int n, double *A, int lda, double *tau) {
    int k = m < n ? m : n;
    for (int i = 0; i < k; i++) {
        double *aii = &A[i + (size_t)i * lda];
        larfg(m - i, aii, &A[(i + 1 < m ? i + 1 : m - 1) + (size_t)i * lda], &tau[i]);
        if (i < n - 1) {
            double keep = *aii;
            *aii = 1.0;
            larf_left(m - i, n - i - 1, aii, tau[i], &A[i + (size_t)(i + 1) * lda], lda);
            *aii = keep;
        }
    }
    return 0;
}

int orc_dorgqr(int m, int n, int k, double *A, int lda, const double *tau) {
    for (int j = k; j < n; j++) {
        for (int l = 0; l < m; l++) A[l + (size_t)j * lda] = 0.0;
        A[j + (size_t)j * lda] = 1.0;
    }
    for (int i = k - 1; i >= 0; i--) {
        double *aii = &A[i + (size_t)i * lda];
        if (i < n - 1) {
            *aii = 1.0;
            larf_left(m - i, n - i - 1, aii, tau[i], &A[i + (size_t)(i + 1) * lda], lda);
        }
        if (i < m - 1) orc_dscal(m - i - 1, -tau[i], &A[i + 1 + (size_t)i * lda]);
        *aii = 1.0 - tau[i];
        for (int l = 0; l < i; l++) A[l + (size_t)i * lda] = 0.0;
    }
    return 0;
}

int orc_dgetrf(int n, double *A, int lda, int *ipiv) {
    int info = 0;
    for (int j = 0; j < n; j++) {
        int p = j;
        double best = fabs(A[j + (size_t)j * lda]);
        for (int i = j + 1; i < n; i++) {
            double a = fabs(A[i + (size_t)j * lda]);
            if (a > best) { best = a; p = i; }
        }
        ipiv[j] = p + 1;                       /* LAPACK's 1-based pivots */
        if (A[p + (size_t)j * lda] != 0.0) {
            if (p != j)
                for (int c = 0; c < n; c++) {
                    double t = A[j + (size_t)c * lda];
                    A[j + (size_t)c * lda] = A[p + (size_t)c * lda];
                    A[p + (size_t)c * lda] = t;
                }
            double r = 1.0 / A[j + (size_t)j * lda];
            for (int i = j + 1; i < n; i++) A[i + (size_t)j * lda] *= r;
        } else if (info == 0) {
            info = j + 1;
        }
        for (int c = j + 1; c < n; c++) {
            double u = A[j + (size_t)c * lda];
            for (int i = j + 1; i < n; i++) A[i + (size_t)c * lda] -= A[i + (size_t)j * lda] * u;
        }
    }
    return info;
}

int orc_dgetrs(int n, int nrhs, const double *LU, int lda, const int *ipiv,
               double *B, int ldb) {
    for (int c = 0; c < nrhs; c++) {
        double *b = B + (size_t)c * ldb;
        for (int i = 0; i < n; i++) {
            int p = ipiv[i] - 1;
            if (p != i) { double t = b[i]; b[i] = b[p]; b[p] = t; }
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

int orc_dgetri(int n, double *A, int lda, const int *ipiv) {
    double *lu = (double *)malloc((size_t)n * n * sizeof(double));
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) lu[i + (size_t)j * n] = A[i + (size_t)j * lda];
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) A[i + (size_t)j * lda] = (i == j) ? 1.0 : 0.0;
    orc_dgetrs(n, n, lu, n, ipiv, A, lda);
    free(lu);
    return 0;
}

int orc_dgesv(int n, int nrhs, double *A, int lda, int *ipiv, double *B, int ldb) {
    int info = orc_dgetrf(n, A, lda, ipiv);
    if (info) return info;
    return orc_dgetrs(n, nrhs, A, lda, ipiv, B, ldb);
}

double orc_wall_seconds(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}
