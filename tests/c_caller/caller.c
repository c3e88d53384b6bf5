/*
 * caller.c -- a C program written against include/bksp.h the way the reference's
 * own driver (tests/test.c:51-310) uses the library: Matrix Market input through
 * read_matrix_market, host arrays, dmat blocks from darray2dmat / dmat_create, and
 * every real solver of that driver with its scenario (b = e1, 16 right-hand sides,
 * tol 1e-14). It includes nothing but bksp.h and the kernel symbols' header and
 * links libbksp_amd.so in place of the reference library, unchanged.
 *
 *   bksp_caller SYMMETRIC.mtx GENERAL.mtx OUTDIR
 *
 * For each solve it appends "name ret n cols" to OUTDIR/results.txt and writes the
 * solution (column-major doubles) to OUTDIR/name.bin; the right-hand-side block goes
 * to OUTDIR/B.bin, so a checker can rerun every scenario on the same data.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "bksp.h"
#include "bksp_kernels.h"

static FILE *g_res;
static const char *g_dir;

static void save(const char *name, int ret, const double *x, int n, int cols) {
    char path[1024];
    snprintf(path, sizeof path, "%s/%s.bin", g_dir, name);
    FILE *f = fopen(path, "wb");
    if (!f || fwrite(x, sizeof(double), (size_t)n * cols, f) != (size_t)n * cols) {
        fprintf(stderr, "cannot write %s\n", path);
        exit(2);
    }
    fclose(f);
    fprintf(g_res, "%s %d %d %d\n", name, ret, n, cols);
    printf("%-14s returned %d\n", name, ret);
}

/* Seeded U(0,1) stream (xorshift64*), in place of the driver's LAPACKE_dlarnv. */
static double unif(unsigned long long *st) {
    unsigned long long x = *st;
    x ^= x >> 12; x ^= x << 25; x ^= x >> 27;
    *st = x;
    return (double)((x * 2685821657736338717ULL) >> 11) * (1.0 / 9007199254740992.0);
}

int main(int argc, char **argv) {
    if (argc < 4) {
        fprintf(stderr, "usage: %s SYMMETRIC.mtx GENERAL.mtx OUTDIR\n", argv[0]);
        return 1;
    }
    g_dir = argv[3];
    char path[1024];
    snprintf(path, sizeof path, "%s/results.txt", g_dir);
    g_res = fopen(path, "w");
    if (!g_res) return 2;

    dspmat *S = read_matrix_market(argv[1]);
    dspmat *G = read_matrix_market(argv[2]);
    printf("%s: %d x %d, %s, ||A||_F = %.15e\n", argv[1], dspmat_rowsize(S), dspmat_colsize(S), S->type,
           sp_norm_f(S));
    printf("%s: %d x %d, %s, ||A||_F = %.15e\n", argv[2], dspmat_rowsize(G), dspmat_colsize(G), G->type,
           sp_norm_f(G));
    fprintf(g_res, "fnorm_sym %.17e\nfnorm_gen %.17e\n", sp_norm_f(S), sp_norm_f(G));

    /* symmetric scenarios: b = e1, tol 1e-14, max_iter 500 */
    const int ns = dspmat_rowsize(S);
    double *bs = calloc((size_t)ns, sizeof(double));
    double *xs = calloc((size_t)ns, sizeof(double));
    bs[0] = 1.0;
    save("cg", bksp_dcg(S, bs, 1.0e-14, 500, xs), xs, ns, 1);
    memset(xs, 0, sizeof(double) * ns);
    save("cr", dcr(S, bs, 1.0e-14, 500, xs), xs, ns, 1);
    const int nu = 16;
    dmat *BI = dmat_create(ns, nu), *XI = dmat_create(ns, nu);
    for (int j = 0; j < nu; j++) dmat_set(BI, j, j, 1.0);
    save("blcg", dblcg(S, BI, 1.0e-14, 500, XI), dmat_value(XI), ns, nu);
    dmat_free(BI);
    dmat_free(XI);

    /* general scenarios */
    const int n = dspmat_rowsize(G), s = 16;
    double *b = calloc((size_t)n, sizeof(double));
    double *x = calloc((size_t)n, sizeof(double));
    b[0] = 1.0;
    save("bicg", dbicg(G, b, 1.0e-14, 500, x), x, n, 1);
    memset(x, 0, sizeof(double) * n);
    save("bicgstab", dbicgstab(G, b, 1.0e-14, 500, x), x, n, 1);
    memset(x, 0, sizeof(double) * n);
    save("bicgstabl", dbicgstabl(G, b, 4, 1.0e-14, 500, x), x, n, 1);

    double *bb = calloc((size_t)n * s, sizeof(double));
    unsigned long long st = 0x9E3779B97F4A7C15ULL;
    for (long long k = 0; k < (long long)n * s; k++) bb[k] = unif(&st);
    for (int j = 0; j < s; j++) bb[(size_t)(n + 1) * j] = 1.0;
    snprintf(path, sizeof path, "%s/B.bin", g_dir);
    FILE *fb = fopen(path, "wb");
    if (!fb || fwrite(bb, sizeof(double), (size_t)n * s, fb) != (size_t)n * s) return 2;
    fclose(fb);
    dmat B;
    darray2dmat(bb, n, s, &B);
    dmat *X = dmat_create(n, s);
    save("blbicgstab", dblbicgstab(G, &B, 1.0e-14, 500, X), dmat_value(X), n, s);
    dmat *XX = dmat_create(n, s);
    save("blbicgstabl", dblbicgstabl(G, &B, 6, 1.0e-14, 500, XX), dmat_value(XX), n, s);

    double sigma[4] = {0.0, 0.2, 0.4, 0.8};
    double *sx = calloc((size_t)n * 5, sizeof(double));
    save("shbicg", dshbicg(G, b, sigma, 4, 1.0e-14, 500, sx), sx, n, 5);
    free(sx);
    sx = calloc((size_t)n * 5, sizeof(double));
    save("shbicgstabl", dshbicgstabl(G, b, 1, sigma, 4, 1.0e-14, 200, sx), sx, n, 5);

    /* the exported kernel symbols a test.c-style residual uses (bksp_internal.h wrappers) */
    double *y = calloc((size_t)n, sizeof(double));
    for (int i = 0; i < n; i++) x[i] = unif(&st);
    snprintf(path, sizeof path, "%s/spmv_x.bin", g_dir);
    fb = fopen(path, "wb");
    if (!fb || fwrite(x, sizeof(double), (size_t)n, fb) != (size_t)n) return 2;
    fclose(fb);
    save("spmv_n", dcsrgemm0('n', n, 1, G->value, G->row, G->col, G->row[n], x, y), y, n, 1);
    save("spmv_t", dcsrgemm0('t', n, 1, G->value, G->row, G->col, G->row[n], x, y), y, n, 1);

    free(y); free(sx); free(bb); free(x); free(b); free(xs); free(bs);
    dmat_free(X);
    dmat_free(XX);
    fclose(g_res);
    return 0;
}
