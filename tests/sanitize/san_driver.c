/*
 * san_driver.c -- runs the host-only C code of the build (Matrix Market reader,
 * seeded generators) and every routine of the CPU oracle on small inputs, so an
 * AddressSanitizer / UndefinedBehaviorSanitizer build of those files reports any
 * out-of-bounds access, leak or undefined operation. Built and run by
 * tests/test_sanitize.py:
 *
 *   san_driver SYMMETRIC.mtx GENERAL.mtx
 *
 * Exit status 0 when every call returned; the sanitizers abort on their first finding.
 */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "bksp.h"
#include "bksp_gen.h"
#include "../../oracle/oracle.h"

static double *dvec(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

/* read_matrix_market's arrays (its type and format strings are literals). */
static void free_dspmat(dspmat *A) {
    free(A->row); free(A->col); free(A->value); free(A->nnz); free(A->row_size); free(A->col_size);
    free(A);
}

int main(int argc, char **argv) {
    if (argc < 3) return 2;
    /* host Matrix Market reader (read_matrix.c semantics) on the reference's fixtures */
    dspmat *S = read_matrix_market(argv[1]);
    dspmat *G = read_matrix_market(argv[2]);
    const int ns = *S->row_size, ng = *G->row_size;
    printf("read %d (%s) and %d (%s)\n", ns, S->type, ng, G->type);
    double *b = dvec(ng), *x = dvec((size_t)ng * 5);
    b[0] = 1.0;
    /* oracle solvers on the fixtures */
    printf("cg %d\n", orc_cg(ns, S->row, S->col, S->value, b, 1e-12, 300, x));
    memset(x, 0, sizeof(double) * ns);
    printf("cr %d\n", orc_cr(ns, S->row, S->col, S->value, b, 1e-12, 300, x));
    memset(x, 0, sizeof(double) * ng);
    printf("bicg %d\n", orc_bicg(ng, G->row, G->col, G->value, b, 1e-12, 300, x));
    memset(x, 0, sizeof(double) * ng);
    printf("bicgstab %d\n", orc_bicgstab(ng, G->row, G->col, G->value, b, 1e-12, 300, x));
    for (int l = 1; l <= 4; l++) {
        memset(x, 0, sizeof(double) * ng);
        printf("bicgstabl(%d) %d\n", l, orc_bicgstabl(ng, G->row, G->col, G->value, b, l, 1e-12, 100, x));
    }
    const double sig[4] = {0.0, 0.2, 0.4, 0.8};
    memset(x, 0, sizeof(double) * ng * 5);
    printf("shbicgstabl %d\n", orc_shbicgstabl(ng, G->row, G->col, G->value, b, 2, sig, 4, 1e-12, 100, x));
    memset(x, 0, sizeof(double) * ng * 5);
    printf("shbicg %d\n", orc_shbicg(ng, G->row, G->col, G->value, b, sig, 4, 1e-12, 300, x));
    const int s = 3;
    double *B = dvec((size_t)ng * s), *X = dvec((size_t)ng * s), tr = 0.0;
    bkh_gen_uniform(41, (long long)ng * s, 0.0, 1.0, B);
    printf("blbicgstabl %d\n", orc_blbicgstabl(ng, G->row, G->col, G->value, s, B, 2, 1e-12, 100, X, &tr));
    memset(X, 0, sizeof(double) * ng * s);
    printf("blbicgstab %d\n", orc_blbicgstab(ng, G->row, G->col, G->value, s, B, 1e-12, 100, X));
    double *Bs = dvec((size_t)ns * s), *Xs = dvec((size_t)ns * s);
    bkh_gen_uniform(42, (long long)ns * s, 0.0, 1.0, Bs);
    printf("blcg %d\n", orc_blcg(ns, S->row, S->col, S->value, s, Bs, 1e-10, 300, Xs, 0));
    /* kernels and helpers */
    double *y = dvec(ng);
    orc_csrgemv('n', ng, G->value, G->row, G->col, b, y);
    orc_csrgemv('t', ng, G->value, G->row, G->col, b, y);
    orc_csrsymv(ns, S->value, S->row, S->col, b, y);
    printf("norms %.6e %.6e\n", orc_sp_norm_f(ns, S->row, S->col, S->value, 1),
           orc_sp_norm_f(ng, G->row, G->col, G->value, 0));
    double R[9], V[30], gam[3] = {0.3, -0.2, 0.1}, hz[4];
    bkh_gen_normal(43, 30, V);
    printf("myqr %d\n", orc_myqr(10, 3, V, R));
    orc_dhorner(gam, 3, 0.4, hz);
    printf("comb %d\n", orc_comb(7, 3));
    /* complex family on a complexified copy of the general fixture */
    double complex *zv = (double complex *)calloc((size_t)G->row[ng], sizeof(double complex));
    double complex *zb = (double complex *)calloc((size_t)ng, sizeof(double complex));
    double complex *zx = (double complex *)calloc((size_t)ng * 3, sizeof(double complex));
    for (int k = 0; k < G->row[ng]; k++) zv[k] = G->value[k] * (1.0 + 0.1 * I);
    zb[0] = 1.0;
    printf("zbicg %d\n", orc_zbicg(ng, G->row, G->col, (const double *)zv, (const double *)zb, 1e-10, 300,
                                   (double *)zx));
    memset(zx, 0, sizeof(double complex) * ng);
    printf("zbicgstabl %d\n", orc_zbicgstabl(ng, G->row, G->col, (const double *)zv, (const double *)zb, 2, 1e-10,
                                             100, (double *)zx));
    const double complex zs[2] = {0.1 + 0.1 * I, 0.5};
    memset(zx, 0, sizeof(double complex) * ng * 3);
    printf("zshbicgstabl %d\n", orc_zshbicgstabl(ng, G->row, G->col, (const double *)zv, (const double *)zb, 2,
                                                 (const double *)zs, 2, 1e-10, 100, (double *)zx));
    /* the seeded generators (every workload's inputs) at small sizes */
    bkh_csr_host m;
    if (bkh_gen_poisson3d(6, 1, &m) == 0) bkh_csr_host_free(&m);
    if (bkh_gen_convdiff2d(8, 10.0, 20.0, &m) == 0) bkh_csr_host_free(&m);
    if (bkh_gen_convdiff3d(6, 0.1, 0.2, 0.4, &m) == 0) bkh_csr_host_free(&m);
    if (bkh_gen_random_rows(404, 300, &m) == 0) bkh_csr_host_free(&m);
    int sizes[8];
    bkh_gen_batch_sizes(404, 8, 10.0, 12.0, sizes);
    double *kappa = dvec(6 * 6 * 6);
    if (bkh_gen_grf(5000, 6, 2.0, 8.0, kappa) == 0 && bkh_gen_fv_diffusion(6, kappa, 1, &m) == 0) {
        /* learned-FSAI oracle on that field: features, forward, loss, backward */
        const int n = m.n;
        double *sqd = dvec(n), *sr = dvec(n), *nl = dvec(n), *ex = dvec(n), *h = dvec(m.nnz), *g = dvec(m.nnz),
               *dg = dvec(m.nnz), *Z = dvec((size_t)2 * n);
        bkh_csr_host full;
        bkh_gen_fv_diffusion(6, kappa, 0, &full);
        orc_fsai_rows(n, full.row_ptr, full.col, full.val, sqd, sr, nl, ex);
        const int np = orc_fsai_nparam();
        float *p = (float *)calloc((size_t)np, sizeof(float));
        double *grad = dvec(np);
        for (int i = 0; i < np; i++) p[i] = 0.01f * (float)((i * 37) % 11 - 5);
        orc_fsai_forward(n, m.row_ptr, m.col, m.val, sqd, sr, nl, ex, p, 1, h);
        orc_fsai_gvals(n, m.row_ptr, m.col, sqd, h, g);
        bkh_gen_uniform(7, 2LL * n, -1.0, 1.0, Z);
        printf("fsai loss %.6f\n", orc_fsai_loss_dg(n, full.row_ptr, full.col, full.val, m.row_ptr, m.col, g, 2, Z, dg));
        orc_fsai_backward(n, m.row_ptr, m.col, m.val, sqd, sr, nl, ex, p, 1, h, dg, grad);
        free(p); free(grad); free(sqd); free(sr); free(nl); free(ex); free(h); free(g); free(dg); free(Z);
        bkh_csr_host_free(&full);
        bkh_csr_host_free(&m);
    }
    free(kappa); free(zv); free(zb); free(zx); free(y); free(B); free(X); free(Bs); free(Xs); free(b); free(x);
    free_dspmat(S);
    free_dspmat(G);
    printf("san_driver done\n");
    return 0;
}
