/*
 * oracle.h -- CPU restatement of the bksp solvers, used ONLY as test infrastructure.
 *
 * This library is the checker for the GPU product. Only tests/, the smoke() entry
 * point and the cpu_baseline leg of bench.py may load it. The product library never
 * links, loads or falls back to it.
 *
 * Every routine restates one function of the reference (aldente39/bksp) call for
 * call, in the reference's operation order, citing the file:line it follows.
 * Matrices are plain CSR arrays (0-based row_ptr[n+1], col[nnz], val[nnz]); a
 * "symmetric" matrix holds ONE stored (lower) triangle, as in the reference.
 * Dense blocks are column-major with leading dimension = rows.
 *
 * Third-party arithmetic (CBLAS / LAPACKE in the reference) is restated from the
 * published Netlib reference algorithms in orc_blas.c; it is pinned by invariants
 * (QR: Q^T Q = I, QR = V; LU: PA = LU; nrm2 against sqrt(sum)), not bit-for-bit
 * against OpenBLAS, whose summation order is not specified.
 */
#ifndef BKSP_ORACLE_H
#define BKSP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- BLAS / LAPACK semantics (orc_blas.c) ---- */
double orc_ddot(int n, const double *x, const double *y);
void   orc_daxpy(int n, double a, const double *x, double *y);
void   orc_dscal(int n, double a, double *x);
void   orc_dcopy(int n, const double *x, double *y);
double orc_dnrm2(int n, const double *x);
/* C = alpha*op(A)*op(B) + beta*C, column-major; ta/tb = 'N' or 'T'. beta==0 overwrites. */
void   orc_dgemm(char ta, char tb, int m, int n, int k, double alpha,
                 const double *A, int lda, const double *B, int ldb,
                 double beta, double *C, int ldc);
int    orc_dgeqrf(int m, int n, double *A, int lda, double *tau);
int    orc_dorgqr(int m, int n, int k, double *A, int lda, const double *tau);
int    orc_dgetrf(int n, double *A, int lda, int *ipiv);
int    orc_dgetri(int n, double *A, int lda, const int *ipiv);
int    orc_dgetrs(int n, int nrhs, const double *LU, int lda, const int *ipiv,
                  double *B, int ldb);
int    orc_dgesv(int n, int nrhs, double *A, int lda, int *ipiv, double *B, int ldb);

/* ---- CSR kernels (orc_matvec.c), restating src/matvec.c ---- */
int orc_dcsrgemm0(char type, int m, int n, const double *A, const int *Ai,
                  const int *Aj, const double *B, double *C);
int orc_dcsrsymm0(char type, int m, int n, const double *A, const int *Ai,
                  const int *Aj, const double *B, double *C);
int orc_dcsrgemm1(char type, double alpha, int m, int n, const double *A,
                  const int *Ai, const int *Aj, const double *B, double beta,
                  double *C);
int orc_dcsrsymm1(char type, double alpha, int m, int n, const double *A,
                  const int *Ai, const int *Aj, const double *B, double beta,
                  double *C);
/* Dispatch wrappers with the semantics of include/bksp_internal.h:35-79. */
void orc_csrsymv(int m, const double *a, const int *ia, const int *ja,
                 const double *x, double *y);
void orc_csrgemv(char trans, int m, const double *a, const int *ia,
                 const int *ja, const double *x, double *y);
/* 0-based CSR in, the 1-based shift of blcg.c:42-49 applied internally. */
void orc_csrmm(char kind, int m, int n, double alpha, const double *val,
               const int *row_ptr0, const int *col0, const double *b,
               double beta, double *c);

/* ---- helpers (orc_util.c) ---- */
double orc_sp_norm_f(int n, const int *rp, const int *ci, const double *v, int symmetric);
int    orc_dhorner(const double *gamma, int l, double sigma, double *out);
int    orc_comb(int n, int r);
int    orc_myqr(int m, int n, double *mat, double *r);

/* ---- solvers (orc_krylov.c, orc_block.c). Return the 0-based break index,
 *      or -1 (NOT_CONVERGED). A symmetric matrix is one stored lower triangle. ---- */
int orc_cg(int n, const int *rp, const int *ci, const double *v,
           const double *b, double tol, int max_iter, double *x);
int orc_cr(int n, const int *rp, const int *ci, const double *v,
           const double *b, double tol, int max_iter, double *x);
int orc_bicg(int n, const int *rp, const int *ci, const double *v,
             const double *b, double tol, int max_iter, double *x);
int orc_bicgstabl(int n, const int *rp, const int *ci, const double *v,
                  const double *b, int l, double tol, int max_iter, double *x);
int orc_shbicgstabl(int n, const int *rp, const int *ci, const double *v,
                    const double *b, int l, const double *sigma, int nsigma,
                    double tol, int max_iter, double *x);
/* src/bicgstab.c:5-72 (BiCGSTAB). */
int orc_bicgstab(int n, const int *rp, const int *ci, const double *v,
                 const double *b, double tol, int max_iter, double *x);
/* src/shbicg.c:5-129 (shifted BiCG); x: n*(nsigma+1), seed first. */
int orc_shbicg(int n, const int *rp, const int *ci, const double *v, const double *b,
               const double *sigma, int nsigma, double tol, int max_iter, double *x);
/* p0_mode 0: calloc reading of blcg.c's uninitialised beta (P0 = 2V);
 * p0_mode 1: corrected P0 = V. */
int orc_blcg(int n, const int *rp, const int *ci, const double *v,
             int s, const double *B, double tol, int max_iter, double *X,
             int p0_mode);
/* Returns the break index; *true_res receives the Frobenius true residual that
 * blbicgstabl.c:215-224 prints. */
int orc_blbicgstabl(int n, const int *rp, const int *ci, const double *v,
                    int s, const double *B, int l, double tol, int max_iter,
                    double *X, double *true_res);
/* src/blbicgstab.c:13-123 (block BiCGSTAB, orthonormalised direction block). */
int orc_blbicgstab(int n, const int *rp, const int *ci, const double *v,
                   int s, const double *B, double tol, int max_iter, double *X);

/* ---- learned FSAI preconditioner (orc_fsai.c); no reference counterpart ---- */
int    orc_fsai_nparam(void);
void   orc_fsai_rows(int n, const int *rp, const int *ci, const double *v, double *sqd, double *s,
                     double *nlow, double *excess);
void   orc_fsai_edge_features(int n, int i, int j, double aij, const double *sqd, const double *s,
                              const double *nlow, const double *excess, float *f);
void   orc_fsai_forward(int n, const int *rpl, const int *cil, const double *vl, const double *sqd,
                        const double *s, const double *nlow, const double *excess,
                        const float *params, int bf16, double *h_out);
void   orc_fsai_gvals(int n, const int *rpl, const int *cil, const double *sqd, const double *h,
                      double *g);
double orc_fsai_loss_dg(int n, const int *rp, const int *ci, const double *v, const int *rpl,
                        const int *cil, const double *g, int P, const double *Z, double *dg);
void   orc_fsai_backward(int n, const int *rpl, const int *cil, const double *vl,
                         const double *sqd, const double *s, const double *nlow,
                         const double *excess, const float *params, int bf16, const double *h,
                         const double *dg, double *grad);

/* ---- complex FP64 family (orc_complex.c); complex values are C99 double complex
 *      (interleaved re, im pairs), passed as double* in this header (orc_complex.c
 *      itself defines them on double complex and skips these declarations). ---- */
#ifndef ORC_COMPLEX_IMPL
void   orc_zdotc_sub(int n, const double *x, const double *y, double *out);
double orc_dznrm2(int n, const double *x);
int    orc_zcsrgemm0(char type, int m, int n, const double *A, const int *Ai, const int *Aj,
                     const double *B, double *C);
int    orc_zcsrgemm1(char type, const double *alpha, int m, int n, const double *A, const int *Ai,
                     const int *Aj, const double *B, const double *beta, double *C);
int    orc_zhorner(const double *gamma, int l, const double *sigma, double *out);
int    orc_zbicg(int n, const int *rp, const int *ci, const double *v, const double *b,
                 double tol, int max_iter, double *x);
int    orc_zbicgstabl(int n, const int *rp, const int *ci, const double *v, const double *b,
                      int l, double tol, int max_iter, double *x);
int    orc_zshbicgstabl(int n, const int *rp, const int *ci, const double *v, const double *b,
                        int l, const double *sigma, int nsigma, double tol, int max_iter, double *x);
int    orc_zqreco(int m, int n, double *mat, double *r);
int    orc_zblbicgstabl(int m, const int *rp, const int *ci, const double *v, int n, const double *B,
                        int l, double tol, int max_iter, double *X, double *true_res);
#endif

/* Time one call of a solver in seconds (CLOCK_MONOTONIC), for the CPU baseline. */
double orc_wall_seconds(void);

#ifdef __cplusplus
}
#endif
#endif
