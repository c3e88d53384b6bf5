/*
 * bksp_hip.h -- extension API of the MI355X-native bksp drop-in.
 *
 * Nothing here exists in the reference: bksp.h's entry points take host arrays,
 * upload them, solve on the GPU and download x. This header exposes the device
 * layer underneath for callers that keep data resident in HBM (benchmarks, the
 * precision/batched/preconditioned extensions) and for tests of single kernels.
 * All pointers named d_* are device pointers obtained from bkh_malloc; all other
 * pointers are host pointers. Every call runs on the calling thread's own HIP
 * stream and is synchronous unless stated otherwise.
 */
#ifndef BKSP_HIP_H
#define BKSP_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Extension status codes (bksp.h keeps -1..-4). */
#define BKH_DEVICE_ERROR  (-5)
#define BKH_BAD_ARGUMENT  (-6)

/* Precision of the Krylov iteration. */
#define BKH_FP64     0   /* plain FP64 */
#define BKH_FP32_IR  1   /* FP32 inner solve + FP64 iterative refinement */

/* Preconditioners of bkh_pcg. */
#define BKH_PREC_NONE   0
#define BKH_PREC_JACOBI 1
#define BKH_PREC_FSAI   2

typedef struct bkh_matrix bkh_matrix;   /* opaque device-resident CSR */

typedef struct {
    int    precision;    /* BKH_FP64 (default) or BKH_FP32_IR */
    int    graph_iters;  /* iterations per captured hipGraph; 0 = library default */
    int    fix_blcg_p0;  /* block CG: 1 = first direction P0 = V (the reference's
                            uninitialised beta gives P0 = 2V, the default here) */
    int    max_inner;    /* FP32-IR: inner iteration cap per refinement pass (0 = default) */
    double inner_tol;    /* FP32-IR: inner relative tolerance (0 = default); for CG the
                            residual reduction between reliable updates (default 0.01) */
    int    path;         /* BiCGSTAB(l): 0 auto, 1 multi-kernel graph, 2 one persistent
                            workgroup (systems up to 16384 rows / 400000 nonzeros).
                            CG FP32-IR: 0 auto (one kernel per iteration on SELL
                            matrices), 1 two kernels per iteration.
                            PCG: 0 auto (the fused loop where the matrices allow),
                            1 separate kernels, 2 fused loop */
} bkh_opts;

typedef struct {
    int    ret;          /* value the bksp.h entry point returns */
    int    status;       /* 1 converged, 2 max_iter reached, 3 breakdown/NaN */
    int    iters;        /* loop index at the break (outer index for BiCGSTAB(l)) */
    int    inner_iters;  /* FP32-IR: inner iterations summed over passes */
    int    passes;       /* FP32-IR: refinement passes */
    double rec_res;      /* final recursive residual ratio */
    double true_res;     /* FP64 true residual ||b - A x|| / ||b|| (device) */
    double time_ms;      /* device time of the solve, HIP events on the solve stream */
} bkh_info;

/* ---- runtime ---- */
int         bkh_init(void);
const char *bkh_last_error(void);
int         bkh_device_name(char *buf, int len);
int         bkh_sync(void);

/* ---- device memory (the calling thread's stream) ---- */
void *bkh_malloc(size_t bytes);
void  bkh_free(void *d_ptr);
int   bkh_h2d(void *d_dst, const void *src, size_t bytes);
int   bkh_d2h(void *dst, const void *d_src, size_t bytes);
int   bkh_d2d(void *d_dst, const void *d_src, size_t bytes);
int   bkh_memset(void *d_ptr, int value, size_t bytes);

/* ---- matrices ----
 * stored_triangle = 1: row_ptr/col/val hold one triangle of a symmetric matrix
 * (bksp.h "symmetric"); the device copy is the expanded full matrix whose rows
 * list entries in the reference's accumulation order (src/matvec.c:100-127).
 * want_transpose = 1 also builds A^T (src/matvec.c:29-46 order).
 * want_fp32 = 1 also keeps FP32 values (BKH_FP32_IR). */
bkh_matrix *bkh_matrix_create(int n, const int *row_ptr, const int *col,
                              const double *val, int stored_triangle,
                              int want_transpose, int want_fp32);
void        bkh_matrix_destroy(bkh_matrix *A);
long long   bkh_matrix_nnz(const bkh_matrix *A);
/* Stored entries of the slice-diagonal copy the CG iteration kernels use (banded
 * matrices whose rows ascend; 0: none; BKH_DIA=0 at creation turns it off). */
long long   bkh_matrix_dia_entries(const bkh_matrix *A);
int         bkh_matrix_rows(const bkh_matrix *A);
int         bkh_matrix_blocks(const bkh_matrix *A);
/* SpMV kernel variant: 0 = workgroup CSR-stream (256 rows / 2048 nonzeros per
 * workgroup), 1 = wave CSR-stream (64 rows / 512 nonzeros per wave). Chosen at
 * creation (1 unless a row exceeds 512 nonzeros); the autotuner may override. */
int         bkh_matrix_set_spmv(bkh_matrix *A, int variant);
/* Autotuner (no reference counterpart: the reference's SpMV/SpMM are fixed loops over
 * CSR, src/matvec.c). Times, on this matrix: the SpMV variants it holds (SELL-64, on its
 * slice-diagonal copy when it has one, with 16- or 32-bit columns; the wave CSR-stream
 * variant at rows per wavefront 16, 32 and 64 with nonzero caps -- the row-bin
 * threshold: a longer row gets a block of its own -- from the matrix's row-length
 * histogram (that many rows of median, 90th-percentile and mean length, rounded up to
 * 64, at least the longest row, at most 512); the workgroup CSR-stream variant) and
 * keeps the fastest; then, for s = 1..max_s (<= 16), bkh_spmm's columns per pass
 * (1, 2, 4, 8, 16 and s, at most s) and keeps the fastest per s (for s > 16,
 * bkh_spmm passes in chunks of the width chosen for 16). Every choice gives bitwise
 * the same products. */
typedef struct {
    int spmv_variant;       /* 0 workgroup CSR-stream, 1 wave CSR-stream, 2 SELL-64, 3 SELL-64 int32 */
    int rows_per_wave;      /* wave variant's blocking in use */
    int nnz_per_wave;
    int candidates;         /* SpMV candidates timed */
    double spmv_ms;         /* the chosen SpMV's time */
    double spmv_ms_default; /* the SpMV variant in use before tuning */
    int spmm_width[17];     /* [s]: columns per bkh_spmm pass, s = 1..max_s */
    double spmm_ms[17];     /* [s]: its time */
    double spmm_ms_full[17];/* [s]: time with all s columns in one pass */
} bkh_matrix_tuning;
int         bkh_matrix_tune(bkh_matrix *A, int max_s, bkh_matrix_tuning *out);

/* ---- kernels on device vectors ----
 * y = alpha*op(A)*x + beta*y, op = 'n' (A) or 't' (A^T). beta == 0 overwrites y.
 * Replaces src/matvec.c:11-53 (dcsrgemm0), :100-127 (dcsrsymm0), :130-172
 * (dcsrgemm1) and :219-246 (dcsrsymm1) for one right-hand side. */
int bkh_spmv(const bkh_matrix *A, char op, double alpha, const double *d_x,
             double beta, double *d_y);
/* FP32 values and vectors (precision extension). */
int bkh_spmv32(const bkh_matrix *A, const float *d_x, float *d_y);
/* Y = alpha*A*X + beta*Y for s column-major columns (ld = rows); replaces the
 * mkl_dcsrmm dispatch of include/bksp_internal.h:63-79. */
int bkh_spmm(const bkh_matrix *A, int s, double alpha, const double *d_X,
             double beta, double *d_Y);
/* Thin QR of an m x s column-major device panel, s <= 16 (CholeskyQR2): d_V <- Q,
 * R (host, s x s column-major, positive diagonal). Replaces src/myqr.c:4-39. */
int bkh_qr(int m, int s, double *d_V, double *R);
/* Thin QR of an m x n column-major device panel (n <= m) with LAPACK's Householder
 * convention, the unblocked dgeqr2 + dorg2r that dgeqrf + dorgqr run for n < 32:
 * d_V <- Q, R (host, n x n column-major, upper triangular; R(k,k) = -sign(alpha_k)
 * times the reflected column norm, so it may be negative, and Q's columns carry the
 * matching signs). Rank-deficient panels factor as LAPACK factors them. Returns 0,
 * -1 (m < 0) or -2 (n < 0 or n > m) as dorgqr's info. Replaces the
 * LAPACKE_dgeqrf/dorgqr pair of src/myqr.c:9,30. */
int bkh_qr_householder(int m, int n, double *d_V, double *R);
/* The same for a complex panel (interleaved re, im): zgeqr2 + zung2r, R with a real
 * diagonal of either sign. Replaces LAPACKE_zgeqrf/zungqr of src/myqr.c:46,67. */
int bkh_zqr_householder(int m, int n, double *d_V, double *R);
/* ||A||_F of host CSR arrays, stored-triangle off-diagonals counted twice when
 * symmetric = 1 (replaces src/norm.c:5-47). */
int bkh_sp_norm_f(int n, const int *row_ptr, const int *col, const double *val,
                  int symmetric, double *out);
/* Deterministic FP64 dot product (result to host). */
int bkh_dot(long long n, const double *d_x, const double *d_y, double *out);
/* G = X^T Y (s x s, column-major, to host) for m x s column-major X, Y, on FP64 MFMA. */
int bkh_gram(int m, int s, const double *d_X, const double *d_Y, double *G);
/* d_AP = A d_P and G = P^T A P (s x s, column-major, to host) for m x s column-major
 * d_P, s <= 16, by the path block CG's iteration takes (src/blcg.c:74-79, its
 * mkl_dcsrmm and the cblas_dgemm after it): one fused SpMM + Gram kernel when the
 * matrix allows it, else the SpMM and the Gram pass. *path: 0 separate, 1 fused,
 * 2 fused on the slice-diagonal copy with per-slice records (pipelined). */
int bkh_spmm_gram(const bkh_matrix *A, int s, const double *d_P, double *d_AP, double *G, int *path);

/* ---- solvers on device vectors (x: initial guess in, solution out) ----
 * The return value is what the matching bksp.h entry point returns. */
int bkh_cg(bkh_matrix *A, const double *d_b, double *d_x, double tol, int max_iter,
           const bkh_opts *opts, bkh_info *info);
/* Allocation-free CG. No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). The plain FP64 iteration of bkh_cg (same
 * kernels, bitwise the same x) with its whole workspace in a static device array of
 * the library (about 34 MB per GPU, held until the process ends): no hipMalloc,
 * hipFree, graph, stream or event is created. A must exist already (bkh_matrix_create
 * allocates), must use the fused iteration (SELL-64, or rows of <= 512 nonzeros) and
 * have at most bkh_cg_noalloc_max_rows() rows (2^20); otherwise BKH_BAD_ARGUMENT.
 * Runs on the null stream; calls that overlap on one GPU share the array and race.
 * info->time_ms is host wall time. */
int bkh_cg_noalloc(const bkh_matrix *A, const double *d_b, double *d_x, double tol, int max_iter,
                   bkh_info *info);
long long bkh_cg_noalloc_max_rows(void);
/* Conjugate residual (replaces src/cr.c:6-80); the loop index starts at 1. */
int bkh_cr(bkh_matrix *A, const double *d_b, double *d_x, double tol, int max_iter,
           const bkh_opts *opts, bkh_info *info);
/* BiCG (replaces src/bicg.c:7-88); A must be created with want_transpose = 1. */
int bkh_bicg(bkh_matrix *A, const double *d_b, double *d_x, double tol, int max_iter,
             const bkh_opts *opts, bkh_info *info);
/* BiCGSTAB (replaces src/bicgstab.c:5-72, step for step). */
int bkh_bicgstab(bkh_matrix *A, const double *d_b, double *d_x, double tol, int max_iter,
                 const bkh_opts *opts, bkh_info *info);
/* BiCGSTAB(l), 1 <= l <= 8 (replaces src/bicgstabl.c:13-138). */
int bkh_bicgstabl(bkh_matrix *A, const double *d_b, int l, double *d_x, double tol,
                  int max_iter, const bkh_opts *opts, bkh_info *info);

/* Block CG with residual QR (replaces src/blcg.c:15-119). B, X: m x s column-major
 * device blocks (dmat layout), 1 <= s <= 16. info->true_res is the Frobenius
 * ||B - A X||_F / ||B||_F. */
int bkh_blcg(bkh_matrix *A, int s, const double *d_B, double *d_X, double tol, int max_iter,
             const bkh_opts *opts, bkh_info *info);

/* Block BiCGSTAB(l) with residual QR (replaces src/blbicgstabl.c:16-252). B, X: m x s
 * column-major device blocks, 1 <= s <= 16, 1 <= l <= 8. Returns the outer iteration
 * index at convergence (||xi||_F / ||B||_F < tol). info->true_res is the Frobenius
 * ||B - A X||_F / ||B||_F. The initial residual is B - A X0. */
int bkh_blbicgstabl(bkh_matrix *A, int s, const double *d_B, int l, double *d_X, double tol,
                    int max_iter, const bkh_opts *opts, bkh_info *info);

/* Block BiCGSTAB with an orthonormalised direction block (replaces
 * src/blbicgstab.c:13-123, step for step; host-driven, for link parity). B, X: m x s
 * column-major device blocks, 1 <= s <= 16. Returns the break index, -1, or -2 when
 * Rs^T V is singular. */
int bkh_blbicgstab(bkh_matrix *A, int s, const double *d_B, double *d_X, double tol, int max_iter);
/* Shifted BiCG (replaces src/shbicg.c:4-130); host-driven, for link parity. A must be
 * created with want_transpose = 1; d_x holds n*(ns+1) entries (seed first). */
int bkh_shbicg(bkh_matrix *A, const double *d_b, const double *sigma, int ns, double *d_x,
               double tol, int max_iter);

/* Shifted BiCGSTAB(l) (replaces src/shbicgstabl.c:7-285): solves A x_0 = b and
 * (A + sigma_k I) x_k = b for ns <= 256 shifts with one Krylov space. d_x holds
 * n*(ns+1) entries: the seed solution, then shift k at d_x + (k+1)*n; the shifted
 * parts should enter as 0. Convergence is tested on the seed only, as in the
 * reference. opts->precision = BKH_FP32_IR stores the per-shift direction vectors
 * in FP32 (seed, x and all recurrences stay FP64). shift_true_res (host, ns+1
 * entries, may be NULL) receives every system's FP64 true residual. */
int bkh_shbicgstabl(bkh_matrix *A, const double *d_b, int l, const double *sigma, int ns,
                    double *d_x, double tol, int max_iter, const bkh_opts *opts,
                    bkh_info *info, double *shift_true_res);

/* ---- complex FP64 family. Complex values are interleaved (re, im) doubles, the
 * layout of double _Complex; matrices are general CSR (0-based). */
typedef struct bkh_zmatrix bkh_zmatrix;
/* val holds 2*nnz doubles. A transposed copy is kept for the conjugate transpose. */
bkh_zmatrix *bkh_zmatrix_create(int n, const int *row_ptr, const int *col, const double *val);
void bkh_zmatrix_destroy(bkh_zmatrix *A);
/* Stored positions of the slice-diagonal copies of A and A^T the complex solvers'
 * SpMV uses (0: none; BKH_DIA=0 at creation turns them off). */
long long bkh_zmatrix_dia_entries(const bkh_zmatrix *A);
/* Rows of A and A^T longer than their 64-row slice's SELL width (at most
 * max(16, twice the slice's mean row length)); their remaining entries form a CSR tail
 * summed after the slice, in stored order. */
long long bkh_zmatrix_split_rows(const bkh_zmatrix *A);
/* y = A x (op 'n') or A^H x (op 't'), each row summed in the order of src/matvec.c:55-98
 * (replaces zcsrgemm0 through mkl_cspblas_zcsrgemv, include/bksp_internal.h:54-59). */
int bkh_zspmv(const bkh_zmatrix *A, char op, const double *d_x, double *d_y);
/* Y = beta Y + alpha op(A) X for s columns (replaces zcsrgemm1, src/matvec.c:174-217). */
int bkh_zspmm1(const bkh_zmatrix *A, char op, const double *alpha, int s, const double *d_X,
               const double *beta, double *d_Y);
/* Replace zbicg (src/bicg.c:91-177), zbicgstabl (src/bicgstabl.c:141-276) and
 * zshbicgstabl (src/shbicgstabl.c:288-592; sigma: 2*ns host doubles, d_x: (ns+1)*n
 * complex entries, seed first, shifts entering as 0). Return values as bksp.h. */
int bkh_zbicg(const bkh_zmatrix *A, const double *d_b, double *d_x, double tol, int max_iter,
              bkh_info *info);
int bkh_zbicgstabl(const bkh_zmatrix *A, const double *d_b, int l, double *d_x, double tol,
                   int max_iter, bkh_info *info);
int bkh_zshbicgstabl(const bkh_zmatrix *A, const double *d_b, int l, const double *sigma, int ns,
                     double *d_x, double tol, int max_iter, bkh_info *info, double *shift_true_res);
/* Replaces zblbicgstabl (src/blbicgstabl.c:255-486): B, X m x s column-major (s <= 16),
 * Frobenius stop test on xi; info->true_res = ||B - A X||_F / ||B||_F. */
int bkh_zblbicgstabl(const bkh_zmatrix *A, int s, const double *d_B, int l, double *d_X, double tol,
                     int max_iter, bkh_info *info);
/* Thin QR of an m x s (s <= 16) complex panel, CholeskyQR2: d_V <- Q, R (host, s x s
 * column-major, real positive diagonal). Replaces zqreco (src/myqr.c:41-76). */
int bkh_zqr(int m, int s, double *d_V, double *R);
/* Average ms of reps back-to-back complex SpMV launches (op 'n' or 't'), HIP events. */
int bkh_zspmv_bench(const bkh_zmatrix *A, char op, int reps, double *avg_ms);

/* Test hook (no reference counterpart): the n-th device allocation the library makes
 * from now on fails as if HBM were exhausted (n <= 0 disarms); every entry point then
 * returns its device-error code. bkh_debug_fail_alloc_pending() is the number of
 * allocations still to go before it (0 once it has fired, or when disarmed). */
void bkh_debug_fail_alloc(int n);
int bkh_debug_fail_alloc_pending(void);
/* Test hook (no reference counterpart): on = 1 runs the block solvers' s <= 8 small
 * dense steps in their LDS form, on = 0 (default) in the register form; the two give
 * bitwise the same results. Retires cached iteration graphs. */
void bkh_debug_fin_lds_form(int on);

/* ---- batches of independent systems (no reference counterpart: the reference
 * solves one system per call, so a batch is that many sequential dbicgstabl calls) */
typedef struct bkh_batch bkh_batch;
/* Stacks ns CSR systems (0-based, square, n_each[s] rows, rows of <= 512 nonzeros)
 * block-diagonally on the device. want_fp32 keeps an FP32 copy for BKH_FP32_IR. */
bkh_batch *bkh_batch_create(int ns, const int *n_each, const int *const *row_ptr,
                            const int *const *col, const double *const *val, int want_fp32);
void bkh_batch_destroy(bkh_batch *B);
long long bkh_batch_rows(const bkh_batch *B);           /* total rows N */
long long bkh_batch_nnz(const bkh_batch *B);
long long bkh_batch_offset(const bkh_batch *B, int s);  /* first row of system s */
/* The SpMV blocking and vector tile size in use (set by bkh_batch_tune). */
void bkh_batch_layout(const bkh_batch *B, int *rows_cap, int *nnz_cap, int *tile_rows);
/* SpMV kernel: 1 = column-chunked (each workgroup stages 8192-column chunks of its
 * system's x in LDS; built at create time when every row's columns ascend, and then
 * the default), 0 = CSR-stream units. Both give bitwise the same y. set returns
 * BKH_BAD_ARGUMENT for 1 on a batch without the chunked copy. */
int bkh_batch_set_spmv_kernel(bkh_batch *B, int chunked);
int bkh_batch_spmv_kernel(const bkh_batch *B);
/* Autotuner: times candidate SpMV blockings (16, 32 or 64 rows per wavefront, each
 * with nonzero caps of that many median, 90th-percentile and mean-length rows of this
 * batch, and 512) and vector tile sizes on this batch (FP32 SpMV if fp32), and the
 * column-chunked kernel against the best blocking, and keeps the fastest. */
int bkh_batch_tune(bkh_batch *B, int fp32, double *best_ms);
/* y = A x over the batch; x, y device vectors of N doubles (floats if fp32). */
int bkh_batch_spmv(const bkh_batch *B, int fp32, const void *d_x, void *d_y);
/* BiCGSTAB(l), 1 <= l <= 4, on every system at once. opts->path: 0 or 1 = one launch
 * per phase over the whole batch, replayed from hipGraphs (the faster path); 2 = the
 * whole solve (every phase of every outer iteration of every system) as one
 * persistent cooperative launch with grid barriers between phases (the same
 * arithmetic: x is bitwise the same). d_b, d_x: device vectors of N entries (system s at its offset); ret (host,
 * ns) receives each system's result as dbicgstabl returns it; true_res (host, ns, may
 * be NULL) each system's FP64 true residual. opts->precision = BKH_FP32_IR runs FP32
 * inner solves (one launch each) inside per-system FP64 refinement (ret = inner
 * iterations used). */
int bkh_batch_bicgstabl(bkh_batch *B, const double *d_b, int l, double *d_x, double tol,
                        int max_iter, const bkh_opts *opts, int *ret, double *true_res,
                        bkh_info *info);
/* Path the last batched solve of the process ran: 1 per-phase graphs, 2 persistent. */
int bkh_batch_last_path(void);

/* ---- learned FSAI preconditioner and preconditioned CG (no reference counterpart:
 * the reference has no preconditioners) ----
 * An edge MLP (8 -> 64 -> 64 -> 1, tanh; FP32 master weights used as BF16 on the
 * matrix cores) maps local features of every edge of tril(A) to the factor G:
 * g_ii = softplus(h + log(e-1)) / sqrt(a_ii), g_ij = h / sqrt(a_jj), so M^-1 = G^T G
 * is SPD for any weights. Parameters are a flat float array of bkh_fsai_nparam()
 * entries: W1[8][64], b1[64], W2[64][64], b2[64], w3[64], b3. */
typedef struct bkh_fsai_model bkh_fsai_model;
typedef struct bkh_fsai_field bkh_fsai_field;
int bkh_fsai_nparam(void);
bkh_fsai_model *bkh_fsai_model_create(const float *params);     /* NULL: all zero (Jacobi) */
void bkh_fsai_model_destroy(bkh_fsai_model *M);
int bkh_fsai_model_get(const bkh_fsai_model *M, float *params);
int bkh_fsai_model_set(bkh_fsai_model *M, const float *params); /* also resets Adam */
/* A field: tril(A) with diagonal (host CSR) of an SPD matrix, and P probe vectors
 * (host, P x n row-major; P = 0 for a field that is only solved on). */
bkh_fsai_field *bkh_fsai_field_create(int n, const int *rp_lower, const int *col_lower,
                                      const double *val_lower, int P, const double *Z);
void bkh_fsai_field_destroy(bkh_fsai_field *F);
bkh_matrix *bkh_fsai_field_matrix(bkh_fsai_field *F);           /* the expanded A */
/* Factor values from the model (g_out: host, nnz(tril A), may be NULL). */
int bkh_fsai_factor(const bkh_fsai_model *M, bkh_fsai_field *F, double *g_out);
/* Loss L = log(T/n) - (2/n) sum log g_ii, T = mean_p (G^T z_p)^T A (G^T z_p), and its
 * parameter gradient (host, may be NULL), from the device forward/backward kernels. */
int bkh_fsai_loss_grad(bkh_fsai_model *M, bkh_fsai_field *F, double *loss, double *grad);
/* One Adam step (lr) on the field's loss; loss before the step to *loss (may be NULL). */
int bkh_fsai_train_step(bkh_fsai_model *M, bkh_fsai_field *F, double lr, double *loss);
/* Average ms of the factor build (MLP forward over all edges), HIP events. */
int bkh_fsai_bench_forward(const bkh_fsai_model *M, bkh_fsai_field *F, int reps, double *avg_ms);
/* PCG on the field's matrix with the factor last built by bkh_fsai_factor. */
int bkh_fsai_pcg(bkh_fsai_field *F, const double *d_b, double *d_x, double tol, int max_iter,
                 const bkh_opts *opts, bkh_info *info);
/* PCG with kind 0 (none: CG) or 1 (Jacobi) on an SPD matrix created with its
 * stored triangle. */
int bkh_pcg(bkh_matrix *A, int kind, const double *d_b, double *d_x, double tol, int max_iter,
            const bkh_opts *opts, bkh_info *info);

/* Timing of the most recent solve on this thread (ms, HIP events). */
double bkh_last_solve_ms(void);
/* Bytes one SpMV with the matrix's current kernel moves (stored entries x value and
 * index bytes, row pointers, x read once, y written once). */
long long bkh_matrix_spmv_bytes(const bkh_matrix *A, int fp32);
/* Average ms per launch of CG's fused kernel (SpMV with the direction update in its
 * gather) on a SELL-layout matrix, HIP events on its stream. */
int bkh_bench_cg_kernel(bkh_matrix *A, int fp32, int reps, double *avg_ms);
/* Average ms per launch of `reps` back-to-back SpMVs (FP64, or FP32 if fp32 = 1),
 * HIP events on the stream the kernels run on. */
int bkh_bench_spmv(const bkh_matrix *A, int fp32, const void *d_x, void *d_y, int reps,
                   double *avg_ms);
/* Average ms of a device copy of `bytes` (a bandwidth reference). */
int bkh_bench_copy(const void *d_a, void *d_b, size_t bytes, int reps, double *avg_ms);

#ifdef __cplusplus
}
#endif
#endif /* BKSP_HIP_H */
