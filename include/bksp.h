/*
 * bksp.h -- public API of the MI355X-native bksp drop-in (real and complex FP64).
 *
 * Same declarations as the reference's public header, so a C caller written
 * against it compiles and links unchanged:
 *   error codes           reference include/bksp.h:4-7 (the trailing ';' is kept,
 *                         so `return NOT_CONVERGED;` works as it does there)
 *   dspmat / dmat         reference include/bksp.h:12-27 (CSR with int* size
 *                         fields; dmat column-major, ld = row_size)
 *   inline accessors      reference include/bksp.h:31-58
 *   utilities             reference include/bksp.h:60-66
 *   solvers               reference include/bksp.h:71-103
 *
 * A "symmetric" dspmat holds ONE stored triangle (the lower one for Matrix Market
 * symmetric files); CG, CR and block CG require type "symmetric".
 * Return values: >= 0 is the 0-based loop index at which the recursive residual
 * fell below tol (outer iterations for the BiCGSTAB(l) family); negative values
 * are the codes below. Device failures use the extension codes of bksp_hip.h.
 *
 * The complex (z*) family (reference include/bksp.h:109-190) is declared in the
 * "complex numbers" section below; values are C99 double _Complex, as there.
 */
#ifndef BKSP_H
#define BKSP_H

#ifdef __cplusplus
extern "C" {
#endif

////////// ERROR CODE //////////

#define NOT_CONVERGED -1;
#define DIVERGENCE -2;
#define INVALID_PARAMETERS -3;
#define INVALID_MATRIX_TYPE -4;

////////// Structs //////////

typedef struct {
    int *row;        /* row_ptr[n+1], 0-based */
    int *col;        /* column index[nnz] */
    int *nnz;
    int *row_size;
    int *col_size;
    double *value;   /* value[nnz] */
    char *format;    /* "csr" */
    char *type;      /* "symmetric" (one stored triangle) or "unsymmetric" */
} dspmat;

typedef struct {
    int row_size;
    int col_size;
    double *value;   /* column-major, value[col * row_size + row] */
} dmat;

////////// Utilities //////////

static inline int dspmat_rowsize(dspmat *mat) { return *mat->row_size; }
static inline int dspmat_colsize(dspmat *mat) { return *mat->col_size; }
static inline int dmat_set(dmat *mat, int m, int n, double num) {
    mat->value[n * mat->row_size + m] = num;
    return 0;
}
static inline double dmat_get(dmat *mat, int m, int n) {
    return mat->value[n * mat->row_size + m];
}
static inline int dmat_rowsize(dmat *mat) { return mat->row_size; }
static inline int dmat_colsize(dmat *mat) { return mat->col_size; }
static inline double *dmat_value(dmat *mat) { return mat->value; }

dspmat *read_matrix_market(char *filename);
int myqr(int m, int n, double *mat, double *r);
dmat *dmat_create(int m, int n);
int darray2dmat(double *arr, int row_size, int col_size, dmat *mat);
int dmat_free(dmat *mat);
int dvec2mat(double *x, int m, dmat *mat);

////////// Krylov subspace methods //////////

int bksp_dcg(dspmat *A, double *b, double tol, int max_iter, double *x);
int dcr(dspmat *A, double *b, double tol, int max_iter, double *x);
int dbicg(dspmat *A, double *b, double tol, int max_iter, double *x);
int dbicgstab(dspmat *A, double *b, double tol, int max_iter, double *x);
int dbicgstabl(dspmat *A, double *b, int l, double tol, int max_iter, double *x);

////////// Block Krylov subspace methods //////////

int dblcg(dspmat *A, dmat *B, double tol, int max_iter, dmat *X);
int dblbicgstab(dspmat *A, dmat *B, double tol, int max_iter, dmat *X);
int dblbicgstabl(dspmat *A, dmat *B, int l, double tol, int max_iter, dmat *X);

////////// Krylov subspace methods for shifted systems //////////

int dshbicg(dspmat *A, double *b, double *sigma, int sigma_size,
            double tol, int max_iter, double *x);
int dshbicgstabl(dspmat *A, double *b, int l, double *sigma,
                 int sigma_size, double tol, int max_iter, double *x);

/* ---- complex numbers (include/bksp.h:109-190 of the reference) ---- */
#ifndef __cplusplus
typedef struct {
    int *row;
    int *col;
    int *nnz;
    int *row_size;
    int *col_size;
    double _Complex *value;
    char *format;
    char *type;
} zspmat;

typedef struct {
    int row_size;
    int col_size;
    double _Complex *value;
} zmat;

static inline int zspmat_rowsize(zspmat *mat) { return *mat->row_size; }
static inline int zspmat_colsize(zspmat *mat) { return *mat->col_size; }
static inline int zmat_set(zmat *mat, int m, int n, double _Complex *num) {
    mat->value[n * mat->row_size + m] = *num;
    return 0;
}
/* returns double as in the reference (bksp.h:147-149): the real part */
static inline double zmat_get(zmat *mat, int m, int n) { return mat->value[n * mat->row_size + m]; }
static inline int zmat_rowsize(zmat *mat) { return mat->row_size; }
static inline int zmat_colsize(zmat *mat) { return mat->col_size; }
static inline double _Complex *zmat_value(zmat *mat) { return mat->value; }

zspmat *zread_matrix_market(char *filename);
zmat *zmat_create(int m, int n);
int zmat_free(zmat *mat);

int zbicg(zspmat *A, double _Complex *b, double tol, int max_iter, double _Complex *x);
int zbicgstabl(zspmat *A, double _Complex *b, int ell, double tol, int max_iter, double _Complex *x);
int zshbicgstabl(zspmat *A, double _Complex *b, int l, double _Complex *sigma,
                 int sigma_size, double tol, int max_iter, double _Complex *x);
int zblbicgstabl(zspmat *A, zmat *b, int ell, double tol, int max_iter, zmat *x);
int zqreco(int m, int n, double _Complex *mat, double _Complex *r);
#endif

#ifdef __cplusplus
}
#endif
#endif /* BKSP_H */
