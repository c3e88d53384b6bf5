/*
 * io.c -- Matrix Market reader of the drop-in (host C).
 *
 * Same contract as the reference's read_matrix_market (src/read_matrix.c:38-173):
 * coordinate real files, 1-based entries converted to a 0-based CSR whose rows keep
 * the file order of their entries (a stable counting sort, :122-173), type
 * "symmetric" when the banner says symmetric (:110-115; the stored triangle is kept
 * as is), format "csr". Like the reference, an unreadable file or banner ends the
 * process with exit(1) (:52-60, :76-77). zread_matrix_market is the complex
 * counterpart (:176-314): "re im" value pairs, the same stable row order.
 */
#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "../../../include/bksp.h"

static void lower(char *s) { for (; *s; s++) *s = (char)tolower((unsigned char)*s); }

dspmat *read_matrix_market(char *filename) {
    FILE *f = fopen(filename, "r");
    if (!f) exit(1);
    char line[1024], tag[64] = "", obj[64] = "", fmt[64] = "", field[64] = "", sym[64] = "";
    if (!fgets(line, sizeof line, f) ||
        sscanf(line, "%63s %63s %63s %63s %63s", tag, obj, fmt, field, sym) != 5) {
        printf("Could not process Matrix Market banner.\n");
        exit(1);
    }
    lower(obj); lower(fmt); lower(field); lower(sym);
    if (strcmp(tag, "%%MatrixMarket") != 0 || strcmp(obj, "matrix") != 0 ||
        strcmp(fmt, "coordinate") != 0 || strcmp(field, "complex") == 0) {
        printf("Could not process Matrix Market banner.\n");
        exit(1);
    }
    (void)field;
    int M = 0, N = 0, nz = 0;
    do {
        if (!fgets(line, sizeof line, f)) exit(1);
    } while (line[0] == '%');
    if (sscanf(line, "%d %d %d", &M, &N, &nz) != 3) exit(1);
    int *row = (int *)malloc(sizeof(int) * (size_t)(nz > 0 ? nz : 1));
    int *col = (int *)malloc(sizeof(int) * (size_t)(nz > 0 ? nz : 1));
    double *val = (double *)malloc(sizeof(double) * (size_t)(nz > 0 ? nz : 1));
    for (int i = 0; i < nz; i++) {
        if (fscanf(f, "%d %d %lg", &row[i], &col[i], &val[i]) != 3) exit(1);
        row[i]--;
        col[i]--;
    }
    fclose(f);
    int *ia = (int *)calloc((size_t)M + 1, sizeof(int));
    int *ja = (int *)malloc(sizeof(int) * (size_t)(nz > 0 ? nz : 1));
    double *a = (double *)malloc(sizeof(double) * (size_t)(nz > 0 ? nz : 1));
    for (int i = 0; i < nz; i++) ia[row[i] + 1]++;
    for (int i = 0; i < M; i++) ia[i + 1] += ia[i];
    int *cur = (int *)malloc(sizeof(int) * (size_t)(M > 0 ? M : 1));
    for (int i = 0; i < M; i++) cur[i] = ia[i];
    for (int i = 0; i < nz; i++) {
        int p = cur[row[i]]++;
        ja[p] = col[i];
        a[p] = val[i];
    }
    free(cur); free(row); free(col); free(val);
    dspmat *mat = (dspmat *)malloc(sizeof(dspmat));
    mat->row = ia;
    mat->col = ja;
    mat->value = a;
    mat->nnz = (int *)malloc(sizeof(int));
    mat->row_size = (int *)malloc(sizeof(int));
    mat->col_size = (int *)malloc(sizeof(int));
    *mat->nnz = nz;
    *mat->row_size = M;
    *mat->col_size = N;
    mat->format = "csr";
    mat->type = (strcmp(sym, "symmetric") == 0) ? "symmetric" : "unsymmetric";
    return mat;
}

zspmat *zread_matrix_market(char *filename) {
    FILE *f = fopen(filename, "r");
    if (!f) exit(1);
    char line[1024], tag[64] = "", obj[64] = "", fmt[64] = "", field[64] = "", sym[64] = "";
    if (!fgets(line, sizeof line, f) ||
        sscanf(line, "%63s %63s %63s %63s %63s", tag, obj, fmt, field, sym) != 5) {
        printf("Could not process Matrix Market banner.\n");
        exit(1);
    }
    lower(obj); lower(fmt); lower(field); lower(sym);
    if (strcmp(tag, "%%MatrixMarket") != 0 || strcmp(obj, "matrix") != 0 || strcmp(fmt, "coordinate") != 0) {
        printf("Could not process Matrix Market banner.\n");
        exit(1);
    }
    int M = 0, N = 0, nz = 0;
    do {
        if (!fgets(line, sizeof line, f)) exit(1);
    } while (line[0] == '%');
    if (sscanf(line, "%d %d %d", &M, &N, &nz) != 3) exit(1);
    int *row = (int *)malloc(sizeof(int) * (size_t)(nz > 0 ? nz : 1));
    int *col = (int *)malloc(sizeof(int) * (size_t)(nz > 0 ? nz : 1));
    double _Complex *val = (double _Complex *)malloc(sizeof(double _Complex) * (size_t)(nz > 0 ? nz : 1));
    for (int i = 0; i < nz; i++) {
        double re = 0.0, im = 0.0;
        if (fscanf(f, "%d %d %lg %lg", &row[i], &col[i], &re, &im) != 4) exit(1);
        row[i]--;
        col[i]--;
        __real__ val[i] = re;
        __imag__ val[i] = im;
    }
    fclose(f);
    int *ia = (int *)calloc((size_t)M + 1, sizeof(int));
    int *ja = (int *)malloc(sizeof(int) * (size_t)(nz > 0 ? nz : 1));
    double _Complex *a = (double _Complex *)malloc(sizeof(double _Complex) * (size_t)(nz > 0 ? nz : 1));
    for (int i = 0; i < nz; i++) ia[row[i] + 1]++;
    for (int i = 0; i < M; i++) ia[i + 1] += ia[i];
    int *cur = (int *)malloc(sizeof(int) * (size_t)(M > 0 ? M : 1));
    for (int i = 0; i < M; i++) cur[i] = ia[i];
    for (int i = 0; i < nz; i++) {
        int p = cur[row[i]]++;
        ja[p] = col[i];
        a[p] = val[i];
    }
    free(cur); free(row); free(col); free(val);
    zspmat *mat = (zspmat *)malloc(sizeof(zspmat));
    mat->row = ia;
    mat->col = ja;
    mat->value = a;
    mat->nnz = (int *)malloc(sizeof(int));
    mat->row_size = (int *)malloc(sizeof(int));
    mat->col_size = (int *)malloc(sizeof(int));
    *mat->nnz = nz;
    *mat->row_size = M;
    *mat->col_size = N;
    mat->format = "csr";
    mat->type = (strcmp(sym, "symmetric") == 0) ? "symmetric" : "unsymmetric";
    return mat;
}
