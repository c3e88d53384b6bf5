/* gen.c -- seeded generators of the benchmark systems (see include/bksp_gen.h). */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "../../../include/bksp_gen.h"

typedef struct { uint64_t s[4]; } rng_t;

static uint64_t splitmix64(uint64_t *x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static void rng_seed(rng_t *r, uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&x);
}
static uint64_t rng_next(rng_t *r) {       /* xoshiro256** */
    uint64_t *s = r->s;
    uint64_t res = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t; s[3] = rotl(s[3], 45);
    return res;
}
static double rng_u01(rng_t *r) { return (double)(rng_next(r) >> 11) * 0x1.0p-53; }

typedef struct { rng_t r; int have; double spare; } nrng_t;
static double rng_normal(nrng_t *g) {
    if (g->have) { g->have = 0; return g->spare; }
    double u1 = 1.0 - rng_u01(&g->r), u2 = rng_u01(&g->r);
    double rad = sqrt(-2.0 * log(u1)), th = 2.0 * M_PI * u2;
    g->spare = rad * sin(th);
    g->have = 1;
    return rad * cos(th);
}

void bkh_csr_host_free(bkh_csr_host *m) {
    if (!m) return;
    free(m->row_ptr); free(m->col); free(m->val);
    m->row_ptr = NULL; m->col = NULL; m->val = NULL;
}

static int csr_alloc(bkh_csr_host *m, int n, long long cap) {
    m->n = n; m->nnz = 0;
    m->row_ptr = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    m->col = (int *)malloc(sizeof(int) * (size_t)(cap > 0 ? cap : 1));
    m->val = (double *)malloc(sizeof(double) * (size_t)(cap > 0 ? cap : 1));
    if (!m->row_ptr || !m->col || !m->val) { bkh_csr_host_free(m); return -1; }
    m->row_ptr[0] = 0;
    return 0;
}

void bkh_gen_uniform(uint64_t seed, long long n, double lo, double hi, double *out) {
    rng_t r;
    rng_seed(&r, seed);
    for (long long i = 0; i < n; i++) out[i] = lo + (hi - lo) * rng_u01(&r);
}

void bkh_gen_normal(uint64_t seed, long long n, double *out) {
    nrng_t g;
    memset(&g, 0, sizeof g);
    rng_seed(&g.r, seed);
    for (long long i = 0; i < n; i++) out[i] = rng_normal(&g);
}

int bkh_gen_convdiff2d(int m, double vx, double vy, bkh_csr_host *out) {
    int n = m * m;
    double h = 1.0 / (m + 1), ih2 = 1.0 / (h * h);
    if (csr_alloc(out, n, 5LL * n)) return -1;
    int k = 0;
    for (int y = 0; y < m; y++)
        for (int x = 0; x < m; x++) {
            int i = x + m * y;
            if (y > 0)     { out->col[k] = i - m; out->val[k++] = -ih2 - vy / (2 * h); }
            if (x > 0)     { out->col[k] = i - 1; out->val[k++] = -ih2 - vx / (2 * h); }
            out->col[k] = i; out->val[k++] = 4 * ih2;
            if (x < m - 1) { out->col[k] = i + 1; out->val[k++] = -ih2 + vx / (2 * h); }
            if (y < m - 1) { out->col[k] = i + m; out->val[k++] = -ih2 + vy / (2 * h); }
            out->row_ptr[i + 1] = k;
        }
    out->nnz = k;
    return 0;
}

/* Shared 7-point builder: coefficient of the neighbour along axis a (0=x,1=y,2=z),
 * direction d (-1 below, +1 above) is off[a][d>0]; diagonal `dg`. */
static int stencil7(int m, double dg, const double off[3][2], int lower, bkh_csr_host *out) {
    long long n = (long long)m * m * m;
    if (n > 0x7fffffffLL) return -1;
    if (csr_alloc(out, (int)n, 7 * n)) return -1;
    long long k = 0, m2 = (long long)m * m;
    for (int z = 0; z < m; z++)
        for (int y = 0; y < m; y++)
            for (int x = 0; x < m; x++) {
                long long i = x + (long long)m * y + m2 * z;
                if (z > 0) { out->col[k] = (int)(i - m2); out->val[k++] = off[2][0]; }
                if (y > 0) { out->col[k] = (int)(i - m);  out->val[k++] = off[1][0]; }
                if (x > 0) { out->col[k] = (int)(i - 1);  out->val[k++] = off[0][0]; }
                out->col[k] = (int)i; out->val[k++] = dg;
                if (!lower) {
                    if (x < m - 1) { out->col[k] = (int)(i + 1);  out->val[k++] = off[0][1]; }
                    if (y < m - 1) { out->col[k] = (int)(i + m);  out->val[k++] = off[1][1]; }
                    if (z < m - 1) { out->col[k] = (int)(i + m2); out->val[k++] = off[2][1]; }
                }
                out->row_ptr[i + 1] = (int)k;
            }
    out->nnz = (int)k;
    return 0;
}

int bkh_gen_poisson3d(int m, int lower, bkh_csr_host *out) {
    const double off[3][2] = {{-1, -1}, {-1, -1}, {-1, -1}};
    return stencil7(m, 6.0, off, lower, out);
}

int bkh_gen_convdiff3d(int m, double px, double py, double pz, bkh_csr_host *out) {
    const double off[3][2] = {{-1 - px, -1 + px}, {-1 - py, -1 + py}, {-1 - pz, -1 + pz}};
    return stencil7(m, 6.0, off, 0, out);
}

static int cmp_int(const void *a, const void *b) {
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

int bkh_gen_random_rows(uint64_t seed, int n, bkh_csr_host *out) {
    if (n < 29) return -1;                  /* needs room for 27 distinct off-diagonals */
    nrng_t g;
    memset(&g, 0, sizeof g);
    rng_seed(&g.r, seed);
    if (csr_alloc(out, n, 28LL * n)) return -1;
    int cols[28];
    double vals[28];
    long long k = 0;
    for (int i = 0; i < n; i++) {
        int cnt = 5 + (int)(rng_u01(&g.r) * 23.0);        /* U{5..27} */
        if (cnt > 27) cnt = 27;
        int got = 0;
        while (got < cnt) {                                /* distinct by rejection */
            int c = (int)(rng_u01(&g.r) * n);
            if (c >= n) c = n - 1;
            if (c == i) continue;
            int dup = 0;
            for (int q = 0; q < got; q++) if (cols[q] == c) { dup = 1; break; }
            if (!dup) cols[got++] = c;
        }
        cols[got++] = i;
        qsort(cols, (size_t)got, sizeof(int), cmp_int);
        double asum = 0;
        for (int q = 0; q < got; q++) {
            if (cols[q] == i) continue;
            vals[q] = rng_normal(&g);
            asum += fabs(vals[q]);
        }
        double f = 1.0 + 0.05 + 0.45 * rng_u01(&g.r);
        for (int q = 0; q < got; q++) {
            out->col[k] = cols[q];
            out->val[k++] = (cols[q] == i) ? asum * f : vals[q];
        }
        out->row_ptr[i + 1] = (int)k;
    }
    out->nnz = (int)k;
    return 0;
}

void bkh_gen_batch_sizes(uint64_t seed, int count, double lo2, double hi2, int *n_out) {
    rng_t r;
    rng_seed(&r, seed);
    for (int i = 0; i < count; i++) n_out[i] = (int)lround(pow(2.0, lo2 + (hi2 - lo2) * rng_u01(&r)));
}

int bkh_gen_grf(uint64_t seed, int m, double sigma, double corr, double *kappa) {
    long long n = (long long)m * m * m;
    double *a = (double *)malloc(sizeof(double) * (size_t)n);
    double *b = (double *)malloc(sizeof(double) * (size_t)n);
    if (!a || !b) { free(a); free(b); return -1; }
    bkh_gen_normal(seed, n, a);
    /* white noise filtered by a Gaussian of std s has covariance exp(-d^2/(4 s^2)),
     * i.e. squared-exponential with length corr when s = corr / sqrt(2) */
    double s = corr / sqrt(2.0);
    int R = (int)ceil(3.0 * s);
    if (R > m / 2 - 1) R = m / 2 - 1;
    double *w = (double *)malloc(sizeof(double) * (2 * R + 1));
    double wsum = 0, w2 = 0;
    for (int q = -R; q <= R; q++) { w[q + R] = exp(-0.5 * q * q / (s * s)); wsum += w[q + R]; }
    for (int q = 0; q <= 2 * R; q++) { w[q] /= wsum; w2 += w[q] * w[q]; }
    long long st[3] = {1, m, (long long)m * m};
    for (int ax = 0; ax < 3; ax++) {
        for (long long i = 0; i < n; i++) {
            long long c = (i / st[ax]) % m, base = i - c * st[ax];
            double acc = 0;
            for (int q = -R; q <= R; q++) {
                long long cc = (c + q + m) % m;
                acc += w[q + R] * a[base + cc * st[ax]];
            }
            b[i] = acc;
        }
        double *t = a; a = b; b = t;
    }
    double sd = sqrt(w2 * w2 * w2);          /* exact std of the filtered field */
    for (long long i = 0; i < n; i++) kappa[i] = exp(sigma * a[i] / sd);
    free(w); free(a); free(b);
    return 0;
}

int bkh_gen_fv_diffusion(int m, const double *kappa, int lower, bkh_csr_host *out) {
    long long n = (long long)m * m * m, m2 = (long long)m * m;
    if (csr_alloc(out, (int)n, 7 * n)) return -1;
    long long k = 0;
    for (int z = 0; z < m; z++)
        for (int y = 0; y < m; y++)
            for (int x = 0; x < m; x++) {
                long long i = x + (long long)m * y + m2 * z;
                int c[3] = {x, y, z};
                long long st[3] = {1, m, m2};
                double ki = kappa[i], diag = 0, lo[3], hi[3];
                for (int a = 0; a < 3; a++) {
                    if (c[a] > 0) { double kj = kappa[i - st[a]]; lo[a] = 2 * ki * kj / (ki + kj); }
                    else lo[a] = 2 * ki;
                    if (c[a] < m - 1) { double kj = kappa[i + st[a]]; hi[a] = 2 * ki * kj / (ki + kj); }
                    else hi[a] = 2 * ki;
                    diag += lo[a] + hi[a];
                }
                for (int a = 2; a >= 0; a--)
                    if (c[a] > 0) { out->col[k] = (int)(i - st[a]); out->val[k++] = -lo[a]; }
                out->col[k] = (int)i; out->val[k++] = diag;
                if (!lower)
                    for (int a = 0; a < 3; a++)
                        if (c[a] < m - 1) { out->col[k] = (int)(i + st[a]); out->val[k++] = -hi[a]; }
                out->row_ptr[i + 1] = (int)k;
            }
    out->nnz = (int)k;
    return 0;
}
