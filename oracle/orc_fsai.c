/*
 * orc_fsai.c -- CPU restatement of the learned FSAI preconditioner, test
 * infrastructure only (the checker of the device kernels in fsai.hip).
 *
 * There is no reference counterpart: the reference has no preconditioners. This
 * file states the model the device implements, in plain C:
 *   features  f(i, j) in R^8 of every edge (i, j), j <= i, of tril(A);
 *   MLP       h = w3 . tanh(W2^T tanh(W1^T f + b1) + b2) + b3 (8 -> 64 -> 64 -> 1);
 *   factor    g_ii = softplus(h + c0) / sqrt(a_ii), g_ij = h / sqrt(a_jj) (j < i),
 *             c0 = log(e - 1), so h = 0 gives G = D^{-1/2} (Jacobi) and M^{-1} =
 *             G^T G is SPD for any parameters;
 *   loss      L = log(T / n) - (2/n) sum_i log g_ii, T = (1/P) sum_p w_p^T A w_p,
 *             w_p = G^T z_p over P probe vectors: the Kaporin condition number of
 *             G A G^T (the classical FSAI objective) estimated with probes;
 *   gradient  dL/dg_ij = (2 / (P T)) sum_p z_p[i] (A w_p)_j - [i == j] (2/n) / g_ii
 *             (orc_fsai_loss_dg), then the MLP backward pass by the chain rule.
 * With bf16 = 1 the forward and backward round the operands the device feeds its
 * BF16 matrix cores (weights, features, hidden activations, hidden gradients) to
 * bfloat16 (round to nearest even) and accumulate in float, as the device does;
 * with bf16 = 0 everything is double (for finite-difference checks).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define H 64
#define NF 8
#define P_W1 0
#define P_B1 (P_W1 + NF * H)
#define P_W2 (P_B1 + H)
#define P_B2 (P_W2 + H * H)
#define P_W3 (P_B2 + H)
#define P_B3 (P_W3 + H)
#define NPARAM (P_B3 + 1)

static const double C0 = 0.54132485461291810;   /* log(e - 1) */

int orc_fsai_nparam(void) { return NPARAM; }

static float bf16r(float x) {
    unsigned u;
    memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    float y;
    memcpy(&y, &u, 4);
    return y;
}

static double rq(double x, int bf16) { return bf16 ? (double)bf16r((float)x) : x; }

/*
 * Per-row data from the full (expanded) symmetric matrix: sqrt(a_ii), the
 * normalised coupling sum s_i = sum_{k != i} |a_ik| / sqrt(a_ii a_kk), the number of
 * strictly-lower entries, and the diagonal excess (a_ii - sum |a_ik|) / a_ii.
 */
void orc_fsai_rows(int n, const int *rp, const int *ci, const double *v, double *sqd, double *s,
                   double *nlow, double *excess) {
    double *d = (double *)malloc(sizeof(double) * (size_t)n);
    for (int i = 0; i < n; i++) {
        d[i] = 0.0;
        for (int k = rp[i]; k < rp[i + 1]; k++)
            if (ci[k] == i) d[i] += v[k];
    }
    for (int i = 0; i < n; i++) {
        double acc = 0.0, asum = 0.0;
        int nl = 0;
        for (int k = rp[i]; k < rp[i + 1]; k++) {
            int j = ci[k];
            if (j == i) continue;
            acc += fabs(v[k]) / sqrt(d[i] * d[j]);
            asum += fabs(v[k]);
            if (j < i) nl++;
        }
        sqd[i] = sqrt(d[i]);
        s[i] = acc;
        nlow[i] = nl;
        excess[i] = (d[i] - asum) / d[i];
    }
    free(d);
}

/* Features of lower edge (i, j, a_ij), as the device computes them (float). */
void orc_fsai_edge_features(int n, int i, int j, double aij, const double *sqd, const double *s,
                            const double *nlow, const double *excess, float *f) {
    const double di = sqd[i] * sqd[i], dj = sqd[j] * sqd[j];
    f[0] = (float)(i == j ? 1.0 : 0.0);
    f[1] = (float)(aij / (sqd[i] * sqd[j]));
    f[2] = (float)(0.25 * log(di / dj));
    f[3] = (float)s[i];
    f[4] = (float)s[j];
    f[5] = (float)(log2(1.0 + (double)(i - j)) / log2(1.0 + (double)n));
    f[6] = (float)(0.25 * nlow[i]);
    f[7] = (float)excess[i];
}

/* MLP forward for one edge; keeps the activations for the backward pass. */
static double mlp_forward(const float *p, const float *f, int bf16, double *a1, double *a2) {
    double x[NF];
    for (int q = 0; q < NF; q++) x[q] = rq(f[q], bf16);
    for (int h = 0; h < H; h++) {
        double z = 0.0;
        for (int q = 0; q < NF; q++) z += x[q] * rq(p[P_W1 + q * H + h], bf16);
        z += p[P_B1 + h];
        a1[h] = bf16 ? (double)tanhf((float)z) : tanh(z);
    }
    for (int o = 0; o < H; o++) {
        double z = 0.0;
        for (int i = 0; i < H; i++) z += rq(a1[i], bf16) * rq(p[P_W2 + i * H + o], bf16);
        z += p[P_B2 + o];
        a2[o] = bf16 ? (double)tanhf((float)z) : tanh(z);
    }
    double h = 0.0;
    for (int o = 0; o < H; o++) h += a2[o] * rq(p[P_W3 + o], bf16);
    return h + p[P_B3];
}

/* h for every lower edge (nnz_l = rpl[n]). */
void orc_fsai_forward(int n, const int *rpl, const int *cil, const double *vl, const double *sqd,
                      const double *s, const double *nlow, const double *excess, const float *params,
                      int bf16, double *h_out) {
    double a1[H], a2[H];
    float f[NF];
    for (int i = 0; i < n; i++)
        for (int k = rpl[i]; k < rpl[i + 1]; k++) {
            orc_fsai_edge_features(n, i, cil[k], vl[k], sqd, s, nlow, excess, f);
            h_out[k] = mlp_forward(params, f, bf16, a1, a2);
        }
}

static double softplus(double x) { return x > 30.0 ? x : log1p(exp(x)); }
static double sigmoid(double x) { return 1.0 / (1.0 + exp(-x)); }

/* g of every lower edge from h (the factor's parametrisation). */
void orc_fsai_gvals(int n, const int *rpl, const int *cil, const double *sqd, const double *h,
                    double *g) {
    for (int i = 0; i < n; i++)
        for (int k = rpl[i]; k < rpl[i + 1]; k++) {
            const int j = cil[k];
            g[k] = (j == i) ? softplus(h[k] + C0) / sqd[i] : h[k] / sqd[j];
        }
}

/*
 * Loss and dL/dg for probes Z (P vectors of n, row-major): w_p = G^T z_p, u_p = A w_p
 * (A as the expanded CSR rp/ci/v), T = (1/P) sum_p w_p.u_p,
 * L = log(T / n) - (2/n) sum_i log g_ii,
 * dL/dg_ij = (2 / (P T)) sum_p z_p[i] u_p[j] - [i == j] (2/n) / g_ii.
 */
double orc_fsai_loss_dg(int n, const int *rp, const int *ci, const double *v, const int *rpl,
                        const int *cil, const double *g, int P, const double *Z, double *dg) {
    const int nnzl = rpl[n];
    double *w = (double *)malloc(sizeof(double) * (size_t)n);
    double *u = (double *)malloc(sizeof(double) * (size_t)n * P);
    double T = 0.0;
    for (int p = 0; p < P; p++) {
        const double *z = Z + (size_t)p * n;
        double *up = u + (size_t)p * n;
        for (int j = 0; j < n; j++) w[j] = 0.0;
        for (int i = 0; i < n; i++)                       /* w = G^T z */
            for (int k = rpl[i]; k < rpl[i + 1]; k++) w[cil[k]] += g[k] * z[i];
        for (int i = 0; i < n; i++) {                     /* u = A w */
            double acc = 0.0;
            for (int k = rp[i]; k < rp[i + 1]; k++) acc += v[k] * w[ci[k]];
            up[i] = acc;
        }
        double t = 0.0;
        for (int i = 0; i < n; i++) t += w[i] * up[i];
        T += t;
    }
    T /= P;
    double slog = 0.0;
    for (int i = 0; i < n; i++)
        for (int k = rpl[i]; k < rpl[i + 1]; k++)
            if (cil[k] == i) slog += log(g[k]);
    const double L = log(T / n) - 2.0 * slog / n;
    if (dg) {
        for (int i = 0; i < n; i++)
            for (int k = rpl[i]; k < rpl[i + 1]; k++) {
                const int j = cil[k];
                double acc = 0.0;
                for (int p = 0; p < P; p++) acc += Z[(size_t)p * n + i] * u[(size_t)p * n + j];
                dg[k] = 2.0 * acc / (P * T) - ((i == j) ? 2.0 / (n * g[k]) : 0.0);
            }
    }
    (void)nnzl;
    free(w);
    free(u);
    return L;
}

/*
 * Parameter gradient: dL/dh from dL/dg through the parametrisation, then the MLP
 * backward pass, summed over all edges into grad[NPARAM] (double).
 */
void orc_fsai_backward(int n, const int *rpl, const int *cil, const double *vl, const double *sqd,
                       const double *s, const double *nlow, const double *excess, const float *p,
                       int bf16, const double *h, const double *dg, double *grad) {
    double a1[H], a2[H], dz2[H], dz1[H];
    float f[NF];
    memset(grad, 0, sizeof(double) * NPARAM);
    for (int i = 0; i < n; i++)
        for (int k = rpl[i]; k < rpl[i + 1]; k++) {
            const int j = cil[k];
            orc_fsai_edge_features(n, i, j, vl[k], sqd, s, nlow, excess, f);
            mlp_forward(p, f, bf16, a1, a2);
            const double dh = (j == i) ? dg[k] * sigmoid(h[k] + C0) / sqd[i] : dg[k] / sqd[j];
            grad[P_B3] += dh;
            for (int o = 0; o < H; o++) {
                grad[P_W3 + o] += a2[o] * dh;
                dz2[o] = rq(p[P_W3 + o], bf16) * dh * (1.0 - a2[o] * a2[o]);
                grad[P_B2 + o] += dz2[o];
            }
            for (int ii = 0; ii < H; ii++) {
                double da = 0.0;
                for (int o = 0; o < H; o++) {
                    grad[P_W2 + ii * H + o] += rq(a1[ii], bf16) * rq(dz2[o], bf16);
                    da += rq(p[P_W2 + ii * H + o], bf16) * rq(dz2[o], bf16);
                }
                dz1[ii] = da * (1.0 - a1[ii] * a1[ii]);
                grad[P_B1 + ii] += dz1[ii];
            }
            for (int q = 0; q < NF; q++)
                for (int hh = 0; hh < H; hh++) grad[P_W1 + q * H + hh] += rq(f[q], bf16) * rq(dz1[hh], bf16);
        }
}
