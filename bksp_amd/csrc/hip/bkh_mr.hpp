// bkh_mr.hpp -- the minimal-residual step of BiCGSTAB(l) from a Gram matrix.
//
// The reference orthogonalises r_1..r_l in place by modified Gram-Schmidt
// (src/bicgstabl.c:84-109, src/shbicgstabl.c:166-191, src/blbicgstabl.c:170-187)
// and then updates x, r_0 with the modified vectors. Here the same recurrences
// run on coefficient vectors: r'_j = sum_k C[j][k] r_k, with inner products taken
// from the Gram matrix G[a][b] = r_a . r_b of the UNMODIFIED vectors (one fused
// pass over the data instead of l^2 passes). The results are re-expressed on the
// unmodified vectors, so the final update is a single fused sweep.
#pragma once
#include "bkh_core.hpp"

#define MR_LMAX 8
#define MR_D (MR_LMAX + 1)

struct MrOut {
    double tau[MR_D][MR_D];   // tau[i][j], 1 <= i < j <= l (reference indexing)
    double sigma[MR_D];       // sigma[j] = |r'_j|^2
    double od[MR_D];          // omega'  (omega_d)
    double om[MR_D];          // omega   (the MR polynomial's coefficients gamma_j)
    double odd[MR_D];         // omega'' (omega_dd)
    double C[MR_D][MR_D];     // r'_j = sum_k C[j][k] r_k
};

__device__ __forceinline__ double gram_dot(const double (&G)[MR_D][MR_D], const double *a,
                                           const double *b, int l) {
    double s = 0.0;
    for (int p = 1; p <= l; p++) {
        if (a[p] == 0.0) continue;
        double t = 0.0;
        for (int q = 1; q <= l; q++) t += b[q] * G[p][q];
        s += a[p] * t;
    }
    return s;
}

// Returns false on a zero/NaN pivot (sigma_j), i.e. a breakdown.
__device__ __forceinline__ bool mr_from_gram(int l, const double (&G)[MR_D][MR_D], MrOut &o) {
    for (int j = 0; j <= l; j++)
        for (int k = 0; k <= l; k++) { o.C[j][k] = (j == k) ? 1.0 : 0.0; o.tau[j][k] = 0.0; }
    for (int j = 1; j <= l; j++) {
        for (int i = 1; i < j; i++) {
            o.tau[i][j] = gram_dot(G, o.C[j], o.C[i], l) / o.sigma[i];
            for (int k = 1; k <= l; k++) o.C[j][k] -= o.tau[i][j] * o.C[i][k];
        }
        o.sigma[j] = gram_dot(G, o.C[j], o.C[j], l);
        double r0rj = 0.0;
        for (int k = 1; k <= l; k++) r0rj += o.C[j][k] * G[0][k];
        o.od[j] = r0rj / o.sigma[j];
        if (!(o.sigma[j] > 0.0) || !(o.od[j] == o.od[j])) return false;
    }
    o.om[l] = o.od[l];
    for (int j = l - 1; j > 0; j--) {
        double t = 0.0;
        for (int i = j + 1; i <= l; i++) t += o.tau[j][i] * o.om[i];
        o.om[j] = o.od[j] - t;
    }
    for (int j = 1; j < l; j++) {
        double t = 0.0;
        for (int i = j + 1; i < l; i++) t += o.tau[j][i] * o.om[i + 1];
        o.odd[j] = o.om[j + 1] + t;
    }
    return true;
}

// Unpacks the upper-triangular Gram list (a <= b, row by row over 0..l).
__device__ __forceinline__ void gram_unpack(int l, const double *tot, double (&G)[MR_D][MR_D]) {
    int idx = 0;
    for (int a = 0; a <= l; a++)
        for (int b = a; b <= l; b++) { G[a][b] = tot[idx]; G[b][a] = tot[idx]; idx++; }
}
