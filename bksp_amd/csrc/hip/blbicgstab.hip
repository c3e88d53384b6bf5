// blbicgstab.hip -- block BiCGSTAB with an orthonormalised direction block
// (replaces src/blbicgstab.c:13-123) for link parity of the bksp.h API. Not on the
// hot path: a host-driven loop over the device layer's SpMM, MFMA Gram, dot and
// Householder QR entry points plus two panel kernels, following blbicgstab.c step
// for step. The s x s systems (Rs^T V) alpha = Rs^T R and (Rs^T V) beta = -Rs^T T
// are solved on the host with the LU of Rs^T V kept between them, as the reference's
// dgesv / dgetrs pair (:69, :94) does; everything of length m runs on the device.
//
// Per iteration (m x s column-major blocks, the dmat layout):
//   V = A P; alpha = (Rs^T V)^-1 Rs^T R; R -= V alpha; T = A R;
//   omega = <T, R> / <T, T>; X += P alpha; X += omega R; R -= omega T;
//   beta = -(Rs^T V)^-1 Rs^T T; P -= omega V; P = P beta + R; P <- Q of P (myqr);
//   error = ||R||_F / ||B||_F.
#include <math.h>
#include <vector>
#include "bkh_host.hpp"

extern "C" int bkh_dot(long long n, const double *x, const double *y, double *out);
extern "C" int bkh_spmm(const bkh_matrix *A, int s, double alpha, const double *X, double beta, double *Y);
extern "C" int bkh_gram(int m, int s, const double *X, const double *Y, double *G);
extern "C" int bkh_qr_householder(int m, int n, double *d_V, double *R);

#define BL_SMAX 16

// Y = a (X M) + b Y for m x s column-major X, Y (Y != X) and the s x s M.
__global__ __launch_bounds__(256) void k_bl1_panel(long long m, int s, double *Y, const double *X,
                                                   const double *M, double a, double b) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < m; i += (long long)gridDim.x * 256) {
        double xr[BL_SMAX];
        for (int k = 0; k < s; k++) xr[k] = X[(size_t)k * m + i];
        for (int j = 0; j < s; j++) {
            double acc = 0.0;
            for (int k = 0; k < s; k++) acc = fma(xr[k], M[j * s + k], acc);
            double *y = Y + (size_t)j * m + i;
            *y = (b == 0.0) ? a * acc : fma(a, acc, b * *y);
        }
    }
}

// y += a x over n entries.
__global__ __launch_bounds__(256) void k_bl1_axpy(long long n, double a, const double *x, double *y) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        y[i] = fma(a, x[i], y[i]);
}

namespace {

// LU with partial pivoting of an s x s column-major matrix (dgetrf); false if singular.
bool lu_factor(int s, double *A, int *piv) {
    for (int k = 0; k < s; k++) {
        int p = k;
        for (int i = k + 1; i < s; i++)
            if (fabs(A[k * s + i]) > fabs(A[k * s + p])) p = i;
        piv[k] = p;
        if (A[k * s + p] == 0.0) return false;
        if (p != k)
            for (int j = 0; j < s; j++) std::swap(A[j * s + k], A[j * s + p]);
        const double inv = 1.0 / A[k * s + k];
        for (int i = k + 1; i < s; i++) A[k * s + i] *= inv;
        for (int j = k + 1; j < s; j++)
            for (int i = k + 1; i < s; i++) A[j * s + i] -= A[k * s + i] * A[j * s + k];
    }
    return true;
}

// Solves (LU) X = B in place for nrhs columns (dgetrs 'N').
void lu_solve(int s, const double *LU, const int *piv, double *B, int nrhs) {
    for (int c = 0; c < nrhs; c++) {
        double *b = B + (size_t)c * s;
        for (int k = 0; k < s; k++) std::swap(b[k], b[piv[k]]);
        for (int k = 0; k < s; k++)
            for (int i = k + 1; i < s; i++) b[i] -= LU[k * s + i] * b[k];
        for (int k = s - 1; k >= 0; k--) {
            b[k] /= LU[k * s + k];
            for (int i = 0; i < k; i++) b[i] -= LU[k * s + i] * b[k];
        }
    }
}

struct DevBuf {
    double *p = nullptr;
    ~DevBuf() {
        if (p) hipFree(p);
    }
};

}  // namespace

// B, X: m x s column-major device blocks, 1 <= s <= 16. Returns the break index or
// -1 (NOT_CONVERGED), -2 (singular Rs^T V or a NaN residual), or a device error.
extern "C" int bkh_blbicgstab(bkh_matrix *A, int s, const double *B, double *X, double tol, int max_iter) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (!A || !B || !X || s < 1 || s > BL_SMAX) return BKH_BAD_ARGUMENT;
    if (max_iter <= 0) return 0;                                   // blbicgstab.c:24,119-122
    const long long m = A->n, ms = m * (long long)s;
    hipStream_t st = t->stream;
    DevBuf base;
    HIP_OK(bkh_dmalloc((void **)&base.p, sizeof(double) * (size_t)(6 * ms + 2 * s * s)));
    double *P = base.p, *R = P + ms, *Rs = R + ms, *V = Rs + ms, *T = V + ms, *W = T + ms;
    double *Md = W + ms;                                           // s x s coefficients
    const int G = vec_grid(ms, 256);
    const int Gm = vec_grid(m, 256, 1);
    int rc;
    // R = B - A X (:49-52), P = R, Rs = R
    if ((rc = bkh_spmm(A, s, -1.0, X, 0.0, R))) return rc;
    hipLaunchKernelGGL(k_bl1_axpy, dim3(G), dim3(256), 0, st, ms, 1.0, B, R);
    LAUNCH_OK();
    HIP_OK(hipMemcpyAsync(P, R, sizeof(double) * ms, hipMemcpyDeviceToDevice, st));
    HIP_OK(hipMemcpyAsync(Rs, R, sizeof(double) * ms, hipMemcpyDeviceToDevice, st));
    double fnb2 = 0.0;
    if ((rc = bkh_dot(ms, B, B, &fnb2))) return rc;
    const double fnb = sqrt(fnb2);
    std::vector<double> lu(s * s), alpha(s * s), beta(s * s), rq(s * s);
    std::vector<int> piv(s);
    double error = 0.0;
    int i;
    for (i = 0; i < max_iter; i++) {
        if ((rc = bkh_spmm(A, s, 1.0, P, 0.0, V))) return rc;                     // :61-62
        if ((rc = bkh_gram((int)m, s, Rs, V, lu.data()))) return rc;              // :65-66
        if ((rc = bkh_gram((int)m, s, Rs, R, alpha.data()))) return rc;           // :67-68
        if (!lu_factor(s, lu.data(), piv.data())) { bkh_set_error("block BiCGSTAB: Rs^T V is singular"); return -2; }
        lu_solve(s, lu.data(), piv.data(), alpha.data(), s);                     // :69
        HIP_OK(hipMemcpyAsync(Md, alpha.data(), sizeof(double) * s * s, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_bl1_panel, dim3(Gm), dim3(256), 0, st, m, s, R, (const double *)V,
                           (const double *)Md, -1.0, 1.0);                        // :72-73 (R -= V alpha)
        LAUNCH_OK();
        if ((rc = bkh_spmm(A, s, 1.0, R, 0.0, T))) return rc;                     // :76-77
        double tr = 0.0, tt = 0.0;
        if ((rc = bkh_dot(ms, T, R, &tr)) || (rc = bkh_dot(ms, T, T, &tt))) return rc;
        const double omega = tr / tt;                                             // :80-81
        hipLaunchKernelGGL(k_bl1_panel, dim3(Gm), dim3(256), 0, st, m, s, X, (const double *)P,
                           (const double *)Md, 1.0, 1.0);                         // :84-85
        hipLaunchKernelGGL(k_bl1_axpy, dim3(G), dim3(256), 0, st, ms, omega, (const double *)R, X);   // :86
        hipLaunchKernelGGL(k_bl1_axpy, dim3(G), dim3(256), 0, st, ms, -omega, (const double *)T, R);  // :89
        LAUNCH_OK();
        if ((rc = bkh_gram((int)m, s, Rs, T, beta.data()))) return rc;            // :92-93
        for (auto &v : beta) v = -v;
        lu_solve(s, lu.data(), piv.data(), beta.data(), s);                      // :94
        HIP_OK(hipMemcpyAsync(Md, beta.data(), sizeof(double) * s * s, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_bl1_axpy, dim3(G), dim3(256), 0, st, ms, -omega, (const double *)V, P);  // :97
        HIP_OK(hipMemcpyAsync(W, R, sizeof(double) * ms, hipMemcpyDeviceToDevice, st));
        hipLaunchKernelGGL(k_bl1_panel, dim3(Gm), dim3(256), 0, st, m, s, W, (const double *)P,
                           (const double *)Md, 1.0, 1.0);                         // :98-100
        LAUNCH_OK();
        std::swap(P, W);                                                          // :101
        if ((rc = bkh_qr_householder((int)m, s, P, rq.data()))) return rc;        // :104
        double rr = 0.0;
        if ((rc = bkh_dot(ms, R, R, &rr))) return rc;
        error = sqrt(rr) / fnb;                                                    // :106
        if (error < tol) break;
        if (!(error == error)) { bkh_set_error("block BiCGSTAB: NaN residual"); return -2; }
    }
    HIP_OK(hipStreamSynchronize(st));
    if (!(error < tol)) return -1;
    return i;
}
