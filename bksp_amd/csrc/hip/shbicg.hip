// shbicg.hip -- shifted BiCG (replaces src/shbicg.c:4-130) for link parity of the
// bksp.h API. Not on the hot path: a host-driven loop over device kernels (one dot
// product round trip per scalar), the seed BiCG and the shifted updates of
// x_k = x + (k+1) n following shbicg.c step for step.
#include <math.h>
#include <vector>
#include "bkh_host.hpp"
#include "bkh_solver_common.hpp"
#include "bkh_spmv.hpp"

extern "C" int bkh_dot(long long n, const double *x, const double *y, double *out);
extern "C" int bkh_spmv(const bkh_matrix *A, char op, double alpha, const double *x, double beta, double *y);

#define GLOOP(i, n) for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < (n); i += (long long)gridDim.x * 256)

// p = r - beta p; sp = sr - beta sp (shbicg.c:56-61)
__global__ __launch_bounds__(256) void k_sb_p(long long n, double beta, const double *r, const double *sr, double *p,
                                              double *sp) {
    GLOOP(i, n) {
        p[i] = -beta * p[i] + r[i];
        sp[i] = -beta * sp[i] + sr[i];
    }
}

// x += alpha p (:70)
__global__ __launch_bounds__(256) void k_sb_x(long long n, double alpha, const double *p, double *x) {
    GLOOP(i, n) x[i] += alpha * p[i];
}

// per shift k: p_k = -beta_k p_k + r / pi_k; x_k += alpha_k p_k (:72-86)
__global__ __launch_bounds__(256) void k_sb_shift(long long n, int ns, const double *coef, const double *r,
                                                  double *ps, double *xs) {
    const int k = blockIdx.y;
    if (k >= ns) return;
    const double mb = coef[3 * k], ip = coef[3 * k + 1], as = coef[3 * k + 2];
    double *pk = ps + (size_t)k * n, *xk = xs + (size_t)k * n;
    GLOOP(i, n) {
        const double v = mb * pk[i] + ip * r[i];
        pk[i] = v;
        xk[i] += as * v;
    }
}

// r -= alpha ap; sr -= alpha atp (:88-91)
__global__ __launch_bounds__(256) void k_sb_r(long long n, double alpha, const double *ap, const double *atp,
                                              double *r, double *sr) {
    GLOOP(i, n) {
        r[i] -= alpha * ap[i];
        sr[i] -= alpha * atp[i];
    }
}

__global__ __launch_bounds__(256) void k_sb_init(long long n, const double *b, double *r, double *sr) {
    GLOOP(i, n) {
        const double v = b[i] - r[i];
        r[i] = v;
        sr[i] = v;
    }
}

// A created with want_transpose; d_x holds n*(ns+1) entries (seed first).
extern "C" int bkh_shbicg(bkh_matrix *A, const double *b, const double *sigma, int ns, double *x, double tol,
                          int max_iter) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (!A || !b || !x || ns < 0 || (ns > 0 && !sigma)) return BKH_BAD_ARGUMENT;
    if (!A->T) { bkh_set_error("shifted BiCG needs the transpose"); return BKH_BAD_ARGUMENT; }
    const long long n = A->n;
    const int G = vec_grid(n, 256);
    const int nsm = ns > 0 ? ns : 1;
    // the front of the workspace stays free for bkh_dot's control block and partials
    const size_t pre = carve_size(sizeof(Ctl)) + carve_size(sizeof(double) * G);
    void *ws = bkh_workspace(pre + 6 * carve_size(sizeof(double) * n) + carve_size(sizeof(double) * n * nsm) +
                             carve_size(sizeof(double) * 3 * nsm));
    if (!ws) return BKH_DEVICE_ERROR;
    Carve cv((char *)ws + pre);
    double *r = cv.take<double>(n), *sr = cv.take<double>(n), *p = cv.take<double>(n), *sp = cv.take<double>(n),
           *ap = cv.take<double>(n), *atp = cv.take<double>(n), *ps = cv.take<double>(n * nsm),
           *coef = cv.take<double>(3 * nsm);
    hipStream_t s = t->stream;
    HIP_OK(hipMemsetAsync(p, 0, sizeof(double) * n, s));
    HIP_OK(hipMemsetAsync(sp, 0, sizeof(double) * n, s));
    HIP_OK(hipMemsetAsync(ps, 0, sizeof(double) * n * nsm, s));
    int rc = bkh_spmv(A, 'n', 1.0, x, 0.0, r);                      // (:24-29)
    if (rc) return rc;
    hipLaunchKernelGGL(k_sb_init, dim3(G), dim3(256), 0, s, n, b, r, sr);
    LAUNCH_OK();
    double fnb = 0.0;
    if ((rc = bkh_dot(n, b, b, &fnb))) return rc;
    fnb = sqrt(fnb);
    double alpha_old = 1.0, rho_old = 1.0, error = 0.0;
    std::vector<double> pi(ns, 1.0), pi_old(ns, 1.0), pi_new(ns), h(3 * nsm);
    int i;
    for (i = 0; i < max_iter; i++) {
        double rho = 0.0;
        if ((rc = bkh_dot(n, r, sr, &rho))) return rc;
        const double beta = -rho / rho_old;
        hipLaunchKernelGGL(k_sb_p, dim3(G), dim3(256), 0, s, n, beta, (const double *)r, (const double *)sr, p, sp);
        LAUNCH_OK();
        if ((rc = bkh_spmv(A, 'n', 1.0, p, 0.0, ap)) || (rc = bkh_spmv(A, 't', 1.0, sp, 0.0, atp))) return rc;
        double apsr = 0.0;
        if ((rc = bkh_dot(n, ap, sr, &apsr))) return rc;
        const double alpha = rho / apsr;
        hipLaunchKernelGGL(k_sb_x, dim3(G), dim3(256), 0, s, n, alpha, (const double *)p, x);
        for (int k = 0; k < ns; k++) {
            pi_new[k] = (1 + alpha * sigma[k]) * pi[k] + alpha * beta / alpha_old * (pi_old[k] - pi[k]);
            const double bs = (pi_old[k] / pi[k]) * (pi_old[k] / pi[k]) * beta;
            h[3 * k] = -bs;
            h[3 * k + 1] = 1.0 / pi[k];
            h[3 * k + 2] = pi[k] / pi_new[k] * alpha;
        }
        if (ns > 0) {
            HIP_OK(hipMemcpyAsync(coef, h.data(), sizeof(double) * 3 * ns, hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(k_sb_shift, dim3(G, ns), dim3(256), 0, s, n, ns, (const double *)coef,
                               (const double *)r, ps, x + n);
        }
        hipLaunchKernelGGL(k_sb_r, dim3(G), dim3(256), 0, s, n, alpha, (const double *)ap, (const double *)atp, r, sr);
        LAUNCH_OK();
        alpha_old = alpha;
        rho_old = rho;
        for (int k = 0; k < ns; k++) { pi_old[k] = pi[k]; pi[k] = pi_new[k]; }
        double rr = 0.0;
        if ((rc = bkh_dot(n, r, r, &rr))) return rc;
        error = sqrt(rr) / fnb;
        if (error < tol) break;
    }
    HIP_OK(hipStreamSynchronize(s));
    if (!(error < tol)) return -1;
    return i;
}
