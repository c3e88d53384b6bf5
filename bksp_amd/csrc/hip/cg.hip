// cg.hip -- conjugate gradients on the device (replaces src/cg.c:6-71).
//
// One iteration is three kernels, captured K at a time into a hipGraph:
//   1. q = A p (CSR-stream SpMV) with the p.q dot in its epilogue; the last
//      workgroup forms alpha = (r.r)/(p.q)                           (cg.c:35-39)
//   2. x += alpha p, r -= alpha q, r.r; the last workgroup forms beta, records
//      the recursive residual ||r||/||b|| and the break index      (cg.c:42-59)
//   3. p = beta p + r                                              (cg.c:53-54)
// Scalars never leave the device; the host reads the control block once per
// graph replay. BKH_FP32_IR runs the same kernels on FP32 values and vectors
// (FP64 scalars and dot accumulation) inside an FP64 refinement loop whose
// residual b - A x and update x += d are FP64.
#include <math.h>
#include "bkh_host.hpp"
#include "bkh_spmv.hpp"
#include "bkh_solver_common.hpp"

enum { CG_RR = 0, CG_ALPHA = 1, CG_BETA = 2 };

struct CgAlphaFin {
    __device__ void operator()(Ctl *c, const double *tot) const {
        double pq = tot[0];
        if (bad(pq)) { ctl_breakdown(c); return; }
        c->s[CG_ALPHA] = c->s[CG_RR] / pq;
    }
};

struct CgBetaFin {
    __device__ void operator()(Ctl *c, const double *tot) const {
        double rr = tot[0];
        c->s[CG_BETA] = rr / c->s[CG_RR];
        c->s[CG_RR] = rr;
        ctl_check(c, rr);
    }
};

// r = b - A x0 (q holds A x0), p = r; r.r and b.b  (cg.c:23-30)
struct CgInitFin {
    __device__ void operator()(Ctl *c, const double *tot) const {
        c->s[CG_RR] = tot[0];
        c->fnb = sqrt(tot[1]);
        if (!(c->fnb > 0.0)) ctl_breakdown(c);
    }
};

template <typename T>
__global__ __launch_bounds__(256) void k_cg_init(long long n, const T *b, const T *q, T *r, T *p,
                                                 Ctl *c, double *part) {
    double v[2] = {0.0, 0.0};
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        T ri = b[i] - q[i];
        r[i] = ri;
        p[i] = ri;
        v[0] += (double)ri * (double)ri;
        v[1] += (double)b[i] * (double)b[i];
    }
    block_partial<2, 256>(v, part);
}

template <typename T>
__global__ __launch_bounds__(256) void k_cg_xr(long long n, T *x, T *r, const T *p, const T *q,
                                               Ctl *c, double *part) {
    if (ctl_stopped(c)) return;
    const T al = (T)c->s[CG_ALPHA];
    double v[1] = {0.0};
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        x[i] = fma(al, p[i], x[i]);
        T ri = fma(-al, q[i], r[i]);
        r[i] = ri;
        v[0] += (double)ri * (double)ri;
    }
    block_partial<1, 256>(v, part);
}

template <typename T>
__global__ __launch_bounds__(256) void k_cg_p(long long n, T *p, const T *r, const Ctl *c) {
    if (ctl_stopped(c)) return;
    const T be = (T)c->s[CG_BETA];
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        p[i] = add_rn(mul_rn(be, p[i]), r[i]);
}

// One CG solve on device vectors of type T (values `val`), x updated in place.
template <typename T, typename VT>
static int cg_core(ThreadCtx *t, const bkh_matrix *A, const VT *val, const T *b, T *x, T *r, T *p,
                   T *q, Ctl *c, double *part, double tol, int max_it, int K) {
    hipStream_t s = t->stream;
    const long long n = A->n;
    const int G = vec_grid(n, 256);
    if (ctl_init(t, c, tol, max_it)) return BKH_DEVICE_ERROR;
    Epi<0, T> e0;
    memset(&e0, 0, sizeof e0);
    launch_spmv<VT, T, 0>(s, A, val, x, q, 1.0, 0.0, e0, nullptr, nullptr, NoFin());
    hipLaunchKernelGGL(k_cg_init<T>, dim3(G), dim3(256), 0, s, n, b, q, r, p, c, part);
    launch_finalize<2>(s, part, G, c, CgInitFin());
    LAUNCH_OK();
    hipGraphExec_t ge = capture_graph(s, K, [&]() {
        launch_spmv_dot_self<VT, T>(s, A, val, p, q, c, part, CgAlphaFin());
        hipLaunchKernelGGL(k_cg_xr<T>, dim3(G), dim3(256), 0, s, n, x, r, p, q, c, part);
        launch_finalize<1>(s, part, G, c, CgBetaFin());
        hipLaunchKernelGGL(k_cg_p<T>, dim3(G), dim3(256), 0, s, n, p, r, c);
    });
    if (!ge) { bkh_set_error("hipGraph capture of the CG iteration failed"); return BKH_DEVICE_ERROR; }
    int rc = run_graph_until_done(t, ge, c, max_it / K + 2);
    hipGraphExecDestroy(ge);
    return rc;
}

extern "C" int bkh_cg(bkh_matrix *A, const double *b, double *x, double tol, int max_iter,
                      const bkh_opts *opts, bkh_info *info) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (!A || !b || !x) return BKH_BAD_ARGUMENT;
    bkh_opts o = opts ? *opts : bkh_opts{};
    bkh_info inf;
    memset(&inf, 0, sizeof inf);
    const long long n = A->n;
    const int K = o.graph_iters > 0 ? o.graph_iters : default_graph_iters(n);
    const int nparts = std::max(spmv_grid(A), vec_grid(n, 256)) * 4;
    if (max_iter <= 0) {                     // cg.c:15,67-70: error = 0 -> returns 0
        inf.ret = 0; inf.status = ST_CONVERGED;
        if (info) *info = inf;
        return 0;
    }
    int ret;
    HIP_OK(hipEventRecord(t->ev0, t->stream));
    if (o.precision == BKH_FP32_IR) {
        if (!A->val32) { bkh_set_error("FP32_IR needs a matrix created with want_fp32"); return BKH_BAD_ARGUMENT; }
        size_t bytes = carve_size(sizeof(Ctl)) * 2 + carve_size(sizeof(double) * nparts) +
                       carve_size(sizeof(double) * n) + 5 * carve_size(sizeof(float) * n);
        void *ws = bkh_workspace(bytes);
        if (!ws) return BKH_DEVICE_ERROR;
        Carve cv(ws);
        Ctl *co = cv.take<Ctl>(1), *ci = cv.take<Ctl>(1);
        double *part = cv.take<double>(nparts);
        double *r64 = cv.take<double>(n);
        float *r32 = cv.take<float>(n), *d32 = cv.take<float>(n), *rr = cv.take<float>(n),
              *pp = cv.take<float>(n), *qq = cv.take<float>(n);
        IrResult ir;
        ret = ir_refine(t, A, b, x, r64, co, part, tol, max_iter, o, &ir,
                        [&](double inner_tol, int inner_max, int *inner_it) -> int {
                            // d32 = 0, solve A d = r32 to inner_tol in FP32
                            int rc2 = cg_core<float, float>(t, A, A->val32, r32, d32, rr, pp, qq, ci,
                                                            part, inner_tol, inner_max, K);
                            if (rc2) return rc2;
                            *inner_it = t->ctl_host->it;
                            return t->ctl_host->status == ST_BREAKDOWN ? -2 : 0;
                        },
                        r32, d32);
        inf.inner_iters = ir.inner_total;
        inf.passes = ir.passes;
        inf.status = ir.status;
        inf.iters = ir.ret;
        inf.rec_res = ir.res;
    } else {
        size_t bytes = carve_size(sizeof(Ctl)) + carve_size(sizeof(double) * nparts) +
                       3 * carve_size(sizeof(double) * n);
        void *ws = bkh_workspace(bytes);
        if (!ws) return BKH_DEVICE_ERROR;
        Carve cv(ws);
        Ctl *c = cv.take<Ctl>(1);
        double *part = cv.take<double>(nparts);
        double *r = cv.take<double>(n), *p = cv.take<double>(n), *q = cv.take<double>(n);
        int rc = cg_core<double, double>(t, A, A->val, b, x, r, p, q, c, part, tol, max_iter, K);
        if (rc) return rc;
        ret = ctl_result(t->ctl_host);
        inf.status = t->ctl_host->status;
        inf.iters = t->ctl_host->result_it;
        inf.rec_res = t->ctl_host->err;
    }
    HIP_OK(hipEventRecord(t->ev1, t->stream));
    HIP_OK(hipEventSynchronize(t->ev1));
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, t->ev0, t->ev1));
    inf.time_ms = ms;
    bkh_note_solve_ms(ms);
    inf.ret = ret;
    if (info) {
        double tr = 0.0;
        if (true_residual(t, A, 0.0, b, x, &tr) == 0) inf.true_res = tr;
        *info = inf;
    }
    return ret;
}
