// cr_bicg.hip -- conjugate residual (replaces src/cr.c:6-80) and BiCG (replaces
// src/bicg.c:7-88) on the device, with the same graph-replay structure as CG.
//
// CR per iteration (loop index from 1, cr.c:37):
//   k_cr_xr   x += alpha p, r -= alpha ap; r.r (kept for the error test)  (cr.c:41-45)
//   SpMV      t = A r, epilogue r.t -> beta                               (:48-52)
//   k_cr_pap  p = beta p + r, ap = beta ap + t; ap.ap -> next alpha, error (:55-68)
// BiCG per iteration:
//   SpMV      ap = A p,  epilogue sp.ap -> alpha                          (bicg.c:37-45)
//   SpMV      atp = A^T sp (transpose built at upload, src/matvec.c:29-46 order)
//   k_bg_upd  x, r, sr updates; r.sr -> beta, r.r -> error                (:48-73)
//   k_bg_p    p = beta p + r, sp = beta sp + sr                           (:62-67)
#include <math.h>
#include "bkh_host.hpp"
#include "bkh_solver_common.hpp"
#include "bkh_spmv.hpp"

#define ROWLOOP(i, n) \
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < (n); i += (long long)gridDim.x * 256)

// ------------------------------------------------------------------- CR

enum { CR_RDOT = 0, CR_ALPHA = 1, CR_BETA = 2, CR_RR = 3 };

struct FnbFin {       // b.b -> ||b||
    __device__ void operator()(Ctl *c, const double *tot) const {
        c->fnb = sqrt(tot[0]);
        if (!(c->fnb > 0.0)) ctl_breakdown(c);
    }
};

struct CrInitFin {    // t = A r0: r_dot_old = r.t, alpha = r_dot_old / (t.t)   (cr.c:30-33, :39)
    __device__ void operator()(Ctl *c, const double *tot) const {
        c->s[CR_RDOT] = tot[0];
        if (bad(tot[1])) { ctl_breakdown(c); return; }
        c->s[CR_ALPHA] = tot[0] / tot[1];
    }
};

struct CrRRFin {
    __device__ void operator()(Ctl *c, const double *tot) const { c->s[CR_RR] = tot[0]; }
};

struct CrBetaFin {
    __device__ void operator()(Ctl *c, const double *tot) const {
        if (bad(c->s[CR_RDOT])) { ctl_breakdown(c); return; }
        c->s[CR_BETA] = tot[0] / c->s[CR_RDOT];
        c->s[CR_RDOT] = tot[0];
    }
};

struct CrAlphaFin {
    __device__ void operator()(Ctl *c, const double *tot) const {
        ctl_check(c, c->s[CR_RR]);
        if (c->done) return;
        if (bad(tot[0])) { ctl_breakdown(c); return; }
        c->s[CR_ALPHA] = c->s[CR_RDOT] / tot[0];
    }
};

__global__ __launch_bounds__(256) void k_cr_init(long long n, const double *b, const double *q,
                                                 double *r, double *p, double *part) {
    double v[1] = {0.0};
    ROWLOOP(i, n) {
        double ri = b[i] - q[i];
        r[i] = ri; p[i] = ri;
        v[0] += b[i] * b[i];
    }
    block_partial<1, 256>(v, part);
}

__global__ __launch_bounds__(256) void k_copy_vec(long long n, const double *a, double *b,
                                                  const Ctl *c) {
    if (ctl_stopped(c)) return;
    ROWLOOP(i, n) b[i] = a[i];
}

__global__ __launch_bounds__(256) void k_cr_xr(long long n, double *x, double *r, const double *p,
                                               const double *ap, const Ctl *c, double *part) {
    if (ctl_stopped(c)) return;
    const double al = c->s[CR_ALPHA];
    double v[1] = {0.0};
    ROWLOOP(i, n) {
        x[i] = fma(al, p[i], x[i]);
        double ri = fma(-al, ap[i], r[i]);
        r[i] = ri;
        v[0] += ri * ri;
    }
    block_partial<1, 256>(v, part);
}

__global__ __launch_bounds__(256) void k_cr_pap(long long n, double *p, double *ap, const double *r,
                                                const double *t, const Ctl *c, double *part) {
    if (ctl_stopped(c)) return;
    const double be = c->s[CR_BETA];
    double v[1] = {0.0};
    ROWLOOP(i, n) {
        p[i] = add_rn(mul_rn(be, p[i]), r[i]);
        double a = add_rn(mul_rn(be, ap[i]), t[i]);
        ap[i] = a;
        v[0] += a * a;
    }
    block_partial<1, 256>(v, part);
}

extern "C" int bkh_cr(bkh_matrix *A, const double *b, double *x, double tol, int max_iter,
                      const bkh_opts *opts, bkh_info *info) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (!A || !b || !x) return BKH_BAD_ARGUMENT;
    bkh_opts o = opts ? *opts : bkh_opts{};
    bkh_info inf;
    memset(&inf, 0, sizeof inf);
    if (max_iter <= 1) {                 // cr.c:37,76-79: the loop starts at i = 1
        inf.ret = 1; inf.status = ST_CONVERGED; inf.iters = 1;
        if (info) *info = inf;
        return 1;
    }
    const long long n = A->n;
    const int G = vec_grid(n, 256);
    const int K = o.graph_iters > 0 ? o.graph_iters : default_graph_iters(n);
    const int nparts = std::max(spmv_grid(A), G) * 2;
    void *ws = bkh_workspace(carve_size(sizeof(Ctl)) + carve_size(sizeof(double) * nparts) +
                             4 * carve_size(sizeof(double) * n));
    if (!ws) return BKH_DEVICE_ERROR;
    Carve cv(ws);
    Ctl *c = cv.take<Ctl>(1);
    double *part = cv.take<double>(nparts);
    double *r = cv.take<double>(n), *p = cv.take<double>(n), *ap = cv.take<double>(n),
           *tv = cv.take<double>(n);
    hipStream_t s = t->stream;
    HIP_OK(hipEventRecord(t->ev0, s));
    if (ctl_init(t, c, tol, max_iter, 1)) return BKH_DEVICE_ERROR;
    Epi<0> e0;
    memset(&e0, 0, sizeof e0);
    launch_spmv<double, double, 0>(s, A, A->val, x, tv, 1.0, 0.0, e0, nullptr, nullptr, NoFin());
    hipLaunchKernelGGL(k_cr_init, dim3(G), dim3(256), 0, s, n, b, tv, r, p, part);
    launch_finalize<1>(s, part, G, c, FnbFin());
    Epi<2> e2 = make_epi<2>({r}, {{0, EPI_Y}, {EPI_Y, EPI_Y}});
    launch_spmv<double, double, 2>(s, A, A->val, r, tv, 1.0, 0.0, e2, c, part, CrInitFin());
    hipLaunchKernelGGL(k_copy_vec, dim3(G), dim3(256), 0, s, n, tv, ap, c);
    LAUNCH_OK();
    hipGraphExec_t ge = capture_cached(t, {0xC200, (uintptr_t)A, (uintptr_t)b, (uintptr_t)x, (uintptr_t)r, (uintptr_t)p,
                                           (uintptr_t)ap, (uintptr_t)tv, (uintptr_t)c, (uintptr_t)part}, s, K, [&]() {
        hipLaunchKernelGGL(k_cr_xr, dim3(G), dim3(256), 0, s, n, x, r, p, ap, c, part);
        launch_finalize<1>(s, part, G, c, CrRRFin());
        launch_spmv_dot_self<double, double>(s, A, A->val, r, tv, c, part, CrBetaFin());
        hipLaunchKernelGGL(k_cr_pap, dim3(G), dim3(256), 0, s, n, p, ap, r, tv, c, part);
        launch_finalize<1>(s, part, G, c, CrAlphaFin());
    });
    if (!ge) { bkh_set_error("hipGraph capture of the CR iteration failed"); return BKH_DEVICE_ERROR; }
    int rc = run_graph_until_done(t, ge, c, max_iter / K + 2);
    if (rc) return rc;
    return finish_solve(t, A, 0.0, b, x, &inf, info);
}

// ------------------------------------------------------------------- BiCG

enum { BG_RDOT = 0, BG_ALPHA = 1, BG_BETA = 2 };

struct BgInitFin {    // r_dot_old = r.sr, ||b||                           (bicg.c:31-32)
    __device__ void operator()(Ctl *c, const double *tot) const {
        c->s[BG_RDOT] = tot[0];
        c->fnb = sqrt(tot[1]);
        if (!(c->fnb > 0.0)) ctl_breakdown(c);
    }
};

struct BgAlphaFin {   // alpha = r_dot_old / (ap . sp)                      (:45)
    __device__ void operator()(Ctl *c, const double *tot) const {
        if (bad(tot[0])) { ctl_breakdown(c); return; }
        c->s[BG_ALPHA] = c->s[BG_RDOT] / tot[0];
    }
};

struct BgBetaFin {    // beta = r.sr / r_dot_old; error = ||r|| / ||b||      (:57-72)
    __device__ void operator()(Ctl *c, const double *tot) const {
        if (bad(c->s[BG_RDOT])) { ctl_breakdown(c); return; }
        c->s[BG_BETA] = tot[0] / c->s[BG_RDOT];
        c->s[BG_RDOT] = tot[0];
        ctl_check(c, tot[1]);
    }
};

__global__ __launch_bounds__(256) void k_bg_init(long long n, const double *b, const double *q,
                                                 double *r, double *sr, double *p, double *sp,
                                                 double *part) {
    double v[2] = {0.0, 0.0};
    ROWLOOP(i, n) {
        double ri = b[i] - q[i];
        r[i] = ri; sr[i] = ri; p[i] = ri; sp[i] = ri;
        v[0] += ri * ri;
        v[1] += b[i] * b[i];
    }
    block_partial<2, 256>(v, part);
}

__global__ __launch_bounds__(256) void k_bg_upd(long long n, double *x, double *r, double *sr,
                                                const double *p, const double *ap,
                                                const double *atp, const Ctl *c, double *part) {
    if (ctl_stopped(c)) return;
    const double al = c->s[BG_ALPHA];
    double v[2] = {0.0, 0.0};
    ROWLOOP(i, n) {
        x[i] = fma(al, p[i], x[i]);
        double ri = fma(-al, ap[i], r[i]);
        double si = fma(-al, atp[i], sr[i]);
        r[i] = ri; sr[i] = si;
        v[0] += ri * si;
        v[1] += ri * ri;
    }
    block_partial<2, 256>(v, part);
}

__global__ __launch_bounds__(256) void k_bg_p(long long n, double *p, double *sp, const double *r,
                                              const double *sr, const Ctl *c) {
    if (ctl_stopped(c)) return;
    const double be = c->s[BG_BETA];
    ROWLOOP(i, n) {
        p[i] = add_rn(mul_rn(be, p[i]), r[i]);
        sp[i] = add_rn(mul_rn(be, sp[i]), sr[i]);
    }
}

extern "C" int bkh_bicg(bkh_matrix *A, const double *b, double *x, double tol, int max_iter,
                        const bkh_opts *opts, bkh_info *info) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (!A || !b || !x) return BKH_BAD_ARGUMENT;
    if (!A->T) { bkh_set_error("BiCG needs a matrix created with want_transpose"); return BKH_BAD_ARGUMENT; }
    bkh_opts o = opts ? *opts : bkh_opts{};
    bkh_info inf;
    memset(&inf, 0, sizeof inf);
    if (max_iter <= 0) {                 // bicg.c:11,84-87
        inf.status = ST_CONVERGED;
        if (info) *info = inf;
        return 0;
    }
    const long long n = A->n;
    const int G = vec_grid(n, 256);
    const int K = o.graph_iters > 0 ? o.graph_iters : default_graph_iters(n);
    const int nparts = std::max(std::max(spmv_grid(A), spmv_grid(A->T)), G) * 2;
    void *ws = bkh_workspace(carve_size(sizeof(Ctl)) + carve_size(sizeof(double) * nparts) +
                             6 * carve_size(sizeof(double) * n));
    if (!ws) return BKH_DEVICE_ERROR;
    Carve cv(ws);
    Ctl *c = cv.take<Ctl>(1);
    double *part = cv.take<double>(nparts);
    double *r = cv.take<double>(n), *sr = cv.take<double>(n), *p = cv.take<double>(n),
           *sp = cv.take<double>(n), *ap = cv.take<double>(n), *atp = cv.take<double>(n);
    hipStream_t s = t->stream;
    HIP_OK(hipEventRecord(t->ev0, s));
    if (ctl_init(t, c, tol, max_iter)) return BKH_DEVICE_ERROR;
    Epi<0> e0;
    memset(&e0, 0, sizeof e0);
    launch_spmv<double, double, 0>(s, A, A->val, x, ap, 1.0, 0.0, e0, nullptr, nullptr, NoFin());
    hipLaunchKernelGGL(k_bg_init, dim3(G), dim3(256), 0, s, n, b, ap, r, sr, p, sp, part);
    launch_finalize<2>(s, part, G, c, BgInitFin());
    LAUNCH_OK();
    hipGraphExec_t ge = capture_cached(t, {0xB1C0, (uintptr_t)A, (uintptr_t)b, (uintptr_t)x, (uintptr_t)r, (uintptr_t)sr,
                                           (uintptr_t)p, (uintptr_t)sp, (uintptr_t)ap, (uintptr_t)atp, (uintptr_t)c,
                                           (uintptr_t)part}, s, K, [&]() {
        launch_spmv_dot_w<double, double>(s, A, A->val, p, ap, sp, c, part, BgAlphaFin());
        Epi<0> ez;
        memset(&ez, 0, sizeof ez);
        launch_spmv<double, double, 0>(s, A->T, A->T->val, sp, atp, 1.0, 0.0, ez, c, nullptr, NoFin());
        hipLaunchKernelGGL(k_bg_upd, dim3(G), dim3(256), 0, s, n, x, r, sr, p, ap, atp, c, part);
        launch_finalize<2>(s, part, G, c, BgBetaFin());
        hipLaunchKernelGGL(k_bg_p, dim3(G), dim3(256), 0, s, n, p, sp, r, sr, c);
    });
    if (!ge) { bkh_set_error("hipGraph capture of the BiCG iteration failed"); return BKH_DEVICE_ERROR; }
    int rc = run_graph_until_done(t, ge, c, max_iter / K + 2);
    if (rc) return rc;
    return finish_solve(t, A, 0.0, b, x, &inf, info);
}

// ------------------------------------------------------------------- BiCGSTAB
// Replaces src/bicgstab.c:5-72 step for step. Per iteration:
//   SpMV      ap = A p,  epilogue rs.ap -> alpha = rho / (rs.ap)         (bicgstab.c:34-36)
//   k_bs_s    s = r - alpha ap                                           (:38-39)
//   SpMV      as = A s,  epilogue as.s, as.as -> omega                   (:41-43)
//   k_bs_upd  x += alpha p, x += omega s, r = s - omega as; rs.r, r.r    (:45-53, :60)
//             -> beta = alpha / omega * rho_new / rho (left to right), error
//   k_bs_p    p = (p - omega ap) * beta + r                              (:55-58)

enum { BS_RHO = 0, BS_ALPHA = 1, BS_OMEGA = 2, BS_BETA = 3 };

struct BsInitFin {    // rho = rs.r (rs = r), ||b||                         (bicgstab.c:28-29)
    __device__ void operator()(Ctl *c, const double *tot) const {
        c->s[BS_RHO] = tot[0];
        c->fnb = sqrt(tot[1]);
        if (!(c->fnb > 0.0)) ctl_breakdown(c);
    }
};

struct BsAlphaFin {
    __device__ void operator()(Ctl *c, const double *tot) const {
        if (bad(tot[0])) { ctl_breakdown(c); return; }
        c->s[BS_ALPHA] = c->s[BS_RHO] / tot[0];
    }
};

struct BsOmegaFin {   // omega = (as.s) / (as.as)
    __device__ void operator()(Ctl *c, const double *tot) const {
        if (bad(tot[1])) { ctl_breakdown(c); return; }
        c->s[BS_OMEGA] = tot[0] / tot[1];
    }
};

struct BsBetaFin {    // beta = alpha / omega * rho_new / rho; error = ||r|| / ||b||
    __device__ void operator()(Ctl *c, const double *tot) const {
        const double om = c->s[BS_OMEGA], rho = c->s[BS_RHO];
        if (bad(om) || bad(rho)) {
            // a zero r ends the solve as converged (lucky breakdown) through ctl_check
            if (tot[1] == 0.0) ctl_check(c, 0.0);
            else ctl_breakdown(c);
            return;
        }
        c->s[BS_BETA] = c->s[BS_ALPHA] / om * tot[0] / rho;
        c->s[BS_RHO] = tot[0];
        ctl_check(c, tot[1]);
    }
};

__global__ __launch_bounds__(256) void k_bs_init(long long n, const double *b, const double *q, double *r,
                                                 double *rs, double *p, double *part) {
    double v[2] = {0.0, 0.0};
    ROWLOOP(i, n) {
        double ri = b[i] - q[i];
        r[i] = ri; rs[i] = ri; p[i] = ri;
        v[0] += ri * ri;
        v[1] += b[i] * b[i];
    }
    block_partial<2, 256>(v, part);
}

__global__ __launch_bounds__(256) void k_bs_s(long long n, const double *r, const double *ap, double *sv,
                                              const Ctl *c) {
    if (ctl_stopped(c)) return;
    const double al = c->s[BS_ALPHA];
    ROWLOOP(i, n) sv[i] = fma(-al, ap[i], r[i]);
}

__global__ __launch_bounds__(256) void k_bs_upd(long long n, double *x, double *r, const double *rs,
                                                const double *p, const double *sv, const double *as,
                                                const Ctl *c, double *part) {
    if (ctl_stopped(c)) return;
    const double al = c->s[BS_ALPHA], om = c->s[BS_OMEGA];
    double v[2] = {0.0, 0.0};
    ROWLOOP(i, n) {
        const double si = sv[i];
        x[i] = fma(om, si, fma(al, p[i], x[i]));
        const double ri = fma(-om, as[i], si);
        r[i] = ri;
        v[0] += rs[i] * ri;
        v[1] += ri * ri;
    }
    block_partial<2, 256>(v, part);
}

__global__ __launch_bounds__(256) void k_bs_p(long long n, double *p, const double *ap, const double *r,
                                              const Ctl *c) {
    if (ctl_stopped(c)) return;
    const double om = c->s[BS_OMEGA], be = c->s[BS_BETA];
    ROWLOOP(i, n) p[i] = add_rn(mul_rn(fma(-om, ap[i], p[i]), be), r[i]);
}

extern "C" int bkh_bicgstab(bkh_matrix *A, const double *b, double *x, double tol, int max_iter,
                            const bkh_opts *opts, bkh_info *info) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (!A || !b || !x) return BKH_BAD_ARGUMENT;
    bkh_opts o = opts ? *opts : bkh_opts{};
    bkh_info inf;
    memset(&inf, 0, sizeof inf);
    if (max_iter <= 0) {                 // bicgstab.c:9,67-71: error = 0 -> returns 0
        inf.status = ST_CONVERGED;
        if (info) *info = inf;
        return 0;
    }
    const long long n = A->n;
    const int G = vec_grid(n, 256);
    const int K = o.graph_iters > 0 ? o.graph_iters : default_graph_iters(n);
    const int nparts = std::max(spmv_grid(A), G) * 2;
    void *ws = bkh_workspace(carve_size(sizeof(Ctl)) + carve_size(sizeof(double) * nparts) +
                             6 * carve_size(sizeof(double) * n));
    if (!ws) return BKH_DEVICE_ERROR;
    Carve cv(ws);
    Ctl *c = cv.take<Ctl>(1);
    double *part = cv.take<double>(nparts);
    double *r = cv.take<double>(n), *rs = cv.take<double>(n), *sv = cv.take<double>(n),
           *p = cv.take<double>(n), *ap = cv.take<double>(n), *as = cv.take<double>(n);
    hipStream_t s = t->stream;
    HIP_OK(hipEventRecord(t->ev0, s));
    if (ctl_init(t, c, tol, max_iter)) return BKH_DEVICE_ERROR;
    Epi<0> e0;
    memset(&e0, 0, sizeof e0);
    launch_spmv<double, double, 0>(s, A, A->val, x, ap, 1.0, 0.0, e0, nullptr, nullptr, NoFin());
    hipLaunchKernelGGL(k_bs_init, dim3(G), dim3(256), 0, s, n, b, ap, r, rs, p, part);
    launch_finalize<2>(s, part, G, c, BsInitFin());
    LAUNCH_OK();
    hipGraphExec_t ge = capture_cached(t, {0xB5AB, (uintptr_t)A, (uintptr_t)b, (uintptr_t)x, (uintptr_t)r,
                                           (uintptr_t)rs, (uintptr_t)sv, (uintptr_t)p, (uintptr_t)ap, (uintptr_t)as,
                                           (uintptr_t)c, (uintptr_t)part}, s, K, [&]() {
        launch_spmv_dot_w<double, double>(s, A, A->val, p, ap, rs, c, part, BsAlphaFin());
        hipLaunchKernelGGL(k_bs_s, dim3(G), dim3(256), 0, s, n, (const double *)r, (const double *)ap, sv,
                           (const Ctl *)c);
        Epi<2> e2 = make_epi<2>({sv}, {{0, EPI_Y}, {EPI_Y, EPI_Y}});
        launch_spmv<double, double, 2>(s, A, A->val, sv, as, 1.0, 0.0, e2, c, part, BsOmegaFin());
        hipLaunchKernelGGL(k_bs_upd, dim3(G), dim3(256), 0, s, n, x, r, (const double *)rs, (const double *)p,
                           (const double *)sv, (const double *)as, (const Ctl *)c, part);
        launch_finalize<2>(s, part, G, c, BsBetaFin());
        hipLaunchKernelGGL(k_bs_p, dim3(G), dim3(256), 0, s, n, p, (const double *)ap, (const double *)r,
                           (const Ctl *)c);
    });
    if (!ge) { bkh_set_error("hipGraph capture of the BiCGSTAB iteration failed"); return BKH_DEVICE_ERROR; }
    int rc = run_graph_until_done(t, ge, c, max_iter / K + 2);
    if (rc) return rc;
    return finish_solve(t, A, 0.0, b, x, &inf, info);
}
