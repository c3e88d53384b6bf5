// bkh_solver_common.hpp -- pieces shared by the solver translation units.
#pragma once
#include <algorithm>
#include "bkh_host.hpp"
#include "bkh_spmv.hpp"

void bkh_note_solve_ms(double ms);

// Iterations per captured graph: long enough to amortise the host check, short
// enough that the no-op replays after convergence cost little.
static inline int default_graph_iters(long long n) {
    if (n >= (1LL << 20)) return 16;
    if (n >= 65536) return 32;
    return 64;
}

// y = A x with the x.y dot product in the epilogue.
template <typename VT, typename T, class Fin>
static inline void launch_spmv_dot_self(hipStream_t s, const bkh_matrix *A, const VT *val,
                                        const T *x, T *y, Ctl *c, double *part, Fin fin) {
    Epi<1, T> e;
    memset(&e, 0, sizeof e);
    e.w[0] = x;
    e.a[0] = 0;
    e.b[0] = EPI_Y;
    launch_spmv<VT, T, 1>(s, A, val, x, y, 1.0, 0.0, e, c, part, fin);
}

// y = A x with the w.y dot product in the epilogue.
template <typename VT, typename T, class Fin>
static inline void launch_spmv_dot_w(hipStream_t s, const bkh_matrix *A, const VT *val,
                                     const T *x, T *y, const T *w, Ctl *c, double *part, Fin fin) {
    Epi<1, T> e;
    memset(&e, 0, sizeof e);
    e.w[0] = w;
    e.a[0] = 0;
    e.b[0] = EPI_Y;
    launch_spmv<VT, T, 1>(s, A, val, x, y, 1.0, 0.0, e, c, part, fin);
}

// FP64 residual r = b - (A + sigma I) x, asynchronously on t's stream;
// c->s[0] = ||r||/||b||, c->s[1] = ||r||, c->s[2] = ||b||. `r` doubles as the
// SpMV output. `part` needs max(spmv_grid(A), 4096) * 2 doubles.
int resid_fp64(ThreadCtx *t, const bkh_matrix *A, double sigma, const double *b, const double *x,
               double *r, Ctl *c, double *part);
// Allocating, synchronous FP64 true residual ||b - (A + sigma I) x|| / ||b||.
int true_residual(ThreadCtx *t, const bkh_matrix *A, double sigma, const double *b,
                  const double *x, double *out);
// r32 = r64 / ||r64|| (||r64|| = c->s[1]), d32 = 0.
int ir_prep(ThreadCtx *t, long long n, const double *r64, const Ctl *c, float *r32, float *d32);
// x += ||r64|| * d32.
int ir_update(ThreadCtx *t, long long n, double *x, const float *d32, const Ctl *c);

// Ends a solve whose start was recorded on t->ev0: fills `inf` from the fetched
// control block (t->ctl_host) and the event time, adds the FP64 true residual
// ||b - (A + sigma I) x|| / ||b|| when `info` is requested, and returns the
// bksp.h return value.
int finish_solve(ThreadCtx *t, const bkh_matrix *A, double sigma, const double *b,
                 const double *x, bkh_info *inf, bkh_info *info);

struct IrResult {
    int ret = -1, status = 0, passes = 0, inner_total = 0;
    double res = 0.0;
};

// FP64 iterative refinement around an FP32 inner solve.
//   loop: r = b - A x (FP64); stop if ||r||/||b|| < tol;
//         solve A d = r/||r|| in FP32 to a relative tolerance; x += ||r|| d (FP64).
// `inner(inner_tol, inner_max, &iters)` solves for d32 given r32 and returns 0,
// -2 (breakdown) or a device error. The return value follows bksp.h: the total
// inner iteration count minus one when converged, -1 or -2 otherwise.
template <class Inner>
int ir_refine(ThreadCtx *t, const bkh_matrix *A, const double *b, double *x, double *r64, Ctl *co,
              double *part, double tol, int max_iter, const bkh_opts &o, IrResult *res,
              Inner inner, float *r32, float *d32) {
    int total = 0;
    const int max_inner = o.max_inner > 0 ? o.max_inner : max_iter;
    if (ctl_init(t, co, 0.0, 1)) return BKH_DEVICE_ERROR;   // the residual finalizers run on co
    for (int pass = 0;; pass++) {
        if (resid_fp64(t, A, 0.0, b, x, r64, co, part)) return BKH_DEVICE_ERROR;
        if (ctl_fetch(t, co)) return BKH_DEVICE_ERROR;
        const double rel = t->ctl_host->s[0];
        res->res = rel;
        res->passes = pass;
        res->inner_total = total;
        if (rel < tol) {
            res->status = ST_CONVERGED;
            res->ret = total > 0 ? total - 1 : 0;
            return res->ret;
        }
        if (!(rel == rel) || rel > 1e300) { res->status = ST_BREAKDOWN; res->ret = -2; return -2; }
        if (total >= max_iter) { res->status = ST_MAXIT; res->ret = -1; return -1; }
        double itol = o.inner_tol > 0 ? o.inner_tol : 1e-5;
        const double need = 0.5 * tol / rel;   // last pass: stop at what is still needed
        if (need > itol) itol = need;
        if (ir_prep(t, A->n, r64, co, r32, d32)) return BKH_DEVICE_ERROR;
        int it = 0;
        int rc = inner(itol, std::min(max_inner, max_iter - total), &it);
        if (rc == BKH_DEVICE_ERROR) return rc;
        total += it;
        if (rc == -2) {
            // an inner breakdown leaves d unusable; stop with the FP64 x as it is
            res->inner_total = total;
            res->status = ST_BREAKDOWN;
            res->ret = -2;
            return -2;
        }
        if (ir_update(t, A->n, x, d32, co)) return BKH_DEVICE_ERROR;
    }
}
