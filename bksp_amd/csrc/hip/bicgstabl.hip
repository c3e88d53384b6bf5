// bicgstabl.hip -- BiCGSTAB(l) on the device (replaces src/bicgstabl.c:13-138).
//
// Per outer iteration, for each BiCG step j = 0..l-1:
//   k_bl_p      p_i = beta p_i + r_i, i <= j                  (bicgstabl.c:64-67)
//   SpMV        p_{j+1} = A p_j, epilogue rs.p_{j+1} -> alpha  (:68-71)
//   k_bl_r      r_i -= alpha p_{i+1}, i <= j; x += alpha p_0   (:73-75, :79)
//   SpMV        r_{j+1} = A r_j, epilogue rs.r_{j+1} -> beta  (:76-77, :60-62)
//               (for j = l-1 the epilogue is the Gram matrix of r_0..r_l instead)
// then the MR step (bkh_mr.hpp) in the last finalize and one fused update
// k_bl_upd of x, r_0, p_0 with r_0.r_0 (error, :122-125) and rs.r_0 (next beta).
// The whole outer iteration is captured K times into one hipGraph.
#include <math.h>
#include "bkh_host.hpp"
#include "bkh_mr.hpp"
#include "bkh_solver_common.hpp"
#include "bkh_spmv.hpp"

enum { BL_G0 = 0, BL_ALPHA = 1, BL_BETA = 2, BL_OML = 3, BL_L = 4, BL_CX = 8, BL_CR = 20, BL_CP = 32 };

// r0 = b - A x0 (in r0), p0 = rs = r0; tot = {rs.r0, b.b}   (bicgstabl.c:44-51)
struct BlInitFin {
    __device__ void operator()(Ctl *c, const double *tot) const {
        c->fnb = sqrt(tot[1]);
        if (!(c->fnb > 0.0)) { ctl_breakdown(c); return; }
        if (tot[0] == 0.0) {                       // x0 already exact
            c->done = 1; c->status = ST_CONVERGED; c->result_it = 0; return;
        }
        // gamma0 = 1, omega[l] = 1 (:20, :38); first outer step: gamma0 = -omega[l]*gamma0,
        // beta = -alpha*gamma1/gamma0 with alpha = 0, gamma0 = gamma1 (:57-62)
        c->s[BL_ALPHA] = 0.0;
        c->s[BL_BETA] = 0.0;
        c->s[BL_G0] = tot[0];
    }
};

struct BlAlphaFin {   // alpha = gamma1 / (rs . p_{j+1})                 (:71)
    __device__ void operator()(Ctl *c, const double *tot) const {
        if (bad(tot[0])) { ctl_breakdown(c); return; }
        c->s[BL_ALPHA] = c->s[BL_G0] / tot[0];
    }
};

struct BlGammaFin {   // gamma1 = rs . r_{j+1}; beta = -alpha gamma1 / gamma0   (:60-62)
    __device__ void operator()(Ctl *c, const double *tot) const {
        double g1 = tot[0], g0 = c->s[BL_G0];
        if (bad(g0)) { ctl_breakdown(c); return; }
        c->s[BL_BETA] = -c->s[BL_ALPHA] * g1 / g0;
        c->s[BL_G0] = g1;
    }
};

// l is a compile-time constant, so the MR step's small arrays stay in registers.
template <int L>
struct BlMrFin {      // Gram of r_0..r_l -> MR coefficients on the unmodified r_k
    static constexpr int l = L;
    __device__ void operator()(Ctl *c, const double *tot) const {
        double G[MR_D][MR_D];
        gram_unpack(l, tot, G);
        if (G[0][0] == 0.0) {                       // r_0 = 0 exactly: solved
            c->done = 1; c->status = ST_CONVERGED; c->result_it = c->it; return;
        }
        MrOut o;
        if (!mr_from_gram(l, G, o)) { ctl_breakdown(c); return; }
        c->s[BL_CX] = o.om[1];                                   // (:113, :119)
        for (int k = 1; k < l; k++) {
            double t = 0.0;
            for (int i = k; i < l; i++) t += o.odd[i] * o.C[i][k];
            c->s[BL_CX + k] = t;
        }
        for (int k = 1; k <= l; k++) {                           // (:114, :118)
            double t = 0.0;
            for (int i = k; i <= l; i++) t += o.od[i] * o.C[i][k];
            c->s[BL_CR + k] = t;
            c->s[BL_CP + k] = o.om[k];                           // (:115, :117)
        }
        c->s[BL_OML] = o.om[l];
    }
};

struct BlUpdFin {     // error check; next outer iteration's first beta  (:122-125, :57-62)
    __device__ void operator()(Ctl *c, const double *tot) const {
        ctl_check(c, tot[0]);
        if (c->done) return;
        double g0 = -c->s[BL_OML] * c->s[BL_G0];
        if (bad(g0)) { ctl_breakdown(c); return; }
        c->s[BL_BETA] = -c->s[BL_ALPHA] * tot[1] / g0;
        c->s[BL_G0] = tot[1];
    }
};

#define ROWLOOP(i, n) \
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < (n); i += (long long)gridDim.x * 256)

template <typename T>
__global__ __launch_bounds__(256) void k_bl_init(long long n, const T *b, T *r0, T *p0, T *rs,
                                                 double *part) {
    double v[2] = {0.0, 0.0};
    ROWLOOP(i, n) {
        T ri = b[i] - r0[i];
        r0[i] = ri; p0[i] = ri; rs[i] = ri;
        v[0] += (double)ri * (double)ri;
        v[1] += (double)b[i] * (double)b[i];
    }
    block_partial<2, 256>(v, part);
}

template <typename T>
__global__ __launch_bounds__(256) void k_bl_p(long long n, int j, T *P, const T *R, const Ctl *c) {
    if (ctl_stopped(c)) return;
    const T be = (T)c->s[BL_BETA];
    ROWLOOP(i, n) {
        for (int q = 0; q <= j; q++) {
            T *p = P + q * n;
            p[i] = add_rn(mul_rn(be, p[i]), R[q * n + i]);
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_bl_r(long long n, int j, const T *P, T *R, T *x,
                                              const Ctl *c) {
    if (ctl_stopped(c)) return;
    const T al = (T)c->s[BL_ALPHA];
    ROWLOOP(i, n) {
        for (int q = 0; q <= j; q++) R[q * n + i] = fma(-al, P[(q + 1) * n + i], R[q * n + i]);
        x[i] = fma(al, P[i], x[i]);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_bl_upd(long long n, int l, T *P, T *R, T *x, const T *rs,
                                                const Ctl *c, double *part) {
    if (ctl_stopped(c)) return;
    double cx[MR_D], cr[MR_D], cp[MR_D];
    for (int k = 0; k <= l; k++) {
        cx[k] = c->s[BL_CX + k]; cr[k] = c->s[BL_CR + k]; cp[k] = c->s[BL_CP + k];
    }
    double v[2] = {0.0, 0.0};
    ROWLOOP(i, n) {
        double xi = (double)x[i], r0 = (double)R[i], p0 = (double)P[i];
        xi += cx[0] * r0;
        for (int k = 1; k <= l; k++) {
            double rk = (double)R[k * n + i];
            if (k < l) xi += cx[k] * rk;
            r0 -= cr[k] * rk;
            p0 -= cp[k] * (double)P[k * n + i];
        }
        x[i] = (T)xi; R[i] = (T)r0; P[i] = (T)p0;
        double r0t = (double)(T)r0;
        v[0] += r0t * r0t;
        v[1] += (double)rs[i] * r0t;
    }
    block_partial<2, 256>(v, part);
}

// SpMV r_{l} = A r_{l-1} with the Gram epilogue of [r_0..r_{l-1}, r_l].
template <int L, typename VT, typename T>
static void launch_gram_spmv(hipStream_t s, const bkh_matrix *A, const VT *val, T *R, long long n,
                             Ctl *c, double *part) {
    constexpr int NV = (L + 1) * (L + 2) / 2;
    Epi<NV, T> e;
    memset(&e, 0, sizeof e);
    for (int k = 0; k < L; k++) e.w[k] = R + k * n;
    int idx = 0;
    for (int a = 0; a <= L; a++)
        for (int b = a; b <= L; b++) {
            e.a[idx] = (unsigned char)(a == L ? EPI_Y : a);
            e.b[idx] = (unsigned char)(b == L ? EPI_Y : b);
            idx++;
        }
    BlMrFin<L> fin;
    launch_spmv<VT, T, NV>(s, A, val, R + (L - 1) * n, R + L * n, 1.0, 0.0, e, c, part, fin);
}

template <typename VT, typename T>
static void gram_spmv(int l, hipStream_t s, const bkh_matrix *A, const VT *val, T *R, long long n,
                      Ctl *c, double *part) {
    switch (l) {
    case 1: launch_gram_spmv<1>(s, A, val, R, n, c, part); break;
    case 2: launch_gram_spmv<2>(s, A, val, R, n, c, part); break;
    case 3: launch_gram_spmv<3>(s, A, val, R, n, c, part); break;
    case 4: launch_gram_spmv<4>(s, A, val, R, n, c, part); break;
    case 5: launch_gram_spmv<5>(s, A, val, R, n, c, part); break;
    case 6: launch_gram_spmv<6>(s, A, val, R, n, c, part); break;
    case 7: launch_gram_spmv<7>(s, A, val, R, n, c, part); break;
    default: launch_gram_spmv<8>(s, A, val, R, n, c, part); break;
    }
}

// One BiCGSTAB(l) solve on vectors of type T; x updated in place. Workspace:
// R = r_0..r_l, P = p_0..p_l, rs (n each).
template <typename T, typename VT>
static int bl_core(ThreadCtx *t, const bkh_matrix *A, const VT *val, int l, const T *b, T *x, T *R,
                   T *P, T *rs, Ctl *c, double *part, double tol, int max_it, int K) {
    hipStream_t s = t->stream;
    const long long n = A->n;
    const int G = vec_grid(n, 256);
    if (ctl_init(t, c, tol, max_it)) return BKH_DEVICE_ERROR;
    Epi<0, T> e0;
    memset(&e0, 0, sizeof e0);
    launch_spmv<VT, T, 0>(s, A, val, x, R, 1.0, 0.0, e0, nullptr, nullptr, NoFin());
    hipLaunchKernelGGL(k_bl_init<T>, dim3(G), dim3(256), 0, s, n, b, R, P, rs, part);
    launch_finalize<2>(s, part, G, c, BlInitFin());
    LAUNCH_OK();
    hipGraphExec_t ge = capture_cached(t, {0xB100, (uintptr_t)A, (uintptr_t)val, (uintptr_t)b, (uintptr_t)x, (uintptr_t)R,
                                           (uintptr_t)P, (uintptr_t)rs, (uintptr_t)c, (uintptr_t)part, (uintptr_t)l,
                                           (uintptr_t)sizeof(T)}, s, K, [&]() {
        for (int j = 0; j < l; j++) {
            hipLaunchKernelGGL(k_bl_p<T>, dim3(G), dim3(256), 0, s, n, j, P, (const T *)R, c);
            launch_spmv_dot_w<VT, T>(s, A, val, P + j * n, P + (j + 1) * n, rs, c, part, BlAlphaFin());
            hipLaunchKernelGGL(k_bl_r<T>, dim3(G), dim3(256), 0, s, n, j, (const T *)P, R, x, c);
            if (j < l - 1)
                launch_spmv_dot_w<VT, T>(s, A, val, R + j * n, R + (j + 1) * n, rs, c, part, BlGammaFin());
            else
                gram_spmv<VT, T>(l, s, A, val, R, n, c, part);
        }
        hipLaunchKernelGGL(k_bl_upd<T>, dim3(G), dim3(256), 0, s, n, l, P, R, x, (const T *)rs, c, part);
        launch_finalize<2>(s, part, G, c, BlUpdFin());
    });
    if (!ge) { bkh_set_error("hipGraph capture of the BiCGSTAB(l) iteration failed"); return BKH_DEVICE_ERROR; }
    int rc = run_graph_until_done(t, ge, c, max_it / K + 2);
    return rc;
}

bool bicgstabl_small(ThreadCtx *t, const bkh_matrix *A, int l, const double *b, double *x, double tol,
                     int max_iter, int *rc);

static int bl_parts(const bkh_matrix *A, int l) {
    int nv = (l + 1) * (l + 2) / 2;
    return std::max(spmv_grid(A), vec_grid(A->n, 256)) * std::max(nv, 2);
}

extern "C" int bkh_bicgstabl(bkh_matrix *A, const double *b, int l, double *x, double tol,
                             int max_iter, const bkh_opts *opts, bkh_info *info) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (!A || !b || !x) return BKH_BAD_ARGUMENT;
    if (l < 1 || l > MR_LMAX) { bkh_set_error("BiCGSTAB(l) supports 1 <= l <= %d", MR_LMAX); return -3; }
    bkh_opts o = opts ? *opts : bkh_opts{};
    bkh_info inf;
    memset(&inf, 0, sizeof inf);
    const long long n = A->n;
    if (max_iter <= 0) {                 // bicgstabl.c:18,134-137: error = 0 -> returns m = 0
        inf.status = ST_CONVERGED;
        if (info) *info = inf;
        return 0;
    }
    const int K = o.graph_iters > 0 ? o.graph_iters : std::max(2, default_graph_iters(n) / (2 * l));
    const int nparts = bl_parts(A, l);
    int ret;
    HIP_OK(hipEventRecord(t->ev0, t->stream));
    if (o.precision == BKH_FP32_IR) {
        if (!A->val32) { bkh_set_error("FP32_IR needs a matrix created with want_fp32"); return BKH_BAD_ARGUMENT; }
        size_t bytes = 2 * carve_size(sizeof(Ctl)) + carve_size(sizeof(double) * nparts) +
                       carve_size(sizeof(double) * n) +
                       (2 * (l + 1) + 3) * carve_size(sizeof(float) * n);
        void *ws = bkh_workspace(bytes);
        if (!ws) return BKH_DEVICE_ERROR;
        Carve cv(ws);
        Ctl *co = cv.take<Ctl>(1), *ci = cv.take<Ctl>(1);
        double *part = cv.take<double>(nparts);
        double *r64 = cv.take<double>(n);
        float *r32 = cv.take<float>(n), *d32 = cv.take<float>(n), *rs = cv.take<float>(n);
        float *R = cv.take<float>((l + 1) * n), *P = cv.take<float>((l + 1) * n);
        IrResult ir;
        ret = ir_refine(t, A, b, x, r64, co, part, tol, max_iter, o, &ir,
                        [&](double itol, int imax, int *iters) -> int {
                            int rc2 = bl_core<float, float>(t, A, A->val32, l, r32, d32, R, P, rs, ci,
                                                            part, itol, imax, K);
                            if (rc2) return rc2;
                            *iters = t->ctl_host->it;
                            return t->ctl_host->status == ST_BREAKDOWN ? -2 : 0;
                        },
                        r32, d32);
        inf.inner_iters = ir.inner_total;
        inf.passes = ir.passes;
        inf.status = ir.status;
        inf.iters = ir.ret;
        inf.rec_res = ir.res;
    } else {
        int src = 0;
        if (o.path != 1 && (o.path == 2 || n <= 16384) && bicgstabl_small(t, A, l, b, x, tol, max_iter, &src)) {
            if (src) return src;
            return finish_solve(t, A, 0.0, b, x, &inf, info);
        }
        size_t bytes = carve_size(sizeof(Ctl)) + carve_size(sizeof(double) * nparts) +
                       (2 * (l + 1) + 1) * carve_size(sizeof(double) * n);
        void *ws = bkh_workspace(bytes);
        if (!ws) return BKH_DEVICE_ERROR;
        Carve cv(ws);
        Ctl *c = cv.take<Ctl>(1);
        double *part = cv.take<double>(nparts);
        double *R = cv.take<double>((l + 1) * n), *P = cv.take<double>((l + 1) * n);
        double *rs = cv.take<double>(n);
        int rc = bl_core<double, double>(t, A, A->val, l, b, x, R, P, rs, c, part, tol, max_iter, K);
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
