// shbicgstabl.hip -- shifted BiCGSTAB(l) on the device (replaces src/shbicgstabl.c:7-285
// with src/util_shift.c:5-27 dhorner and :55-63 comb).
//
// The seed system (sigma = 0) runs BiCGSTAB(l) as in bicgstabl.hip. The per-shift
// work of each BiCG step runs in one kernel over all shifts (k_sh_shift_step): a
// 2-D grid of (rows x groups of SH_SPT shifts) in which every thread reads the
// seed vectors of its row and issues the loads of its shifts' p_shift vectors
// together (shbicgstabl.c:111-134, :146-153); the seed's own r/x update
// (:107-108, :137-139) follows in k_sh_seed_r, so every shift thread reads the
// pre-update r. The MR step's per-shift updates (:202-237, k_sh_shift_upd) and
// the seed update (:194-199, :240-242, k_sh_seed_upd) are ordered the same way.
// The pi/theta/psi recurrences and the Horner re-expansion are FP64 scalar work
// in the finalize kernels, one shift per thread. x holds the seed solution first,
// then shift k at x + (k+1) n (shbicgstabl.c:62). PT is the storage type of
// p_shift (double, or float for the reduced-traffic variant).
#include <math.h>
#include "bkh_host.hpp"
#include "bkh_mr.hpp"
#include "bkh_solver_common.hpp"
#include "bkh_spmv.hpp"

#define SH_MAXS 256

enum { SB_G0 = 0, SB_ALPHA = 1, SB_BETA = 2, SB_OML = 3, SB_AOLD = 4, SB_CX = 8, SB_CR = 20, SB_CP = 32 };

struct ShState {
    int ns, l;
    double sigma[SH_MAXS], pi[SH_MAXS], pi_old[SH_MAXS], theta[SH_MAXS];
    // per-step coefficients of k_sh_step
    double cA[SH_MAXS], cB[SH_MAXS], cC[SH_MAXS], cD[SH_MAXS], cE[SH_MAXS], ak[SH_MAXS];
    // MR-step coefficients of k_sh_upd
    double xc[SH_MAXS][MR_LMAX], psi_inv[SH_MAXS], om[MR_D];
};

// r0 = b - A x0, p0 = rs = r0 (:32-42); pi = pi_old = theta = 1 (:60-66)
struct ShInitFin {
    ShState *st;
    __device__ void operator()(Ctl *c, const double *tot) const {
        c->fnb = sqrt(tot[1]);
        if (!(c->fnb > 0.0)) { ctl_breakdown(c); return; }
        if (tot[0] == 0.0) { c->done = 1; c->status = ST_CONVERGED; c->result_it = 0; return; }
        c->s[SB_ALPHA] = 0.0;
        c->s[SB_BETA] = 0.0;
        c->s[SB_G0] = tot[0];
        c->s[SB_AOLD] = 1.0;
        for (int k = 0; k < st->ns; k++) { st->pi[k] = 1.0; st->pi_old[k] = 1.0; st->theta[k] = 1.0; }
    }
};

// alpha = rho / (rs . p_{j+1}) (:105); per-shift pi_new, beta_k, alpha_k (:113-121)
// and the coefficients of this step's fused update, one shift per thread; then
// pi_old <- pi, pi <- pi_new and alpha_old <- alpha (:156-161).
__global__ __launch_bounds__(256) void k_sh_alpha_fin(const double *part, int nblocks, Ctl *c,
                                                      ShState *st) {
    if (ctl_stopped(c)) return;
    double acc[1];
    sum_partials<1>(part, nblocks, acc);
    __shared__ double al_s, be_s, ao_s;
    __shared__ int ok;
    if (threadIdx.x == 0) {
        ok = !bad(acc[0]);
        if (!ok) {
            ctl_breakdown(c);
        } else {
            al_s = c->s[SB_G0] / acc[0];
            be_s = -c->s[SB_BETA];                  // shbicgstabl.c's beta (p = -beta p + r)
            ao_s = c->s[SB_AOLD];
            c->s[SB_ALPHA] = al_s;
            c->s[SB_AOLD] = al_s;
        }
    }
    __syncthreads();
    if (!ok) return;
    const double al = al_s, be = be_s, aold = ao_s;
    for (int k = threadIdx.x; k < st->ns; k += blockDim.x) {
        const double pi = st->pi[k], po = st->pi_old[k], sg = st->sigma[k], th = st->theta[k];
        const double pn = (1.0 + al * sg) * pi + al * be / aold * (po - pi);
        const double bk = (po / pi) * (po / pi) * be;
        const double ak = pi / pn * al;
        st->cA[k] = -bk;
        st->cB[k] = 1.0 / (pi * th);
        st->cC[k] = 1.0 / pi;
        st->cD[k] = -1.0 / pn;
        st->cE[k] = 1.0 / (th * ak);
        st->ak[k] = ak;
        st->pi_old[k] = pi;
        st->pi[k] = pn;
        if (!(pn == pn) || pn == 0.0) ctl_breakdown(c);
    }
}

struct ShGammaFin {   // rho = r_{j+1} . rs; beta = alpha rho / rho_old (:91-92)
    __device__ void operator()(Ctl *c, const double *tot) const {
        double g1 = tot[0], g0 = c->s[SB_G0];
        if (bad(g0)) { ctl_breakdown(c); return; }
        c->s[SB_BETA] = -c->s[SB_ALPHA] * g1 / g0;
        c->s[SB_G0] = g1;
    }
};

__device__ __forceinline__ double binom(int n, int r) {
    double c = 1.0;
    for (int i = 1; i <= r; i++) c = c * (n - r + i) / i;
    return c;
}

// Seed MR step (Gram -> coefficients, :166-199) in thread 0, then, one shift per
// thread, the Horner re-expansion of the MR polynomial (util_shift.c:5-27), the
// theta update and the x_k / p_shift[k][0] coefficients (:202-237), re-expressed
// on the unmodified r_0..r_{l-1}.
template <int L>
__global__ __launch_bounds__(256) void k_sh_mr_fin(const double *part, int nblocks, Ctl *c, ShState *st) {
    if (ctl_stopped(c)) return;
    constexpr int NV = (L + 1) * (L + 2) / 2;
    double acc[NV];
    sum_partials<NV>(part, nblocks, acc);
    __shared__ MrOut o;
    __shared__ int ok;
    const int l = L;
    if (threadIdx.x == 0) {
        double G[MR_D][MR_D];
        gram_unpack(l, acc, G);
        ok = 1;
        if (G[0][0] == 0.0) {
            c->done = 1; c->status = ST_CONVERGED; c->result_it = c->it; ok = 0;
        } else if (!mr_from_gram(l, G, o)) {
            ctl_breakdown(c); ok = 0;
        } else {
            c->s[SB_CX] = o.om[1];
            for (int k = 1; k < l; k++) {
                double t = 0.0;
                for (int i = k; i < l; i++) t += o.odd[i] * o.C[i][k];
                c->s[SB_CX + k] = t;
            }
            for (int k = 1; k <= l; k++) {
                double t = 0.0;
                for (int i = k; i <= l; i++) t += o.od[i] * o.C[i][k];
                c->s[SB_CR + k] = t;
                c->s[SB_CP + k] = o.om[k];
                st->om[k] = o.om[k];
            }
            c->s[SB_OML] = o.om[l];
        }
    }
    __syncthreads();
    if (!ok) return;
    // shbicgstabl.c indexes omega[0..l-1] = om[1..l] and tau_s[i][j] = tau[i+1][j+1]
    for (int k = threadIdx.x; k < st->ns; k += blockDim.x) {
        double out[MR_D], oms[MR_LMAX], osd[MR_LMAX], osdd[MR_LMAX];
        const double sg = st->sigma[k];
        out[l] = -o.om[l];                                       // dhorner
        for (int i = l - 1; i > 0; i--) out[i] = -sg * out[i + 1] - o.om[i];
        out[0] = -sg * out[1] + 1.0;
        for (int i = 1; i < l; i++)
            for (int j = l - 1; j >= i; j--) out[j] += -sg * out[j + 1];
        for (int i = 1; i <= l; i++) out[i] = -out[i] / out[0];
        const double psi = out[0];
        for (int i = 0; i < l; i++) oms[i] = out[i + 1];
        const double e = st->theta[k] * st->pi[k];
        st->theta[k] *= psi;
        for (int j = 0; j < l; j++) {                            // omega_shift_d (:212-218)
            double t = 0.0;
            for (int i = j; i < l; i++) {
                double mij = (i == j) ? 1.0 : pow(sg, (double)(i - j)) * binom(i, j);
                t += mij * oms[i];
            }
            osd[j] = t;
        }
        for (int j = 0; j < l - 1; j++) {                        // omega_shift_dd (:220-226)
            double t = 0.0;
            for (int i = j + 1; i < l - 1; i++) t += o.tau[j + 1][i + 1] * osd[i + 1];
            osdd[j] = osd[j + 1] + t;
        }
        st->xc[k][0] = osd[0] / e;                               // (:229-234)
        for (int kk = 1; kk < l; kk++) {
            double t = 0.0;
            for (int i = kk; i < l; i++) t += osdd[i - 1] * o.C[i][kk];
            st->xc[k][kk] = t / e;
        }
        st->psi_inv[k] = 1.0 / psi;
        if (!(psi == psi) || psi == 0.0 || !(e == e) || e == 0.0) ctl_breakdown(c);
    }
}

struct ShUpdFin {     // seed error check; next outer iteration's beta (:244-247, :88-92)
    __device__ void operator()(Ctl *c, const double *tot) const {
        ctl_check(c, tot[0]);
        if (c->done) return;
        double g0 = -c->s[SB_OML] * c->s[SB_G0];
        if (bad(g0)) { ctl_breakdown(c); return; }
        c->s[SB_BETA] = -c->s[SB_ALPHA] * tot[1] / g0;
        c->s[SB_G0] = tot[1];
    }
};

#define ROWLOOP(i, n) \
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < (n); i += (long long)gridDim.x * 256)

__global__ __launch_bounds__(256) void k_sh_init(long long n, const double *b, double *r0, double *p0,
                                                 double *rs, double *part) {
    double v[2] = {0.0, 0.0};
    ROWLOOP(i, n) {
        double ri = b[i] - r0[i];
        r0[i] = ri; p0[i] = ri; rs[i] = ri;
        v[0] += ri * ri;
        v[1] += b[i] * b[i];
    }
    block_partial<2, 256>(v, part);
}

__global__ __launch_bounds__(256) void k_sh_p(long long n, int j, double *P, const double *R, const Ctl *c) {
    if (ctl_stopped(c)) return;
    const double be = c->s[SB_BETA];
    ROWLOOP(i, n) {
        for (int q = 0; q <= j; q++) P[q * n + i] = add_rn(mul_rn(be, P[q * n + i]), R[q * n + i]);
    }
}

// Seed part of BiCG step j: r_q -= alpha p_{q+1} (q <= j), x += alpha p_0
// (:107-108, :137-139). Runs after the shift kernel, which reads the old r.
__global__ __launch_bounds__(256) void k_sh_seed_r(long long n, int j, const double *P, double *R,
                                                   double *x, const Ctl *c) {
    if (ctl_stopped(c)) return;
    const double al = c->s[SB_ALPHA];
    ROWLOOP(i, n) {
        for (int q = 0; q <= j; q++) R[q * n + i] = fma(-al, P[(q + 1) * n + i], R[q * n + i]);
        x[i] = fma(al, P[i], x[i]);
    }
}

#define SH_SPT 8   // shifts per thread

// Shift part of BiCG step J for shifts [blockIdx.y*SH_SPT, +SH_SPT): p_shift[k][0..J+1]
// and x_k (:124-133, :146-153). r_J after the seed update is recomputed here with
// the seed kernel's exact operation, so every thread reads the pre-update r.
template <int J, typename PT>
__global__ __launch_bounds__(256) void k_sh_shift_step(long long n, int l, const double *P, const double *R,
                                                       double *x, PT *PS, const ShState *st, const Ctl *c) {
    if (ctl_stopped(c)) return;
    const double al = c->s[SB_ALPHA];
    const int ns = st->ns, l1 = l + 1, k0 = blockIdx.y * SH_SPT;
    ROWLOOP(i, n) {
        double rold[J + 1];
#pragma unroll
        for (int q = 0; q <= J; q++) rold[q] = R[q * n + i];
        const double rjn = fma(-al, P[(J + 1) * n + i], rold[J]);
        double pv[SH_SPT][J + 1], xv[SH_SPT];
#pragma unroll
        for (int u = 0; u < SH_SPT; u++) {
            const int k = k0 + u;
            if (k < ns) {
                const PT *ps = PS + (size_t)k * l1 * n;
#pragma unroll
                for (int q = 0; q <= J; q++) pv[u][q] = (double)ps[q * n + i];
                xv[u] = x[(size_t)(k + 1) * n + i];
            }
        }
#pragma unroll
        for (int u = 0; u < SH_SPT; u++) {
            const int k = k0 + u;
            if (k < ns) {
                PT *ps = PS + (size_t)k * l1 * n;
                const double cA = st->cA[k], cB = st->cB[k];
                double p0 = 0.0, pj = 0.0;
#pragma unroll
                for (int q = 0; q <= J; q++) {
                    const PT v = (PT)fma(cB, rold[q], mul_rn(cA, pv[u][q]));
                    ps[q * n + i] = v;
                    if (q == 0) p0 = (double)v;
                    if (q == J) pj = (double)v;
                }
                double t = mul_rn(rold[J], st->cC[k]);
                t = fma(st->cD[k], rjn, t);
                t = mul_rn(t, st->cE[k]);
                t = fma(-st->sigma[k], pj, t);
                ps[(J + 1) * n + i] = (PT)t;
                x[(size_t)(k + 1) * n + i] = fma(st->ak[k], p0, xv[u]);
            }
        }
    }
}

// Seed MR update (x, r_0, p_0; r_0.r_0 and rs.r_0), after the shift update.
__global__ __launch_bounds__(256) void k_sh_seed_upd(long long n, int l, double *P, double *R, double *x,
                                                     const double *rs, const Ctl *c, double *part) {
    if (ctl_stopped(c)) return;
    double cx[MR_D], cr[MR_D], cp[MR_D];
    for (int k = 0; k <= l; k++) {
        cx[k] = c->s[SB_CX + k]; cr[k] = c->s[SB_CR + k]; cp[k] = c->s[SB_CP + k];
    }
    double v[2] = {0.0, 0.0};
    ROWLOOP(i, n) {
        double r0 = R[i], xi = x[i] + cx[0] * r0, p0 = P[i];
        for (int k = 1; k <= l; k++) {
            const double rk = R[k * n + i];
            if (k < l) xi += cx[k] * rk;
            r0 -= cr[k] * rk;
            p0 -= cp[k] * P[k * n + i];
        }
        x[i] = xi; R[i] = r0; P[i] = p0;
        v[0] += r0 * r0;
        v[1] += rs[i] * r0;
    }
    block_partial<2, 256>(v, part);
}

// Shift part of the MR step (:202-237): x_k += sum_q xc[k][q] r_q (original r),
// p_shift[k][0] = (p_shift[k][0] - sum_i omega_i p_shift[k][i]) / psi_k.
template <int L, typename PT>
__global__ __launch_bounds__(256) void k_sh_shift_upd(long long n, const double *R, double *x, PT *PS,
                                                      const ShState *st, const Ctl *c) {
    if (ctl_stopped(c)) return;
    const int ns = st->ns, k0 = blockIdx.y * SH_SPT;
    ROWLOOP(i, n) {
        double r[L];
#pragma unroll
        for (int q = 0; q < L; q++) r[q] = R[q * n + i];
        double pv[SH_SPT][L + 1], xv[SH_SPT];
#pragma unroll
        for (int u = 0; u < SH_SPT; u++) {
            const int k = k0 + u;
            if (k < ns) {
                const PT *ps = PS + (size_t)k * (L + 1) * n;
#pragma unroll
                for (int q = 0; q <= L; q++) pv[u][q] = (double)ps[q * n + i];
                xv[u] = x[(size_t)(k + 1) * n + i];
            }
        }
#pragma unroll
        for (int u = 0; u < SH_SPT; u++) {
            const int k = k0 + u;
            if (k < ns) {
                double xk = xv[u];
#pragma unroll
                for (int q = 0; q < L; q++) xk += st->xc[k][q] * r[q];
                x[(size_t)(k + 1) * n + i] = xk;
                double p = pv[u][0];
#pragma unroll
                for (int q = 1; q <= L; q++) p -= st->om[q] * pv[u][q];
                PS[(size_t)k * (L + 1) * n + i] = (PT)(p * st->psi_inv[k]);
            }
        }
    }
}

__global__ void k_zero_pt(float *p, long long n) { ROWLOOP(i, n) p[i] = 0.0f; }
__global__ void k_zero_pd(double *p, long long n) { ROWLOOP(i, n) p[i] = 0.0; }

// SpMV r_L = A r_{L-1} with the Gram epilogue of [r_0..r_L], then the MR finalize.
template <int L>
static void sh_gram_spmv(hipStream_t s, const bkh_matrix *A, double *R, long long n, Ctl *c,
                         double *part, ShState *st) {
    constexpr int NV = (L + 1) * (L + 2) / 2;
    Epi<NV, double> e;
    memset(&e, 0, sizeof e);
    for (int k = 0; k < L; k++) e.w[k] = R + k * n;
    int idx = 0;
    for (int a = 0; a <= L; a++)
        for (int b = a; b <= L; b++) {
            e.a[idx] = (unsigned char)(a == L ? EPI_Y : a);
            e.b[idx] = (unsigned char)(b == L ? EPI_Y : b);
            idx++;
        }
    const int G = launch_spmv_kernel<double, double, NV>(s, A, A->val, R + (L - 1) * n, R + L * n, 1.0, 0.0,
                                                         e, c, part, NoFin());
    double *r2 = bkh_red2();
    const int nb2 = (G + RED_CHUNK - 1) / RED_CHUNK;
    hipLaunchKernelGGL(k_part_reduce, dim3(nb2), dim3(256), 0, s, (const double *)part, G, NV, r2);
    hipLaunchKernelGGL((k_sh_mr_fin<L>), dim3(1), dim3(256), 0, s, (const double *)r2, nb2, c, st);
}

// p_{j+1} = A p_j with rs.p_{j+1}, then the per-shift alpha finalize.
static void sh_alpha_spmv(hipStream_t s, const bkh_matrix *A, const double *x, double *y, const double *rs,
                          Ctl *c, double *part, ShState *st) {
    Epi<1, double> e;
    memset(&e, 0, sizeof e);
    e.w[0] = rs;
    e.a[0] = 0;
    e.b[0] = EPI_Y;
    const int G = launch_spmv_kernel<double, double, 1>(s, A, A->val, x, y, 1.0, 0.0, e, c, part, NoFin());
    hipLaunchKernelGGL(k_sh_alpha_fin, dim3(1), dim3(256), 0, s, (const double *)part, G, c, st);
}

template <typename PT>
static void sh_shift_step(int j, dim3 grid, hipStream_t s, long long n, int l, const double *P,
                          const double *R, double *x, PT *PS, const ShState *st, const Ctl *c) {
    switch (j) {
#define CASE(J_) case J_: hipLaunchKernelGGL((k_sh_shift_step<J_, PT>), grid, dim3(256), 0, s, n, l, P, R, x, PS, st, c); break;
        CASE(0) CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6)
        default: hipLaunchKernelGGL((k_sh_shift_step<7, PT>), grid, dim3(256), 0, s, n, l, P, R, x, PS, st, c); break;
#undef CASE
    }
}

template <typename PT>
static void sh_shift_upd(int l, dim3 grid, hipStream_t s, long long n, const double *R, double *x, PT *PS,
                         const ShState *st, const Ctl *c) {
    switch (l) {
#define CASE(L_) case L_: hipLaunchKernelGGL((k_sh_shift_upd<L_, PT>), grid, dim3(256), 0, s, n, R, x, PS, st, c); break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7)
        default: hipLaunchKernelGGL((k_sh_shift_upd<8, PT>), grid, dim3(256), 0, s, n, R, x, PS, st, c); break;
#undef CASE
    }
}

static void sh_gram_dispatch(int l, hipStream_t s, const bkh_matrix *A, double *R, long long n, Ctl *c,
                             double *part, ShState *st) {
    switch (l) {
    case 1: sh_gram_spmv<1>(s, A, R, n, c, part, st); break;
    case 2: sh_gram_spmv<2>(s, A, R, n, c, part, st); break;
    case 3: sh_gram_spmv<3>(s, A, R, n, c, part, st); break;
    case 4: sh_gram_spmv<4>(s, A, R, n, c, part, st); break;
    case 5: sh_gram_spmv<5>(s, A, R, n, c, part, st); break;
    case 6: sh_gram_spmv<6>(s, A, R, n, c, part, st); break;
    case 7: sh_gram_spmv<7>(s, A, R, n, c, part, st); break;
    default: sh_gram_spmv<8>(s, A, R, n, c, part, st); break;
    }
}

template <typename PT>
static int sh_core(ThreadCtx *t, const bkh_matrix *A, int l, const double *b, double *x, double *R,
                   double *P, double *rs, PT *PS, ShState *st, const double *sigma, int ns, Ctl *c,
                   double *part, double tol, int max_it, int K) {
    hipStream_t s = t->stream;
    const long long n = A->n;
    const int G = vec_grid(n, 256);
    const dim3 g2(G, (ns + SH_SPT - 1) / SH_SPT > 0 ? (ns + SH_SPT - 1) / SH_SPT : 1);
    if (ctl_init(t, c, tol, max_it)) return BKH_DEVICE_ERROR;
    {
        ShState *h = new ShState();
        memset(h, 0, sizeof *h);
        h->ns = ns;
        h->l = l;
        for (int k = 0; k < ns; k++) h->sigma[k] = sigma[k];
        hipError_t e = hipMemcpyAsync(st, h, sizeof *h, hipMemcpyHostToDevice, s);
        hipError_t e2 = hipStreamSynchronize(s);
        delete h;
        if (e != hipSuccess || e2 != hipSuccess) { bkh_set_error("ShState upload failed"); return BKH_DEVICE_ERROR; }
    }
    const long long nps = (long long)ns * (l + 1) * n;      // p_shift starts at 0 (:53 calloc)
    if (sizeof(PT) == 4)
        hipLaunchKernelGGL(k_zero_pt, dim3(vec_grid(nps, 256)), dim3(256), 0, s, (float *)PS, nps);
    else
        hipLaunchKernelGGL(k_zero_pd, dim3(vec_grid(nps, 256)), dim3(256), 0, s, (double *)PS, nps);
    Epi<0> e0;
    memset(&e0, 0, sizeof e0);
    launch_spmv<double, double, 0>(s, A, A->val, x, R, 1.0, 0.0, e0, nullptr, nullptr, NoFin());
    hipLaunchKernelGGL(k_sh_init, dim3(G), dim3(256), 0, s, n, b, R, P, rs, part);
    launch_finalize<2>(s, part, G, c, ShInitFin{st});
    LAUNCH_OK();
    hipGraphExec_t ge = capture_cached(t, {0x5B00, (uintptr_t)A, (uintptr_t)b, (uintptr_t)x, (uintptr_t)R, (uintptr_t)P,
                                           (uintptr_t)rs, (uintptr_t)PS, (uintptr_t)st, (uintptr_t)c, (uintptr_t)part,
                                           (uintptr_t)ns, (uintptr_t)l, (uintptr_t)sizeof(PT)}, s, K, [&]() {
        for (int j = 0; j < l; j++) {
            hipLaunchKernelGGL(k_sh_p, dim3(G), dim3(256), 0, s, n, j, P, (const double *)R, c);
            sh_alpha_spmv(s, A, P + j * n, P + (j + 1) * n, rs, c, part, st);
            if (ns > 0) sh_shift_step<PT>(j, g2, s, n, l, (const double *)P, (const double *)R, x, PS, st, c);
            hipLaunchKernelGGL(k_sh_seed_r, dim3(G), dim3(256), 0, s, n, j, (const double *)P, R, x, c);
            if (j < l - 1)
                launch_spmv_dot_w<double, double>(s, A, A->val, R + j * n, R + (j + 1) * n, rs, c, part,
                                                  ShGammaFin());
            else
                sh_gram_dispatch(l, s, A, R, n, c, part, st);
        }
        if (ns > 0) sh_shift_upd<PT>(l, g2, s, n, (const double *)R, x, PS, st, c);
        hipLaunchKernelGGL(k_sh_seed_upd, dim3(G), dim3(256), 0, s, n, l, P, R, x, (const double *)rs, c, part);
        launch_finalize<2>(s, part, G, c, ShUpdFin());
    });
    if (!ge) { bkh_set_error("hipGraph capture of the shifted BiCGSTAB(l) iteration failed"); return BKH_DEVICE_ERROR; }
    int rc = run_graph_until_done(t, ge, c, max_it / K + 2);
    return rc;
}

extern "C" int bkh_shbicgstabl(bkh_matrix *A, const double *b, int l, const double *sigma, int ns,
                               double *x, double tol, int max_iter, const bkh_opts *opts,
                               bkh_info *info, double *shift_true_res) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (!A || !b || !x || (ns > 0 && !sigma)) return BKH_BAD_ARGUMENT;
    if (l < 1 || l > MR_LMAX || ns < 0 || ns > SH_MAXS) {
        bkh_set_error("shifted BiCGSTAB(l) supports 1 <= l <= %d and up to %d shifts", MR_LMAX, SH_MAXS);
        return -3;
    }
    bkh_opts o = opts ? *opts : bkh_opts{};
    bkh_info inf;
    memset(&inf, 0, sizeof inf);
    if (max_iter <= 0) {                 // shbicgstabl.c:12, 280-284
        inf.status = ST_CONVERGED;
        if (info) *info = inf;
        return 0;
    }
    const long long n = A->n;
    const int K = o.graph_iters > 0 ? o.graph_iters : std::max(2, default_graph_iters(n) / (2 * l));
    const int nv = (l + 1) * (l + 2) / 2;
    const int nparts = std::max(spmv_grid(A), vec_grid(n, 256)) * std::max(nv, 2);
    const bool p32 = (o.precision == BKH_FP32_IR);
    const size_t psb = (size_t)ns * (l + 1) * n * (p32 ? sizeof(float) : sizeof(double));
    size_t bytes = carve_size(sizeof(Ctl)) + carve_size(sizeof(ShState)) + carve_size(sizeof(double) * nparts) +
                   (2 * (l + 1) + 1) * carve_size(sizeof(double) * n) + carve_size(psb);
    void *ws = bkh_workspace(bytes);
    if (!ws) return BKH_DEVICE_ERROR;
    Carve cv(ws);
    Ctl *c = cv.take<Ctl>(1);
    ShState *st = cv.take<ShState>(1);
    double *part = cv.take<double>(nparts);
    double *R = cv.take<double>((l + 1) * n), *P = cv.take<double>((l + 1) * n), *rs = cv.take<double>(n);
    char *PS = cv.take<char>(psb);
    HIP_OK(hipEventRecord(t->ev0, t->stream));
    int rc = p32 ? sh_core<float>(t, A, l, b, x, R, P, rs, (float *)PS, st, sigma, ns, c, part, tol, max_iter, K)
                 : sh_core<double>(t, A, l, b, x, R, P, rs, (double *)PS, st, sigma, ns, c, part, tol, max_iter, K);
    if (rc) return rc;
    int ret = finish_solve(t, A, 0.0, b, x, &inf, info);
    if (shift_true_res) {
        shift_true_res[0] = inf.true_res;
        for (int k = 0; k < ns; k++) {
            double tr = 0.0;
            if (true_residual(t, A, sigma[k], b, x + (size_t)(k + 1) * n, &tr)) return BKH_DEVICE_ERROR;
            shift_true_res[k + 1] = tr;
        }
        if (!info) {
            double tr = 0.0;
            if (true_residual(t, A, 0.0, b, x, &tr)) return BKH_DEVICE_ERROR;
            shift_true_res[0] = tr;
        }
    }
    return ret;
}
