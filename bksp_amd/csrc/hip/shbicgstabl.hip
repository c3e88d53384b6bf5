// shbicgstabl.hip -- shifted BiCGSTAB(l) on the device (replaces src/shbicgstabl.c:7-285
// with src/util_shift.c:5-27 dhorner and :55-63 comb).
//
// The seed system (sigma = 0) runs BiCGSTAB(l) as in bicgstabl.hip. The per-shift
// vectors are stored SHIFT-INTERLEAVED: p_shift[q] and x_k as n x SP row-major blocks
// (SP = the shift count rounded up to 4; padded shifts carry zero coefficients and
// stay zero), so a thread owns one row and V = 16/sizeof(PT) consecutive shifts and
// every access is one contiguous 16-byte vector per thread -- a wave streams
// 64 x 16 contiguous bytes per vector instead of touching 8 shifts x (j+2) separate
// n-strided arrays. The per-shift work of BiCG step j is one kernel over all shifts
// (k_sh_step, shbicgstabl.c:111-134, :146-153); the seed's own r/x update (:107-108,
// :137-139) follows in k_sh_seed_r, so every shift thread reads the pre-update r.
// The MR step's per-shift updates (:202-237, k_sh_upd) and the seed update
// (:194-199, :240-242, k_sh_seed_upd) are ordered the same way. The pi/theta/psi
// recurrences and the Horner re-expansion are FP64 scalar work in the finalize
// kernels, one shift per thread. Each element follows the reference's operation
// order, so the FP64 path matches the oracle's iterates to rounding.
//
// BKH_FP32_IR: p_shift and x_k are FP32 (the seed, its scalars and the pi/theta/psi
// recurrences stay FP64). FP32 storage limits the shifted residuals to ~kappa*eps32,
// so every shift is then refined in FP64: r_k = b - (A + sigma_k I) x_k (FP64, all
// shifts in one kernel), and (A + sigma_k I) d_k = r_k is solved for all shifts at
// once by a multi-shift BiCGSTAB in FP32 (per-shift scalars, shared SpMV sweeps over
// the interleaved block), x_k += d_k in FP64, until every shift meets tol.
#include <math.h>
#include "bkh_host.hpp"
#include "bkh_mr.hpp"
#include "bkh_solver_common.hpp"
#include "bkh_spmv.hpp"

#define SH_MAXS 256
#define SH_KCHUNK 16        // shifts per workgroup row of the fused per-shift pass

enum { SB_G0 = 0, SB_ALPHA = 1, SB_BETA = 2, SB_OML = 3, SB_AOLD = 4, SB_CX = 8, SB_CR = 20, SB_CP = 32 };

struct ShState {
    int ns, l;
    double sigma[SH_MAXS], pi[SH_MAXS], pi_old[SH_MAXS], theta[SH_MAXS];
    // per-step coefficients of k_sh_step
    double cA[SH_MAXS], cB[SH_MAXS], cC[SH_MAXS], cD[SH_MAXS], cE[SH_MAXS], ak[SH_MAXS];
    // MR-step coefficients of k_sh_upd
    double xc[SH_MAXS][MR_LMAX], psi_inv[SH_MAXS], om[MR_D];
    // every step's coefficients (cA, cB, cC, cD, cE, ak of step j at ck[k][j]) for the
    // fused per-shift pass, shift-major so a shift's are read with scalar loads
    double ck[SH_MAXS][MR_LMAX][6];
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
                                                      ShState *st, int j) {
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
        double *ck = st->ck[k][j];
        ck[0] = st->cA[k]; ck[1] = st->cB[k]; ck[2] = st->cC[k]; ck[3] = st->cD[k]; ck[4] = st->cE[k]; ck[5] = ak;
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
// snap (fused per-shift pass): r_q before this update goes to snap[j(j+1)/2 + q].
__global__ __launch_bounds__(256) void k_sh_seed_r(long long n, int j, const double *P, double *R,
                                                   double *x, const Ctl *c, double *snap) {
    if (ctl_stopped(c)) return;
    const double al = c->s[SB_ALPHA];
    double *sj = snap ? snap + (size_t)(j * (j + 1) / 2) * n : nullptr;
    ROWLOOP(i, n) {
        for (int q = 0; q <= j; q++) {
            const double r = R[q * n + i];
            if (sj) sj[q * n + i] = r;
            R[q * n + i] = fma(-al, P[(q + 1) * n + i], r);
        }
        x[i] = fma(al, P[i], x[i]);
    }
}

// ------------------------------------------------ shift-interleaved per-shift work
//
// Launch geometry of the interleaved kernels: a thread owns (row i, shift group gq)
// with gq = tid % G (G = SP / V groups of V shifts); the grid-stride loop advances
// by whole rows because the thread count is a multiple of G, so each thread loads
// its shifts' coefficients once.
struct ShGrid {
    int SP, V, G, blocks;
};

static int gcd_i(int a, int b) { while (b) { int t = a % b; a = b; b = t; } return a; }

static ShGrid sh_grid(long long n, int SP, int V, int cap = 8192) {
    ShGrid g;
    g.SP = SP;
    g.V = V;
    g.G = SP / V;
    const int m = g.G / gcd_i(g.G, 256);                      // blocks * 256 % G == 0
    long long want = (n * g.G + 255) / 256;
    if (want > cap) want = cap;
    want = (want + m - 1) / m * m;
    g.blocks = (int)std::max<long long>(m, want);
    return g;
}

// Shift part of BiCG step J (:124-133, :146-153) for the thread's V shifts:
// p_shift[k][0..J] = -beta_k p_shift[k][q] + r_q / (pi_k theta_k), p_shift[k][J+1]
// from r_J before and after the seed update, x_k += alpha_k p_shift[k][0]. r_J after
// the seed update is recomputed here with the seed kernel's exact operation, so
// every thread reads the pre-update r. PS: (l+1) interleaved n x SP blocks; XS: n x SP.
template <int J, typename PT>
__global__ __launch_bounds__(256) void k_sh_step(long long n, int G, const double *P, const double *R,
                                                 PT *XS, PT *PS, const ShState *st, const Ctl *c) {
    if (ctl_stopped(c)) return;
    constexpr int V = 16 / sizeof(PT);
    typedef PT pvec __attribute__((ext_vector_type(V)));
    const double al = c->s[SB_ALPHA];
    const long long tid = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long rstride = (long long)gridDim.x * 256 / G;
    const int k0 = (int)(tid % G) * V;
    const size_t SP = (size_t)G * V;
    double cA[V], cB[V], cC[V], cD[V], cE[V], ak[V], sg[V];
#pragma unroll
    for (int u = 0; u < V; u++) {
        const int k = k0 + u;
        cA[u] = st->cA[k]; cB[u] = st->cB[k]; cC[u] = st->cC[k]; cD[u] = st->cD[k];
        cE[u] = st->cE[k]; ak[u] = st->ak[k]; sg[u] = st->sigma[k];
    }
    for (long long i = tid / G; i < n; i += rstride) {
        double rold[J + 1];
#pragma unroll
        for (int q = 0; q <= J; q++) rold[q] = R[q * n + i];
        const double rjn = fma(-al, P[(J + 1) * n + i], rold[J]);
        pvec pv[J + 1];
#pragma unroll
        for (int q = 0; q <= J; q++) pv[q] = *(const pvec *)(PS + ((size_t)q * n + i) * SP + k0);
        const pvec xv = *(const pvec *)(XS + (size_t)i * SP + k0);
        pvec out[J + 2], xo;
#pragma unroll
        for (int u = 0; u < V; u++) {
            double p0 = 0.0, pj = 0.0;
#pragma unroll
            for (int q = 0; q <= J; q++) {
                const PT v = (PT)fma(cB[u], rold[q], mul_rn(cA[u], (double)pv[q][u]));
                out[q][u] = v;
                if (q == 0) p0 = (double)v;
                if (q == J) pj = (double)v;
            }
            double t = mul_rn(rold[J], cC[u]);
            t = fma(cD[u], rjn, t);
            t = mul_rn(t, cE[u]);
            t = fma(-sg[u], pj, t);
            out[J + 1][u] = (PT)t;
            xo[u] = (PT)fma(ak[u], p0, (double)xv[u]);
        }
#pragma unroll
        for (int q = 0; q <= J + 1; q++) *(pvec *)(PS + ((size_t)q * n + i) * SP + k0) = out[q];
        *(pvec *)(XS + (size_t)i * SP + k0) = xo;
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

// Shift part of the MR step (:202-237) for the thread's V shifts: x_k += sum_q
// xc[k][q] r_q (original r), p_shift[k][0] = (p_shift[k][0] - sum_i omega_i
// p_shift[k][i]) / psi_k.
template <int L, typename PT>
__global__ __launch_bounds__(256) void k_sh_upd(long long n, int G, const double *R, PT *XS, PT *PS,
                                                const ShState *st, const Ctl *c) {
    if (ctl_stopped(c)) return;
    constexpr int V = 16 / sizeof(PT);
    typedef PT pvec __attribute__((ext_vector_type(V)));
    const long long tid = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long rstride = (long long)gridDim.x * 256 / G;
    const int k0 = (int)(tid % G) * V;
    const size_t SP = (size_t)G * V;
    double xc[V][L], pinv[V], om[L + 1];
#pragma unroll
    for (int u = 0; u < V; u++) {
#pragma unroll
        for (int q = 0; q < L; q++) xc[u][q] = st->xc[k0 + u][q];
        pinv[u] = st->psi_inv[k0 + u];
    }
#pragma unroll
    for (int q = 1; q <= L; q++) om[q] = st->om[q];
    for (long long i = tid / G; i < n; i += rstride) {
        double r[L];
#pragma unroll
        for (int q = 0; q < L; q++) r[q] = R[q * n + i];
        pvec pv[L + 1];
#pragma unroll
        for (int q = 0; q <= L; q++) pv[q] = *(const pvec *)(PS + ((size_t)q * n + i) * SP + k0);
        const pvec xv = *(const pvec *)(XS + (size_t)i * SP + k0);
        pvec xo, po;
#pragma unroll
        for (int u = 0; u < V; u++) {
            double xk = (double)xv[u];
#pragma unroll
            for (int q = 0; q < L; q++) xk += xc[u][q] * r[q];
            xo[u] = (PT)xk;
            double p = (double)pv[0][u];
#pragma unroll
            for (int q = 1; q <= L; q++) p -= om[q] * (double)pv[q][u];
            po[u] = (PT)(p * pinv[u]);
        }
        *(pvec *)(XS + (size_t)i * SP + k0) = xo;
        *(pvec *)(PS + (size_t)i * SP + k0) = po;
    }
}

// Every per-shift vector operation of one outer iteration in one pass: the l BiCG
// steps of k_sh_step (shbicgstabl.c:111-134, :146-153) and the MR update of k_sh_upd
// (:202-237). p_shift[k][1..l] live only within an outer iteration (step j writes
// p_shift[k][j+1] before step j+1 reads it), so a thread keeps them in registers and
// reads and writes only p_shift[k][0] and x_k, shift-major (shift k at k n, the
// reference's own x layout). The seed's r_q at the start of step j come from the
// snapshots k_sh_seed_r wrote (snap[j(j+1)/2 + q]); r_j after step j's update is
// snap[(j+1)(j+2)/2 + j], or the final r_{l-1}; the MR update uses the final r_0..r_{l-1}.
// Each element goes through k_sh_step's and k_sh_upd's operations in order with the
// same roundings to PT, so the iterates are bitwise those of the per-step kernels, while
// each shift moves 4 n PT per outer iteration instead of (l + 1)(l + 4) n. A workgroup
// row (blockIdx.y) takes kchunk shifts; a shift's coefficients are uniform loads.
template <int L, typename PT>
__global__ __launch_bounds__(256) void k_sh_fused(long long n, int ns, int kchunk, const double *R,
                                                  const double *snap, PT *XSM, PT *PSM, const ShState *st,
                                                  const Ctl *c) {
    if (ctl_stopped(c)) return;
    constexpr int NS = L * (L + 1) / 2;
    constexpr int CW = 6 * L + L + 2;                    // a shift's coefficients: steps, xc, 1/psi, sigma
    constexpr int U = 4;                                 // shifts in flight per thread
    __shared__ double cks[SH_KCHUNK][CW];
    const int kb = (int)blockIdx.y * kchunk, ke = min(ns, kb + kchunk);
    for (int e = threadIdx.x; e < (ke - kb) * CW; e += 256) {
        const int k = kb + e / CW, w = e % CW;
        cks[e / CW][w] = w < 6 * L ? st->ck[k][w / 6][w % 6]
                       : w < 7 * L ? st->xc[k][w - 6 * L] : (w == 7 * L ? st->psi_inv[k] : st->sigma[k]);
    }
    double om[L + 1];
#pragma unroll
    for (int q = 1; q <= L; q++) om[q] = st->om[q];
    __syncthreads();
    ROWLOOP(i, n) {
        double sn[NS], rf[L];
#pragma unroll
        for (int e = 0; e < NS; e++) sn[e] = snap[(size_t)e * n + i];
#pragma unroll
        for (int q = 0; q < L; q++) rf[q] = R[(size_t)q * n + i];
        for (int k0 = kb; k0 < ke; k0 += U) {
            PT pl[U], xl[U];
#pragma unroll
            for (int u = 0; u < U; u++) {                // loads of U shifts in flight together
                const int k = min(k0 + u, ke - 1);
                pl[u] = PSM[(size_t)k * n + i];
                xl[u] = XSM[(size_t)k * n + i];
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                if (k0 + u >= ke) break;
                const double *cf = cks[k0 + u - kb];
                const double sg = cf[7 * L + 1];
                PT p[L + 1];
                p[0] = pl[u];
                PT x = xl[u];
#pragma unroll
                for (int j = 0; j < L; j++) {
                    const double cA = cf[6 * j], cB = cf[6 * j + 1], cC = cf[6 * j + 2], cD = cf[6 * j + 3],
                                 cE = cf[6 * j + 4], ak = cf[6 * j + 5];
                    const int s0 = j * (j + 1) / 2;
                    const double rjn = (j + 1 < L) ? sn[(j + 1) * (j + 2) / 2 + j] : rf[L - 1];
                    PT out[L + 1];
                    double p0 = 0.0, pj = 0.0;
#pragma unroll
                    for (int q = 0; q <= j; q++) {
                        const PT v = (PT)fma(cB, sn[s0 + q], mul_rn(cA, (double)p[q]));
                        out[q] = v;
                        if (q == 0) p0 = (double)v;
                        if (q == j) pj = (double)v;
                    }
                    double t = mul_rn(sn[s0 + j], cC);
                    t = fma(cD, rjn, t);
                    t = mul_rn(t, cE);
                    t = fma(-sg, pj, t);
                    out[j + 1] = (PT)t;
                    x = (PT)fma(ak, p0, (double)x);
#pragma unroll
                    for (int q = 0; q <= j + 1; q++) p[q] = out[q];
                }
                double xk = (double)x;
#pragma unroll
                for (int q = 0; q < L; q++) xk += cf[6 * L + q] * rf[q];
                double pp = (double)p[0];
#pragma unroll
                for (int q = 1; q <= L; q++) pp -= om[q] * (double)p[q];
                const int k = k0 + u;
                XSM[(size_t)k * n + i] = (PT)xk;
                PSM[(size_t)k * n + i] = (PT)(pp * cf[7 * L]);
            }
        }
    }
}

template <typename PT>
__global__ __launch_bounds__(256) void k_sh_zero(PT *p, long long n) { ROWLOOP(i, n) p[i] = (PT)0; }

// Tiled transposes between shift-major blocks (shift k at xm + k n: the reference's x
// layout from x + n) and the interleaved n x SP block (32 x 32 tiles through LDS, both
// sides coalesced).
template <typename ST, typename PT>
__global__ __launch_bounds__(256) void k_xs_in(long long n, int ns, int SP, const ST *xm, PT *XS) {
    __shared__ double tile[32][33];
    const long long i0 = (long long)blockIdx.x * 32;
    const int k0 = blockIdx.y * 32, tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
#pragma unroll
    for (int jj = 0; jj < 4; jj++) {
        const int k = k0 + ty + 8 * jj;
        const long long i = i0 + tx;
        tile[ty + 8 * jj][tx] = (k < ns && i < n) ? (double)xm[(size_t)k * n + i] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int jj = 0; jj < 4; jj++) {
        const long long i = i0 + ty + 8 * jj;
        const int k = k0 + tx;
        if (i < n && k < SP) XS[(size_t)i * SP + k] = (PT)tile[tx][ty + 8 * jj];
    }
}

template <typename PT>
__global__ __launch_bounds__(256) void k_xs_out(long long n, int ns, int SP, const PT *XS, double *x) {
    __shared__ double tile[32][33];
    const long long i0 = (long long)blockIdx.x * 32;
    const int k0 = blockIdx.y * 32, tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
#pragma unroll
    for (int jj = 0; jj < 4; jj++) {
        const long long i = i0 + ty + 8 * jj;
        const int k = k0 + tx;
        tile[ty + 8 * jj][tx] = (i < n && k < SP) ? (double)XS[(size_t)i * SP + k] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int jj = 0; jj < 4; jj++) {
        const int k = k0 + ty + 8 * jj;
        const long long i = i0 + tx;
        if (k < ns && i < n) x[(size_t)(k + 1) * n + i] = tile[tx][ty + 8 * jj];
    }
}

// ---------------------------------------- multi-shift FP64 residual and refinement
//
// Per-shift state of the multi-shift BiCGSTAB that refines the FP32 shifted
// solutions: (A + sigma_k I) d_k = r_k for every shift k at once, FP32 vectors in the
// interleaved layout, FP64 scalars and dot accumulation, one shared sweep over A per
// SpMV. A shift that has met its target (or broken down) is frozen: alpha = omega =
// beta = 0 leave its d and r unchanged while the others continue.
struct MsCol {
    double rho, alpha, omega, beta, r0n2, tgt2, rn2, rel;
    int act, it, pad[2];
};
#define MS_NB 2048          // workgroup cap of the multi-shift kernels (partials per shift)

// Multi-shift kernels: thread t of the grid serves shift group t % G (G divides 256)
// of rows t / G, t / G + stride, ... (the grid is a multiple of G threads), so all
// workgroups sweep one window of rows together and the neighbour rows an SpMV
// gathers are shared through L2/MALL while they are hot.
#define MS_ROWS_SETUP                                                              \
    const long long tid_ = (long long)blockIdx.x * 256 + threadIdx.x;              \
    const long long rstride_ = (long long)gridDim.x * 256 / G;                     \
    const int k0 = (int)(threadIdx.x % G) * 4;                                     \
    const size_t SP = (size_t)G * 4
#define MS_ROWS(i) for (long long i = tid_ / G; i < n; i += rstride_)
// A group of 4 shifts whose shifts are all frozen keeps its d and r whatever the
// kernels do (alpha = omega = 0), so its rows are skipped (no loads or stores); its
// partials stay 0, which the finalize ignores for inactive shifts.
#define MS_LIVE(mc) ((mc)[k0].act | (mc)[k0 + 1].act | (mc)[k0 + 2].act | (mc)[k0 + 3].act)
#define MS_LIVE_ROWS(i, mc) if (MS_LIVE(mc)) MS_ROWS(i)

// Sums NV x 4 per-shift values over the threads of one shift group (gq = lane % G,
// G a power of two <= 64) and writes part[(blockIdx.x * NV + v) * SP + k], fixed order.
template <int NV>
__device__ __forceinline__ void ms_block_partial(double (&v)[NV][4], int G, int SP, double *part) {
    __shared__ double sh[4][64 * 4 * NV];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int off = G; off < 64; off <<= 1) {
#pragma unroll
        for (int a = 0; a < NV; a++)
#pragma unroll
            for (int u = 0; u < 4; u++) v[a][u] += __shfl_xor(v[a][u], off, 64);
    }
    if (lane < G) {
#pragma unroll
        for (int a = 0; a < NV; a++)
#pragma unroll
            for (int u = 0; u < 4; u++) sh[w][(lane * 4 + u) * NV + a] = v[a][u];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < G * 4 * NV; e += 256) {
        const double t = ((sh[0][e] + sh[1][e]) + sh[2][e]) + sh[3][e];
        part[((size_t)blockIdx.x * NV + (e % NV)) * SP + e / NV] = t;
    }
}

// r_k = b - (A + sigma_k I) x_k in FP64 for all shifts (XS: n x SP, FP64); the FP32
// copy goes to R32 and RH when given; partials of |r_k|^2.
template <typename VT>
__global__ __launch_bounds__(256) void k_ms_resid(long long n, int G, const int *__restrict__ rp,
                                                  const int *__restrict__ ci, const VT *__restrict__ val,
                                                  const double *b, const double *XS, const ShState *st,
                                                  float *R32, float *RH, double *part) {
    typedef double dvec __attribute__((ext_vector_type(4)));
    typedef float fvec __attribute__((ext_vector_type(4)));
    MS_ROWS_SETUP;
    double sg[4], dv[1][4] = {{0.0, 0.0, 0.0, 0.0}};
#pragma unroll
    for (int u = 0; u < 4; u++) sg[u] = st->sigma[k0 + u];
    MS_ROWS(i) {
        dvec acc = {0.0, 0.0, 0.0, 0.0};
        for (int e = rp[i]; e < rp[i + 1]; e++) {
            const double a = (double)val[e];
            const dvec xv = *(const dvec *)(XS + (size_t)ci[e] * SP + k0);
            acc += a * xv;
        }
        const dvec xi = *(const dvec *)(XS + (size_t)i * SP + k0);
        const double bi = b[i];
        fvec r32;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const double r = bi - acc[u] - sg[u] * xi[u];
            r32[u] = (float)r;
            dv[0][u] += r * r;
        }
        if (R32) {
            *(fvec *)(R32 + (size_t)i * SP + k0) = r32;
            *(fvec *)(RH + (size_t)i * SP + k0) = r32;
        }
    }
    ms_block_partial<1>(dv, G, (int)SP, part);
}

// K0: p = r + beta (p - omega v) (own rows; v of the previous iteration).
__global__ __launch_bounds__(256) void k_ms_k0(long long n, int G, const float *R, float *Pv, const float *Vv,
                                               const MsCol *mc, const Ctl *c) {
    if (ctl_stopped(c)) return;
    typedef float fvec __attribute__((ext_vector_type(4)));
    MS_ROWS_SETUP;
    fvec be, om;
#pragma unroll
    for (int u = 0; u < 4; u++) { be[u] = (float)mc[k0 + u].beta; om[u] = (float)mc[k0 + u].omega; }
    MS_LIVE_ROWS(i, mc) {
        const size_t o = (size_t)i * SP + k0;
        *(fvec *)(Pv + o) = *(const fvec *)(R + o) + be * (*(const fvec *)(Pv + o) - om * *(const fvec *)(Vv + o));
    }
}

// Y = (A + sigma_k I) X for every shift (one gathered interleaved vector), with the
// per-shift partials W.Y (NV = 1) or Y.X, Y.Y (NV = 2). K1: v = (A + sigma) p, r^.v;
// K2: t = (A + sigma) s, t.s and t.t.
template <typename VT, int NV>
__global__ __launch_bounds__(256) void k_ms_spmv(long long n, int G, const int *__restrict__ rp,
                                                 const int *__restrict__ ci, const VT *__restrict__ val,
                                                 const float *X, float *Y, const float *W, const ShState *st,
                                                 const MsCol *mc, const Ctl *c, double *part) {
    if (ctl_stopped(c)) return;
    typedef float fvec __attribute__((ext_vector_type(4)));
    MS_ROWS_SETUP;
    fvec sg;
#pragma unroll
    for (int u = 0; u < 4; u++) sg[u] = (float)st->sigma[k0 + u];
    double dv[NV][4];
#pragma unroll
    for (int a = 0; a < NV; a++)
#pragma unroll
        for (int u = 0; u < 4; u++) dv[a][u] = 0.0;
    MS_LIVE_ROWS(i, mc) {
        fvec acc = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int e = rp[i]; e < rp[i + 1]; e++)
            acc += (float)val[e] * *(const fvec *)(X + (size_t)ci[e] * SP + k0);
        const size_t o = (size_t)i * SP + k0;
        const fvec xi = *(const fvec *)(X + o);
        const fvec y = acc + sg * xi;
        *(fvec *)(Y + o) = y;
        if (NV == 1) {
            const fvec w = *(const fvec *)(W + o);
#pragma unroll
            for (int u = 0; u < 4; u++) dv[0][u] += (double)w[u] * (double)y[u];
        } else {
#pragma unroll
            for (int u = 0; u < 4; u++) {
                dv[0][u] += (double)y[u] * (double)xi[u];
                dv[NV - 1][u] += (double)y[u] * (double)y[u];
            }
        }
    }
    ms_block_partial<NV>(dv, G, (int)SP, part);
}

// s = r - alpha v.
__global__ __launch_bounds__(256) void k_ms_s(long long n, int G, const float *R, const float *Vv, float *S,
                                              const MsCol *mc, const Ctl *c) {
    if (ctl_stopped(c)) return;
    typedef float fvec __attribute__((ext_vector_type(4)));
    MS_ROWS_SETUP;
    fvec al;
#pragma unroll
    for (int u = 0; u < 4; u++) al[u] = (float)mc[k0 + u].alpha;
    MS_LIVE_ROWS(i, mc) {
        const size_t o = (size_t)i * SP + k0;
        *(fvec *)(S + o) = *(const fvec *)(R + o) - al * *(const fvec *)(Vv + o);
    }
}

// K3: d += alpha p + omega s; r = s - omega t; r^.r and r.r partials.
__global__ __launch_bounds__(256) void k_ms_k3(long long n, int G, float *D, float *R, const float *Pv,
                                               const float *S, const float *T, const float *RH, const MsCol *mc,
                                               const Ctl *c, double *part) {
    if (ctl_stopped(c)) return;
    typedef float fvec __attribute__((ext_vector_type(4)));
    MS_ROWS_SETUP;
    fvec al, om;
#pragma unroll
    for (int u = 0; u < 4; u++) { al[u] = (float)mc[k0 + u].alpha; om[u] = (float)mc[k0 + u].omega; }
    double dv[2][4] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
    MS_LIVE_ROWS(i, mc) {
        const size_t o = (size_t)i * SP + k0;
        const fvec sv = *(const fvec *)(S + o);
        *(fvec *)(D + o) = *(const fvec *)(D + o) + al * *(const fvec *)(Pv + o) + om * sv;
        const fvec rn = sv - om * *(const fvec *)(T + o);
        *(fvec *)(R + o) = rn;
        const fvec rh = *(const fvec *)(RH + o);
#pragma unroll
        for (int u = 0; u < 4; u++) {
            dv[0][u] += (double)rh[u] * (double)rn[u];
            dv[1][u] += (double)rn[u] * (double)rn[u];
        }
    }
    ms_block_partial<2>(dv, G, (int)SP, part);
}

// One workgroup per shift: fixed-order sum of the shift's partials, then `op`.
template <int NV, class Op>
__global__ __launch_bounds__(256) void k_ms_fin(const double *part, int nb, int SP, MsCol *mc, const Ctl *c, Op op) {
    if (c && ctl_stopped(c)) return;
    __shared__ double sh[NV * 4];
    const int k = blockIdx.x;
    double acc[NV];
#pragma unroll
    for (int a = 0; a < NV; a++) acc[a] = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) {
#pragma unroll
        for (int a = 0; a < NV; a++) acc[a] += part[((size_t)b * NV + a) * SP + k];
    }
    block_sum<NV, 256>(acc, sh);
    if (threadIdx.x == 0) op(mc[k], acc, k);
}

struct MsStartOp {      // a pass: r0 = b - (A + sigma I) x, targets, r^ = r0
    int ns;
    double tol, fnb, itol_min;
    __device__ void operator()(MsCol &m, const double *a, int k) const {
        const double rel = sqrt(a[0]) / fnb;
        m.rel = rel;
        m.rn2 = a[0];
        m.r0n2 = a[0];
        m.rho = a[0];
        m.alpha = 0.0; m.omega = 0.0; m.beta = 0.0; m.it = 0;
        m.act = (k < ns && rel == rel && !(rel < tol) && a[0] > 0.0) ? 1 : 0;
        // half of tol: the pass's FP64 true residual check starts another pass for a
        // shift the recurrence misjudged (96^3 convection-diffusion, 64 shifts: 24 inner
        // iterations, 124 ms, against 29 and 133 ms with a quarter)
        double itol = 0.5 * tol / rel;
        if (itol < itol_min) itol = itol_min;
        if (itol > 0.5) itol = 0.5;
        m.tgt2 = itol * itol * a[0];
    }
};
struct MsAlphaOp {
    __device__ void operator()(MsCol &m, const double *a, int) const {
        if (!m.act) { m.alpha = 0.0; return; }
        if (bad(a[0])) { m.act = 0; m.alpha = 0.0; m.beta = 0.0; return; }
        m.alpha = m.rho / a[0];
    }
};
struct MsOmegaOp {
    __device__ void operator()(MsCol &m, const double *a, int) const {
        if (!m.act) { m.omega = 0.0; return; }
        m.omega = (a[1] > 0.0 && a[1] == a[1]) ? a[0] / a[1] : 0.0;
    }
};
struct MsRhoOp {
    int max_it;
    __device__ void operator()(MsCol &m, const double *a, int) const {
        if (!m.act) { m.alpha = 0.0; m.omega = 0.0; m.beta = 0.0; return; }
        m.rn2 = a[1];
        m.it++;
        if (a[1] <= m.tgt2 || m.it >= max_it || m.omega == 0.0 || bad(a[0]) || !(a[1] == a[1])) {
            m.act = 0; m.beta = 0.0;
            return;
        }
        m.beta = (a[0] / m.rho) * (m.alpha / m.omega);
        m.rho = a[0];
    }
};

// Stops the pass (Ctl::done) when no shift is active; counts iterations in Ctl::it.
__global__ __launch_bounds__(64) void k_ms_check(const MsCol *mc, int SP, Ctl *c) {
    if (ctl_stopped(c)) return;
    int any = 0;
    for (int k = threadIdx.x; k < SP; k += 64) any |= mc[k].act;
    for (int o = 32; o >= 1; o >>= 1) any |= __shfl_xor(any, o, 64);
    if (threadIdx.x == 0) {
        c->it++;
        if (!any) { c->done = 1; c->status = ST_CONVERGED; }
    }
}

__global__ __launch_bounds__(256) void k_ms_apply(long long m, double *X, const float *D) {
    ROWLOOP(i, m) X[i] += (double)D[i];
}

template <typename PT>
__global__ __launch_bounds__(256) void k_to_f64(long long m, const PT *a, double *b) {
    ROWLOOP(i, m) b[i] = (double)a[i];
}

template <typename TI, typename TO>
__global__ __launch_bounds__(256) void k_conv(long long m, const TI *a, TO *b) {
    ROWLOOP(i, m) b[i] = (TO)a[i];
}

struct MsBufs {
    float *R, *RH, *D, *P, *V, *S, *T;
};

// Per-shift FP64 true residuals |b - (A + sigma_k I) x_k| / |b| of the interleaved
// FP64 solutions; `rel` (host, SP entries).
static int ms_true_res(ThreadCtx *t, const bkh_matrix *A, const double *b, const double *XS64, int SP,
                       const ShState *st, MsCol *mc, double *part, double fnb, int ns, double *rel) {
    hipStream_t s = t->stream;
    const ShGrid g = sh_grid(A->n, SP, 4, MS_NB);
    hipLaunchKernelGGL((k_ms_resid<double>), dim3(g.blocks), dim3(256), 0, s, (long long)A->n, g.G, A->rp, A->ci,
                       (const double *)A->val, b, XS64, st, (float *)nullptr, (float *)nullptr, part);
    hipLaunchKernelGGL((k_ms_fin<1, MsStartOp>), dim3(SP), dim3(256), 0, s, (const double *)part, g.blocks, SP, mc,
                       (const Ctl *)nullptr, MsStartOp{ns, 0.0, fnb, 1e-6});
    LAUNCH_OK();
    std::vector<MsCol> h(SP);
    HIP_OK(hipMemcpyAsync(h.data(), mc, sizeof(MsCol) * SP, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    for (int k = 0; k < SP; k++) rel[k] = h[k].rel;
    return 0;
}

// FP64 refinement of every shift's solution XS64 (n x SP interleaved, FP64) with the
// multi-shift FP32 BiCGSTAB; at most max_pass passes. Returns 0 or a device error.
static int ms_refine(ThreadCtx *t, const bkh_matrix *A, const double *b, double *XS64, int SP, int ns,
                     const ShState *st, MsCol *mc, Ctl *c, double *part, const MsBufs &B, double fnb,
                     double tol, int max_inner, int max_pass, int *passes, int *inner_total) {
    hipStream_t s = t->stream;
    const long long n = A->n, m = n * SP;
    const ShGrid g = sh_grid(n, SP, 4, MS_NB);
    const float *v32 = A->val32;
    const int Kg = 8;
    *passes = 0;
    *inner_total = 0;
    hipGraphExec_t ge = capture_cached(t, {0x5B5F, (uintptr_t)A, (uintptr_t)XS64, (uintptr_t)SP, (uintptr_t)st,
                                           (uintptr_t)mc, (uintptr_t)c, (uintptr_t)part, (uintptr_t)B.R,
                                           (uintptr_t)B.P, (uintptr_t)max_inner}, s, 1, [&]() {
        for (int j = 0; j < Kg; j++) {
            hipLaunchKernelGGL(k_ms_k0, dim3(g.blocks), dim3(256), 0, s, n, g.G, (const float *)B.R, B.P,
                               (const float *)B.V, (const MsCol *)mc, (const Ctl *)c);
            if (v32)
                hipLaunchKernelGGL((k_ms_spmv<float, 1>), dim3(g.blocks), dim3(256), 0, s, n, g.G, A->rp, A->ci, v32,
                                   (const float *)B.P, B.V, (const float *)B.RH, st, (const MsCol *)mc, (const Ctl *)c, part);
            else
                hipLaunchKernelGGL((k_ms_spmv<double, 1>), dim3(g.blocks), dim3(256), 0, s, n, g.G, A->rp, A->ci,
                                   (const double *)A->val, (const float *)B.P, B.V, (const float *)B.RH, st,
                                   (const MsCol *)mc, (const Ctl *)c, part);
            hipLaunchKernelGGL((k_ms_fin<1, MsAlphaOp>), dim3(SP), dim3(256), 0, s, (const double *)part, g.blocks,
                               SP, mc, (const Ctl *)c, MsAlphaOp());
            hipLaunchKernelGGL(k_ms_s, dim3(g.blocks), dim3(256), 0, s, n, g.G, (const float *)B.R,
                               (const float *)B.V, B.S, (const MsCol *)mc, (const Ctl *)c);
            if (v32)
                hipLaunchKernelGGL((k_ms_spmv<float, 2>), dim3(g.blocks), dim3(256), 0, s, n, g.G, A->rp, A->ci, v32,
                                   (const float *)B.S, B.T, (const float *)nullptr, st, (const MsCol *)mc, (const Ctl *)c, part);
            else
                hipLaunchKernelGGL((k_ms_spmv<double, 2>), dim3(g.blocks), dim3(256), 0, s, n, g.G, A->rp, A->ci,
                                   (const double *)A->val, (const float *)B.S, B.T, (const float *)nullptr, st,
                                   (const MsCol *)mc, (const Ctl *)c, part);
            hipLaunchKernelGGL((k_ms_fin<2, MsOmegaOp>), dim3(SP), dim3(256), 0, s, (const double *)part, g.blocks,
                               SP, mc, (const Ctl *)c, MsOmegaOp());
            hipLaunchKernelGGL(k_ms_k3, dim3(g.blocks), dim3(256), 0, s, n, g.G, B.D, B.R, (const float *)B.P,
                               (const float *)B.S, (const float *)B.T, (const float *)B.RH, (const MsCol *)mc,
                               (const Ctl *)c, part);
            hipLaunchKernelGGL((k_ms_fin<2, MsRhoOp>), dim3(SP), dim3(256), 0, s, (const double *)part, g.blocks,
                               SP, mc, (const Ctl *)c, MsRhoOp{max_inner});
            hipLaunchKernelGGL(k_ms_check, dim3(1), dim3(64), 0, s, (const MsCol *)mc, SP, c);
        }
    });
    if (!ge) { bkh_set_error("hipGraph capture of the multi-shift refinement failed"); return BKH_DEVICE_ERROR; }
    for (int pass = 0; pass < max_pass; pass++) {
        if (ctl_init(t, c, 0.0, 1 << 30)) return BKH_DEVICE_ERROR;
        hipLaunchKernelGGL((k_ms_resid<double>), dim3(g.blocks), dim3(256), 0, s, n, g.G, A->rp, A->ci,
                           (const double *)A->val, b, (const double *)XS64, st, B.R, B.RH, part);
        hipLaunchKernelGGL((k_ms_fin<1, MsStartOp>), dim3(SP), dim3(256), 0, s, (const double *)part, g.blocks, SP,
                           mc, (const Ctl *)nullptr, MsStartOp{ns, tol, fnb, 1e-6});
        // D, P, V = 0 (the first K0 then forms p = r)
        hipLaunchKernelGGL(k_sh_zero<float>, dim3(vec_grid(m, 256)), dim3(256), 0, s, B.D, m);
        hipLaunchKernelGGL(k_sh_zero<float>, dim3(vec_grid(m, 256)), dim3(256), 0, s, B.P, m);
        hipLaunchKernelGGL(k_sh_zero<float>, dim3(vec_grid(m, 256)), dim3(256), 0, s, B.V, m);
        hipLaunchKernelGGL(k_ms_check, dim3(1), dim3(64), 0, s, (const MsCol *)mc, SP, c);
        LAUNCH_OK();
        if (ctl_fetch(t, c)) return BKH_DEVICE_ERROR;
        if (t->ctl_host->done) break;                      // every shift meets tol
        if (ctl_init(t, c, 0.0, 1 << 30)) return BKH_DEVICE_ERROR;
        if (run_graph_until_done(t, ge, c, max_inner / Kg + 2)) return BKH_DEVICE_ERROR;
        *inner_total += t->ctl_host->it;
        hipLaunchKernelGGL(k_ms_apply, dim3(vec_grid(m, 256)), dim3(256), 0, s, m, XS64, (const float *)B.D);
        LAUNCH_OK();
        *passes = pass + 1;
    }
    return 0;
}

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
                          Ctl *c, double *part, ShState *st, int j) {
    Epi<1, double> e;
    memset(&e, 0, sizeof e);
    e.w[0] = rs;
    e.a[0] = 0;
    e.b[0] = EPI_Y;
    const int G = launch_spmv_kernel<double, double, 1>(s, A, A->val, x, y, 1.0, 0.0, e, c, part, NoFin());
    hipLaunchKernelGGL(k_sh_alpha_fin, dim3(1), dim3(256), 0, s, (const double *)part, G, c, st, j);
}

template <typename PT>
static void sh_shift_step(int j, const ShGrid &g, hipStream_t s, long long n, const double *P, const double *R,
                          PT *XS, PT *PS, const ShState *st, const Ctl *c) {
    switch (j) {
#define CASE(J_) case J_: hipLaunchKernelGGL((k_sh_step<J_, PT>), dim3(g.blocks), dim3(256), 0, s, n, g.G, P, R, XS, PS, st, c); break;
        CASE(0) CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6)
        default: hipLaunchKernelGGL((k_sh_step<7, PT>), dim3(g.blocks), dim3(256), 0, s, n, g.G, P, R, XS, PS, st, c); break;
#undef CASE
    }
}

template <typename PT>
static void sh_shift_upd(int l, const ShGrid &g, hipStream_t s, long long n, const double *R, PT *XS, PT *PS,
                         const ShState *st, const Ctl *c) {
    switch (l) {
#define CASE(L_) case L_: hipLaunchKernelGGL((k_sh_upd<L_, PT>), dim3(g.blocks), dim3(256), 0, s, n, g.G, R, XS, PS, st, c); break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7)
        default: hipLaunchKernelGGL((k_sh_upd<8, PT>), dim3(g.blocks), dim3(256), 0, s, n, g.G, R, XS, PS, st, c); break;
#undef CASE
    }
}

// The fused per-shift pass (k_sh_fused): rows x chunks of SH_KCHUNK shifts.
template <typename PT>
static void sh_shift_fused(int l, hipStream_t s, long long n, int ns, const double *R, const double *snap, PT *XSM,
                           PT *PSM, const ShState *st, const Ctl *c) {
    const dim3 grid((unsigned)vec_grid(n, 256, 1), (unsigned)((ns + SH_KCHUNK - 1) / SH_KCHUNK));
    switch (l) {
#define CASE(L_) case L_: hipLaunchKernelGGL((k_sh_fused<L_, PT>), grid, dim3(256), 0, s, n, ns, SH_KCHUNK, R, snap, XSM, PSM, st, c); break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7)
        default: hipLaunchKernelGGL((k_sh_fused<8, PT>), grid, dim3(256), 0, s, n, ns, SH_KCHUNK, R, snap, XSM, PSM, st, c); break;
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

// Shifted BiCGSTAB(l). Unfused (snap == nullptr): the shifts' p_shift and x_k in
// interleaved PT blocks, the seed in x[0, n), shift k's solution in XS (n x SP),
// starting from x's shifted parts (the reference's x + (k+1) n layout). Fused (snap: the
// seed's r snapshots, l(l+1)/2 n doubles): p_shift[k][0] in PS and x_k in XS, both
// shift-major (XS is x + n itself when PT is double), one k_sh_fused pass per outer
// iteration.
template <typename PT>
static int sh_core(ThreadCtx *t, const bkh_matrix *A, int l, const double *b, double *x, double *R,
                   double *P, double *rs, PT *PS, PT *XS, int SP, ShState *st, const double *sigma, int ns,
                   Ctl *c, double *part, double tol, int max_it, int K, double *snap) {
    const bool fused = snap != nullptr;
    hipStream_t s = t->stream;
    const long long n = A->n;
    const int G = vec_grid(n, 256);
    const ShGrid g = sh_grid(n, SP, 16 / (int)sizeof(PT));
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
    if (ns > 0 && fused) {
        const long long nps = (long long)ns * n;                // p_shift[k][0] starts at 0 (:53 calloc)
        hipLaunchKernelGGL(k_sh_zero<PT>, dim3(vec_grid(nps, 256)), dim3(256), 0, s, PS, nps);
        if ((const void *)XS != (const void *)(x + n))
            hipLaunchKernelGGL((k_conv<double, PT>), dim3(vec_grid(nps, 256)), dim3(256), 0, s, nps,
                               (const double *)(x + n), XS);
    } else if (ns > 0) {
        const long long nps = (long long)SP * (l + 1) * n;     // p_shift starts at 0 (:53 calloc)
        hipLaunchKernelGGL(k_sh_zero<PT>, dim3(vec_grid(nps, 256)), dim3(256), 0, s, PS, nps);
        hipLaunchKernelGGL((k_xs_in<double, PT>), dim3((unsigned)((n + 31) / 32), (unsigned)((SP + 31) / 32)), dim3(256),
                           0, s, n, ns, SP, (const double *)(x + n), XS);
    }
    Epi<0> e0;
    memset(&e0, 0, sizeof e0);
    launch_spmv<double, double, 0>(s, A, A->val, x, R, 1.0, 0.0, e0, nullptr, nullptr, NoFin());
    hipLaunchKernelGGL(k_sh_init, dim3(G), dim3(256), 0, s, n, b, R, P, rs, part);
    launch_finalize<2>(s, part, G, c, ShInitFin{st});
    LAUNCH_OK();
    hipGraphExec_t ge = capture_cached(t, {0x5B01, (uintptr_t)A, (uintptr_t)b, (uintptr_t)x, (uintptr_t)R, (uintptr_t)P,
                                           (uintptr_t)rs, (uintptr_t)PS, (uintptr_t)XS, (uintptr_t)SP, (uintptr_t)st,
                                           (uintptr_t)c, (uintptr_t)part, (uintptr_t)ns, (uintptr_t)l,
                                           (uintptr_t)sizeof(PT), (uintptr_t)snap}, s, K, [&]() {
        for (int j = 0; j < l; j++) {
            hipLaunchKernelGGL(k_sh_p, dim3(G), dim3(256), 0, s, n, j, P, (const double *)R, c);
            sh_alpha_spmv(s, A, P + j * n, P + (j + 1) * n, rs, c, part, st, j);
            if (ns > 0 && !fused) sh_shift_step<PT>(j, g, s, n, (const double *)P, (const double *)R, XS, PS, st, c);
            hipLaunchKernelGGL(k_sh_seed_r, dim3(G), dim3(256), 0, s, n, j, (const double *)P, R, x, c,
                               (ns > 0 && fused) ? snap : (double *)nullptr);
            if (j < l - 1)
                launch_spmv_dot_w<double, double>(s, A, A->val, R + j * n, R + (j + 1) * n, rs, c, part,
                                                  ShGammaFin());
            else
                sh_gram_dispatch(l, s, A, R, n, c, part, st);
        }
        if (ns > 0 && fused) sh_shift_fused<PT>(l, s, n, ns, (const double *)R, (const double *)snap, XS, PS, st, c);
        else if (ns > 0) sh_shift_upd<PT>(l, g, s, n, (const double *)R, XS, PS, st, c);
        hipLaunchKernelGGL(k_sh_seed_upd, dim3(G), dim3(256), 0, s, n, l, P, R, x, (const double *)rs, c, part);
        launch_finalize<2>(s, part, G, c, ShUpdFin());
    });
    if (!ge) { bkh_set_error("hipGraph capture of the shifted BiCGSTAB(l) iteration failed"); return BKH_DEVICE_ERROR; }
    int rc = run_graph_until_done(t, ge, c, max_it / K + 2);
    return rc;
}

static int pow2_at_least(int v) {
    int p = 4;
    while (p < v) p <<= 1;
    return p;
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
    const int SP = pow2_at_least(std::max(ns, 1));        // interleaved width (power of two, >= 4)
    const size_t pts = p32 ? sizeof(float) : sizeof(double);
    const size_t blk = (size_t)SP * n;                     // one interleaved block
    const bool shifts = ns > 0;
    // the fused per-shift pass unless BKH_SH_UNFUSED=1 (the per-step kernels, bitwise the same)
    const char *uf = getenv("BKH_SH_UNFUSED");
    const bool fused = shifts && !(uf && uf[0] == '1');
    size_t bytes = carve_size(sizeof(Ctl)) * 2 + carve_size(sizeof(ShState)) + carve_size(sizeof(double) * nparts) +
                   (2 * (l + 1) + 1) * carve_size(sizeof(double) * n) + carve_size(sizeof(MsCol) * SP) +
                   carve_size(sizeof(double) * (size_t)MS_NB * 2 * SP);
    if (fused) bytes += carve_size(sizeof(double) * n * (size_t)(l * (l + 1) / 2));
    if (shifts) {
        bytes += carve_size(pts * blk * (l + 1)) + carve_size(pts * blk);       // p_shift, x_k
        if (p32) bytes += carve_size(sizeof(double) * blk) + 7 * carve_size(sizeof(float) * blk);
    }
    void *ws = bkh_workspace(bytes);
    if (!ws) return BKH_DEVICE_ERROR;
    Carve cv(ws);
    Ctl *c = cv.take<Ctl>(1), *c2 = cv.take<Ctl>(1);
    ShState *st = cv.take<ShState>(1);
    double *part = cv.take<double>(nparts);
    double *R = cv.take<double>((l + 1) * n), *P = cv.take<double>((l + 1) * n), *rs = cv.take<double>(n);
    MsCol *mc = cv.take<MsCol>(SP);
    double *part2 = cv.take<double>((size_t)MS_NB * 2 * SP);
    char *PS = shifts ? cv.take<char>(pts * blk * (l + 1)) : nullptr;
    char *XS = shifts ? cv.take<char>(pts * blk) : nullptr;
    double *XS64 = p32 ? nullptr : (double *)XS;
    double *snap = fused ? cv.take<double>(n * (size_t)(l * (l + 1) / 2)) : nullptr;
    // fused: x_k shift-major, in x itself (FP64) or in a shift-major FP32 copy
    char *XW = (fused && !p32) ? (char *)(x + n) : XS;
    MsBufs B;
    memset(&B, 0, sizeof B);
    if (shifts && p32) {
        XS64 = cv.take<double>(blk);
        B.R = cv.take<float>(blk); B.RH = cv.take<float>(blk); B.D = cv.take<float>(blk);
        B.P = cv.take<float>(blk); B.V = cv.take<float>(blk); B.S = cv.take<float>(blk); B.T = cv.take<float>(blk);
    }
    hipStream_t s = t->stream;
    HIP_OK(hipEventRecord(t->ev0, s));
    int rc = p32 ? sh_core<float>(t, A, l, b, x, R, P, rs, (float *)PS, (float *)XW, SP, st, sigma, ns, c, part, tol,
                                  max_iter, K, snap)
                 : sh_core<double>(t, A, l, b, x, R, P, rs, (double *)PS, (double *)XW, SP, st, sigma, ns, c, part,
                                   tol, max_iter, K, snap);
    if (rc) return rc;
    const double fnb = t->ctl_host->fnb;
    if (shifts && p32) {
        // FP64 refinement of every shift (the FP32 p_shift limits them to ~kappa eps32)
        if (fused)
            hipLaunchKernelGGL((k_xs_in<float, double>), dim3((unsigned)((n + 31) / 32), (unsigned)((SP + 31) / 32)),
                               dim3(256), 0, s, n, ns, SP, (const float *)XS, XS64);
        else
            hipLaunchKernelGGL(k_to_f64<float>, dim3(vec_grid((long long)blk, 256)), dim3(256), 0, s, (long long)blk,
                               (const float *)XS, XS64);
        LAUNCH_OK();
        if (fnb > 0.0) {
            const int max_inner = o.max_inner > 0 ? o.max_inner : 1000;
            if (ms_refine(t, A, b, XS64, SP, ns, st, mc, c2, part2, B, fnb, tol, max_inner, 4, &inf.passes,
                          &inf.inner_iters))
                return BKH_DEVICE_ERROR;
        }
    }
    if (shifts && (p32 || !fused)) {
        hipLaunchKernelGGL(k_xs_out<double>, dim3((unsigned)((n + 31) / 32), (unsigned)((SP + 31) / 32)), dim3(256), 0,
                           s, n, ns, SP, (const double *)XS64, x);
        LAUNCH_OK();
    } else if (shifts && shift_true_res) {
        // fused FP64: the solutions are already in x; the per-shift residual kernel reads
        // the interleaved block
        hipLaunchKernelGGL((k_xs_in<double, double>), dim3((unsigned)((n + 31) / 32), (unsigned)((SP + 31) / 32)),
                           dim3(256), 0, s, n, ns, SP, (const double *)(x + n), XS64);
        LAUNCH_OK();
    }
    if (ctl_fetch(t, c)) return BKH_DEVICE_ERROR;          // the seed's control block for finish_solve
    int ret = finish_solve(t, A, 0.0, b, x, &inf, info);
    if (shift_true_res) {
        if (info) shift_true_res[0] = inf.true_res;
        else if (true_residual(t, A, 0.0, b, x, &shift_true_res[0])) return BKH_DEVICE_ERROR;
        if (shifts) {
            std::vector<double> rel(SP);
            if (fnb > 0.0 && ms_true_res(t, A, b, XS64, SP, st, mc, part2, fnb, ns, rel.data())) return BKH_DEVICE_ERROR;
            for (int k = 0; k < ns; k++) shift_true_res[k + 1] = fnb > 0.0 ? rel[k] : 0.0;
        }
    }
    return ret;
}
