// blbicgstabl.hip -- block BiCGSTAB(l) on the device (replaces src/blbicgstabl.c:16-252).
//
// Blocks are m x s (s <= 16), row-major in the device layer; the s x s algebra
// (dgesv, dgetrf/dgetri, the QR's R factor) runs in one workgroup on matrices kept in
// device memory (BbMats), so no scalar or small matrix crosses PCIe during a solve.
// One outer iteration (captured in a hipGraph, replayed until the control block
// says done) follows the reference step for step:
//   BiCG part, j = 0..l-1                                          (:96-164)
//     Gram Rs^T V_j (MFMA)  -> gamma1; beta = -alpha gamma0^-1 gamma1   (:101-109)
//     P_i = V_i + P_i beta, i <= j                                   (:112-117)
//     P_{j+1} = A P_j (SpMM); Gram Rs^T P_{j+1} -> alpha = (.)^-1 gamma1 (:117-124)
//     tmp = P_0 alpha, X += tmp xi; V_0 -= A tmp (SpMM)               (:128-132, :137-139)
//     V_0 = Q t by CholeskyQR2; xi = t xi                             (:135, :142-145)
//     V_i = (V_i - P_{i+1} alpha) t^-1, i = 1..j                      (:147-157)
//     V_{j+1} = A V_j (SpMM)                                         (:159-160)
//   MR part: R_i = V_i xi; Frobenius Gram of R_0..R_l in one pass; the MGS
//     recurrences of :170-187 in coefficient space (bkh_mr.hpp)
//   update: tmp = sum omega_i V_{i-1}, P_0 -= sum omega_i P_i; V_0 -= A tmp;
//     X += tmp xi; error = ||xi||_F / ||B||_F                         (:198-212)
// The QR's R has a positive diagonal (CholeskyQR2, first pass shifted if the Gram is
// numerically singular) where LAPACK's Householder R may carry signs; Rs, gamma,
// alpha and xi change by the same diagonal sign matrix and X does not.
// Deliberate differences: the initial residual is B - A X0 (the reference forms
// B + A X0, :83-86, identical for the X0 = 0 every caller passes), and a NaN error
// returns DIVERGENCE instead of a non-negative index (:210, :248-251).
#include <math.h>
#include "bkh_block.hpp"
#include "bkh_host.hpp"
#include "bkh_mr.hpp"
#include "bkh_small.hpp"
#include "bkh_solver_common.hpp"

struct BbMats {
    double gamma0[SM_N], gamma1[SM_N], alpha[SM_N], beta[SM_N], xi[SM_N];
    double tinv[SM_N], at[SM_N], R1[SM_N], R1i[SM_N], R2i[SM_N];
    double omega[MR_D];
    double fnb2;
    int m, s, l;
};

enum { BB_BETA = 0, BB_ALPHA = 1, BB_QR1 = 2, BB_QR2 = 3, BB_QR2_INIT = 4 };

// One workgroup: s x s Gram from the reduced partials, then the small dense step.
template <int S, int OP>
__global__ __launch_bounds__(256) void k_bb_fin(const double *part, int nblocks, BbMats *mt, Ctl *c,
                                                int j) {
    if (ctl_stopped(c)) return;
    __shared__ double G[SM_N], T1[SM_N], T2[SM_N], T3[SM_N], work[2 * SM_N];
    __shared__ int flag[2];
    __shared__ double red;
    const int t = threadIdx.x;
    if (t < S * S) {
        double acc = 0.0;
        for (int b = 0; b < nblocks; b++) acc += part[(size_t)b * S * S + t];
        G[t] = acc;
    }
    __syncthreads();
    if (OP == BB_BETA) {
        // gamma0 <- -omega_l gamma0 at the start of an outer iteration (:96)
        if (t < S * S) T3[t] = (j == 0) ? -mt->omega[mt->l] * mt->gamma0[t] : mt->gamma0[t];
        __syncthreads();
        // beta = -alpha * gamma0^-1 gamma1 (dgesv + dgemm, :104-108); gamma0 = gamma1
        if (!sm_inverse(S, T3, T1, work, flag)) { if (t == 0) ctl_breakdown(c); return; }
        sm_mul(S, T1, 0, G, 0, T2);
        sm_copy(S, mt->alpha, T1);
        sm_mul(S, T1, 0, T2, 0, T3);
        if (t < S * S) {
            mt->beta[t] = -T3[t];
            mt->gamma1[t] = G[t];
            mt->gamma0[t] = G[t];
        }
        return;
    }
    if (OP == BB_ALPHA) {
        // alpha = (Rs^T P_{j+1})^-1 gamma1 (dgesv, :119-124)
        if (!sm_inverse(S, G, T1, work, flag)) { if (t == 0) ctl_breakdown(c); return; }
        sm_copy(S, mt->gamma1, T2);
        sm_mul(S, T1, 0, T2, 0, T3);
        if (t < S * S) mt->alpha[t] = T3[t];
        return;
    }
    // CholeskyQR2 of V_0 (myqr, :135): pass 1 with a diagonal shift on failure
    bool ok = sm_chol_upper(S, G, 0.0, T1, flag);
    if (!ok && OP == BB_QR1) {
        if (t == 0) {
            double tr = 0.0;
            for (int k = 0; k < S; k++) tr += G[k + k * S];
            red = 11.0 * ((double)mt->m * S + S * (S + 1)) * 1.1102230246251565e-16 * tr;
        }
        __syncthreads();
        ok = sm_chol_upper(S, G, red, T1, flag);
    }
    if (!ok) { if (t == 0) ctl_breakdown(c); return; }
    sm_triu_inv(S, T1, T2);                       // T1 = R, T2 = R^-1
    if (OP == BB_QR1) {
        if (t < S * S) { mt->R1[t] = T1[t]; mt->R1i[t] = T2[t]; }
        return;
    }
    // pass 2: t = R2 R1, t^-1 = R1^-1 R2^-1
    if (t < S * S) mt->R2i[t] = T2[t];
    sm_copy(S, mt->R1, work);
    sm_mul(S, T1, 0, work, 0, T3);                // T3 = t
    sm_copy(S, mt->R1i, work);
    sm_mul(S, work, 0, T2, 0, G);                 // G = t^-1
    if (OP == BB_QR2_INIT) {                      // V_0 = Q, xi = R (:86-87)
        if (t < S * S) mt->xi[t] = T3[t];
        return;
    }
    sm_copy(S, mt->xi, work);
    sm_mul(S, T3, 0, work, 0, T1);                // xi <- t xi (:142-145)
    sm_copy(S, mt->alpha, work);
    sm_mul(S, work, 0, G, 0, T2);                 // at = alpha t^-1
    if (t < S * S) {
        mt->xi[t] = T1[t];
        mt->tinv[t] = G[t];
        mt->at[t] = T2[t];
    }
}

template <int S, int OP>
static void launch_bb_fin(hipStream_t s, const double *part, int nblocks, BbMats *mt, Ctl *c, int j) {
    double *r2 = bkh_red2();
    const int nb2 = (nblocks + RED_CHUNK - 1) / RED_CHUNK;
    hipLaunchKernelGGL(k_part_reduce, dim3(nb2), dim3(256), 0, s, part, nblocks, S * S, r2);
    hipLaunchKernelGGL((k_bb_fin<S, OP>), dim3(1), dim3(256), 0, s, (const double *)r2, nb2, mt, c, j);
}

#define TILE_LOOP(row0, m)                                                                    \
    for (long long row0 = ((long long)blockIdx.x * 4 + (threadIdx.x >> 6)) * 16; row0 < (m); \
         row0 += (long long)gridDim.x * 64)

// P_i = V_i + P_i beta, i = 0..j (blocks of m*S doubles, consecutive).
template <int S>
__global__ __launch_bounds__(256) void k_bb_pupd(long long m, double *P, const double *V, int j,
                                                 const BbMats *mt, const Ctl *c) {
    if (ctl_stopped(c)) return;
    __shared__ double Bm[SM_N];
    for (int e = threadIdx.x; e < S * S; e += 256) Bm[e] = mt->beta[e];
    __syncthreads();
    const long long ms = m * S;
    TILE_LOOP(row0, m) {
        for (int i = 0; i <= j; i++) {
            mf4 pt = tile_mul<S>(P + i * ms, m, row0, Bm) + tile_load<S>(V + i * ms, m, row0);
            tile_store<S>(P + i * ms, m, row0, pt);
        }
    }
}

// tmp = P_0 alpha (:128-129). X += tmp xi follows in k_bb_xacc on the stored tmp, so
// the solution and the residual (V_0 -= A tmp) see the same rounded tmp.
template <int S>
__global__ __launch_bounds__(256) void k_bb_tmp(long long m, const double *P0, double *tmp,
                                                const BbMats *mt, const Ctl *c) {
    if (ctl_stopped(c)) return;
    __shared__ double Am[SM_N];
    for (int e = threadIdx.x; e < S * S; e += 256) Am[e] = mt->alpha[e];
    __syncthreads();
    TILE_LOOP(row0, m) tile_store<S>(tmp, m, row0, tile_mul<S>(P0, m, row0, Am));
}

// V_0 = V_0 R1^-1 and its Gram (CholeskyQR pass 1 -> pass 2 input).
template <int S>
__global__ __launch_bounds__(256) void k_bb_vr(long long m, double *V0, const BbMats *mt, const Ctl *c,
                                               double *part) {
    if (ctl_stopped(c)) return;
    __shared__ double Rm[SM_N], sh[4 * 256];
    for (int e = threadIdx.x; e < S * S; e += 256) Rm[e] = mt->R1i[e];
    __syncthreads();
    mf4 acc = {0.0, 0.0, 0.0, 0.0};
    TILE_LOOP(row0, m) {
        mf4 vt = tile_mul<S>(V0, m, row0, Rm);
        tile_store<S>(V0, m, row0, vt);
        acc = tile_gram(vt, vt, acc);
    }
    gram_block_partial<S>(acc, sh, part);
}

// V_0 = V_0 R2^-1 (= Q); V_i = V_i t^-1 - P_{i+1} (alpha t^-1), i = 1..j (:147-157).
template <int S>
__global__ __launch_bounds__(256) void k_bb_vupd(long long m, double *V, const double *P, int j,
                                                 const BbMats *mt, const Ctl *c) {
    if (ctl_stopped(c)) return;
    __shared__ double Rm[SM_N], Tm[SM_N], Am[SM_N];
    for (int e = threadIdx.x; e < S * S; e += 256) { Rm[e] = mt->R2i[e]; Tm[e] = mt->tinv[e]; Am[e] = mt->at[e]; }
    __syncthreads();
    const long long ms = m * S;
    TILE_LOOP(row0, m) {
        tile_store<S>(V, m, row0, tile_mul<S>(V, m, row0, Rm));
        for (int i = 1; i <= j; i++) {
            mf4 vt = tile_mul<S>(V + i * ms, m, row0, Tm) - tile_mul<S>(P + (i + 1) * ms, m, row0, Am);
            tile_store<S>(V + i * ms, m, row0, vt);
        }
    }
}

// R_i = V_i xi, i = 0..l (:166-168).
template <int S>
__global__ __launch_bounds__(256) void k_bb_r(long long m, const double *V, double *R, int l,
                                              const BbMats *mt, const Ctl *c) {
    if (ctl_stopped(c)) return;
    __shared__ double Xm[SM_N];
    for (int e = threadIdx.x; e < S * S; e += 256) Xm[e] = mt->xi[e];
    __syncthreads();
    const long long ms = m * S;
    TILE_LOOP(row0, m) {
        for (int i = 0; i <= l; i++) tile_store<S>(R + i * ms, m, row0, tile_mul<S>(V + i * ms, m, row0, Xm));
    }
}

// X += tmp xi (:206-207).
template <int S>
__global__ __launch_bounds__(256) void k_bb_xacc(long long m, const double *tmp, double *X, const BbMats *mt,
                                                 const Ctl *c) {
    if (ctl_stopped(c)) return;
    __shared__ double Xm[SM_N];
    for (int e = threadIdx.x; e < S * S; e += 256) Xm[e] = mt->xi[e];
    __syncthreads();
    TILE_LOOP(row0, m) {
        tile_store<S>(X, m, row0, tile_load<S>(X, m, row0) + tile_mul<S>(tmp, m, row0, Xm));
    }
}

// Frobenius Gram <R_a, R_b>, 0 <= a <= b <= L, over the flat ms-long blocks.
template <int L>
__global__ __launch_bounds__(256) void k_bb_mgram(long long ms, const double *R, const Ctl *c, double *part) {
    if (ctl_stopped(c)) return;
    constexpr int NG = (L + 1) * (L + 2) / 2;
    double g[NG];
#pragma unroll
    for (int q = 0; q < NG; q++) g[q] = 0.0;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < ms; e += (long long)gridDim.x * 256) {
        double rv[L + 1];
#pragma unroll
        for (int a = 0; a <= L; a++) rv[a] = R[a * ms + e];
        int q = 0;
#pragma unroll
        for (int a = 0; a <= L; a++)
#pragma unroll
            for (int b = a; b <= L; b++) g[q++] += rv[a] * rv[b];
    }
    block_partial<NG, 256>(g, part);
}

// MR step (:170-190) from the Gram; omega_1..omega_l into mt.
template <int L>
struct BbMrFin {
    BbMats *mt;
    __device__ void operator()(Ctl *c, const double *tot) const {
        double G[MR_D][MR_D];
        gram_unpack(L, tot, G);
        MrOut o;
        if (!mr_from_gram(L, G, o)) { ctl_breakdown(c); return; }
        for (int i = 1; i <= L; i++) mt->omega[i] = o.om[i];
    }
};

// tmp = sum_i omega_i V_{i-1}; P_0 -= sum_i omega_i P_i (:198-203).
__global__ __launch_bounds__(256) void k_bb_upd(long long ms, int l, const double *V, double *P, double *tmp,
                                                const BbMats *mt, const Ctl *c) {
    if (ctl_stopped(c)) return;
    __shared__ double om[MR_D];
    if (threadIdx.x <= l) om[threadIdx.x] = mt->omega[threadIdx.x];
    __syncthreads();
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < ms; e += (long long)gridDim.x * 256) {
        double tv = 0.0, p0 = P[e];
        for (int i = 1; i <= l; i++) {
            tv += om[i] * V[(i - 1) * ms + e];
            p0 -= om[i] * P[i * ms + e];
        }
        tmp[e] = tv;
        P[e] = p0;
    }
}

// error = ||xi||_F / ||B||_F; stop on error < tol, NaN or max_iter (:209-212).
__global__ void k_bb_check(const BbMats *mt, Ctl *c) {
    if (ctl_stopped(c) || threadIdx.x != 0) return;
    const int s = mt->s;
    double f2 = 0.0;
    for (int k = 0; k < s * s; k++) f2 += mt->xi[k] * mt->xi[k];
    ctl_check(c, f2);
}

// Initial state: gamma0 = I, omega_l = 1, alpha = 0 (calloc, :31-41).
__global__ void k_bb_init(BbMats *mt, int m, int s, int l) {
    const int t = threadIdx.x;
    if (t < SM_N) {
        mt->gamma0[t] = (t < s * s && t % s == t / s) ? 1.0 : 0.0;
        mt->alpha[t] = 0.0;
        mt->xi[t] = 0.0;
    }
    if (t < MR_D) mt->omega[t] = (t == l) ? 1.0 : 0.0;
    if (t == 0) { mt->m = m; mt->s = s; mt->l = l; }
}

// R = B - AX, ||B||^2 partials (AX in R on entry).
__global__ __launch_bounds__(256) void k_bb_resid(long long ms, const double *B, double *R, double *part) {
    double v[1] = {0.0};
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < ms; e += (long long)gridDim.x * 256) {
        R[e] = B[e] - R[e];
        v[0] += B[e] * B[e];
    }
    block_partial<1, 256>(v, part);
}

struct BbFnbFin {
    BbMats *mt;
    __device__ void operator()(Ctl *c, const double *tot) const {
        mt->fnb2 = tot[0];
        c->fnb = sqrt(tot[0]);
        if (!(c->fnb > 0.0)) ctl_breakdown(c);
    }
};

__global__ __launch_bounds__(256) void k_bb_copy2(long long ms, const double *V0, double *P0, double *Rs) {
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < ms; e += (long long)gridDim.x * 256) {
        P0[e] = V0[e];
        Rs[e] = V0[e];
    }
}

// On the SELL-64 / slice-diagonal copies when the matrix has them (launch_spmm): the
// same stored-order row sums as the CSR-stream form.
template <int S>
static void bb_spmm(hipStream_t s, const bkh_matrix *A, const double *X, double *Y, double alpha, double beta,
                    const Ctl *c) {
    launch_spmm<S>(s, A, A->val, X, Y, alpha, beta, c, true);
}

template <int L>
static void bb_mr(hipStream_t s, long long ms, const double *R, double *part, int Gv, BbMats *mt, Ctl *c) {
    constexpr int NG = (L + 1) * (L + 2) / 2;
    hipLaunchKernelGGL((k_bb_mgram<L>), dim3(Gv), dim3(256), 0, s, ms, R, (const Ctl *)c, part);
    launch_finalize<NG>(s, part, Gv, c, BbMrFin<L>{mt});
}

template <int S>
static int blbicgstabl_dispatch(ThreadCtx *t, const bkh_matrix *A, int l, const double *B, double *X,
                                double tol, int max_iter, int K, bkh_info *inf, bool want_true) {
    const long long m = A->n, ms = m * S;
    const int G = block_grid(m), Gv = vec_grid(ms, 256);
    constexpr int NGMAX = (MR_LMAX + 1) * (MR_LMAX + 2) / 2;
    const int nparts = std::max(G * S * S, Gv * NGMAX);
    // blocks: Bt, Xt, Rs, tmp, P_0..P_l, V_0..V_l, R_0..R_l
    const int nblk = 4 + 3 * (l + 1);
    size_t bytes = carve_size(sizeof(Ctl)) + carve_size(sizeof(BbMats)) + carve_size(sizeof(double) * nparts) +
                   (size_t)nblk * carve_size(sizeof(double) * ms);
    void *ws = bkh_workspace(bytes);
    if (!ws) return BKH_DEVICE_ERROR;
    Carve cv(ws);
    Ctl *c = cv.take<Ctl>(1);
    BbMats *mt = cv.take<BbMats>(1);
    double *part = cv.take<double>(nparts);
    double *Bt = cv.take<double>(ms), *Xt = cv.take<double>(ms), *Rs = cv.take<double>(ms),
           *tmp = cv.take<double>(ms);
    double *P = cv.take<double>((l + 1) * ms), *V = cv.take<double>((l + 1) * ms), *R = cv.take<double>((l + 1) * ms);
    hipStream_t s = t->stream;
    HIP_OK(hipEventRecord(t->ev0, s));
    hipLaunchKernelGGL((k_cm2rm<double, double>), dim3(Gv), dim3(256), 0, s, m, S, B, Bt);
    hipLaunchKernelGGL((k_cm2rm<double, double>), dim3(Gv), dim3(256), 0, s, m, S, (const double *)X, Xt);
    if (ctl_init(t, c, tol, max_iter)) return BKH_DEVICE_ERROR;
    hipLaunchKernelGGL(k_bb_init, dim3(1), dim3(SM_N), 0, s, mt, (int)m, S, l);
    // V_0 = B - A X0, ||B||_F (:83-86, :92), [V_0, xi] = qr(V_0), P_0 = Rs = V_0 (:87-89)
    bb_spmm<S>(s, A, Xt, V, 1.0, 0.0, nullptr);
    hipLaunchKernelGGL(k_bb_resid, dim3(Gv), dim3(256), 0, s, ms, (const double *)Bt, V, part);
    launch_finalize<1>(s, part, Gv, c, BbFnbFin{mt});
    hipLaunchKernelGGL((k_gram<S, double>), dim3(G), dim3(256), 0, s, m, (const double *)V, (const double *)V,
                       (const Ctl *)c, part);
    launch_bb_fin<S, BB_QR1>(s, part, G, mt, c, 0);
    hipLaunchKernelGGL((k_bb_vr<S>), dim3(G), dim3(256), 0, s, m, V, (const BbMats *)mt, (const Ctl *)c, part);
    launch_bb_fin<S, BB_QR2_INIT>(s, part, G, mt, c, 0);
    hipLaunchKernelGGL((k_bb_vupd<S>), dim3(G), dim3(256), 0, s, m, V, (const double *)P, 0, (const BbMats *)mt,
                       (const Ctl *)c);
    hipLaunchKernelGGL(k_bb_copy2, dim3(Gv), dim3(256), 0, s, ms, (const double *)V, P, Rs);
    LAUNCH_OK();
    hipGraphExec_t ge = capture_cached(t, {0xBB00, (uintptr_t)A, (uintptr_t)Xt, (uintptr_t)Rs, (uintptr_t)tmp, (uintptr_t)P,
                                           (uintptr_t)V, (uintptr_t)R, (uintptr_t)mt, (uintptr_t)c, (uintptr_t)part,
                                           (uintptr_t)l, (uintptr_t)S}, s, K, [&]() {
        for (int j = 0; j < l; j++) {
            double *Vj = V + j * ms, *Pj = P + j * ms;
            hipLaunchKernelGGL((k_gram<S, double>), dim3(G), dim3(256), 0, s, m, (const double *)Rs,
                               (const double *)Vj, (const Ctl *)c, part);
            launch_bb_fin<S, BB_BETA>(s, part, G, mt, c, j);
            hipLaunchKernelGGL((k_bb_pupd<S>), dim3(G), dim3(256), 0, s, m, P, (const double *)V, j,
                               (const BbMats *)mt, (const Ctl *)c);
            bb_spmm<S>(s, A, Pj, Pj + ms, 1.0, 0.0, c);
            hipLaunchKernelGGL((k_gram<S, double>), dim3(G), dim3(256), 0, s, m, (const double *)Rs,
                               (const double *)(Pj + ms), (const Ctl *)c, part);
            launch_bb_fin<S, BB_ALPHA>(s, part, G, mt, c, j);
            hipLaunchKernelGGL((k_bb_tmp<S>), dim3(G), dim3(256), 0, s, m, (const double *)P, tmp,
                               (const BbMats *)mt, (const Ctl *)c);
            hipLaunchKernelGGL((k_bb_xacc<S>), dim3(G), dim3(256), 0, s, m, (const double *)tmp, Xt,
                               (const BbMats *)mt, (const Ctl *)c);
            bb_spmm<S>(s, A, tmp, V, -1.0, 1.0, c);
            hipLaunchKernelGGL((k_gram<S, double>), dim3(G), dim3(256), 0, s, m, (const double *)V,
                               (const double *)V, (const Ctl *)c, part);
            launch_bb_fin<S, BB_QR1>(s, part, G, mt, c, j);
            hipLaunchKernelGGL((k_bb_vr<S>), dim3(G), dim3(256), 0, s, m, V, (const BbMats *)mt, (const Ctl *)c,
                               part);
            launch_bb_fin<S, BB_QR2>(s, part, G, mt, c, j);
            hipLaunchKernelGGL((k_bb_vupd<S>), dim3(G), dim3(256), 0, s, m, V, (const double *)P, j,
                               (const BbMats *)mt, (const Ctl *)c);
            bb_spmm<S>(s, A, Vj, Vj + ms, 1.0, 0.0, c);
        }
        hipLaunchKernelGGL((k_bb_r<S>), dim3(G), dim3(256), 0, s, m, (const double *)V, R, l, (const BbMats *)mt,
                           (const Ctl *)c);
        switch (l) {
#define CASE(L_) case L_: bb_mr<L_>(s, ms, R, part, Gv, mt, c); break;
            CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
        }
        hipLaunchKernelGGL(k_bb_upd, dim3(Gv), dim3(256), 0, s, ms, l, (const double *)V, P, tmp,
                           (const BbMats *)mt, (const Ctl *)c);
        bb_spmm<S>(s, A, tmp, V, -1.0, 1.0, c);
        hipLaunchKernelGGL((k_bb_xacc<S>), dim3(G), dim3(256), 0, s, m, (const double *)tmp, Xt,
                           (const BbMats *)mt, (const Ctl *)c);
        hipLaunchKernelGGL(k_bb_check, dim3(1), dim3(64), 0, s, (const BbMats *)mt, c);
    });
    if (!ge) { bkh_set_error("hipGraph capture of the block BiCGSTAB(l) iteration failed"); return BKH_DEVICE_ERROR; }
    int rc = run_graph_until_done(t, ge, c, max_iter / K + 2);
    if (rc) return rc;
    Ctl hc = *t->ctl_host;
    hipLaunchKernelGGL((k_rm2cm<double, double>), dim3(Gv), dim3(256), 0, s, m, S, (const double *)Xt, X);
    LAUNCH_OK();
    HIP_OK(hipEventRecord(t->ev1, s));
    HIP_OK(hipEventSynchronize(t->ev1));
    float msec = 0.f;
    HIP_OK(hipEventElapsedTime(&msec, t->ev0, t->ev1));
    inf->time_ms = msec;
    bkh_note_solve_ms(msec);
    inf->ret = ctl_result(&hc);
    inf->status = hc.status;
    inf->iters = (hc.status == ST_MAXIT) ? max_iter : hc.result_it;
    inf->rec_res = hc.err;
    if (want_true) {
        // ||B - A X||_F / ||B||_F (:216-224)
        if (ctl_init(t, c, 0.0, 1)) return BKH_DEVICE_ERROR;
        bb_spmm<S>(s, A, Xt, tmp, 1.0, 0.0, nullptr);
        hipLaunchKernelGGL(k_bb_resid, dim3(Gv), dim3(256), 0, s, ms, (const double *)Bt, tmp, part);
        hipLaunchKernelGGL(k_bb_resid, dim3(Gv), dim3(256), 0, s, ms, (const double *)tmp, R, part + Gv);
        LAUNCH_OK();
        std::vector<double> h((size_t)Gv * 2);
        HIP_OK(hipMemcpyAsync(h.data(), part, sizeof(double) * Gv * 2, hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        double bb = 0.0, rr = 0.0;
        for (int i = 0; i < Gv; i++) { bb += h[i]; rr += h[Gv + i]; }
        inf->true_res = sqrt(rr) / sqrt(bb);
    }
    return inf->ret;
}

extern "C" int bkh_blbicgstabl(bkh_matrix *A, int s, const double *B, int l, double *X, double tol,
                               int max_iter, const bkh_opts *opts, bkh_info *info) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (!A || !B || !X) return BKH_BAD_ARGUMENT;
    if (s < 1 || s > SM_MAX) { bkh_set_error("block BiCGSTAB(l) supports 1 <= s <= %d", SM_MAX); return -3; }
    if (l < 1 || l > MR_LMAX) { bkh_set_error("block BiCGSTAB(l) supports 1 <= l <= %d", MR_LMAX); return -3; }
    bkh_opts o = opts ? *opts : bkh_opts{};
    bkh_info inf;
    memset(&inf, 0, sizeof inf);
    if (max_iter <= 0) {                 // error = 0 < tol -> returns iter = 0 (:98, :248-251)
        inf.status = ST_CONVERGED;
        if (info) *info = inf;
        return 0;
    }
    const int K = o.graph_iters > 0 ? o.graph_iters : 1;
    int ret;
    switch (s) {
#define CASE(S_) case S_: ret = blbicgstabl_dispatch<S_>(t, A, l, B, X, tol, max_iter, K, &inf, true); break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
        CASE(9) CASE(10) CASE(11) CASE(12) CASE(13) CASE(14) CASE(15) CASE(16)
#undef CASE
    default: return -3;
    }
    if (info) *info = inf;
    return ret;
}
