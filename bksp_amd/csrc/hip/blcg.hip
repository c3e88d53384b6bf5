// blcg.hip -- block CG with residual QR on the device (replaces src/blcg.c:15-119).
//
// Blocks are m x s (s <= 16), row-major inside the device layer. One iteration:
//   SpMM      AP = A P                                                (blcg.c:74-75)
//   k_gram    P^T AP on MFMA -> finalize: alpha = (P^T AP)^-1, W = alpha C  (:78-86)
//   k_xv      X += P W, V -= AP alpha, Gram V^T V (MFMA)              (:87-92)
//   k_vr      CholQR pass 1: V = V R1^-1, Gram V^T V                  (:95, myqr.c)
//   finalize  R2 = chol, beta = R2 R1, C = beta C, ||C||_F/||B||_F     (:98-105)
//   k_vp      CholQR pass 2: V = V R2^-1, and the next P = V + P beta^T  (:68-71)
// The QR of the residual block is CholeskyQR2 (two Gram + triangular-solve passes,
// the first with a Fukaya-style diagonal shift if the Gram matrix is numerically
// singular). Its R has a positive diagonal where LAPACK's Householder R may have
// negative entries; X is invariant under that sign choice (Q -> QD, R -> DR gives
// P -> PD, alpha -> D alpha D, C -> DC, so P alpha C is unchanged).
// First direction: blcg.c leaves beta's off-diagonal uninitialised (malloc) and sets
// its diagonal to 1, so P0 = V + P beta^T; this build uses the calloc reading,
// P0 = 2V, unless bkh_opts.fix_blcg_p0 asks for P0 = V.
#include <math.h>
#include "bkh_block.hpp"
#include "bkh_host.hpp"
#include "bkh_small.hpp"
#include "bkh_spmv.hpp"
#include "bkh_solver_common.hpp"

struct BlkMats {
    double C[SM_N], alpha[SM_N], W[SM_N], R1[SM_N], R1i[SM_N], R2i[SM_N], R12i[SM_N], betaT[SM_N];
    double work[2 * SM_N];
    double fnb2;
    int m, s;
};

enum { FIN_QR1 = 0, FIN_QR2_INIT = 1, FIN_QR2_ITER = 2, FIN_ALPHA = 3, FIN_FNB = 4 };

__global__ __launch_bounds__(256) void k_sh_fill_f(float *p, long long m) {
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < m; e += (long long)gridDim.x * 256) p[e] = 0.0f;
}

// Partial rows the fin kernel sums itself; more go through k_part_reduce first. The
// 1024-thread fin kernels (s <= 8) take up to FIN_DIRECT_WIDE rows (16 groups of
// threads, 32 rows each at most).
#define FIN_DIRECT 256
#define FIN_DIRECT_WIDE 512

// Fixed-order sum of nblocks rows of E partials by a BS-thread workgroup: thread
// t < Q*E (Q = BS/E groups) sums rows q, q+Q, q+2Q, ... of entry e = t % E, q = t / E,
// then the Q group sums are added in order. The rows of a group are loaded RMAX at
// a time into registers before they are added, so their loads (from another XCD's
// writes, i.e. from memory) are in flight together instead of in dependent batches.
// red: Q*E <= BS doubles of LDS; out[e], e < E, valid in every thread on return.
template <int E, int BS>
__device__ __forceinline__ void fin_sum_parts(const double *part, int nblocks, double *red, double *out) {
    constexpr int Q = BS / E;
    constexpr int RMAX = (FIN_DIRECT + Q - 1) / Q < 64 ? (FIN_DIRECT + Q - 1) / Q : 64;
    const int t = threadIdx.x;
    if (t < Q * E) {
        const int e = t % E;
        double acc = 0.0;
        for (int b0 = t / E; b0 < nblocks; b0 += RMAX * Q) {
            double v[RMAX];
#pragma unroll
            for (int u = 0; u < RMAX; u++) {
                const int b = b0 + u * Q;
                v[u] = b < nblocks ? part[(size_t)b * E + e] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < RMAX; u++)
                if (b0 + u * Q < nblocks) acc += v[u];
        }
        red[t] = acc;
    }
    __syncthreads();
    if (t < E) {
        double v = red[t];
        for (int q = 1; q < Q; q++) v += red[q * E + t];
        out[t] = v;
    }
    __syncthreads();
}

// The small dense steps for s <= 8 on the first wave, the matrices held by columns in
// registers (bkh_small.hpp's rg_* forms, bitwise the LDS forms): G is the summed
// Gram in LDS. 128^3 block CG FP32-IR, s = 8: the three fins 28.5 -> 26.5 us per
// iteration (each launch costs ~4.4 us before any work; the rest is the partial sums).
template <int S, int OP>
__device__ __forceinline__ void blk_fin_regs(const double *G, BlkMats *mt, Ctl *c) {
    const int j = (int)(threadIdx.x & 63);
    const int jc = j < S ? j : S - 1;               // lanes past the columns load column S - 1
    const bool own = j < S;
    double g[S];
#pragma unroll
    for (int i = 0; i < S; i++) g[i] = G[i + jc * S];
    if constexpr (OP == FIN_ALPHA) {
        // alpha = (P^T A P)^-1 (dgetrf/dgetri, blcg.c:81-82), W = alpha C (:85-86)
        double m[S];
#pragma unroll
        for (int i = 0; i < S; i++) m[i] = own ? g[i] : ((i == j - S) ? 1.0 : 0.0);
        if (!rg_inverse<S>(m)) { if (j == 0) ctl_breakdown(c); return; }
        if (j >= S && j < 2 * S) {
#pragma unroll
            for (int i = 0; i < S; i++) mt->alpha[i + (j - S) * S] = m[i];
        }
        double cc[S], w[S];
#pragma unroll
        for (int i = 0; i < S; i++) cc[i] = mt->C[i + jc * S];
        rg_mul<S>(m, S, cc, w);
        if (own) {
#pragma unroll
            for (int i = 0; i < S; i++) mt->W[i + j * S] = w[i];
        }
        return;
    } else {
        // CholQR pass: G = R^T R; pass 1 retries with a diagonal shift when G is
        // numerically singular (shifted CholeskyQR, Fukaya et al. 2020)
        double r[S];
#pragma unroll
        for (int i = 0; i < S; i++) r[i] = g[i];
        bool ok = rg_chol_upper<S>(r, 0.0);
        if (!ok && OP == FIN_QR1) {
            double tr = 0.0;
#pragma unroll
            for (int k = 0; k < S; k++) tr += G[k + k * S];
            const double shift = 11.0 * ((double)mt->m * S + S * (S + 1)) * 1.1102230246251565e-16 * tr;
#pragma unroll
            for (int i = 0; i < S; i++) r[i] = g[i];
            ok = rg_chol_upper<S>(r, shift);
        }
        if (!ok) { if (j == 0) ctl_breakdown(c); return; }
        double ri[S];
        rg_triu_inv<S>(r, ri);
        if constexpr (OP == FIN_QR1) {
            if (own) {
#pragma unroll
                for (int i = 0; i < S; i++) { mt->R1[i + j * S] = r[i]; mt->R1i[i + j * S] = ri[i]; }
            }
            return;
        } else {
            // pass 2: full R of the QR = R2 R1; R1^-1 R2^-1 for the row-stream vp
            double r1[S], r1i[S], t2[S], t1[S];
#pragma unroll
            for (int i = 0; i < S; i++) { r1[i] = mt->R1[i + jc * S]; r1i[i] = mt->R1i[i + jc * S]; }
            rg_mul<S>(r, 0, r1, t2);
            rg_mul<S>(r1i, 0, ri, t1);
            if (own) {
#pragma unroll
                for (int i = 0; i < S; i++) { mt->R2i[i + j * S] = ri[i]; mt->R12i[i + j * S] = t1[i]; }
            }
            if constexpr (OP == FIN_QR2_INIT) {                    // R0 = V C  (blcg.c:58)
                if (own) {
#pragma unroll
                    for (int i = 0; i < S; i++) mt->C[i + j * S] = t2[i];
                }
                return;
            } else {
                // C = beta C; error = ||C||_F / ||B||_F (blcg.c:98-105); next P uses beta^T
                double cc[S], bc[S];
#pragma unroll
                for (int i = 0; i < S; i++) cc[i] = mt->C[i + jc * S];
                rg_mul<S>(t2, 0, cc, bc);
                if (own) {
#pragma unroll
                    for (int i = 0; i < S; i++) { mt->C[i + j * S] = bc[i]; mt->betaT[j + i * S] = t2[i]; }
                }
                double f2 = 0.0;
#pragma unroll
                for (int jj = 0; jj < S; jj++)
#pragma unroll
                    for (int i = 0; i < S; i++) {
                        const double v = rg_bcast(bc[i], jj);
                        f2 += v * v;
                    }
                if (j == 0) ctl_check(c, f2);
            }
        }
    }
}

// One workgroup: sums the s x s Gram partials of `nblocks` workgroups in a fixed
// order, then runs the block solver's small dense step `op`.
template <int S, int OP, int BS, bool REG = true>
__global__ __launch_bounds__(BS) void k_blk_fin(const double *part, int nblocks, BlkMats *mt,
                                                Ctl *c) {
    if (ctl_stopped(c)) return;
    __shared__ double G[SM_N], R[SM_N], Ri[SM_N], T1[SM_N], T2[SM_N], work[BS > 2 * SM_N ? BS : 2 * SM_N];
    __shared__ int flag[2];
    __shared__ double red;
    const int t = threadIdx.x;
    if (OP == FIN_FNB) {
        fin_sum_parts<1, BS>(part, nblocks, work, G);
        if (t == 0) {
            mt->fnb2 = G[0];
            c->fnb = sqrt(G[0]);
            if (!(c->fnb > 0.0)) ctl_breakdown(c);
        }
        return;
    }
    fin_sum_parts<S * S, BS>(part, nblocks, work, G);
    // s <= 8: the small dense step runs on the first wave alone, with wave barriers
    // instead of workgroup barriers (block CG s = 8 at 128^3: 350.3 -> 348.5 ms FP64,
    // 192.9 -> 190.2 ms FP32, the same iterates)
    constexpr bool W = S * S <= 64;
    if (W && t >= 64) return;
    if constexpr (W && REG) {
        blk_fin_regs<S, OP>(G, mt, c);
        return;
    }
    if (OP == FIN_ALPHA) {
        // alpha = (P^T A P)^-1 (dgetrf/dgetri, blcg.c:81-82), W = alpha C (:85-86)
        if (!sm_inverse<W>(S, G, Ri, work, flag)) { if (t == 0) ctl_breakdown(c); return; }
        sm_copy<W>(S, Ri, mt->alpha);
        sm_copy<W>(S, mt->C, T1);
        sm_mul<W>(S, Ri, 0, T1, 0, T2);
        sm_copy<W>(S, T2, mt->W);
        return;
    }
    // CholQR pass: G = R^T R. Pass 1 retries with a diagonal shift when G is
    // numerically singular (shifted CholeskyQR, Fukaya et al. 2020).
    bool ok = sm_chol_upper<W>(S, G, 0.0, R, flag);
    if (!ok && OP == FIN_QR1) {
        double tr = 0.0;
        if (t == 0) {
            for (int k = 0; k < S; k++) tr += G[k + k * S];
            red = 11.0 * ((double)mt->m * S + S * (S + 1)) * 1.1102230246251565e-16 * tr;
        }
        sm_sync<W>();
        ok = sm_chol_upper<W>(S, G, red, R, flag);
    }
    if (!ok) { if (t == 0) ctl_breakdown(c); return; }
    sm_triu_inv<W>(S, R, Ri);
    if (OP == FIN_QR1) {
        sm_copy<W>(S, R, mt->R1);
        sm_copy<W>(S, Ri, mt->R1i);
        return;
    }
    // pass 2: full R of the QR = R2 R1
    sm_copy<W>(S, mt->R1, T1);
    sm_mul<W>(S, R, 0, T1, 0, T2);                  // T2 = beta = R2 R1
    sm_copy<W>(S, Ri, mt->R2i);
    sm_copy<W>(S, mt->R1i, G);
    sm_mul<W>(S, G, 0, Ri, 0, T1);                 // R1^-1 R2^-1 for the row-stream vp
    sm_copy<W>(S, T1, mt->R12i);
    if (OP == FIN_QR2_INIT) {                     // R0 = V C  (blcg.c:58)
        sm_copy<W>(S, T2, mt->C);
        return;
    }
    // C = beta C; error = ||C||_F / ||B||_F (blcg.c:98-105); next P uses beta^T
    sm_copy<W>(S, mt->C, T1);
    sm_mul<W>(S, T2, 0, T1, 0, R);                  // R <- beta C
    sm_copy<W>(S, R, mt->C);
    if (t < S * S) mt->betaT[(t / S) + (t % S) * S] = T2[t];
    const double f2 = sm_fro2<W>(S, R, &red);
    if (t == 0) ctl_check(c, f2);
}

// Test hook: 1 runs the s <= 8 small dense steps in their LDS form instead of the
// register form (the two are bitwise the same); cached graphs are retired.
static std::atomic<int> g_fin_lds_form{0};
extern "C" void bkh_debug_fin_lds_form(int on) {
    g_fin_lds_form.store(on ? 1 : 0);
    bkh_graph_epoch_bump();
}

template <int S, int OP>
static void launch_blk_fin(hipStream_t s, const double *part, int nblocks, BlkMats *mt, Ctl *c) {
    const int E = (OP == FIN_FNB) ? 1 : S * S;
    // s <= 8: 1024 threads, each summing a quarter as many partial rows as with 256
    // (block CG s = 8 at 128^3: 349.7 -> 346.6 ms FP64, 193.3 -> 190.1 ms FP32); larger s
    // keep 256 (their 64 partial loads per thread would not fit 128 registers)
    constexpr int FBS = S <= 8 ? 1024 : 256;
    if (nblocks <= (S <= 8 ? FIN_DIRECT_WIDE : FIN_DIRECT)) {
        if (g_fin_lds_form)
            hipLaunchKernelGGL((k_blk_fin<S, OP, FBS, false>), dim3(1), dim3(FBS), 0, s, part, nblocks, mt, c);
        else
            hipLaunchKernelGGL((k_blk_fin<S, OP, FBS>), dim3(1), dim3(FBS), 0, s, part, nblocks, mt, c);
        return;
    }
    double *r2 = bkh_red2();
    const int nb2 = (nblocks + RED_CHUNK - 1) / RED_CHUNK;
    hipLaunchKernelGGL(k_part_reduce, dim3(nb2), dim3(256), 0, s, part, nblocks, E, r2);
    hipLaunchKernelGGL((k_blk_fin<S, OP, 256>), dim3(1), dim3(256), 0, s, (const double *)r2, nb2, mt, c);
}

template <typename T>
__global__ __launch_bounds__(256) void k_blk_resid(long long ms, const T *B, const T *AX, T *V,
                                                   double *part) {
    double v[1] = {0.0};
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < ms; e += (long long)gridDim.x * 256) {
        const T bv = B[e];                 // read once: V may alias B (and AX)
        V[e] = bv - AX[e];
        v[0] += (double)bv * (double)bv;
    }
    block_partial<1, 256>(v, part);
}

// Loads an s x s matrix (column-major) into LDS.
__device__ __forceinline__ void load_small(double *dst, const double *src, int n) {
    for (int t = threadIdx.x; t < n; t += blockDim.x) dst[t] = src[t];
}

// Panel kernels: each wave streams 16-row tiles through its LDS tiles (stage_fetch /
// stage_store: 16-byte vectors on every lane), multiplies on FP64 MFMA and, where the
// next step needs it, accumulates the Gram matrix of the stored result.
#define BLK_TILES __shared__ double tbs[4][2][16 * TS_LD]; \
    double *t0 = tbs[threadIdx.x >> 6][0], *t1 = tbs[threadIdx.x >> 6][1]

// X += P W; V -= AP alpha; Gram V^T V of the new V.
template <int S, typename T>
__global__ __launch_bounds__(256) void k_blk_xv(long long m, T *X, const T *P, T *V, const T *AP,
                                                const BlkMats *mt, const Ctl *c, double *part) {
    if (ctl_stopped(c)) return;
    __shared__ double Wm[SM_N], Am[SM_N], sh[4 * 256];
    BLK_TILES;
    load_small(Wm, mt->W, S * S);
    load_small(Am, mt->alpha, S * S);
    __syncthreads();
    const int w = threadIdx.x >> 6;
    const long long step = (long long)gridDim.x * 64;
    long long row0 = ((long long)blockIdx.x * 4 + w) * 16;
    mf4 acc = {0.0, 0.0, 0.0, 0.0};
    StageRegs<S, T> rp, rx, ra, rv;
    if (row0 < m) {
        stage_load<S>(P, m, row0, rp);
        stage_load<S>(X, m, row0, rx);
        stage_load<S>(AP, m, row0, ra);
        stage_load<S>(V, m, row0, rv);
    }
    for (; row0 < m; row0 += step) {
        const long long nx = row0 + step;
        wave_lds_sync();
        stage_put<S>(rp, t0);
        stage_put<S>(rx, t1);
        if (nx < m) { stage_load<S>(P, m, nx, rp); stage_load<S>(X, m, nx, rx); }
        wave_lds_sync();
        const mf4 xt = ts_load_d<S>(t1) + ts_mul<S>(t0, Wm);
        wave_lds_sync();
        ts_put_d<S>(t1, xt);
        stage_put<S>(ra, t0);
        if (nx < m) stage_load<S>(AP, m, nx, ra);
        wave_lds_sync();
        stage_store<S>(X, m, row0, t1);
        const mf4 dv = ts_mul<S>(t0, Am);
        wave_lds_sync();
        stage_put<S>(rv, t1);
        if (nx < m) stage_load<S>(V, m, nx, rv);
        wave_lds_sync();
        const mf4 vt = as_stored<T>(ts_load_d<S>(t1) - dv);
        wave_lds_sync();
        ts_put_d<S>(t1, vt);
        wave_lds_sync();
        stage_store<S>(V, m, row0, t1);
        acc = tile_gram(vt, vt, acc);
    }
    gram_block_partial<S>(acc, sh, part);
}

// V = V R1^-1; Gram V^T V.
template <int S, typename T>
__global__ __launch_bounds__(256) void k_blk_vr(long long m, T *V, const BlkMats *mt, const Ctl *c,
                                                double *part) {
    if (ctl_stopped(c)) return;
    __shared__ double Rm[SM_N], sh[4 * 256];
    BLK_TILES;
    (void)t1;
    load_small(Rm, mt->R1i, S * S);
    __syncthreads();
    const int w = threadIdx.x >> 6;
    const long long step = (long long)gridDim.x * 64;
    long long row0 = ((long long)blockIdx.x * 4 + w) * 16;
    mf4 acc = {0.0, 0.0, 0.0, 0.0};
    StageRegs<S, T> rv;
    if (row0 < m) stage_load<S>(V, m, row0, rv);
    for (; row0 < m; row0 += step) {
        wave_lds_sync();
        stage_put<S>(rv, t0);
        if (row0 + step < m) stage_load<S>(V, m, row0 + step, rv);
        wave_lds_sync();
        const mf4 vt = as_stored<T>(ts_mul<S>(t0, Rm));
        wave_lds_sync();
        ts_put_d<S>(t0, vt);
        wave_lds_sync();
        stage_store<S>(V, m, row0, t0);
        acc = tile_gram(vt, vt, acc);
    }
    gram_block_partial<S>(acc, sh, part);
}

// V = V R2^-1 (k_blk_vr stored V R1^-1); P = V + P beta^T (or P = pscale V at the start).
template <int S, typename T>
__global__ __launch_bounds__(256) void k_blk_vp(long long m, T *V, T *P, const BlkMats *mt,
                                                double pscale, const Ctl *c) {
    if (ctl_stopped(c)) return;
    __shared__ double Rm[SM_N], Bm[SM_N];
    BLK_TILES;
    load_small(Rm, mt->R2i, S * S);
    load_small(Bm, mt->betaT, S * S);
    __syncthreads();
    const int w = threadIdx.x >> 6;
    const long long step = (long long)gridDim.x * 64;
    long long row0 = ((long long)blockIdx.x * 4 + w) * 16;
    const bool withp = (pscale == 0.0);
    StageRegs<S, T> rv, rp;
    if (row0 < m) {
        stage_load<S>(V, m, row0, rv);
        if (withp) stage_load<S>(P, m, row0, rp);
    }
    for (; row0 < m; row0 += step) {
        const long long nx = row0 + step;
        wave_lds_sync();
        stage_put<S>(rv, t0);
        if (withp) stage_put<S>(rp, t1);
        if (nx < m) {
            stage_load<S>(V, m, nx, rv);
            if (withp) stage_load<S>(P, m, nx, rp);
        }
        wave_lds_sync();
        const mf4 vt = as_stored<T>(ts_mul<S>(t0, Rm));
        const mf4 pt = withp ? vt + ts_mul<S>(t1, Bm) : vt * pscale;
        wave_lds_sync();
        ts_put_d<S>(t0, vt);
        ts_put_d<S>(t1, pt);
        wave_lds_sync();
        stage_store<S>(V, m, row0, t0);
        stage_store<S>(P, m, row0, t1);
    }
}

// ----------------------------------------------- row panel kernels (s <= 8)
//
// A lane owns a row of the m x S row-major blocks: the products with the s x s
// matrices (column-major, read with uniform addresses) run in FP64 with explicit
// fma, results are rounded to T as stored, and the Gram matrix of the stored
// result is accumulated per lane in FP64 (upper triangle, S(S+1)/2 values), then
// reduced once per workgroup (wave shuffles, LDS across the 8 waves) and written
// as a full s x s partial. Streaming kernels without LDS phases: for s <= 8 the
// FP64 flops per row are far below what the HBM time of its 2-6 row vectors
// allows. 512-thread workgroups (up to 256 VGPRs per lane: no spills of the FP64
// Gram accumulators), at most FIN_DIRECT of them, so the fin kernel
// sums the partials itself.
#define RP_BS 512

static inline int rp_grid(long long m) {
    long long g = (m + RP_BS - 1) / RP_BS;
    return (int)std::max(1LL, std::min(g, (long long)FIN_DIRECT));
}

// A row of S values of T in registers (vector loads of 16, 8 or sizeof(T) bytes).
template <int S, typename T>
struct RowRaw {
    static constexpr int B = S * (int)sizeof(T);
    static constexpr int NE = (B % 16 == 0) ? 16 / (int)sizeof(T) : (B % 8 == 0) ? 8 / (int)sizeof(T) : 1;
    T v[S];
    __device__ __forceinline__ void load(const T *__restrict__ p) {
        if constexpr (NE == 1) {
#pragma unroll
            for (int j = 0; j < S; j++) v[j] = p[j];
        } else {
            typedef T vt __attribute__((ext_vector_type(NE)));
#pragma unroll
            for (int q = 0; q < S / NE; q++) {
                const vt w = *(const vt *)(p + q * NE);
#pragma unroll
                for (int u = 0; u < NE; u++) v[q * NE + u] = w[u];
            }
        }
    }
    __device__ __forceinline__ void get(double (&r)[S]) const {
#pragma unroll
        for (int j = 0; j < S; j++) r[j] = (double)v[j];
    }
};

template <int S, typename T>
__device__ __forceinline__ void row_store(T *__restrict__ p, const double (&r)[S]) {
    constexpr int NE = RowRaw<S, T>::NE;
    if constexpr (NE == 1) {
#pragma unroll
        for (int j = 0; j < S; j++) p[j] = (T)r[j];
    } else {
        typedef T vt __attribute__((ext_vector_type(NE)));
#pragma unroll
        for (int q = 0; q < S / NE; q++) {
            vt v;
#pragma unroll
            for (int u = 0; u < NE; u++) v[u] = (T)r[q * NE + u];
            *(vt *)(p + q * NE) = v;
        }
    }
}

// o = r M (M column-major S x S), rounded to T when `round`.
template <int S, typename T, bool ROUND>
__device__ __forceinline__ void row_mul(const double (&r)[S], const double *__restrict__ M, double (&o)[S]) {
#pragma unroll
    for (int j = 0; j < S; j++) {
        double a = 0.0;
#pragma unroll
        for (int k = 0; k < S; k++) a = fma(r[k], M[k + j * S], a);
        o[j] = ROUND ? (double)(T)a : a;
    }
}

template <int S>
__device__ __forceinline__ void row_gram(const double (&x)[S], const double (&y)[S], double (&g)[S * (S + 1) / 2]) {
    int q = 0;
#pragma unroll
    for (int i = 0; i < S; i++)
#pragma unroll
        for (int j = i; j < S; j++, q++) g[q] = fma(x[i], y[j], g[q]);
}

// Workgroup total of the per-lane upper-triangle Grams, written as the full
// (mirrored) s x s partial of this workgroup.
template <int S>
__device__ __forceinline__ void row_gram_partial(double (&g)[S * (S + 1) / 2], double *part) {
    constexpr int NG = S * (S + 1) / 2, NW = RP_BS / 64;
    __shared__ double sh[NW][NG];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NG; q++) g[q] = wave_sum(g[q]);
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < NG; q++) sh[w][q] = g[q];
    }
    __syncthreads();
    const int t = threadIdx.x;
    if (t < NG) {
        double v = sh[0][t];
        for (int k = 1; k < NW; k++) v += sh[k][t];
        int i = 0, r = t;                      // t -> (i, j), i <= j, row-wise upper triangle
        while (r >= S - i) { r -= S - i; i++; }
        const int j = i + r;
        double *pb = part + (size_t)blockIdx.x * S * S;
        pb[i + j * S] = v;
        pb[j + i * S] = v;
    }
}

// Each lane's next row is loaded before the current one is processed (register
// double buffering), so a wave keeps its loads in flight through the arithmetic.
#define RP_LOOP_BEGIN                                                         \
    const long long step = (long long)gridDim.x * RP_BS;                      \
    long long row = (long long)blockIdx.x * RP_BS + threadIdx.x;

// X += P W; V -= AP alpha; Gram V^T V of the new V.
template <int S, typename T>
__global__ __launch_bounds__(RP_BS) void k_blkr_xv(long long m, T *__restrict__ X, const T *__restrict__ P,
                                                   T *__restrict__ V, const T *__restrict__ AP,
                                                   const BlkMats *__restrict__ mt, const Ctl *c,
                                                   double *__restrict__ part) {
    if (ctl_stopped(c)) return;
    double g[S * (S + 1) / 2];
#pragma unroll
    for (int q = 0; q < S * (S + 1) / 2; q++) g[q] = 0.0;
    RP_LOOP_BEGIN
    RowRaw<S, T> rp, rx, ra, rv;
    if (row < m) { rp.load(P + row * S); rx.load(X + row * S); ra.load(AP + row * S); rv.load(V + row * S); }
    for (; row < m; row += step) {
        double p[S], x[S], a[S], v[S], d[S];
        rp.get(p); rx.get(x); ra.get(a); rv.get(v);
        const long long nr = row + step;
        if (nr < m) { rp.load(P + nr * S); rx.load(X + nr * S); ra.load(AP + nr * S); rv.load(V + nr * S); }
        row_mul<S, T, false>(p, mt->W, d);
#pragma unroll
        for (int j = 0; j < S; j++) x[j] += d[j];
        row_store<S>(X + row * S, x);
        row_mul<S, T, false>(a, mt->alpha, d);
#pragma unroll
        for (int j = 0; j < S; j++) v[j] = (double)(T)(v[j] - d[j]);
        row_store<S>(V + row * S, v);
        row_gram<S>(v, v, g);
    }
    row_gram_partial<S>(g, part);
}

// CholeskyQR2 pass 1 without its store: the Gram of V R1^-1 (each row rounded to T as
// a stored copy would be). k_blkr_vp then applies the product R1^-1 R2^-1 (formed by
// the pass-2 finalize) to the same V in one step, so the intermediate block is never
// written and the panel kernel does one s x s product per row instead of two.
template <int S, typename T>
__global__ __launch_bounds__(RP_BS) void k_blkr_vr(long long m, const T *__restrict__ V,
                                                   const BlkMats *__restrict__ mt, const Ctl *c,
                                                   double *__restrict__ part) {
    if (ctl_stopped(c)) return;
    double g[S * (S + 1) / 2];
#pragma unroll
    for (int q = 0; q < S * (S + 1) / 2; q++) g[q] = 0.0;
    RP_LOOP_BEGIN
    RowRaw<S, T> rv;
    if (row < m) rv.load(V + row * S);
    for (; row < m; row += step) {
        double v[S], o[S];
        rv.get(v);
        const long long nr = row + step;
        if (nr < m) rv.load(V + nr * S);
        row_mul<S, T, true>(v, mt->R1i, o);
        row_gram<S>(o, o, g);
    }
    row_gram_partial<S>(g, part);
}

// V = V R1^-1 R2^-1 (V as k_blkr_vr read it); P = V + P beta^T (or P = pscale V at the
// start).
template <int S, typename T>
__global__ __launch_bounds__(RP_BS) void k_blkr_vp(long long m, T *__restrict__ V, T *__restrict__ P,
                                                   const BlkMats *__restrict__ mt, double pscale, const Ctl *c) {
    if (ctl_stopped(c)) return;
    const bool withp = (pscale == 0.0);
    RP_LOOP_BEGIN
    RowRaw<S, T> rv, rp;
    if (row < m) { rv.load(V + row * S); if (withp) rp.load(P + row * S); }
    for (; row < m; row += step) {
        double v[S], o[S], p[S];
        rv.get(v);
        if (withp) rp.get(p);
        const long long nr = row + step;
        if (nr < m) { rv.load(V + nr * S); if (withp) rp.load(P + nr * S); }
        row_mul<S, T, true>(v, mt->R12i, o);
        row_store<S>(V + row * S, o);
        if (withp) {
            double d[S];
            row_mul<S, T, false>(p, mt->betaT, d);
#pragma unroll
            for (int j = 0; j < S; j++) o[j] += d[j];
        } else {
#pragma unroll
            for (int j = 0; j < S; j++) o[j] *= pscale;
        }
        row_store<S>(P + row * S, o);
    }
}

// G = X^T Y (upper triangle, mirrored: X^T Y is symmetric where block CG uses it,
// P^T A P and V^T V).
template <int S, typename T>
__global__ __launch_bounds__(RP_BS) void k_blkr_gram(long long m, const T *__restrict__ X, const T *__restrict__ Y,
                                                     const Ctl *c, double *__restrict__ part) {
    if (c && ctl_stopped(c)) return;
    double g[S * (S + 1) / 2];
#pragma unroll
    for (int q = 0; q < S * (S + 1) / 2; q++) g[q] = 0.0;
    const bool same = (X == Y);
    RP_LOOP_BEGIN
    RowRaw<S, T> rx, ry;
    if (row < m) { rx.load(X + row * S); if (!same) ry.load(Y + row * S); }
    for (; row < m; row += step) {
        double x[S], y[S];
        rx.get(x);
        if (!same) ry.get(y);
        const long long nr = row + step;
        if (nr < m) { rx.load(X + nr * S); if (!same) ry.load(Y + nr * S); }
        if (same) row_gram<S>(x, x, g);
        else row_gram<S>(x, y, g);
    }
    row_gram_partial<S>(g, part);
}

// ------------------------------------ lane-pair row kernels (FP64, s = 8)
//
// Two lanes per row, each loading and storing one 32-byte half of it (a wave's
// accesses then cover 32 whole rows; a lane per 64-byte FP64 row strides its 16-byte
// loads across four times as many cache lines). The halves are swapped between the
// pair with DPP, both lanes form the full-row products (uniform matrix reads), each
// keeps its own columns, and each accumulates the Gram columns of its half against
// the whole row: G[i][j], j in its half, all i (the mirrored entries are the same
// products summed in the same order, so the partial is exactly symmetric).
__device__ __forceinline__ double pair_swap(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0xB1, 0xF, 0xF, false);   // quad_perm(1,0,3,2)
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0xB1, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

template <int S>
__device__ __forceinline__ void pair_full(const double (&mine)[S / 2], bool hi, double (&full)[S]) {
#pragma unroll
    for (int u = 0; u < S / 2; u++) {
        const double o = pair_swap(mine[u]);
        full[u] = hi ? o : mine[u];
        full[S / 2 + u] = hi ? mine[u] : o;
    }
}

template <int S>
__device__ __forceinline__ void pair_mul(const double (&r)[S], const double *__restrict__ M, bool hi, bool round32,
                                         double (&o)[S / 2]) {
    double f[S];
    row_mul<S, double, false>(r, M, f);
#pragma unroll
    for (int u = 0; u < S / 2; u++) {
        const double v = hi ? f[S / 2 + u] : f[u];
        o[u] = round32 ? (double)(float)v : v;
    }
}

template <int S>
__device__ __forceinline__ void pair_gram(const double (&full)[S], const double (&mine)[S / 2], double (&g)[S][S / 2]) {
#pragma unroll
    for (int i = 0; i < S; i++)
#pragma unroll
        for (int u = 0; u < S / 2; u++) g[i][u] = fma(full[i], mine[u], g[i][u]);
}

template <int S>
__device__ __forceinline__ void pair_gram_partial(double (&g)[S][S / 2], bool hi, double *part) {
    constexpr int NW = RP_BS / 64;
    __shared__ double sh[NW][S * S];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < S; i++)
#pragma unroll
        for (int u = 0; u < S / 2; u++) {
            double v = g[i][u];
#pragma unroll
            for (int o = 32; o >= 2; o >>= 1) v += __shfl_xor(v, o, 64);
            g[i][u] = v;
        }
    if (lane < 2) {
#pragma unroll
        for (int i = 0; i < S; i++)
#pragma unroll
            for (int u = 0; u < S / 2; u++) sh[w][i + ((hi ? S / 2 : 0) + u) * S] = g[i][u];
    }
    __syncthreads();
    const int t = threadIdx.x;
    if (t < S * S) {
        double v = sh[0][t];
        for (int k = 1; k < NW; k++) v += sh[k][t];
        part[(size_t)blockIdx.x * S * S + t] = v;
    }
}

#define RP2_LOOP_BEGIN                                                        \
    const long long step = (long long)gridDim.x * (RP_BS / 2);                \
    const bool hi = (threadIdx.x & 1) != 0;                                   \
    const int off = hi ? S / 2 : 0;                                           \
    long long row = (long long)blockIdx.x * (RP_BS / 2) + (threadIdx.x >> 1);

template <int S>
__global__ __launch_bounds__(RP_BS) void k_blkr2_xv(long long m, double *__restrict__ X, const double *__restrict__ P,
                                                    double *__restrict__ V, const double *__restrict__ AP,
                                                    const BlkMats *__restrict__ mt, const Ctl *c,
                                                    double *__restrict__ part) {
    if (ctl_stopped(c)) return;
    double g[S][S / 2];
#pragma unroll
    for (int i = 0; i < S; i++)
#pragma unroll
        for (int u = 0; u < S / 2; u++) g[i][u] = 0.0;
    RP2_LOOP_BEGIN
    RowRaw<S / 2, double> rp, rx, ra, rv;
    if (row < m) {
        rp.load(P + row * S + off); rx.load(X + row * S + off); ra.load(AP + row * S + off); rv.load(V + row * S + off);
    }
    for (; row < m; row += step) {
        double p[S / 2], x[S / 2], a[S / 2], v[S / 2], pf[S], af[S], vf[S], d[S / 2];
        rp.get(p); rx.get(x); ra.get(a); rv.get(v);
        const long long nr = row + step;
        if (nr < m) {
            rp.load(P + nr * S + off); rx.load(X + nr * S + off); ra.load(AP + nr * S + off); rv.load(V + nr * S + off);
        }
        pair_full<S>(p, hi, pf);
        pair_mul<S>(pf, mt->W, hi, false, d);
#pragma unroll
        for (int u = 0; u < S / 2; u++) x[u] += d[u];
        row_store<S / 2>(X + row * S + off, x);
        pair_full<S>(a, hi, af);
        pair_mul<S>(af, mt->alpha, hi, false, d);
#pragma unroll
        for (int u = 0; u < S / 2; u++) v[u] = v[u] - d[u];
        row_store<S / 2>(V + row * S + off, v);
        pair_full<S>(v, hi, vf);
        pair_gram<S>(vf, v, g);
    }
    pair_gram_partial<S>(g, hi, part);
}

// Pass 1 of CholeskyQR2 as k_blkr_vr (Gram of V R1^-1, nothing stored).
template <int S>
__global__ __launch_bounds__(RP_BS) void k_blkr2_vr(long long m, const double *__restrict__ V,
                                                    const BlkMats *__restrict__ mt, const Ctl *c,
                                                    double *__restrict__ part) {
    if (ctl_stopped(c)) return;
    double g[S][S / 2];
#pragma unroll
    for (int i = 0; i < S; i++)
#pragma unroll
        for (int u = 0; u < S / 2; u++) g[i][u] = 0.0;
    RP2_LOOP_BEGIN
    RowRaw<S / 2, double> rv;
    if (row < m) rv.load(V + row * S + off);
    for (; row < m; row += step) {
        double v[S / 2], vf[S], o[S / 2], of[S];
        rv.get(v);
        const long long nr = row + step;
        if (nr < m) rv.load(V + nr * S + off);
        pair_full<S>(v, hi, vf);
        pair_mul<S>(vf, mt->R1i, hi, false, o);
        pair_full<S>(o, hi, of);
        pair_gram<S>(of, o, g);
    }
    pair_gram_partial<S>(g, hi, part);
}

template <int S>
__global__ __launch_bounds__(RP_BS) void k_blkr2_vp(long long m, double *__restrict__ V, double *__restrict__ P,
                                                    const BlkMats *__restrict__ mt, double pscale, const Ctl *c) {
    if (ctl_stopped(c)) return;
    const bool withp = (pscale == 0.0);
    RP2_LOOP_BEGIN
    RowRaw<S / 2, double> rv, rp;
    if (row < m) { rv.load(V + row * S + off); if (withp) rp.load(P + row * S + off); }
    for (; row < m; row += step) {
        double v[S / 2], vf[S], o[S / 2], p[S / 2];
        rv.get(v);
        if (withp) rp.get(p);
        const long long nr = row + step;
        if (nr < m) { rv.load(V + nr * S + off); if (withp) rp.load(P + nr * S + off); }
        pair_full<S>(v, hi, vf);
        pair_mul<S>(vf, mt->R12i, hi, false, o);         // V R1^-1 R2^-1
        row_store<S / 2>(V + row * S + off, o);
        if (withp) {
            double pf[S], d[S / 2];
            pair_full<S>(p, hi, pf);
            pair_mul<S>(pf, mt->betaT, hi, false, d);
#pragma unroll
            for (int u = 0; u < S / 2; u++) o[u] += d[u];
        } else {
#pragma unroll
            for (int u = 0; u < S / 2; u++) o[u] *= pscale;
        }
        row_store<S / 2>(P + row * S + off, o);
    }
}

template <int S>
__global__ __launch_bounds__(RP_BS) void k_blkr2_gram(long long m, const double *__restrict__ X,
                                                      const double *__restrict__ Y, const Ctl *c,
                                                      double *__restrict__ part) {
    if (c && ctl_stopped(c)) return;
    double g[S][S / 2];
#pragma unroll
    for (int i = 0; i < S; i++)
#pragma unroll
        for (int u = 0; u < S / 2; u++) g[i][u] = 0.0;
    const bool same = (X == Y);
    RP2_LOOP_BEGIN
    RowRaw<S / 2, double> rx, ry;
    if (row < m) { rx.load(X + row * S + off); if (!same) ry.load(Y + row * S + off); }
    for (; row < m; row += step) {
        double x[S / 2], y[S / 2], xf[S];
        rx.get(x);
        if (!same) ry.get(y);
        const long long nr = row + step;
        if (nr < m) { rx.load(X + nr * S + off); if (!same) ry.load(Y + nr * S + off); }
        pair_full<S>(x, hi, xf);
        if (same) pair_gram<S>(xf, x, g);
        else pair_gram<S>(xf, y, g);
    }
    pair_gram_partial<S>(g, hi, part);
}

// Launchers: lane-pair kernels for FP64 s = 8, row kernels for other s <= 8,
// LDS-tile MFMA kernels above; each returns the number of partial rows it wrote
// (for launch_blk_fin).
template <typename T, int S>
constexpr bool panel_pair() { return S == 8 && sizeof(T) == sizeof(double); }

static inline int rp2_grid(long long m) {
    return (int)std::max(1LL, std::min((m + RP_BS / 2 - 1) / (RP_BS / 2), (long long)FIN_DIRECT));
}

template <int S, typename T>
static int panel_gram(hipStream_t s, long long m, const T *X, const T *Y, const Ctl *c, double *part) {
    if constexpr (panel_pair<T, S>()) {
        const int G = rp2_grid(m);
        hipLaunchKernelGGL((k_blkr2_gram<S>), dim3(G), dim3(RP_BS), 0, s, m, X, Y, c, part);
        return G;
    } else if constexpr (S <= 8) {
        const int G = rp_grid(m);
        hipLaunchKernelGGL((k_blkr_gram<S, T>), dim3(G), dim3(RP_BS), 0, s, m, X, Y, c, part);
        return G;
    } else {
        const int G = block_grid(m);
        hipLaunchKernelGGL((k_gram_st<S, T>), dim3(G), dim3(256), 0, s, m, X, Y, c, part);
        return G;
    }
}

template <int S, typename T>
static int panel_xv(hipStream_t s, long long m, T *X, const T *P, T *V, const T *AP, const BlkMats *mt,
                    const Ctl *c, double *part) {
    if constexpr (panel_pair<T, S>()) {
        const int G = rp2_grid(m);
        hipLaunchKernelGGL((k_blkr2_xv<S>), dim3(G), dim3(RP_BS), 0, s, m, X, P, V, AP, mt, c, part);
        return G;
    } else if constexpr (S <= 8) {
        const int G = rp_grid(m);
        hipLaunchKernelGGL((k_blkr_xv<S, T>), dim3(G), dim3(RP_BS), 0, s, m, X, P, V, AP, mt, c, part);
        return G;
    } else {
        const int G = block_grid(m);
        hipLaunchKernelGGL((k_blk_xv<S, T>), dim3(G), dim3(256), 0, s, m, X, P, V, AP, mt, c, part);
        return G;
    }
}

template <int S, typename T>
static int panel_vr(hipStream_t s, long long m, T *V, const BlkMats *mt, const Ctl *c, double *part) {
    if constexpr (panel_pair<T, S>()) {
        const int G = rp2_grid(m);
        hipLaunchKernelGGL((k_blkr2_vr<S>), dim3(G), dim3(RP_BS), 0, s, m, V, mt, c, part);
        return G;
    } else if constexpr (S <= 8) {
        const int G = rp_grid(m);
        hipLaunchKernelGGL((k_blkr_vr<S, T>), dim3(G), dim3(RP_BS), 0, s, m, V, mt, c, part);
        return G;
    } else {
        const int G = block_grid(m);
        hipLaunchKernelGGL((k_blk_vr<S, T>), dim3(G), dim3(256), 0, s, m, V, mt, c, part);
        return G;
    }
}

template <int S, typename T>
static void panel_vp(hipStream_t s, long long m, T *V, T *P, const BlkMats *mt, double pscale, const Ctl *c) {
    // no partials: up to 4 FIN_DIRECT workgroups
    if constexpr (panel_pair<T, S>()) {
        const int G = (int)std::max(1LL, std::min((m + RP_BS / 2 - 1) / (RP_BS / 2), 4LL * FIN_DIRECT));
        hipLaunchKernelGGL((k_blkr2_vp<S>), dim3(G), dim3(RP_BS), 0, s, m, V, P, mt, pscale, c);
    } else if constexpr (S <= 8) {
        const int G = (int)std::max(1LL, std::min((m + RP_BS - 1) / RP_BS, 4LL * FIN_DIRECT));
        hipLaunchKernelGGL((k_blkr_vp<S, T>), dim3(G), dim3(RP_BS), 0, s, m, V, P, mt, pscale, c);
    } else {
        hipLaunchKernelGGL((k_blk_vp<S, T>), dim3(block_grid(m)), dim3(256), 0, s, m, V, P, mt, pscale, c);
    }
}

// Y = alpha A X + beta Y: the SELL-64 kernel when the matrix has that copy in the
// values' precision, else CSR-stream. dia: the solvers' own products (finite X) may
// run on the slice-diagonal copy.
// AP = A P with the Gram P^T AP fused on MFMA (k_spmm_gram) when the matrix has a
// SELL-64 copy in the values' precision and the blocks allow 16-byte rows: returns the
// partial rows written (at most FIN_DIRECT), or 0 when the caller must run the SpMM
// and the Gram pass separately. BKH_GRAM_FUSE=0 in the environment keeps them apart.
template <int S, typename VT, typename T>
static int launch_spmm_gram(hipStream_t s, const bkh_matrix *A, const VT *val, const T *P, T *AP, const Ctl *c,
                            double *part) {
    constexpr int LPR = spmm_lpr<S, T>();
    if constexpr (LPR == 0 || S > 16) {
        return 0;
    } else {
        static const bool off = [] { const char *e = getenv("BKH_GRAM_FUSE"); return e && e[0] == '0'; }();
        if (off || A->nsl == 0 || A->variant != SPMV_SELL || (((uintptr_t)P | (uintptr_t)AP) & 15)) return 0;
        // FP64 rows of s <= 8: eight entries per row in flight (88 registers, one 16-wave
        // workgroup per CU, 256 workgroups): 120.5 us per s = 8 launch at 128^3 against
        // 129.3 us for four entries at 8 waves per SIMD (512 workgroups; 1024: 129.2),
        // 154-157 us with two slices per thread per pass. Otherwise four entries.
        constexpr bool WIDE = sizeof(T) == 8 && S <= 8;
        constexpr int BSG = 1024, UNR = 1, KB = WIDE ? 8 : 4, MINW = WIDE ? 4 : (S <= 8 ? 8 : 4);
        constexpr int SPB = UNR * BSG / (64 * LPR);                  // slices per pass
        const int npass = (A->nsl + SPB - 1) / SPB;
        const int G0 = std::max(1, std::min(WIDE ? 256 : S <= 8 ? FIN_DIRECT_WIDE : FIN_DIRECT, npass));
        const int chunk = ((npass + G0 - 1) / G0) * SPB;           // slices per workgroup
        const int G = (A->nsl + chunk - 1) / chunk;
        const VT *dv = nullptr, *sv = nullptr;
        if constexpr (sizeof(VT) == sizeof(float)) {
            if (val == (const VT *)A->val32) { dv = (const VT *)A->dval32; sv = (const VT *)A->sval32; }
        } else {
            if (val == (const VT *)A->val) { dv = (const VT *)A->dval; sv = (const VT *)A->sval; }
        }
        // the pipelined slice-diagonal kernel (values in the blocks' precision, every slice
        // with <= 8 positions and an offset 0, byte offsets of the block within 32 bits);
        // 256 workgroups of 1024 threads at every precision
        if constexpr (std::is_same<VT, T>::value) {
            static const bool pipe_off = [] { const char *e = getenv("BKH_DIA_PIPE"); return e && e[0] == '0'; }();
            if (!pipe_off && A->dia_entries > 0 && dv && A->dmeta && A->dia_maxpos <= 8 &&
                (unsigned long long)A->n * S * sizeof(T) < (1ULL << 32)) {
                constexpr int BSP = 1024, SPBP = BSP / 64 / LPR;
                const int np = (A->nsl + SPBP - 1) / SPBP;
                const int g0 = std::max(1, std::min(256, np));
                const int ch = ((np + g0 - 1) / g0) * SPBP;
                const int gp = (A->nsl + ch - 1) / ch;
                if (A->dia_maxpos <= 7)
                    hipLaunchKernelGGL((k_spmm_gram_dia<S, T, BSP, 7, 4>), dim3(gp), dim3(BSP), 0, s, (int)A->n,
                                       A->nsl, ch, (const DiaSlice *)A->dmeta, (const T *)dv, P, AP, c, part);
                else
                    hipLaunchKernelGGL((k_spmm_gram_dia<S, T, BSP, 8, 4>), dim3(gp), dim3(BSP), 0, s, (int)A->n,
                                       A->nsl, ch, (const DiaSlice *)A->dmeta, (const T *)dv, P, AP, c, part);
                return gp;
            }
        }
        if (A->dia_entries > 0 && dv)
            hipLaunchKernelGGL((k_spmm_gram<S, VT, T, int, true, BSG, UNR, KB, MINW>), dim3(G), dim3(BSG), 0, s, (int)A->n, A->nsl,
                               chunk, A->rp, A->dso, dv, (const int *)A->doff, P, AP, c, part);
        else if (sv && A->scd)
            hipLaunchKernelGGL((k_spmm_gram<S, VT, T, short, false, BSG, UNR, KB, MINW>), dim3(G), dim3(BSG), 0, s, (int)A->n,
                               A->nsl, chunk, A->rp, A->sso, sv, (const short *)A->scd, P, AP, c, part);
        else if (sv && A->scol)
            hipLaunchKernelGGL((k_spmm_gram<S, VT, T, int, false, BSG, UNR, KB, MINW>), dim3(G), dim3(BSG), 0, s, (int)A->n,
                               A->nsl, chunk, A->rp, A->sso, sv, (const int *)A->scol, P, AP, c, part);
        else
            return 0;
        return G;
    }
}

// Block CG on row-major device blocks: Bt (m x s), Xt (in/out).
template <int S, typename VT, typename T>
static int blcg_core(ThreadCtx *t, const bkh_matrix *A, const VT *val, const T *Bt, T *Xt, T *V,
                     T *P, T *AP, BlkMats *mt, Ctl *c, double *part, double tol, int max_it,
                     int p0_mode, int K) {
    hipStream_t s = t->stream;
    const long long m = A->n;
    const int Gv = vec_grid(m * S, 256);
    if (ctl_init(t, c, tol, max_it)) return BKH_DEVICE_ERROR;
    BlkMats h;
    memset(&h, 0, sizeof h);
    h.m = (int)m;
    h.s = S;
    HIP_OK(hipMemcpyAsync(mt, &h, sizeof h, hipMemcpyHostToDevice, s));
    // R0 = B - A X0 (blcg.c:54-57), V, C = qr(R0) (:58), fnb = ||B||_F (:60)
    launch_spmm<S>(s, A, val, (const T *)Xt, AP, 1.0, 0.0, c, true);
    hipLaunchKernelGGL(k_blk_resid<T>, dim3(Gv), dim3(256), 0, s, m * S, Bt, (const T *)AP, V, part);
    launch_blk_fin<S, FIN_FNB>(s, part, Gv, mt, c);
    launch_blk_fin<S, FIN_QR1>(s, part, panel_gram<S, T>(s, m, V, V, c, part), mt, c);
    launch_blk_fin<S, FIN_QR2_INIT>(s, part, panel_vr<S, T>(s, m, V, mt, c, part), mt, c);
    // P = V (then P = V + P beta^T with beta = I: P0 = 2V, blcg.c:59-71)
    panel_vp<S, T>(s, m, V, P, mt, p0_mode ? 1.0 : 2.0, c);
    LAUNCH_OK();
    hipGraphExec_t ge = capture_cached(t, {0xB1C6, (uintptr_t)A, (uintptr_t)val, (uintptr_t)Bt, (uintptr_t)Xt, (uintptr_t)V,
                                           (uintptr_t)P, (uintptr_t)AP, (uintptr_t)mt, (uintptr_t)c, (uintptr_t)part,
                                           (uintptr_t)S, (uintptr_t)sizeof(T)}, s, K, [&]() {
        int ng = launch_spmm_gram<S>(s, A, val, (const T *)P, AP, c, part);
        if (!ng) {
            launch_spmm<S>(s, A, val, (const T *)P, AP, 1.0, 0.0, c, true);
            ng = panel_gram<S, T>(s, m, P, AP, c, part);
        }
        launch_blk_fin<S, FIN_ALPHA>(s, part, ng, mt, c);
        launch_blk_fin<S, FIN_QR1>(s, part, panel_xv<S, T>(s, m, Xt, P, V, AP, mt, c, part), mt, c);
        launch_blk_fin<S, FIN_QR2_ITER>(s, part, panel_vr<S, T>(s, m, V, mt, c, part), mt, c);
        panel_vp<S, T>(s, m, V, P, mt, 0.0, c);
    });
    if (!ge) { bkh_set_error("hipGraph capture of the block CG iteration failed"); return BKH_DEVICE_ERROR; }
    int rc = run_graph_until_done(t, ge, c, max_it / K + 2);
    return rc;
}

// Frobenius true residual ||B - A X||_F / ||B||_F of row-major blocks.
template <int S>
static int block_true_res(ThreadCtx *t, const bkh_matrix *A, const double *Bt, const double *Xt,
                          double *tmp, double *part, Ctl *c, double *out) {
    hipStream_t s = t->stream;
    const long long m = A->n;
    const int Gv = vec_grid(m * S, 256);
    if (ctl_init(t, c, 0.0, 1)) return BKH_DEVICE_ERROR;
    launch_spmm<S>(s, A, A->val, Xt, tmp, 1.0, 0.0, (const Ctl *)nullptr, true);
    hipLaunchKernelGGL(k_blk_resid<double>, dim3(Gv), dim3(256), 0, s, m * S, Bt, (const double *)tmp, tmp, part);
    // ||R||^2 from a second pass (tmp now holds R)
    hipLaunchKernelGGL(k_blk_resid<double>, dim3(Gv), dim3(256), 0, s, m * S, (const double *)tmp,
                       (const double *)tmp, tmp + 0, part + Gv);
    LAUNCH_OK();
    std::vector<double> h((size_t)Gv * 2);
    HIP_OK(hipMemcpyAsync(h.data(), part, sizeof(double) * Gv * 2, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    double bb = 0.0, rr = 0.0;
    for (int i = 0; i < Gv; i++) { bb += h[i]; rr += h[Gv + i]; }
    *out = sqrt(rr) / sqrt(bb);
    return 0;
}

// R = B - AX in FP64 (AX holds A X on entry, R may alias it), its FP32 copy, and
// partials of |R|_F^2 and |B|_F^2.
__global__ __launch_bounds__(256) void k_blk_resid32(long long ms, const double *B, const double *AX, double *R,
                                                     float *R32, double *part) {
    double v[2] = {0.0, 0.0};
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < ms; e += (long long)gridDim.x * 256) {
        const double r = B[e] - AX[e];
        R[e] = r;
        R32[e] = (float)r;
        v[0] += r * r;
        v[1] += B[e] * B[e];
    }
    block_partial<2, 256>(v, part);
}

__global__ __launch_bounds__(256) void k_blk_addx(long long ms, double *X, const float *D) {
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < ms; e += (long long)gridDim.x * 256)
        X[e] += (double)D[e];
}

struct BlkResFin {     // s[0] = |R|_F / |B|_F, s[1] = |R|_F, s[2] = |B|_F
    __device__ void operator()(Ctl *c, const double *tot) const {
        c->s[1] = sqrt(tot[0]);
        c->s[2] = sqrt(tot[1]);
        c->s[0] = c->s[2] > 0.0 ? c->s[1] / c->s[2] : 0.0;
    }
};

// BKH_FP32_IR block CG: restarted FP64 iterative refinement around the FP32 block CG
// (FP32 values and blocks; FP64 Gram accumulation, small s x s algebra and
// convergence test): R = B - A X (FP64), solve A D = R to a relative inner
// tolerance with block CG in FP32, X += D (FP64), until |B - A X|_F / |B|_F < tol.
// Each pass starts a new block Krylov space from R with the same first-direction
// rule as the FP64 path (bkh_opts.fix_blcg_p0).
template <int S>
static int blcg_ir(ThreadCtx *t, bkh_matrix *A, const double *B, double *X, double tol, int max_iter,
                   const bkh_opts &o, bkh_info *inf) {
    if (!A->val32) { bkh_set_error("FP32_IR needs a matrix created with want_fp32"); return BKH_BAD_ARGUMENT; }
    const long long m = A->n, ms = m * S;
    const int G = block_grid(m), Gv = vec_grid(ms, 256);
    const int K = o.graph_iters > 0 ? o.graph_iters : std::max(2, default_graph_iters(m) / 4);
    const int nparts = std::max(G * S * S, 2 * Gv);
    size_t bytes = 2 * carve_size(sizeof(Ctl)) + carve_size(sizeof(BlkMats)) + carve_size(sizeof(double) * nparts) +
                   3 * carve_size(sizeof(double) * ms) + 5 * carve_size(sizeof(float) * ms);
    void *ws = bkh_workspace(bytes);
    if (!ws) return BKH_DEVICE_ERROR;
    Carve cv(ws);
    Ctl *c = cv.take<Ctl>(1), *co = cv.take<Ctl>(1);
    BlkMats *mt = cv.take<BlkMats>(1);
    double *part = cv.take<double>(nparts);
    double *Bt = cv.take<double>(ms), *Xt = cv.take<double>(ms), *R = cv.take<double>(ms);
    float *R32 = cv.take<float>(ms), *D32 = cv.take<float>(ms), *V32 = cv.take<float>(ms), *P32 = cv.take<float>(ms),
          *AP32 = cv.take<float>(ms);
    hipStream_t s = t->stream;
    HIP_OK(hipEventRecord(t->ev0, s));
    hipLaunchKernelGGL((k_cm2rm<double, double>), dim3(Gv), dim3(256), 0, s, m, S, B, Bt);
    hipLaunchKernelGGL((k_cm2rm<double, double>), dim3(Gv), dim3(256), 0, s, m, S, (const double *)X, Xt);
    LAUNCH_OK();
    const int max_inner = o.max_inner > 0 ? o.max_inner : max_iter;
    int total = 0, passes = 0, rc = 0, st = ST_MAXIT;
    double rel = 0.0;
    for (;;) {
        if (ctl_init(t, co, 0.0, 1)) return BKH_DEVICE_ERROR;
        launch_spmm<S>(s, A, A->val, (const double *)Xt, R, 1.0, 0.0, (const Ctl *)nullptr, true);
        hipLaunchKernelGGL(k_blk_resid32, dim3(Gv), dim3(256), 0, s, ms, (const double *)Bt, (const double *)R, R,
                           R32, part);
        launch_finalize<2>(s, part, Gv, co, BlkResFin());
        LAUNCH_OK();
        if (ctl_fetch(t, co)) return BKH_DEVICE_ERROR;
        rel = t->ctl_host->s[0];
        if (rel < tol) { st = ST_CONVERGED; break; }
        if (!(rel == rel) || rel > 1e300) { st = ST_BREAKDOWN; break; }
        if (total >= max_iter) { st = ST_MAXIT; break; }
        double itol = o.inner_tol > 0 ? o.inner_tol : 1e-5;
        itol = std::max(itol, 0.5 * tol / rel);
        hipLaunchKernelGGL((k_sh_fill_f), dim3(Gv), dim3(256), 0, s, D32, ms);
        LAUNCH_OK();
        rc = blcg_core<S, float, float>(t, A, A->val32, (const float *)R32, D32, V32, P32, AP32, mt, c, part, itol,
                                        std::min(max_inner, max_iter - total), o.fix_blcg_p0, K);
        if (rc) return rc;
        const Ctl hc = *t->ctl_host;
        total += hc.it;
        passes++;
        if (hc.status == ST_BREAKDOWN) { st = ST_BREAKDOWN; break; }
        hipLaunchKernelGGL(k_blk_addx, dim3(Gv), dim3(256), 0, s, ms, Xt, (const float *)D32);
        LAUNCH_OK();
    }
    hipLaunchKernelGGL((k_rm2cm<double, double>), dim3(Gv), dim3(256), 0, s, m, S, (const double *)Xt, X);
    LAUNCH_OK();
    HIP_OK(hipEventRecord(t->ev1, s));
    HIP_OK(hipEventSynchronize(t->ev1));
    float msec = 0.f;
    HIP_OK(hipEventElapsedTime(&msec, t->ev0, t->ev1));
    inf->time_ms = msec;
    bkh_note_solve_ms(msec);
    inf->status = st;
    inf->inner_iters = total;
    inf->passes = passes;
    inf->iters = total > 0 ? total - 1 : 0;
    inf->rec_res = rel;
    inf->true_res = rel;
    inf->ret = st == ST_CONVERGED ? inf->iters : (st == ST_BREAKDOWN ? -2 : -1);
    return inf->ret;
}

__global__ void k_blk_setfnb(Ctl *c, const Ctl *co) {
    if (threadIdx.x == 0) c->fnb = co->s[2];
}

__global__ void k_blk_resume(Ctl *c, double target) {
    if (threadIdx.x == 0) { c->done = 0; c->status = ST_RUNNING; c->tol = target; }
}

__global__ __launch_bounds__(256) void k_blk_fold(long long ms, double *X, float *X32) {
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < ms; e += (long long)gridDim.x * 256) {
        X[e] += (double)X32[e];
        X32[e] = 0.0f;
    }
}

// BKH_FP32_IR block CG with reliable updates (the block form of the CG driver): the
// FP32 block CG runs on a correction X32 (FP64 Gram accumulation and s x s algebra,
// its recursive residual measured against |B|_F); whenever |C|_F has fallen by a
// factor delta, the host folds X += X32 and replaces the residual basis by the QR of
// the FP64 true residual B - A X (V, C), then continues with the SAME directions P
// (P = V + P beta^T with the last step's beta), so the block Krylov information is
// kept across replacements.
template <int S>
static int blcg_ru(ThreadCtx *t, bkh_matrix *A, const double *B, double *X, double tol, int max_iter,
                   const bkh_opts &o, bkh_info *inf) {
    if (!A->val32) { bkh_set_error("FP32_IR needs a matrix created with want_fp32"); return BKH_BAD_ARGUMENT; }
    const long long m = A->n, ms = m * S;
    const int G = block_grid(m), Gv = vec_grid(ms, 256);
    const int K = o.graph_iters > 0 ? o.graph_iters : std::max(2, default_graph_iters(m) / 4);
    const int nparts = std::max(G * S * S, 2 * Gv);
    const double delta = o.inner_tol > 0 ? o.inner_tol : 0.01;
    size_t bytes = 2 * carve_size(sizeof(Ctl)) + carve_size(sizeof(BlkMats)) + carve_size(sizeof(double) * nparts) +
                   3 * carve_size(sizeof(double) * ms) + 4 * carve_size(sizeof(float) * ms);
    void *ws = bkh_workspace(bytes);
    if (!ws) return BKH_DEVICE_ERROR;
    Carve cv(ws);
    Ctl *c = cv.take<Ctl>(1), *co = cv.take<Ctl>(1);
    BlkMats *mt = cv.take<BlkMats>(1);
    double *part = cv.take<double>(nparts);
    double *Bt = cv.take<double>(ms), *Xt = cv.take<double>(ms), *R = cv.take<double>(ms);
    float *X32 = cv.take<float>(ms), *V32 = cv.take<float>(ms), *P32 = cv.take<float>(ms), *AP32 = cv.take<float>(ms);
    hipStream_t s = t->stream;
    const float *val = A->val32;
    HIP_OK(hipEventRecord(t->ev0, s));
    hipLaunchKernelGGL((k_cm2rm<double, double>), dim3(Gv), dim3(256), 0, s, m, S, B, Bt);
    hipLaunchKernelGGL((k_cm2rm<double, double>), dim3(Gv), dim3(256), 0, s, m, S, (const double *)X, Xt);
    hipLaunchKernelGGL((k_sh_fill_f), dim3(Gv), dim3(256), 0, s, X32, ms);
    BlkMats h;
    memset(&h, 0, sizeof h);
    h.m = (int)m;
    h.s = S;
    HIP_OK(hipMemcpyAsync(mt, &h, sizeof h, hipMemcpyHostToDevice, s));
    if (ctl_init(t, c, 1.0, max_iter)) return BKH_DEVICE_ERROR;
    LAUNCH_OK();
    hipGraphExec_t ge = capture_cached(t, {0xB1C7, (uintptr_t)A, (uintptr_t)X32, (uintptr_t)V32, (uintptr_t)P32,
                                           (uintptr_t)AP32, (uintptr_t)mt, (uintptr_t)c, (uintptr_t)part,
                                           (uintptr_t)S}, s, K, [&]() {
        int ng = launch_spmm_gram<S>(s, A, val, (const float *)P32, AP32, c, part);
        if (!ng) {
            launch_spmm<S>(s, A, val, (const float *)P32, AP32, 1.0, 0.0, c, true);
            ng = panel_gram<S, float>(s, m, P32, AP32, c, part);
        }
        launch_blk_fin<S, FIN_ALPHA>(s, part, ng, mt, c);
        launch_blk_fin<S, FIN_QR1>(s, part, panel_xv<S, float>(s, m, X32, P32, V32, AP32, mt, c, part), mt, c);
        launch_blk_fin<S, FIN_QR2_ITER>(s, part, panel_vr<S, float>(s, m, V32, mt, c, part), mt, c);
        panel_vp<S, float>(s, m, V32, P32, mt, 0.0, c);
    });
    if (!ge) { bkh_set_error("hipGraph capture of the FP32 block CG iteration failed"); return BKH_DEVICE_ERROR; }
    int passes = 0, st = ST_MAXIT, used = 0;
    double rel = 0.0;
    for (;;) {
        // X += X32 (FP64); R = B - A X in FP64, its FP32 copy straight into V32
        if (passes > 0) hipLaunchKernelGGL(k_blk_fold, dim3(Gv), dim3(256), 0, s, ms, Xt, X32);
        if (ctl_init(t, co, 0.0, 1)) return BKH_DEVICE_ERROR;
        launch_spmm<S>(s, A, A->val, (const double *)Xt, R, 1.0, 0.0, (const Ctl *)nullptr, true);
        hipLaunchKernelGGL(k_blk_resid32, dim3(Gv), dim3(256), 0, s, ms, (const double *)Bt, (const double *)R, R,
                           V32, part);
        launch_finalize<2>(s, part, Gv, co, BlkResFin());
        LAUNCH_OK();
        if (ctl_fetch(t, co)) return BKH_DEVICE_ERROR;
        rel = t->ctl_host->s[0];
        if (rel < tol) { st = ST_CONVERGED; break; }
        if (!(rel == rel) || rel > 1e300) { st = ST_BREAKDOWN; break; }
        if (passes > 0) {
            if (ctl_fetch(t, c)) return BKH_DEVICE_ERROR;
            used = t->ctl_host->it;
            if (t->ctl_host->status == ST_BREAKDOWN) { st = ST_BREAKDOWN; break; }
            if (used >= max_iter) { st = ST_MAXIT; break; }
        }
        // the new residual basis: V, C = qr(R); then P = V + P beta^T (the first pass:
        // P = V or 2V as the FP64 path's first direction)
        hipLaunchKernelGGL(k_blk_resume, dim3(1), dim3(64), 0, s, c, std::max(0.5 * tol, delta * rel));
        if (passes == 0) hipLaunchKernelGGL(k_blk_setfnb, dim3(1), dim3(64), 0, s, c, (const Ctl *)co);
        launch_blk_fin<S, FIN_QR1>(s, part, panel_gram<S, float>(s, m, V32, V32, c, part), mt, c);
        launch_blk_fin<S, FIN_QR2_INIT>(s, part, panel_vr<S, float>(s, m, V32, mt, c, part), mt, c);
        panel_vp<S, float>(s, m, V32, P32, mt, passes == 0 ? (o.fix_blcg_p0 ? 1.0 : 2.0) : 0.0, c);
        LAUNCH_OK();
        if (run_graph_until_done(t, ge, c, (max_iter - used) / K + 2)) return BKH_DEVICE_ERROR;
        passes++;
    }
    if (ctl_fetch(t, c)) return BKH_DEVICE_ERROR;
    const int total = t->ctl_host->it;
    hipLaunchKernelGGL((k_rm2cm<double, double>), dim3(Gv), dim3(256), 0, s, m, S, (const double *)Xt, X);
    LAUNCH_OK();
    HIP_OK(hipEventRecord(t->ev1, s));
    HIP_OK(hipEventSynchronize(t->ev1));
    float msec = 0.f;
    HIP_OK(hipEventElapsedTime(&msec, t->ev0, t->ev1));
    inf->time_ms = msec;
    bkh_note_solve_ms(msec);
    inf->status = st;
    inf->inner_iters = total;
    inf->passes = passes;
    inf->iters = total > 0 ? total - 1 : 0;
    inf->rec_res = rel;
    inf->true_res = rel;
    inf->ret = st == ST_CONVERGED ? inf->iters : (st == ST_BREAKDOWN ? -2 : -1);
    return inf->ret;
}

template <int S>
static int blcg_dispatch(ThreadCtx *t, bkh_matrix *A, const double *B, double *X, double tol,
                         int max_iter, const bkh_opts &o, bkh_info *inf, bool want_true) {
    const long long m = A->n;
    const int G = block_grid(m);
    const int Gv = vec_grid(m * S, 256);
    const int K = o.graph_iters > 0 ? o.graph_iters : std::max(2, default_graph_iters(m) / 4);
    const int nparts = std::max(G * S * S, 2 * Gv);
    size_t bytes = carve_size(sizeof(Ctl)) + carve_size(sizeof(BlkMats)) + carve_size(sizeof(double) * nparts) +
                   5 * carve_size(sizeof(double) * m * S);
    void *ws = bkh_workspace(bytes);
    if (!ws) return BKH_DEVICE_ERROR;
    Carve cv(ws);
    Ctl *c = cv.take<Ctl>(1);
    BlkMats *mt = cv.take<BlkMats>(1);
    double *part = cv.take<double>(nparts);
    double *Bt = cv.take<double>(m * S), *Xt = cv.take<double>(m * S), *V = cv.take<double>(m * S),
           *P = cv.take<double>(m * S), *AP = cv.take<double>(m * S);
    hipStream_t s = t->stream;
    const int Gt = vec_grid(m * S, 256);
    HIP_OK(hipEventRecord(t->ev0, s));
    hipLaunchKernelGGL((k_cm2rm<double, double>), dim3(Gt), dim3(256), 0, s, m, S, B, Bt);
    hipLaunchKernelGGL((k_cm2rm<double, double>), dim3(Gt), dim3(256), 0, s, m, S, (const double *)X, Xt);
    LAUNCH_OK();
    int rc = blcg_core<S, double, double>(t, A, A->val, Bt, Xt, V, P, AP, mt, c, part, tol, max_iter,
                                          o.fix_blcg_p0, K);
    if (rc) return rc;
    Ctl hc = *t->ctl_host;
    hipLaunchKernelGGL((k_rm2cm<double, double>), dim3(Gt), dim3(256), 0, s, m, S, (const double *)Xt, X);
    LAUNCH_OK();
    HIP_OK(hipEventRecord(t->ev1, s));
    HIP_OK(hipEventSynchronize(t->ev1));
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, t->ev0, t->ev1));
    inf->time_ms = ms;
    bkh_note_solve_ms(ms);
    inf->ret = ctl_result(&hc);
    inf->status = hc.status;
    inf->iters = hc.result_it;
    inf->rec_res = hc.err;
    if (want_true) {
        double tr = 0.0;
        if (block_true_res<S>(t, A, Bt, Xt, V, part, c, &tr) == 0) inf->true_res = tr;
    }
    return inf->ret;
}

extern "C" int bkh_blcg(bkh_matrix *A, int s, const double *B, double *X, double tol, int max_iter,
                        const bkh_opts *opts, bkh_info *info) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (!A || !B || !X) return BKH_BAD_ARGUMENT;
    if (s < 1 || s > SM_MAX) { bkh_set_error("block CG supports 1 <= s <= %d", SM_MAX); return -3; }
    bkh_opts o = opts ? *opts : bkh_opts{};
    bkh_info inf;
    memset(&inf, 0, sizeof inf);
    if (max_iter <= 0) {                 // blcg.c:31,115-118: error = 0 -> returns 0
        inf.status = ST_CONVERGED;
        if (info) *info = inf;
        return 0;
    }
    int ret;
    if (o.precision == BKH_FP32_IR) {
        switch (s) {
#define CASE(S_) case S_: ret = (o.path == 1) ? blcg_ir<S_>(t, A, B, X, tol, max_iter, o, &inf) \
                                             : blcg_ru<S_>(t, A, B, X, tol, max_iter, o, &inf); break;
            CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
            CASE(9) CASE(10) CASE(11) CASE(12) CASE(13) CASE(14) CASE(15) CASE(16)
#undef CASE
        default: return -3;
        }
        if (ret < -2) return ret;
        if (info) *info = inf;
        return ret;
    }
    switch (s) {
#define CASE(S_) case S_: ret = blcg_dispatch<S_>(t, A, B, X, tol, max_iter, o, &inf, info != nullptr); break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
        CASE(9) CASE(10) CASE(11) CASE(12) CASE(13) CASE(14) CASE(15) CASE(16)
#undef CASE
    default: return -3;
    }
    if (info) *info = inf;
    return ret;
}

// Thin QR of an m x s column-major device panel (CholeskyQR2): V <- Q, R (host,
// s x s column-major, upper, positive diagonal). Replaces src/myqr.c:4-39, whose
// LAPACK Householder R may differ from this R by the signs of its rows.
extern "C" int bkh_qr(int m, int s, double *V, double *R) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (!V || !R || m < 1) return BKH_BAD_ARGUMENT;
    if (s < 1 || s > SM_MAX) { bkh_set_error("QR supports 1 <= s <= %d", SM_MAX); return -3; }
    const int G = block_grid(m);
    size_t bytes = carve_size(sizeof(Ctl)) + carve_size(sizeof(BlkMats)) +
                   carve_size(sizeof(double) * G * SM_N) + 2 * carve_size(sizeof(double) * (size_t)m * s);
    void *ws = bkh_workspace(bytes);
    if (!ws) return BKH_DEVICE_ERROR;
    Carve cv(ws);
    Ctl *c = cv.take<Ctl>(1);
    BlkMats *mt = cv.take<BlkMats>(1);
    double *part = cv.take<double>((size_t)G * SM_N);
    double *Vt = cv.take<double>((size_t)m * s), *Pt = cv.take<double>((size_t)m * s);
    hipStream_t st = t->stream;
    if (ctl_init(t, c, 0.0, 1)) return BKH_DEVICE_ERROR;
    BlkMats h;
    memset(&h, 0, sizeof h);
    h.m = m;
    h.s = s;
    HIP_OK(hipMemcpyAsync(mt, &h, sizeof h, hipMemcpyHostToDevice, st));
    const int Gt = vec_grid((long long)m * s, 256);
    hipLaunchKernelGGL((k_cm2rm<double, double>), dim3(Gt), dim3(256), 0, st, (long long)m, s, (const double *)V, Vt);
    switch (s) {
#define CASE(S_)                                                                                            \
    case S_:                                                                                                \
        hipLaunchKernelGGL((k_gram<S_, double>), dim3(G), dim3(256), 0, st, (long long)m, (const double *)Vt, \
                           (const double *)Vt, c, part);                                                     \
        launch_blk_fin<S_, FIN_QR1>(st, part, G, mt, c);             \
        hipLaunchKernelGGL((k_blk_vr<S_, double>), dim3(G), dim3(256), 0, st, (long long)m, Vt, mt, c, part); \
        launch_blk_fin<S_, FIN_QR2_INIT>(st, part, G, mt, c);        \
        hipLaunchKernelGGL((k_blk_vp<S_, double>), dim3(G), dim3(256), 0, st, (long long)m, Vt, Pt, mt, 1.0, c); \
        break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
        CASE(9) CASE(10) CASE(11) CASE(12) CASE(13) CASE(14) CASE(15) CASE(16)
#undef CASE
    }
    hipLaunchKernelGGL((k_rm2cm<double, double>), dim3(Gt), dim3(256), 0, st, (long long)m, s, (const double *)Vt, V);
    LAUNCH_OK();
    if (ctl_fetch(t, c)) return BKH_DEVICE_ERROR;
    if (t->ctl_host->status == ST_BREAKDOWN) { bkh_set_error("QR: rank-deficient panel"); return -2; }
    HIP_OK(hipMemcpyAsync(R, mt->C, sizeof(double) * s * s, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    return 0;
}

// Y = alpha A X + beta Y for s column-major columns (the mkl_dcsrmm dispatch of
// include/bksp_internal.h:63-79): s <= 16 through the row-major SpMM kernel, wider
// blocks column by column through the SpMV.
static_assert(SM_MAX == 16, "bkh_matrix::spmm_w holds widths for s = 1..16");

// Y = alpha A X + beta Y over s column-major columns, w columns per pass: each chunk is
// turned row-major, multiplied by the S = w SpMM kernel (one sweep over A per chunk) and
// turned back. A column's products do not depend on w (every row sums in stored order).
int spmm_chunked(ThreadCtx *t, const bkh_matrix *A, int s, int w, double alpha, const double *X, double beta,
                 double *Y) {
    hipStream_t st = t->stream;
    const long long m = A->n;
    w = std::max(1, std::min(w, std::min(s, SM_MAX)));
    void *ws = bkh_workspace(2 * carve_size(sizeof(double) * m * w));
    if (!ws) return BKH_DEVICE_ERROR;
    Carve cv(ws);
    double *Xt = cv.take<double>(m * w), *Yt = cv.take<double>(m * w);
    for (int j = 0; j < s; j += w) {
        const int wj = std::min(w, s - j);
        const double *Xj = X + (size_t)j * m;
        double *Yj = Y + (size_t)j * m;
        const int Gt = vec_grid(m * wj, 256);
        hipLaunchKernelGGL((k_cm2rm<double, double>), dim3(Gt), dim3(256), 0, st, m, wj, Xj, Xt);
        if (beta != 0.0)
            hipLaunchKernelGGL((k_cm2rm<double, double>), dim3(Gt), dim3(256), 0, st, m, wj, (const double *)Yj, Yt);
        switch (wj) {
#define CASE(S_) case S_: launch_spmm<S_>(st, A, A->val, (const double *)Xt, Yt, alpha, beta, (const Ctl *)nullptr); break;
            CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
            CASE(9) CASE(10) CASE(11) CASE(12) CASE(13) CASE(14) CASE(15) CASE(16)
#undef CASE
        }
        hipLaunchKernelGGL((k_rm2cm<double, double>), dim3(Gt), dim3(256), 0, st, m, wj, (const double *)Yt, Yj);
    }
    LAUNCH_OK();
    return 0;
}

extern "C" int bkh_spmm(const bkh_matrix *A, int s, double alpha, const double *X, double beta, double *Y) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (!A || !X || !Y || s < 1) return BKH_BAD_ARGUMENT;
    // columns per pass: the autotuner's choice for this s (for s > 16, for 16), else all
    // of them up to 16
    const int tuned = A->spmm_w[std::min(s, SM_MAX)];
    const int rc = spmm_chunked(t, A, s, tuned > 0 ? tuned : SM_MAX, alpha, X, beta, Y);
    if (rc) return rc;
    HIP_OK(hipStreamSynchronize(t->stream));
    return 0;
}

// G = X^T Y (s x s, column-major, to host) for m x s column-major X, Y, on FP64 MFMA;
// partials summed in a fixed order.
extern "C" int bkh_gram(int m, int s, const double *X, const double *Y, double *G) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (!X || !Y || !G || m < 1) return BKH_BAD_ARGUMENT;
    if (s < 1 || s > SM_MAX) { bkh_set_error("Gram supports 1 <= s <= %d", SM_MAX); return -3; }
    hipStream_t st = t->stream;
    const int Gb = block_grid(m), Gt = vec_grid((long long)m * s, 256);
    void *ws = bkh_workspace(2 * carve_size(sizeof(double) * (size_t)m * s) + carve_size(sizeof(double) * Gb * SM_N));
    if (!ws) return BKH_DEVICE_ERROR;
    Carve cv(ws);
    double *Xt = cv.take<double>((size_t)m * s), *Yt = cv.take<double>((size_t)m * s);
    double *part = cv.take<double>((size_t)Gb * SM_N);
    hipLaunchKernelGGL((k_cm2rm<double, double>), dim3(Gt), dim3(256), 0, st, (long long)m, s, X, Xt);
    hipLaunchKernelGGL((k_cm2rm<double, double>), dim3(Gt), dim3(256), 0, st, (long long)m, s, Y, Yt);
    switch (s) {
#define CASE(S_) case S_: hipLaunchKernelGGL((k_gram<S_, double>), dim3(Gb), dim3(256), 0, st, (long long)m, \
                                             (const double *)Xt, (const double *)Yt, (const Ctl *)nullptr, part); break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
        CASE(9) CASE(10) CASE(11) CASE(12) CASE(13) CASE(14) CASE(15) CASE(16)
#undef CASE
    }
    double *r2 = bkh_red2();
    const int nb2 = (Gb + RED_CHUNK - 1) / RED_CHUNK;
    hipLaunchKernelGGL(k_part_reduce, dim3(nb2), dim3(256), 0, st, (const double *)part, Gb, s * s, r2);
    LAUNCH_OK();
    std::vector<double> h((size_t)nb2 * s * s);
    HIP_OK(hipMemcpyAsync(h.data(), r2, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    for (int e = 0; e < s * s; e++) {
        double acc = 0.0;
        for (int b = 0; b < nb2; b++) acc += h[(size_t)b * s * s + e];
        G[e] = acc;
    }
    return 0;
}

// AP = A P and G = P^T AP (s x s, column-major, to host) for m x s column-major P, by
// the path the block CG iteration takes: the fused SpMM + Gram kernel when the matrix
// allows it, else the SpMM and the Gram pass. Partial rows summed in order on the host.
// Returns 0, or 1 when the fused kernel ran and 2 when the pipelined slice-diagonal
// form of it ran (through *path), for tests that compare the paths.
template <int S>
static int spmm_gram_host(ThreadCtx *t, const bkh_matrix *A, const double *P, double *AP, double *G, int *path) {
    hipStream_t st = t->stream;
    const long long m = A->n;
    const int Gt = vec_grid(m * S, 256), Gb = block_grid((int)m);
    const int nrows = std::max(Gb, 1024);
    void *ws = bkh_workspace(2 * carve_size(sizeof(double) * m * S) + carve_size(sizeof(double) * nrows * SM_N));
    if (!ws) return BKH_DEVICE_ERROR;
    Carve cv(ws);
    double *Pt = cv.take<double>(m * S), *APt = cv.take<double>(m * S);
    double *part = cv.take<double>((size_t)nrows * SM_N);
    hipLaunchKernelGGL((k_cm2rm<double, double>), dim3(Gt), dim3(256), 0, st, m, S, P, Pt);
    int ng = launch_spmm_gram<S>(st, A, A->val, (const double *)Pt, APt, (const Ctl *)nullptr, part);
    *path = 0;
    if (ng) {
        *path = (A->dmeta && A->dia_maxpos <= 8 && A->dia_entries > 0) ? 2 : 1;
    } else {
        launch_spmm<S>(st, A, A->val, (const double *)Pt, APt, 1.0, 0.0, (const Ctl *)nullptr, true);
        ng = panel_gram<S, double>(st, m, Pt, APt, (const Ctl *)nullptr, part);
    }
    hipLaunchKernelGGL((k_rm2cm<double, double>), dim3(Gt), dim3(256), 0, st, m, S, (const double *)APt, AP);
    LAUNCH_OK();
    std::vector<double> h((size_t)ng * S * S);
    HIP_OK(hipMemcpyAsync(h.data(), part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    for (int e = 0; e < S * S; e++) {
        double acc = 0.0;
        for (int b = 0; b < ng; b++) acc += h[(size_t)b * S * S + e];
        G[e] = acc;
    }
    return 0;
}

extern "C" int bkh_spmm_gram(const bkh_matrix *A, int s, const double *P, double *AP, double *G, int *path) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (!A || !P || !AP || !G || !path) return BKH_BAD_ARGUMENT;
    if (s < 1 || s > SM_MAX) { bkh_set_error("SpMM + Gram supports 1 <= s <= %d", SM_MAX); return BKH_BAD_ARGUMENT; }
    switch (s) {
#define CASE(S_) case S_: return spmm_gram_host<S_>(t, A, P, AP, G, path);
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
        CASE(9) CASE(10) CASE(11) CASE(12) CASE(13) CASE(14) CASE(15) CASE(16)
#undef CASE
    }
    return BKH_BAD_ARGUMENT;
}
