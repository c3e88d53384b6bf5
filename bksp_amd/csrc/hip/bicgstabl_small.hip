// bicgstabl_small.hip -- BiCGSTAB(l) for small systems in ONE persistent workgroup.
//
// At n of a few thousand (the 64x64 convection-diffusion case) an outer iteration
// of the multi-kernel path is ~14 dependent launches whose latency dwarfs the
// arithmetic. Here a single 1024-thread workgroup runs the whole solve with
// workgroup barriers between phases: the same steps as bicgstabl.hip
// (src/bicgstabl.c:54-126), the SpMV as a sequential row sum per thread (the
// reference's order, src/matvec.c:22-25), dot products as workgroup tree
// reductions, the MR step from the fused Gram matrix (bkh_mr.hpp) in thread 0.
// Vectors live in global memory (L2-resident at these sizes). The loop is bounded
// by max_iter, so the kernel always terminates.
#include <math.h>
#include <stdlib.h>
#include "bkh_host.hpp"
#include "bkh_mr.hpp"
#include "bkh_solver_common.hpp"

#define SMALL_BS 1024
#define SMALL_MAX_N 16384
#define SMALL_MAX_NNZ 400000

template <int NV>
__device__ __forceinline__ void wg_sum(double (&v)[NV], double *sh) {
    // sh: NV * 16 doubles; result broadcast to every thread
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; i++) v[i] = wave_sum(v[i]);
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < NV; i++) sh[i * 16 + w] = v[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NV; i++) {
        double t = sh[i * 16];
        for (int k = 1; k < 16; k++) t += sh[i * 16 + k];
        v[i] = t;
    }
    __syncthreads();
}

__device__ __forceinline__ void wg_spmv(int n, const int *rp, const int *ci, const double *val,
                                        const double *x, double *y) {
    for (int i = threadIdx.x; i < n; i += SMALL_BS) {
        double s = 0.0;
        for (int k = rp[i]; k < rp[i + 1]; k++) s = add_rn(s, mul_rn(val[k], x[ci[k]]));
        y[i] = s;
    }
}

template <int L>
__global__ __launch_bounds__(SMALL_BS) void k_bicgstabl_small(int n, const int *rp, const int *ci,
                                                              const double *val, const double *b,
                                                              double *x, double *R, double *P,
                                                              double *rs, Ctl *c) {
    constexpr int NG = (L + 1) * (L + 2) / 2;
    __shared__ double sh[NG * 16 + 32];
    __shared__ double coef[3 * MR_D + 4];
    __shared__ int flags[2];   // [0]: stop, [1]: status
    const int tid = threadIdx.x;
    int max_it = c->max_it;
    const double tol = c->tol;
    // r0 = b - A x0; p0 = rs = r0 (bicgstabl.c:44-51)
    wg_spmv(n, rp, ci, val, x, R);
    __syncthreads();
    double v2[2] = {0.0, 0.0};
    for (int i = tid; i < n; i += SMALL_BS) {
        double ri = b[i] - R[i];
        R[i] = ri; P[i] = ri; rs[i] = ri;
        v2[0] += ri * ri;
        v2[1] += b[i] * b[i];
    }
    wg_sum<2>(v2, sh);
    const double fnb = sqrt(v2[1]);
    double alpha = 0.0, beta = 0.0, g0 = v2[0], oml = 1.0, err = 0.0;
    int status = ST_MAXIT, it = 0, result = 0;
    if (!(fnb > 0.0)) { status = ST_BREAKDOWN; max_it = 0; }
    else if (v2[0] == 0.0) { status = ST_CONVERGED; result = 0; }
    for (it = 0; it < max_it && status != ST_CONVERGED; it++) {
        bool stop = false;
        for (int j = 0; j < L && !stop; j++) {
            for (int i = tid; i < n; i += SMALL_BS)                // (:64-67)
                for (int q = 0; q <= j; q++) P[q * n + i] = add_rn(mul_rn(beta, P[q * n + i]), R[q * n + i]);
            __syncthreads();
            wg_spmv(n, rp, ci, val, P + j * n, P + (j + 1) * n);   // (:68-69)
            __syncthreads();
            double d1[1] = {0.0};
            for (int i = tid; i < n; i += SMALL_BS) d1[0] += rs[i] * P[(j + 1) * n + i];
            wg_sum<1>(d1, sh);
            if (bad(d1[0])) { status = ST_BREAKDOWN; stop = true; break; }
            alpha = g0 / d1[0];                                    // (:71)
            for (int i = tid; i < n; i += SMALL_BS) {              // (:73-75, :79)
                for (int q = 0; q <= j; q++) R[q * n + i] = fma(-alpha, P[(q + 1) * n + i], R[q * n + i]);
                x[i] = fma(alpha, P[i], x[i]);
            }
            __syncthreads();
            wg_spmv(n, rp, ci, val, R + j * n, R + (j + 1) * n);   // (:76-77)
            __syncthreads();
            if (j < L - 1) {
                double d2[1] = {0.0};
                for (int i = tid; i < n; i += SMALL_BS) d2[0] += rs[i] * R[(j + 1) * n + i];
                wg_sum<1>(d2, sh);
                if (bad(g0)) { status = ST_BREAKDOWN; stop = true; break; }
                beta = -alpha * d2[0] / g0;                        // (:60-62)
                g0 = d2[0];
            }
        }
        if (stop) break;
        // MR step from the Gram matrix of r_0..r_L (:84-109)
        double gv[NG];
#pragma unroll
        for (int q = 0; q < NG; q++) gv[q] = 0.0;
        for (int i = tid; i < n; i += SMALL_BS) {
            double rv[L + 1];
#pragma unroll
            for (int q = 0; q <= L; q++) rv[q] = R[q * n + i];
            int idx = 0;
#pragma unroll
            for (int a = 0; a <= L; a++)
#pragma unroll
                for (int bb = a; bb <= L; bb++) gv[idx++] += rv[a] * rv[bb];
        }
        wg_sum<NG>(gv, sh);
        if (tid == 0) {
            double G[MR_D][MR_D];
            gram_unpack(L, gv, G);
            MrOut o;
            flags[0] = 0;
            if (G[0][0] == 0.0) {
                flags[0] = 1; flags[1] = ST_CONVERGED;
            } else if (!mr_from_gram(L, G, o)) {
                flags[0] = 1; flags[1] = ST_BREAKDOWN;
            } else {
                coef[0] = o.om[1];
                for (int k = 1; k < L; k++) {
                    double t = 0.0;
                    for (int i = k; i < L; i++) t += o.odd[i] * o.C[i][k];
                    coef[k] = t;
                }
                for (int k = 1; k <= L; k++) {
                    double t = 0.0;
                    for (int i = k; i <= L; i++) t += o.od[i] * o.C[i][k];
                    coef[MR_D + k] = t;
                    coef[2 * MR_D + k] = o.om[k];
                }
                coef[3 * MR_D] = o.om[L];
            }
        }
        __syncthreads();
        if (flags[0]) {
            status = flags[1];
            result = it;
            break;
        }
        double v3[2] = {0.0, 0.0};
        for (int i = tid; i < n; i += SMALL_BS) {                  // (:113-120)
            double xi = x[i], r0 = R[i], p0 = P[i];
            xi += coef[0] * r0;
            for (int k = 1; k <= L; k++) {
                const double rk = R[k * n + i];
                if (k < L) xi += coef[k] * rk;
                r0 -= coef[MR_D + k] * rk;
                p0 -= coef[2 * MR_D + k] * P[k * n + i];
            }
            x[i] = xi; R[i] = r0; P[i] = p0;
            v3[0] += r0 * r0;
            v3[1] += rs[i] * r0;
        }
        wg_sum<2>(v3, sh);
        oml = coef[3 * MR_D];
        err = sqrt(v3[0]) / fnb;                                   // (:122-125)
        if (err < tol) { status = ST_CONVERGED; result = it; break; }
        if (!(err == err) || err > 1e300) { status = ST_BREAKDOWN; result = it; break; }
        const double gg = -oml * g0;                               // (:57-62)
        if (bad(gg)) { status = ST_BREAKDOWN; result = it; break; }
        beta = -alpha * v3[1] / gg;
        g0 = v3[1];
    }
    if (tid == 0) {
        c->fnb = fnb;
        c->err = err;
        c->status = status;
        c->result_it = (status == ST_MAXIT) ? (it > 0 ? it - 1 : 0) : result;
        c->it = it;
        c->done = 1;
    }
}

// ------------------------------------------------ register-resident variant
//
// For n <= RPT * 1024 with few Krylov vectors: thread t owns rows t, t + 1024, ...
// and keeps ITS entries of every Krylov vector (r_0..r_l, p_0..p_l, rs, x) in
// registers for the whole solve. An SpMV stages its input vector in LDS (two buffers
// used alternately, so one barrier per SpMV suffices: a buffer is rewritten only after
// an intervening barrier has ordered every read of it) and each thread gathers its
// rows from LDS in the reference's order; the matrix streams from L2. Dot products
// are wave shuffles plus one LDS exchange (again double-buffered, one barrier). The
// arithmetic of every element is that of k_bicgstabl_small.

template <int NV, int BS>
__device__ __forceinline__ void reg_sum(double (&v)[NV], double *sh) {
    constexpr int REG_NW = BS / 64;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; i++) v[i] = wave_sum(v[i]);
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < NV; i++) sh[i * REG_NW + w] = v[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NV; i++) {
        double t = sh[i * REG_NW];
#pragma unroll
        for (int k = 1; k < REG_NW; k++) t += sh[i * REG_NW + k];
        v[i] = t;
    }
}

// Row r's entries: CSR [rb, re), or on the SELL-64 copy (so != null) entry k at
// so[r / 64] + 64 k + r % 64 -- contiguous across a wave. The first REG_W entries of
// every row are loaded together (predicated), then summed in stored order.
#define REG_W 6
template <int RPT, int REG_BS>
__device__ __forceinline__ void reg_spmv(int n, const int (&rb)[RPT], const int (&re)[RPT], const int *ci,
                                         const double *val, const int *so, const double *sval, const int *scol,
                                         const double (&in)[RPT], double (&out)[RPT], double *xs) {
#pragma unroll
    for (int q = 0; q < RPT; q++) {
        const int i = threadIdx.x + q * REG_BS;
        if (i < n) xs[i] = in[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < RPT; q++) {
        const int i = threadIdx.x + q * REG_BS;
        const int len = re[q] - rb[q];
        double a[REG_W];
        int cc[REG_W];
        if (so) {
            const int base = i < n ? so[i >> 6] + (i & 63) : 0;
#pragma unroll
            for (int k = 0; k < REG_W; k++) {
                a[k] = k < len ? sval[base + 64 * k] : 0.0;
                cc[k] = k < len ? scol[base + 64 * k] : 0;
            }
        } else {
#pragma unroll
            for (int k = 0; k < REG_W; k++) {
                a[k] = k < len ? val[rb[q] + k] : 0.0;
                cc[k] = k < len ? ci[rb[q] + k] : 0;
            }
        }
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < REG_W; k++)
            if (k < len) s = add_rn(s, mul_rn(a[k], xs[cc[k]]));
        for (int k = REG_W; k < len; k++) {
            const int e = so ? so[i >> 6] + (i & 63) + 64 * k : rb[q] + k;
            s = add_rn(s, mul_rn(so ? sval[e] : val[e], xs[so ? scol[e] : ci[e]]));
        }
        out[q] = s;
    }
}

// The same SpMV with the matrix held in registers (KD > 0): on the slice-diagonal copy a
// wave's rows of one q lie in one slice, so the slice's positions are uniform (KD of
// them, absent entries 0, column clamped into [0, n)) and each thread keeps its rows'
// KD values in registers for the whole solve; an SpMV is then the LDS staging, one
// barrier and KD LDS reads per row. Every row sums its positions in ascending offset
// order, i.e. its stored order; a 0 value adds +-0 to a sum that starts at +0 and so
// never holds -0, so for finite vectors the sums are bitwise those of reg_spmv.
// xs: the staging buffer with DIA_HALO zeros before entry 0 and zeros from entry n to
// its end, so a position's column needs no clamp (every |offset| <= DIA_HALO; an
// absent position reads a zero or a finite entry and adds +-0 either way).
#define DIA_HALO 128
template <int RPT, int REG_BS, int KD>
__device__ __forceinline__ void dia_reg_spmv(int n, const double (&dv)[RPT][KD], const int (&doff)[RPT][KD],
                                             const double (&in)[RPT], double (&out)[RPT], double *xs) {
#pragma unroll
    for (int q = 0; q < RPT; q++) {
        const int i = threadIdx.x + q * REG_BS;
        if (i < n) xs[i] = in[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < RPT; q++) {
        const int i = threadIdx.x + q * REG_BS;
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < KD; k++) s = add_rn(s, mul_rn(dv[q][k], xs[i + doff[q][k]]));
        out[q] = s;
    }
}

template <int L, int RPT, int REG_BS, int KD = 0>
__global__ __launch_bounds__(REG_BS) void k_bicgstabl_reg(int n, const int *rp, const int *ci, const double *val,
                                                          const int *so, const double *sval, const int *scol,
                                                          const double *b, double *xg, Ctl *c,
                                                          const DiaSlice *meta = nullptr, const double *dval = nullptr,
                                                          int nsl = 0) {
    constexpr int NG = (L + 1) * (L + 2) / 2;
    constexpr int REG_NW = REG_BS / 64;
    __shared__ double xs_mem[2][RPT * REG_BS + (KD > 0 ? 2 * DIA_HALO : 0)];
    __shared__ double shr[2][REG_NW * (NG > 2 ? NG : 2)];
    __shared__ double coef[3 * MR_D + 4];
    __shared__ int flags[2];
    __shared__ MrOut o;                                    // thread 0's MR-step workspace
    __shared__ double Gs[MR_D][MR_D];
    // rs and x are touched once per step: they live in LDS (each thread only ever
    // reads its own slots, so no barrier guards them), the other vectors in registers
    __shared__ double rs_s[RPT * REG_BS], x_s[RPT * REG_BS];
    const int tid = threadIdx.x;
    // slice-diagonal form: entry 0 of each staging buffer sits DIA_HALO zeros in, and
    // the tail past n is zero
    constexpr int XOFF = KD > 0 ? DIA_HALO : 0;
    double *const xs0 = &xs_mem[0][XOFF], *const xs1 = &xs_mem[1][XOFF];
    if constexpr (KD > 0) {
        for (int e = tid; e < RPT * REG_BS + 2 * DIA_HALO; e += REG_BS)
            if (e < DIA_HALO || e >= DIA_HALO + n) { xs_mem[0][e] = 0.0; xs_mem[1][e] = 0.0; }
    }
    int rb[RPT], re[RPT];
    double R[L + 1][RPT], P[L + 1][RPT];
    constexpr int KDA = KD > 0 ? KD : 1;
    double dv[RPT][KDA];
    int doff[RPT][KDA];
#define RS(q) rs_s[(q) * REG_BS + tid]
#define XX(q) x_s[(q) * REG_BS + tid]
#pragma unroll
    for (int q = 0; q < RPT; q++) {
        const int i = tid + q * REG_BS;
        const bool live = i < n;
        rb[q] = live ? rp[i] : 0;
        re[q] = live ? rp[i + 1] : 0;
        XX(q) = live ? xg[i] : 0.0;
        if constexpr (KD > 0) {
            // this wave's rows of q form one slice (REG_BS is a multiple of 64)
            const int sl = __builtin_amdgcn_readfirstlane(min((tid >> 6) + q * (REG_BS / 64), nsl - 1));
            const DiaSlice m = meta[sl];
#pragma unroll
            for (int k = 0; k < KD; k++) {
                doff[q][k] = m.o[k];
                dv[q][k] = (live && k < m.len) ? dval[m.s0 + 64 * k + (i & 63)] : 0.0;
            }
        }
    }
    auto spmv = [&](const double (&in)[RPT], double (&out)[RPT], double *xsb) {
        if constexpr (KD > 0) dia_reg_spmv<RPT, REG_BS, KDA>(n, dv, doff, in, out, xsb);
        else reg_spmv<RPT, REG_BS>(n, rb, re, ci, val, so, sval, scol, in, out, xsb);
    };
    int sx = 0, sr = 0;                                   // alternating LDS buffers
    const int max_it = c->max_it;
    const double tol = c->tol;
    // r0 = b - A x0; p0 = rs = r0 (bicgstabl.c:44-51)
    {
        double ax[RPT], x0[RPT];
#pragma unroll
        for (int q = 0; q < RPT; q++) x0[q] = XX(q);
        spmv(x0, ax, sx ? xs1 : xs0);
        sx ^= 1;
        double v2[2] = {0.0, 0.0};
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            const int i = tid + q * REG_BS;
            const double bi = i < n ? b[i] : 0.0;
            const double ri = i < n ? bi - ax[q] : 0.0;
            R[0][q] = ri; P[0][q] = ri; RS(q) = ri;
            v2[0] += ri * ri;
            v2[1] += bi * bi;
        }
        reg_sum<2, REG_BS>(v2, shr[sr]);
        sr ^= 1;
        double fnb = sqrt(v2[1]);
        double alpha = 0.0, beta = 0.0, g0 = v2[0], oml = 1.0, err = 0.0;
        int status = ST_MAXIT, it = 0, result = 0;
        int mit = max_it;
        if (!(fnb > 0.0)) { status = ST_BREAKDOWN; mit = 0; }
        else if (v2[0] == 0.0) { status = ST_CONVERGED; result = 0; }
        for (it = 0; it < mit && status != ST_CONVERGED; it++) {
            bool stop = false;
#pragma unroll
            for (int j = 0; j < L; j++) {
                if (!stop) {
#pragma unroll
                    for (int q = 0; q < RPT; q++)               // (:64-67)
#pragma unroll
                        for (int a = 0; a <= j; a++) P[a][q] = add_rn(mul_rn(beta, P[a][q]), R[a][q]);
                    spmv(P[j], P[j + 1], sx ? xs1 : xs0);   // (:68-69)
                    sx ^= 1;
                    double d1[1] = {0.0};
#pragma unroll
                    for (int q = 0; q < RPT; q++) d1[0] += RS(q) * P[j + 1][q];
                    reg_sum<1, REG_BS>(d1, shr[sr]);
                    sr ^= 1;
                    if (bad(d1[0])) {
                        status = ST_BREAKDOWN;
                        stop = true;
                    } else {
                        alpha = g0 / d1[0];                            // (:71)
#pragma unroll
                        for (int q = 0; q < RPT; q++) {                // (:73-75, :79)
#pragma unroll
                            for (int a = 0; a <= j; a++) R[a][q] = fma(-alpha, P[a + 1][q], R[a][q]);
                            XX(q) = fma(alpha, P[0][q], XX(q));
                        }
                        spmv(R[j], R[j + 1], sx ? xs1 : xs0);   // (:76-77)
                        sx ^= 1;
                        if (j < L - 1) {
                            double d2[1] = {0.0};
#pragma unroll
                            for (int q = 0; q < RPT; q++) d2[0] += RS(q) * R[j + 1][q];
                            reg_sum<1, REG_BS>(d2, shr[sr]);
                            sr ^= 1;
                            if (bad(g0)) {
                                status = ST_BREAKDOWN;
                                stop = true;
                            } else {
                                beta = -alpha * d2[0] / g0;            // (:60-62)
                                g0 = d2[0];
                            }
                        }
                    }
                }
            }
            if (stop) break;
            // MR step from the Gram matrix of r_0..r_L (:84-109)
            double gv[NG];
#pragma unroll
            for (int a = 0; a < NG; a++) gv[a] = 0.0;
#pragma unroll
            for (int q = 0; q < RPT; q++) {
                int idx = 0;
#pragma unroll
                for (int a = 0; a <= L; a++)
#pragma unroll
                    for (int bb = a; bb <= L; bb++) gv[idx++] += R[a][q] * R[bb][q];
            }
            reg_sum<NG, REG_BS>(gv, shr[sr]);
            sr ^= 1;
            if (tid == 0) {
                gram_unpack(L, gv, Gs);
                flags[0] = 0;
                if (Gs[0][0] == 0.0) {
                    flags[0] = 1; flags[1] = ST_CONVERGED;
                } else if (!mr_from_gram(L, Gs, o)) {
                    flags[0] = 1; flags[1] = ST_BREAKDOWN;
                } else {
                    coef[0] = o.om[1];
                    for (int k = 1; k < L; k++) {
                        double t = 0.0;
                        for (int i = k; i < L; i++) t += o.odd[i] * o.C[i][k];
                        coef[k] = t;
                    }
                    for (int k = 1; k <= L; k++) {
                        double t = 0.0;
                        for (int i = k; i <= L; i++) t += o.od[i] * o.C[i][k];
                        coef[MR_D + k] = t;
                        coef[2 * MR_D + k] = o.om[k];
                    }
                    coef[3 * MR_D] = o.om[L];
                }
            }
            __syncthreads();
            if (flags[0]) {
                status = flags[1];
                result = it;
                break;
            }
            double v3[2] = {0.0, 0.0};
#pragma unroll
            for (int q = 0; q < RPT; q++) {                             // (:113-120)
                double xi = XX(q), r0 = R[0][q], p0 = P[0][q];
                xi += coef[0] * r0;
#pragma unroll
                for (int k = 1; k <= L; k++) {
                    const double rk = R[k][q];
                    if (k < L) xi += coef[k] * rk;
                    r0 -= coef[MR_D + k] * rk;
                    p0 -= coef[2 * MR_D + k] * P[k][q];
                }
                XX(q) = xi; R[0][q] = r0; P[0][q] = p0;
                v3[0] += r0 * r0;
                v3[1] += RS(q) * r0;
            }
            oml = coef[3 * MR_D];
            reg_sum<2, REG_BS>(v3, shr[sr]);
            sr ^= 1;
            err = sqrt(v3[0]) / fnb;                                    // (:122-125)
            if (err < tol) { status = ST_CONVERGED; result = it; break; }
            if (!(err == err) || err > 1e300) { status = ST_BREAKDOWN; result = it; break; }
            const double gg = -oml * g0;                                // (:57-62)
            if (bad(gg)) { status = ST_BREAKDOWN; result = it; break; }
            beta = -alpha * v3[1] / gg;
            g0 = v3[1];
        }
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            const int i = tid + q * REG_BS;
            if (i < n) xg[i] = XX(q);
        }
        if (tid == 0) {
            c->fnb = fnb;
            c->err = err;
            c->status = status;
            c->result_it = (status == ST_MAXIT) ? (it > 0 ? it - 1 : 0) : result;
            c->it = it;
            c->done = 1;
        }
    }
}
#undef RS
#undef XX

// Launch shapes that hold the vectors without register spills (from the compiler's
// resource report): 1024 threads while a thread's share is small, else 512 threads
// with up to 256 registers per lane. Larger systems take the global-memory kernel.
template <int L>
static bool reg_launch(hipStream_t s, int n, const bkh_matrix *A, const double *b, double *x, Ctl *c) {
    const bool sell = A->nsl > 0 && A->sval && A->scol;
    const int *so = sell ? A->sso : nullptr;
    const double *sv = sell ? A->sval : nullptr;
    const int *sc = sell ? A->scol : nullptr;
#define REG_GO(RPT_, BS_) \
    { hipLaunchKernelGGL((k_bicgstabl_reg<L, RPT_, BS_>), dim3(1), dim3(BS_), 0, s, n, A->rp, A->ci, A->val, so, sv, sc, \
                         b, x, c); return true; }
    // matrix values in registers on the slice-diagonal copy (<= 5 or <= 7 positions per
    // slice): 64x64 convection-diffusion BiCGSTAB(2) 1.94 -> 0.70 ms, 31 -> 11.4 us per
    // outer iteration (streamed from L2 into the one CU, the matrix cost 5-7 us per SpMV)
#define DIA_GO(RPT_, BS_, KD_) \
    { hipLaunchKernelGGL((k_bicgstabl_reg<L, RPT_, BS_, KD_>), dim3(1), dim3(BS_), 0, s, n, A->rp, A->ci, A->val, so, sv, \
                         sc, b, x, c, (const DiaSlice *)A->dmeta, (const double *)A->dval, A->nsl); return true; }
    const bool dia = A->dmeta && A->dia_entries > 0 && A->dval && A->dia_maxoff <= DIA_HALO &&
                     !getenv("BKH_SMALL_NODIA");
    if constexpr (L <= 2) {
        if (dia && A->dia_maxpos <= 5) {
            if (n <= 1024) DIA_GO(1, 1024, 5)
            if (n <= 2048) DIA_GO(4, 512, 5)
            if (n <= 4096) DIA_GO(8, 512, 5)
        }
        if (dia && A->dia_maxpos <= 7) {
            if (n <= 1024) DIA_GO(1, 1024, 7)
            if (n <= 2048) DIA_GO(4, 512, 7)
            if (n <= 4096) DIA_GO(8, 512, 7)
        }
    }
#undef DIA_GO
    if constexpr (L <= 2) {
        if (n <= 1024) REG_GO(1, 1024)
        if (n <= 2048) REG_GO(2, 1024)
        if (n <= 4096) REG_GO(8, 512)
    } else if constexpr (L == 3) {
        if (n <= 1024) REG_GO(1, 1024)
        if (n <= 2048) REG_GO(4, 512)
    } else {
        if (n <= 1024) REG_GO(2, 512)
    }
#undef REG_GO
    return false;
}

// Returns true (and the solve's result in t->ctl_host) if the small path ran.
bool bicgstabl_small(ThreadCtx *t, const bkh_matrix *A, int l, const double *b, double *x, double tol,
                     int max_iter, int *rc) {
    if (A->n > SMALL_MAX_N || A->nnz > SMALL_MAX_NNZ || l < 1 || l > 8) return false;
    const long long n = A->n;
    void *ws = bkh_workspace(carve_size(sizeof(Ctl)) + (2 * (l + 1) + 1) * carve_size(sizeof(double) * n));
    if (!ws) { *rc = BKH_DEVICE_ERROR; return true; }
    Carve cv(ws);
    Ctl *c = cv.take<Ctl>(1);
    double *R = cv.take<double>((l + 1) * n), *P = cv.take<double>((l + 1) * n), *rs = cv.take<double>(n);
    hipStream_t s = t->stream;
    *rc = ctl_init(t, c, tol, max_iter);
    if (*rc) return true;
    bool reg = false;
    if (!getenv("BKH_SMALL_GLOBAL")) {
        switch (l) {
        case 1: reg = reg_launch<1>(s, (int)n, A, b, x, c); break;
        case 2: reg = reg_launch<2>(s, (int)n, A, b, x, c); break;
        case 3: reg = reg_launch<3>(s, (int)n, A, b, x, c); break;
        case 4: reg = reg_launch<4>(s, (int)n, A, b, x, c); break;
        default: break;
        }
    }
    if (!reg) switch (l) {
#define CASE(L_) case L_: hipLaunchKernelGGL((k_bicgstabl_small<L_>), dim3(1), dim3(SMALL_BS), 0, s, (int)n, \
                     (const int *)A->rp, (const int *)A->ci, (const double *)A->val, b, x, R, P, rs, c); break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
    }
    if (hipGetLastError() != hipSuccess) { bkh_set_error("small BiCGSTAB(l) launch failed"); *rc = BKH_DEVICE_ERROR; return true; }
    *rc = ctl_fetch(t, c);
    return true;
}
