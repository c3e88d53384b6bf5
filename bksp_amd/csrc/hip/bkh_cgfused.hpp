// bkh_cgfused.hpp -- CG in two kernels per iteration, with no finalize kernels.
//
// Iteration k of cg.c:32-60 becomes
//   A(k): p_k = r_k + beta_k p_{k-1} folded into the SpMV's gather (cg.c:53-54),
//         q = A p_k (cg.c:35-36), partial sums of p_k.q (cg.c:39);
//   B(k): x += alpha_k p_k, r -= alpha_k q (cg.c:42-48), partial sums of r.r.
// Each kernel needs one scalar of the kernel before it: instead of a one-workgroup
// finalize kernel between them (a launch plus a serial reduction on the critical
// path), every wave of the consumer sums the producer's partials itself, all in the
// same fixed order, so every wave holds the same bits (the butterfly below adds the
// same two operands in both lanes of each pair). A(k) also runs the convergence test
// of iteration k-1 (cg.c:56-59) and stops the solve; a one-wave tail kernel at the
// end of a graph replay runs the test of the replay's last iteration and advances
// Ctl::base, the loop index of the replay's first iteration.
//
// p alternates between two buffers by the parity of k; r.r of iteration k lives in
// Ctl::s[CGF_RR + (k & 1)], written by block 0 of A(k) and read by B(k) and A(k+1)
// (which writes the other slot), so no kernel reads a slot that the same kernel
// writes. Partial counts are capped at CGF_NPART so the redundant sums stay short.
//
// Column indices are int32, or 16-bit deltas from the wave block's first row
// (6 bytes per FP32 nonzero instead of 8) when the matrix has them.
#pragma once
#include "bkh_core.hpp"
#include "bkh_host.hpp"

#define CGF_NPART 2048
#define CGF_WPB 8                     // waves per workgroup of A
enum { CGF_RR = 8, CGF_FRESH = 10 };  // Ctl::s slots (8, 9: r.r by parity)

// Sum of np partials (np <= CGF_NPART) in a fixed order, valid in every lane.
__device__ __forceinline__ double wave_sum_parts(const double *part, int np) {
    const int lane = threadIdx.x & 63;
    double v[CGF_NPART / 64];
#pragma unroll
    for (int m = 0; m < CGF_NPART / 64; m++) {
        const int i = lane + 64 * m;
        v[m] = i < np ? part[i] : 0.0;
    }
    double a = 0.0;
#pragma unroll
    for (int m = 0; m < CGF_NPART / 64; m++) a += v[m];
    return wave_sum(a);
}

// The same sum formed by a whole workgroup of BS threads (each reads np/BS partials,
// then a fixed-order workgroup reduction), broadcast to every thread: every workgroup
// of one kernel computes identical bits, and the partial array is read once per
// workgroup instead of once per wave.
template <int BS>
__device__ __forceinline__ double block_sum_parts(const double *part, int np) {
    __shared__ double sh[BS / 64 + 1];
    double a = 0.0;
    for (int i = threadIdx.x; i < np; i += BS) a += part[i];
    a = wave_sum(a);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[w] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = sh[0];
#pragma unroll
        for (int q = 1; q < BS / 64; q++) t += sh[q];
        sh[BS / 64] = t;
    }
    __syncthreads();
    return sh[BS / 64];
}

// Convergence test of iteration k-1 given r_k.r_k (the reference's test, cg.c:56-59;
// NaN/Inf stops with ST_BREAKDOWN). Returns the status; workgroup 0 records it.
__device__ __forceinline__ int cgf_check(Ctl *c, int k, double rr, bool writer) {
    const double err = sqrt(rr) / c->fnb;
    int st = ST_RUNNING;
    if (err < c->tol) st = ST_CONVERGED;
    else if (!(err == err) || err > 1.0e300) st = ST_BREAKDOWN;
    else if (k >= c->max_it) st = ST_MAXIT;
    if (writer) {
        c->err = err;
        c->it = k;
        if (st != ST_RUNNING) {
            c->status = st;
            c->result_it = k - 1;
            c->done = 1;
        }
    }
    return st;
}

// A(k). Workgroup g streams the contiguous range of `chunk` wave blocks starting
// at g*chunk (g from the XCD-aware remap), its waves taking every CGF_WPB-th.
template <typename VT, typename T, bool D16>
__global__ __launch_bounds__(64 * CGF_WPB) void k_cgf_a(int nwblk, int chunk, const int *__restrict__ rp,
                                                       const int *__restrict__ wblk, const int *__restrict__ ci,
                                                       const short *__restrict__ cd, const VT *__restrict__ val,
                                                       const T *r, T *P0, T *P1, T *q, Ctl *c, int j,
                                                       const double *partB, int npB, double *partA) {
    if (ctl_stopped(c)) return;
    const int k = c->base + j;
    const bool writer = blockIdx.x == 0 && threadIdx.x == 0;
    double rr;
    if (k == c->first) {
        rr = c->s[CGF_RR + (k & 1)];
    } else {
        rr = block_sum_parts<64 * CGF_WPB>(partB, npB);
        if (cgf_check(c, k, rr, writer) != ST_RUNNING) return;
        if (writer) c->s[CGF_RR + (k & 1)] = rr;
    }
    const bool fresh = k == c->first && c->s[CGF_FRESH] != 0.0;
    const T be = fresh ? (T)0 : (T)(rr / c->s[CGF_RR + ((k + 1) & 1)]);
    const T *pold = (k & 1) ? P1 : P0;
    T *pnew = (k & 1) ? P0 : P1;

    __shared__ T prod_all[CGF_WPB][SPMV_WCAP];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    T *prod = prod_all[w];
    const int g = xcd_block(blockIdx.x, gridDim.x);
    const int b_end = min(nwblk, (g + 1) * chunk);
    double dv[1] = {0.0};
    constexpr int VEC = 16 / sizeof(VT);
    constexpr int STEPS = SPMV_WCAP / (VEC * 64) + 1;
    typedef VT vvec __attribute__((ext_vector_type(VEC)));
    typedef int ivec __attribute__((ext_vector_type(VEC)));
    typedef short svec __attribute__((ext_vector_type(VEC)));
    for (int b = g * chunk + w; b < b_end; b += CGF_WPB) {
        const int r0 = wblk[b], r1 = wblk[b + 1];
        const int k0 = rp[r0], k1 = rp[r1];
        const int kb = k0 & ~(VEC - 1);
        vvec vv[STEPS];
        ivec cc[STEPS];
#pragma unroll
        for (int s = 0; s < STEPS; s++) {
            const int kk = kb + VEC * (lane + s * 64);
            if (kk < k1) {
                vv[s] = *(const vvec *)(val + kk);
                if (D16) {
                    const svec s16 = *(const svec *)(cd + kk);
#pragma unroll
                    for (int u = 0; u < VEC; u++) cc[s][u] = r0 + (int)s16[u];
                } else {
                    cc[s] = *(const ivec *)(ci + kk);
                }
            }
        }
#pragma unroll
        for (int s = 0; s < STEPS; s++) {
            const int kk = kb + VEC * (lane + s * 64);
#pragma unroll
            for (int u = 0; u < VEC; u++) {
                if (kk + u >= k0 && kk + u < k1) {
                    const int col = cc[s][u];
                    const T pc = add_rn(mul_rn(be, pold[col]), r[col]);
                    prod[kk + u - k0] = mul_rn((T)vv[s][u], pc);
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int row = r0 + lane;
        if (row < r1) {
            const int s0 = rp[row] - k0, s1 = rp[row + 1] - k0;
            T sum = (T)0;
            for (int kk = s0; kk < s1; kk++) sum = add_rn(sum, prod[kk]);
            q[row] = sum;
            const T pn = add_rn(mul_rn(be, pold[row]), r[row]);
            pnew[row] = pn;
            dv[0] += (double)pn * (double)sum;
        }
        // the next block's products overwrite prod: every lane has read its rows
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    block_partial<1, 64 * CGF_WPB>(dv, partA);
}

// Calls f(e, column, live) for the stored entries of row `row`, lane l of slice sl
// (sl the same on every lane of the wave), in the row's order; an entry with live false
// is to be added as 0. SELL-64: the row's length from rp, IT int columns or short
// (column - row), every entry live. Slice-diagonal (DIA, IT int): the slice's width w
// from so, column = row + the slice's offset of the position (uniform, scalar loads),
// clamped into [0, n) for the value-0 positions a row does not have. With GRP = 8 the
// positions run in groups of 8 with unconditional loads (so a group's loads are all in
// flight at once; FP32 rows are latency-bound): positions past w read the last entry
// of the slice with live false, and the 8 offsets of a group are read as one block (the
// offset table has 8 entries of padding past its end; an offset read past w only picks
// a clamped column). GRP = 0: one position at a time, unrolled by 4, all live.
template <typename IT, bool DIA, int GRP = 0, typename F>
__device__ __forceinline__ void sell_row_entries(int n, int row, int l, int sl, const int *__restrict__ rp,
                                                 const int *__restrict__ so, const IT *__restrict__ col, F &&f) {
    static_assert(GRP == 0 || GRP == 8, "position groups of 8 (the offset table's padding)");
    if constexpr (DIA && GRP == 0) {
        const int s0 = so[sl], w = (so[sl + 1] - s0) >> 6;
        const int *o = (const int *)col + (s0 >> 6);
        const int base = s0 + l;
#pragma unroll 4
        for (int kk = 0; kk < w; kk++) f(base + 64 * kk, min(max(row + o[kk], 0), n - 1), true);
    } else if constexpr (DIA) {
        const int s0 = so[sl], w = (so[sl + 1] - s0) >> 6;
        const int *o = (const int *)col + (s0 >> 6);
        const int base = s0 + l;
        for (int k0 = 0; k0 < w; k0 += 8) {
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int kk = min(k0 + u, w - 1);
                f(base + 64 * kk, min(max(row + o[k0 + u], 0), n - 1), k0 + u < w);
            }
        }
    } else {
        const int len = rp[row + 1] - rp[row];
        const int base = so[sl] + l;
#pragma unroll 4
        for (int kk = 0; kk < len; kk++) {
            const int e = base + 64 * kk;
            f(e, sizeof(IT) == 2 ? row + (int)col[e] : (int)col[e], true);
        }
    }
}

// A(k) on the SELL-64 layout (or its slice-diagonal copy, DIA): a lane owns a row and
// forms p_k at each column of it in the gather; workgroup g owns `chunk` consecutive
// slices, its waves taking every fourth. IT: int columns or short (column - row).
template <typename VT, typename T, typename IT, int BS, bool DIA = false>
__global__ __launch_bounds__(BS) void k_cgf_a_sell(int n, int nsl, int chunk, const int *__restrict__ rp,
                                                    const int *__restrict__ so, const VT *__restrict__ val,
                                                    const IT *__restrict__ col, const T *r, T *P0, T *P1, T *q,
                                                    Ctl *c, int j, const double *partB, int npB, double *partA) {
    if (ctl_stopped(c)) return;
    const int k = c->base + j;
    const bool writer = blockIdx.x == 0 && threadIdx.x == 0;
    const T *pold = (k & 1) ? P1 : P0;
    T *pnew = (k & 1) ? P0 : P1;
    const int g = xcd_block(blockIdx.x, gridDim.x);
    const int l = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int send = min(nsl, (g + 1) * chunk);
    double rr;
    if (k == c->first) {
        rr = c->s[CGF_RR + (k & 1)];
    } else {
        rr = block_sum_parts<BS>(partB, npB);
        if (cgf_check(c, k, rr, writer) != ST_RUNNING) return;
        if (writer) c->s[CGF_RR + (k & 1)] = rr;
    }
    const bool fresh = k == c->first && c->s[CGF_FRESH] != 0.0;
    const T be = fresh ? (T)0 : (T)(rr / c->s[CGF_RR + ((k + 1) & 1)]);
    double dv[1] = {0.0};
    for (int sl = g * chunk + w; sl < send; sl += BS / 64) {
        const int row = sl * 64 + l;
        if (row >= n) continue;
        T acc = (T)0;
        sell_row_entries<IT, DIA>(n, row, l, sl, rp, so, col, [&](int e, int cc, bool live) {
            const T pc = add_rn(mul_rn(be, pold[cc]), r[cc]);
            const T v = (T)val[e];
            acc = add_rn(acc, mul_rn(live ? v : (T)0, pc));
        });
        q[row] = acc;
        const T pn = add_rn(mul_rn(be, pold[row]), r[row]);
        pnew[row] = pn;
        dv[0] += (double)pn * (double)acc;
    }
    block_partial<1, BS>(dv, partA);
}

// B(k): alpha_k = (r_k.r_k) / (p_k.q); x += alpha p_k, r -= alpha q; r.r partials.
template <typename T, bool VEC = false>
__global__ __launch_bounds__(512) void k_cgf_b(long long n, T *x, T *r, const T *P0, const T *P1,
                                               const T *q, Ctl *c, int j, const double *partA, int npA,
                                               double *partB) {
    if (ctl_stopped(c)) return;
    const int k = c->base + j;
    const T *p = (k & 1) ? P0 : P1;
    const long long tid = (long long)blockIdx.x * 512 + threadIdx.x, stride = (long long)gridDim.x * 512;
    constexpr int NV = VEC ? 16 / (int)sizeof(T) : 1;
    typedef T tv __attribute__((ext_vector_type(NV)));
    // VEC: each thread's first vectors of x, r, p, q are loaded before kernel A's
    // partials are summed, so their latency overlaps that prologue
    tv x0, r0, p0, q0;
    if (VEC && tid < n / NV) {
        x0 = ((const tv *)x)[tid]; r0 = ((const tv *)r)[tid];
        p0 = ((const tv *)p)[tid]; q0 = ((const tv *)q)[tid];
    }
    const double pq = block_sum_parts<512>(partA, npA);
    if (bad(pq)) {                                    // breakdown (zero or non-finite p.q)
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            c->status = ST_BREAKDOWN; c->result_it = k; c->done = 1;
        }
        return;
    }
    const T al = (T)(c->s[CGF_RR + (k & 1)] / pq);
    double v[1] = {0.0};
    if constexpr (VEC) {
        // 16-byte vectors (every array 16-byte aligned), the odd last entry by thread 0
        const long long nv = n / NV;
        for (long long iv = tid; iv < nv; iv += stride) {
            tv xv, rv, pv, qv;
            if (iv == tid) {
                xv = x0; rv = r0; pv = p0; qv = q0;
            } else {
                xv = ((const tv *)x)[iv]; rv = ((const tv *)r)[iv];
                pv = ((const tv *)p)[iv]; qv = ((const tv *)q)[iv];
            }
#pragma unroll
            for (int u = 0; u < NV; u++) {
                xv[u] = fma(al, pv[u], xv[u]);
                rv[u] = fma(-al, qv[u], rv[u]);
                v[0] += (double)rv[u] * (double)rv[u];
            }
            ((tv *)x)[iv] = xv;
            ((tv *)r)[iv] = rv;
        }
        for (long long i = nv * NV + tid; i < n; i += stride) {
            x[i] = fma(al, p[i], x[i]);
            const T ri = fma(-al, q[i], r[i]);
            r[i] = ri;
            v[0] += (double)ri * (double)ri;
        }
    } else {
        for (long long i = tid; i < n; i += stride) {
            x[i] = fma(al, p[i], x[i]);
            const T ri = fma(-al, q[i], r[i]);
            r[i] = ri;
            v[0] += (double)ri * (double)ri;
        }
    }
    block_partial<1, 512>(v, partB);
}

// End of a graph replay of K iterations: test of iteration base+K-1, base += K.
[[maybe_unused]] static __global__ __launch_bounds__(64) void k_cgf_tail(Ctl *c, int K, const double *partB, int npB) {
    if (ctl_stopped(c)) return;
    const int k = c->base + K;
    const double rr = wave_sum_parts(partB, npB);
    const bool writer = threadIdx.x == 0;
    if (cgf_check(c, k, rr, writer) != ST_RUNNING) return;
    if (writer) c->base = k;
}

// ------------------------------------------- one kernel per iteration (FP32 inner)
//
// The FP32 inner iteration of the reliable-update solve runs as ONE kernel per
// iteration. Kernel k reads r_{k-1}, q_{k-1}, p_{k-1} (buffers of parity (k-1)&1)
// and the partial sums (p.q, r.q, q.q) of kernel k-1, and
//   alpha_{k-1} = rho_{k-1} / (p_{k-1}.q_{k-1})
//   rho_k       = rho_{k-1} - 2 alpha r_{k-1}.q_{k-1} + alpha^2 q_{k-1}.q_{k-1}
//                 (= |r_{k-1} - alpha q_{k-1}|^2, FP64 sums of the stored FP32 values)
//   beta_{k-1}  = rho_k / rho_{k-1}
//   r_k = r_{k-1} - alpha q_{k-1}         (formed in the gather at every column)
//   q_k = A r_k + beta q_{k-1},  p_k = r_k + beta p_{k-1}   (q_k = A p_k in exact arithmetic)
//   x += alpha p_{k-1}  (the deferred update of iteration k-1)
// and writes r_k, p_k, q_k (parity k&1) and its partials (p_k.q_k, r_k.q_k,
// q_k.q_k). Only r and q are gathered (the SpMV multiplies r_k), one global
// synchronisation per iteration. The first kernel of a pass (k == Ctl::first: alpha = 0,
// x and r hold the replaced values) forms q_k = A p_k explicitly from p_k = r + beta
// p_{k-1} in its gather, so the recursively updated q starts every pass consistent;
// the periodic FP64 residual replacement of the reliable-update driver bounds the
// drift of r and rho in between.
// Ctl::s slots: alpha_{k-1} at a stop; the alpha of the x update an even kernel
// deferred to the next (odd) one, 0 when none is pending
enum { CG1_ALPHA = 12, CG1_ALPHAD = 13 };

// Fixed-order sum of np rows of 3 partials by a workgroup of BS threads,
// broadcast to every thread.
template <int BS>
__device__ __forceinline__ void block_sum_parts3(const double *part, int np, double (&o)[3]) {
    __shared__ double sh[3 * (BS / 64) + 3];
    double a[3] = {0.0, 0.0, 0.0};
    // rows i and i + BS loaded together (one memory round trip for up to 2 BS rows),
    // then added in row order
    for (int i = threadIdx.x; i < np; i += 2 * BS) {
        const bool h = i + BS < np;
        double u[3], t[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int v = 0; v < 3; v++) u[v] = part[3 * i + v];
        if (h) {
#pragma unroll
            for (int v = 0; v < 3; v++) t[v] = part[3 * (i + BS) + v];
        }
#pragma unroll
        for (int v = 0; v < 3; v++) {
            a[v] += u[v];
            if (h) a[v] += t[v];
        }
    }
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int v = 0; v < 3; v++) a[v] = wave_sum(a[v]);
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int v = 0; v < 3; v++) sh[v * (BS / 64) + w] = a[v];
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double t = sh[threadIdx.x * (BS / 64)];
#pragma unroll
        for (int q = 1; q < BS / 64; q++) t += sh[threadIdx.x * (BS / 64) + q];
        sh[3 * (BS / 64) + threadIdx.x] = t;
    }
    __syncthreads();
#pragma unroll
    for (int v = 0; v < 3; v++) o[v] = sh[3 * (BS / 64) + v];
}

// Scalars of kernel k (or of the tail check of iteration k). Returns false when the
// solve stops at k (converged, max_it, breakdown); workgroup 0's thread 0 records.
template <int BS>
__device__ __forceinline__ bool cg1_scalars(Ctl *c, int k, const double *partPrev, int np, bool writer,
                                            double &alpha, double &beta) {
    if (k == c->first) {
        alpha = 0.0;
        const double rho = c->s[CGF_RR + (k & 1)];
        beta = c->s[CGF_FRESH] != 0.0 ? 0.0 : rho / c->s[CGF_RR + ((k + 1) & 1)];
        return true;
    }
    double d[3];
    block_sum_parts3<BS>(partPrev, np, d);
    const double rho_prev = c->s[CGF_RR + ((k + 1) & 1)];
    if (bad(d[0])) {
        if (writer) {
            c->status = ST_BREAKDOWN; c->result_it = k - 1; c->it = k; c->done = 1;
            c->s[CG1_ALPHA] = 0.0;
        }
        return false;
    }
    // the recurrence must use the alpha the FP32 update applies: with the unrounded
    // FP64 alpha, rho drifts from |r|^2 by ~1e-4 relative within a pass and beta's
    // error costs ~10 % more iterations (reproduced in a CPU model of the iteration)
    alpha = (double)(float)(rho_prev / d[0]);
    double rho = rho_prev - 2.0 * alpha * d[1] + alpha * alpha * d[2];
    if (rho < 0.0) rho = 0.0;
    if (cgf_check(c, k, rho, writer) != ST_RUNNING) {
        if (writer) c->s[CG1_ALPHA] = alpha;
        return false;
    }
    if (writer) c->s[CGF_RR + (k & 1)] = rho;
    beta = rho / rho_prev;
    return true;
}

// Kernel k on the SELL-64 layout or its slice-diagonal copy (DIA; lane per row,
// workgroup g owns `chunk` slices). R, Q, P hold two buffers each (parity), n entries
// apart; x32 is updated in place every second kernel: kernel k owes x the update
// alpha_{k-1} p_{k-1}; an even kernel defers it (alpha in CG1_ALPHAD) and the next odd
// kernel applies both, reading p_{k-2} from the buffer it is about to overwrite (the
// same two roundings in the same order, one x read and write per two iterations).
template <typename VT, typename IT, int BS, bool DIA = false>
__global__ __launch_bounds__(BS) void k_cg1_sell(int n, int nsl, int chunk, const int *__restrict__ rp,
                                                  const int *__restrict__ so, const VT *__restrict__ val,
                                                  const IT *__restrict__ col, float *x, float *R, float *Q,
                                                  float *P, Ctl *c, int j, const double *partPrev,
                                                  int np, double *partOut) {
    if (ctl_stopped(c)) return;
    const int k = c->base + j;
    const bool writer = blockIdx.x == 0 && threadIdx.x == 0;
    const bool start = (k == c->first);            // q = A p explicitly
    const size_t po = (size_t)((k + 1) & 1) * n, pn = (size_t)(k & 1) * n;
    const float *rp_ = R + po, *qp_ = Q + po, *pp_ = P + po;
    float *rn_ = R + pn, *qn_ = Q + pn, *pn_ = P + pn;
    const int g = xcd_block(blockIdx.x, gridDim.x);
    const int l = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int send = min(nsl, (g + 1) * chunk);
    const int sl0 = g * chunk + w;
    // DIA: the wave's first slice has its first eight positions (values, and the
    // gathered r and q -- or p at a pass's first kernel) loaded before the previous
    // kernel's partials are summed, so their latency overlaps that prologue.
    float v0[8], r0[8], g0[8];
    int w0 = 0;
    if constexpr (DIA) {
        const int row = sl0 * 64 + l;
        if (sl0 < send && row < n) {
            const int s0 = so[sl0];
            w0 = (so[sl0 + 1] - s0) >> 6;
            if (w0 > 0) {
                const int *o = (const int *)col + (s0 >> 6);
                const float *gs = start ? pp_ : qp_;
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int cc = min(max(row + o[u], 0), n - 1);
                    v0[u] = (float)val[s0 + l + 64 * min(u, w0 - 1)];
                    r0[u] = rp_[cc];
                    g0[u] = gs[cc];
                }
            }
        }
    }
    double ad, bd;
    if (!cg1_scalars<BS>(c, k, partPrev, np, writer, ad, bd)) return;
    const float al = (float)ad, be = (float)bd;
    const bool odd = (k & 1) != 0;
    // alpha_{k-2} of the update kernel k-1 deferred (none at a pass's first kernel)
    const float ald = (odd && !start) ? (float)c->s[CG1_ALPHAD] : 0.0f;
    if (!odd && writer) c->s[CG1_ALPHAD] = ad;
    double dv[3] = {0.0, 0.0, 0.0};
    for (int sl = sl0; sl < send; sl += BS / 64) {
        const int row = sl * 64 + l;
        if (row >= n) continue;
        float acc = 0.0f;
        if (DIA && sl == sl0 && w0 > 0) {
            // the prefetched positions, then the slice's positions from 8 on
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const float gv = start ? fmaf(be, g0[u], r0[u]) : fmaf(-al, g0[u], r0[u]);
                acc = add_rn(acc, mul_rn(u < w0 ? v0[u] : 0.0f, gv));
            }
            const int s0 = so[sl], wd = (so[sl + 1] - s0) >> 6;
            const int *o = (const int *)col + (s0 >> 6);
            for (int k0 = 8; k0 < wd; k0 += 8) {
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int kk = min(k0 + u, wd - 1);
                    const int cc = min(max(row + o[k0 + u], 0), n - 1);
                    const float v = (float)val[s0 + l + 64 * kk];
                    const float gv = start ? fmaf(be, pp_[cc], rp_[cc]) : fmaf(-al, qp_[cc], rp_[cc]);
                    acc = add_rn(acc, mul_rn(k0 + u < wd ? v : 0.0f, gv));
                }
            }
        } else if (start)
            sell_row_entries<IT, DIA, 8>(n, row, l, sl, rp, so, col, [&](int e, int cc, bool live) {
                const float v = (float)val[e];
                acc = add_rn(acc, mul_rn(live ? v : 0.0f, fmaf(be, pp_[cc], rp_[cc])));
            });
        else
            sell_row_entries<IT, DIA, 8>(n, row, l, sl, rp, so, col, [&](int e, int cc, bool live) {
                const float v = (float)val[e];
                acc = add_rn(acc, mul_rn(live ? v : 0.0f, fmaf(-al, qp_[cc], rp_[cc])));
            });
        const float pold = pp_[row], qold = qp_[row];
        const float rr = fmaf(-al, qold, rp_[row]);
        const float pnew = fmaf(be, pold, rr);
        const float qnew = start ? acc : fmaf(be, qold, acc);
        if (odd) {
            float xv = x[row];
            if (ald != 0.0f) xv = fmaf(ald, pn_[row], xv);   // p_{k-2}, before p_k replaces it
            x[row] = fmaf(al, pold, xv);
        }
        rn_[row] = rr;
        pn_[row] = pnew;
        qn_[row] = qnew;
        dv[0] += (double)pnew * (double)qnew;
        dv[1] += (double)rr * (double)qnew;
        dv[2] += (double)qnew * (double)qnew;
    }
    block_partial<3, BS>(dv, partOut);
}

// End of a graph replay of K single-kernel iterations: test of iteration base+K.
[[maybe_unused]] static __global__ __launch_bounds__(256) void k_cg1_tail(Ctl *c, int K, const double *partPrev, int np) {
    if (ctl_stopped(c)) return;
    const int k = c->base + K;
    double ad, bd;
    if (!cg1_scalars<256>(c, k, partPrev, np, threadIdx.x == 0, ad, bd)) return;
    if (threadIdx.x == 0) c->base = k;
}
