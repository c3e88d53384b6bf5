// bkh_small.hpp -- small dense s x s FP64 kernels (s <= 16) executed by one
// 256-thread workgroup on column-major matrices in LDS or global memory.
// They replace the reference's LAPACKE calls on block-size matrices:
//   dgetrf/dgetri (blcg.c:81-82, blbicgstabl.c:147-148)  -> sm_inverse
//   dgesv (blbicgstabl.c:105, :124)                       -> sm_inverse + sm_mul
//   dgeqrf+dorgqr R factor (myqr.c)                       -> sm_chol_upper (CholQR)
// Every routine ends with a barrier so results are visible to the block. W = true:
// the routine runs on the first wave only (the other waves have left; s <= 8, so every
// index fits in 64 lanes) and its barriers are wave barriers with LDS fences.
#pragma once
#include "bkh_core.hpp"

#define SM_MAX 16
#define SM_N (SM_MAX * SM_MAX)

template <bool W>
__device__ __forceinline__ void sm_sync() {
    if constexpr (W) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

// C = op(A) * op(B) (s x s); ta/tb: 0 = as stored, 1 = transposed. C may not alias A/B.
template <bool W = false>
__device__ __forceinline__ void sm_mul(int s, const double *A, int ta, const double *B, int tb,
                                       double *C) {
    const int t = threadIdx.x;
    if (t < s * s) {
        const int i = t % s, j = t / s;
        double acc = 0.0;
        for (int k = 0; k < s; k++) {
            double a = ta ? A[k + i * s] : A[i + k * s];
            double b = tb ? B[j + k * s] : B[k + j * s];
            acc = fma(a, b, acc);
        }
        C[i + j * s] = acc;
    }
    sm_sync<W>();
}

template <bool W = false>
__device__ __forceinline__ void sm_copy(int s, const double *A, double *B) {
    const int t = threadIdx.x;
    if (t < s * s) B[t] = A[t];
    sm_sync<W>();
}

template <bool W = false>
__device__ __forceinline__ void sm_eye(int s, double *A, double diag) {
    const int t = threadIdx.x;
    if (t < s * s) A[t] = (t % s == t / s) ? diag : 0.0;
    sm_sync<W>();
}

// Upper Cholesky factor R of SPD G (G = R^T R). `shift` is added to the diagonal
// first. Returns false (uniformly) if a pivot is not positive.
template <bool W = false>
__device__ __forceinline__ bool sm_chol_upper(int s, const double *G, double shift, double *R,
                                              int *flag) {
    const int t = threadIdx.x;
    if (t == 0) *flag = 1;
    if (t < s * s) {
        const int i = t % s, j = t / s;
        R[t] = (i <= j) ? G[t] + ((i == j) ? shift : 0.0) : 0.0;
    }
    sm_sync<W>();
    for (int k = 0; k < s; k++) {
        if (t == 0) {
            double d = R[k + k * s];
            if (!(d > 0.0)) *flag = 0;
            R[k + k * s] = d > 0.0 ? sqrt(d) : 1.0;
        }
        sm_sync<W>();
        const double rkk = R[k + k * s];
        if (t > k && t < s) R[k + t * s] /= rkk;                 // row k of R
        sm_sync<W>();
        if (t < s * s) {                                          // trailing update
            const int i = t % s, j = t / s;
            if (i > k && j >= i) R[i + j * s] -= R[k + i * s] * R[k + j * s];
        }
        sm_sync<W>();
    }
    return *flag != 0;
}

// Inverse of an upper-triangular R (column j by back substitution, thread j).
template <bool W = false>
__device__ __forceinline__ void sm_triu_inv(int s, const double *R, double *Ri) {
    const int t = threadIdx.x;
    if (t < s * s) Ri[t] = 0.0;
    sm_sync<W>();
    if (t < s) {
        const int j = t;
        for (int i = j; i >= 0; i--) {
            double acc = (i == j) ? 1.0 : 0.0;
            for (int k = i + 1; k <= j; k++) acc -= R[i + k * s] * Ri[k + j * s];
            Ri[i + j * s] = acc / R[i + i * s];
        }
    }
    sm_sync<W>();
}

// General inverse by Gauss-Jordan elimination with partial pivoting (works on a
// copy in `work`, 2*s*s doubles). Returns false (uniformly) if singular.
template <bool W = false>
__device__ __forceinline__ bool sm_inverse(int s, const double *A, double *Ai, double *work,
                                           int *piv) {
    const int t = threadIdx.x;
    double *M = work;   // s x 2s augmented [A | I], column-major with ld s
    if (t < s * s) {
        M[t] = A[t];
        M[s * s + t] = (t % s == t / s) ? 1.0 : 0.0;
    }
    if (t == 0) piv[1] = 1;
    sm_sync<W>();
    for (int k = 0; k < s; k++) {
        if (t == 0) {
            int p = k;
            double best = fabs(M[k + k * s]);
            for (int i = k + 1; i < s; i++)
                if (fabs(M[i + k * s]) > best) { best = fabs(M[i + k * s]); p = i; }
            piv[0] = p;
            if (!(best > 0.0)) piv[1] = 0;
        }
        sm_sync<W>();
        const int p = piv[0];
        if (p != k && t < 2 * s) {
            double a = M[k + t * s], b = M[p + t * s];
            M[k + t * s] = b; M[p + t * s] = a;
        }
        sm_sync<W>();
        const double inv = piv[1] ? 1.0 / M[k + k * s] : 0.0;
        if (t < 2 * s) M[k + t * s] *= inv;
        sm_sync<W>();
        // eliminate column k from the other rows: thread (i, c)
        for (int e = t; e < s * 2 * s; e += W ? 64 : (int)blockDim.x) {
            const int i = e % s, c = e / s;
            if (i != k && c != k) M[i + c * s] -= M[i + k * s] * M[k + c * s];
        }
        sm_sync<W>();
        if (t < s && t != k) M[t + k * s] = 0.0;
        sm_sync<W>();
    }
    if (t < s * s) Ai[t] = M[s * s + t];
    sm_sync<W>();
    return piv[1] != 0;
}

// Frobenius norm squared (thread 0 result, broadcast through `red`).
template <bool W = false>
__device__ __forceinline__ double sm_fro2(int s, const double *A, double *red) {
    if (threadIdx.x == 0) {
        double acc = 0.0;
        for (int k = 0; k < s * s; k++) acc += A[k] * A[k];
        *red = acc;
    }
    sm_sync<W>();
    return *red;
}

// ---- register forms for s <= 8 on one wave: lane j holds column j of an s x s
// column-major matrix in `double v[S]` (lanes >= s carry values nobody reads), and an
// entry of another column arrives by a readlane broadcast instead of an LDS round trip
// and a barrier. Each routine performs the operations of its LDS counterpart above in
// the same order with the same roundings, so the results are bitwise the same.
__device__ __forceinline__ double rg_bcast(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// Column j (this lane) of the upper Cholesky factor of G (sm_chol_upper): in, column j
// of G; out, column j of R. Returns false (uniformly) if a pivot is not positive.
template <int S>
__device__ __forceinline__ bool rg_chol_upper(double (&r)[S], double shift) {
    const int j = (int)(threadIdx.x & 63);
#pragma unroll
    for (int i = 0; i < S; i++) r[i] = (i <= j) ? r[i] + ((i == j) ? shift : 0.0) : 0.0;
    bool ok = true;
#pragma unroll
    for (int k = 0; k < S; k++) {
        const double d = rg_bcast(r[k], k);
        if (!(d > 0.0)) ok = false;
        const double rk = d > 0.0 ? sqrt(d) : 1.0;
        if (j == k) r[k] = rk;
        if (j > k) r[k] = r[k] / rk;
#pragma unroll
        for (int i = k + 1; i < S; i++) {
            const double rki = rg_bcast(r[k], i);
            if (i <= j) r[i] = r[i] - rki * r[k];
        }
    }
    return ok;
}

// Column j of R^-1 for upper-triangular R held by columns (sm_triu_inv).
template <int S>
__device__ __forceinline__ void rg_triu_inv(const double (&r)[S], double (&ri)[S]) {
    const int j = (int)(threadIdx.x & 63);
#pragma unroll
    for (int i = S - 1; i >= 0; i--) {
        double acc = (i == j) ? 1.0 : 0.0;
#pragma unroll
        for (int k = i + 1; k < S; k++) {
            const double rik = rg_bcast(r[i], k);
            if (k <= j) acc -= rik * ri[k];
        }
        ri[i] = (i <= j) ? acc / rg_bcast(r[i], i) : 0.0;
    }
}

// Column j of C = A B (sm_mul with ta = tb = 0: fma chain over k): a, b column j of A
// and B on this lane; A's columns are read from lanes a0 .. a0 + S - 1.
template <int S>
__device__ __forceinline__ void rg_mul(const double (&a)[S], int a0, const double (&b)[S], double (&c)[S]) {
#pragma unroll
    for (int i = 0; i < S; i++) {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < S; k++) acc = fma(rg_bcast(a[i], a0 + k), b[k], acc);
        c[i] = acc;
    }
}

// Gauss-Jordan with partial pivoting on [A | I] (sm_inverse): lane c < 2S holds
// column c of the augmented matrix; on return lanes S .. 2S - 1 hold A^-1's columns.
// Returns false (uniformly) if a pivot was zero.
template <int S>
__device__ __forceinline__ bool rg_inverse(double (&m)[S]) {
    const int c = (int)(threadIdx.x & 63);
    bool ok = true;
#pragma unroll
    for (int k = 0; k < S; k++) {
        int p = k;
        double best = fabs(rg_bcast(m[k], k));
#pragma unroll
        for (int i = k + 1; i < S; i++) {
            const double v = fabs(rg_bcast(m[i], k));
            if (v > best) { best = v; p = i; }
        }
        if (!(best > 0.0)) ok = false;
        // row swap: p is uniform, so each candidate is a scalar branch over fixed
        // registers (no dynamically indexed array)
#pragma unroll
        for (int i = k + 1; i < S; i++)
            if (i == p) {
                const double tmp = m[k];
                m[k] = m[i];
                m[i] = tmp;
            }
        const double inv = ok ? 1.0 / rg_bcast(m[k], k) : 0.0;
        m[k] *= inv;
#pragma unroll
        for (int i = 0; i < S; i++) {
            if (i == k) continue;
            const double mik = rg_bcast(m[i], k);
            if (c != k) m[i] -= mik * m[k];
        }
        if (c == k) {
#pragma unroll
            for (int i = 0; i < S; i++)
                if (i != k) m[i] = 0.0;
        }
    }
    return ok;
}
