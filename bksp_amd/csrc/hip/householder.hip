// householder.hip -- thin QR of a tall-skinny panel with LAPACK's Householder
// convention, for the exported myqr / zqreco (reference src/myqr.c:4-39, 41-76:
// LAPACKE_dgeqrf + dorgqr, zgeqrf + zungqr).
//
// For fewer than 32 columns LAPACK's dgeqrf/dorgqr run their unblocked forms
// dgeqr2/dorg2r (the block size is 32), which this file follows step for step:
//
//   factor, column p = 0..n-1 (dgeqr2 / dlarfg / dlarf):
//     x = V(p+1:m, p), alpha = V(p,p), xnorm = ||x||;
//     xnorm == 0 (and Im alpha == 0): tau = 0, H = I;
//     else beta = -sign(Re alpha) |(alpha, xnorm)|, tau = (beta - alpha)/beta,
//          v = [1; x / (alpha - beta)], V(p,p) = beta;
//     the columns to the right get H^H: C -= conj(tau) v (v^H C).
//   R = upper triangle of V (diagonal = the betas; negative where LAPACK's is).
//   form Q, column p = n-1..0 (dorg2r / zung2r):
//     C = V(p:m, p+1:n) -= tau v (v^H C); V(p+1:m, p) *= -tau; V(p,p) = 1 - tau;
//     V(0:p-1, p) = 0.
//
// Each step is two passes over the rows: a reduction kernel forms v^H C for up to
// 16 columns per launch (per-workgroup partials, summed in a fixed order by a
// one-workgroup kernel that also runs the scalar algebra on the device), then an
// update kernel applies the reflector. The panel is column-major with ld = m, as
// myqr receives it. Rank-deficient and ill-conditioned panels are handled as
// LAPACK handles them: the reflectors stay orthogonal whatever the column norms.
// Differences from LAPACK: ||x|| is the square root of a tree-summed sum of
// squares (dnrm2 scales to avoid overflow beyond ~1e154), and dlarfg's rescaling
// loop for |beta| < safmin/eps (~1e-292) is not done.
#include <algorithm>
#include <type_traits>
#include "bkh_host.hpp"

namespace {

struct Cx {
    double re, im;
};

__host__ __device__ inline double re_of(double a) { return a; }
__host__ __device__ inline double re_of(Cx a) { return a.re; }
__host__ __device__ inline double im_of(double) { return 0.0; }
__host__ __device__ inline double im_of(Cx a) { return a.im; }
__device__ inline double conj_of(double a) { return a; }
__device__ inline Cx conj_of(Cx a) { return {a.re, -a.im}; }
__device__ inline double mul(double a, double b) { return a * b; }
__device__ inline Cx mul(Cx a, Cx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ inline double add(double a, double b) { return a + b; }
__device__ inline Cx add(Cx a, Cx b) { return {a.re + b.re, a.im + b.im}; }
__device__ inline double sub(double a, double b) { return a - b; }
__device__ inline Cx sub(Cx a, Cx b) { return {a.re - b.re, a.im - b.im}; }
template <class T> __device__ inline T from_parts(double r, double i);
template <> __device__ inline double from_parts<double>(double r, double) { return r; }
template <> __device__ inline Cx from_parts<Cx>(double r, double i) { return {r, i}; }

// conj(a) * b, the term of v^H C.
__device__ inline double cmul(double a, double b) { return a * b; }
__device__ inline Cx cmul(Cx a, Cx b) { return {a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re}; }

// LAPACK dlapy2 / dlapy3: sqrt of a sum of squares without destructive overflow.
__device__ inline double lapy2(double x, double y) {
    const double xa = fabs(x), ya = fabs(y);
    const double w = fmax(xa, ya), z = fmin(xa, ya);
    if (z == 0.0 || w > 1.79769313486231571e308) return w;
    const double q = z / w;
    return w * sqrt(1.0 + q * q);
}
__device__ inline double lapy3(double x, double y, double z) {
    const double xa = fabs(x), ya = fabs(y), za = fabs(z);
    const double w = fmax(xa, fmax(ya, za));
    if (w == 0.0) return xa + ya + za;
    const double a = xa / w, b = ya / w, c = za / w;
    return w * sqrt(a * a + b * b + c * c);
}

// 1 / d for a complex d by Smith's algorithm (what zladiv computes, up to rounding).
__device__ inline Cx recip(Cx d) {
    if (fabs(d.re) >= fabs(d.im)) {
        const double r = d.im / d.re, den = d.re + d.im * r;
        return {1.0 / den, -r / den};
    }
    const double r = d.re / d.im, den = d.re * r + d.im;
    return {r / den, -1.0 / den};
}
__device__ inline double recip(double d) { return 1.0 / d; }

constexpr int HH_NC = 16;     // columns per reduction launch
constexpr int HH_BS = 256;

// s_j = sum_{i > p} conj(x_i) V(i, j) for j in [j0, jn), x = V(:, p); one partial per
// workgroup (2 doubles per column: real, imaginary).
template <class T>
__global__ __launch_bounds__(HH_BS) void k_hh_dots(int m, const T *V, int p, int j0, int jn, double *part) {
    double ar[HH_NC], ai[HH_NC];
#pragma unroll
    for (int c = 0; c < HH_NC; c++) { ar[c] = 0.0; ai[c] = 0.0; }
    const size_t ld = (size_t)m;
    for (long long i = (long long)p + 1 + (long long)blockIdx.x * HH_BS + threadIdx.x; i < m;
         i += (long long)gridDim.x * HH_BS) {
        const T x = V[(size_t)p * ld + i];
#pragma unroll
        for (int c = 0; c < HH_NC; c++) {
            if (j0 + c < jn) {
                const T t = cmul(x, V[(size_t)(j0 + c) * ld + i]);
                ar[c] += re_of(t);
                ai[c] += im_of(t);
            }
        }
    }
    __shared__ double sh[2 * HH_NC * (HH_BS / BKH_WAVE)];
    double v[2 * HH_NC];
#pragma unroll
    for (int c = 0; c < HH_NC; c++) { v[2 * c] = ar[c]; v[2 * c + 1] = ai[c]; }
    block_sum<2 * HH_NC, HH_BS>(v, sh);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int e = 0; e < 2 * HH_NC; e++) part[(size_t)blockIdx.x * 2 * HH_NC + e] = v[e];
    }
}

// Scalar state of the factorization, device-resident: tau[n], coef[n] (the row
// vector conj(tau) v^H C, or tau v^H C while forming Q) and the current reflector's
// scale 1/(alpha - beta).
template <class T>
struct HhState {
    T *tau;
    T *coef;
    T *scale;
};

// Sums the partials of one reduction launch in a fixed order and runs the step's
// scalar algebra. form_q = 0: factor step (dlarfg when the chunk starts at column
// p); form_q = 1: Q-forming step.
template <class T>
__global__ __launch_bounds__(HH_BS) void k_hh_fin(int m, T *V, int p, int j0, int jn, const double *part,
                                                   int nblocks, HhState<T> st, int form_q) {
    double acc[2 * HH_NC];
    __shared__ double sh[2 * HH_NC * 4];
#pragma unroll
    for (int e = 0; e < 2 * HH_NC; e++) acc[e] = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += HH_BS) {
#pragma unroll
        for (int e = 0; e < 2 * HH_NC; e++) acc[e] += part[(size_t)b * 2 * HH_NC + e];
    }
    block_sum<2 * HH_NC, HH_BS>(acc, sh);
    if (threadIdx.x != 0) return;
    const size_t ld = (size_t)m;
    T *colp = V + (size_t)p * ld;
    if (!form_q) {
        if (j0 == p) {
            // dlarfg / zlarfg on alpha = V(p,p), x = V(p+1:m, p)
            const T alpha = colp[p];
            const double ar = re_of(alpha), ai = im_of(alpha);
            const double xnorm = sqrt(acc[0]);
            if (xnorm == 0.0 && ai == 0.0) {
                st.tau[p] = from_parts<T>(0.0, 0.0);
                *st.scale = from_parts<T>(1.0, 0.0);       // x stays as it is
            } else {
                // dlarfg takes |(alpha, xnorm)|, zlarfg |(Re alpha, Im alpha, xnorm)|
                const double h = std::is_same<T, Cx>::value ? lapy3(ar, ai, xnorm) : lapy2(ar, xnorm);
                const double beta = -copysign(h, ar);
                st.tau[p] = from_parts<T>((beta - ar) / beta, -ai / beta);
                *st.scale = recip(sub(alpha, from_parts<T>(beta, 0.0)));
                colp[p] = from_parts<T>(beta, 0.0);
            }
        }
        const T ctau = conj_of(st.tau[p]), sc = conj_of(*st.scale);
        for (int j = (j0 == p ? p + 1 : j0); j < jn; j++) {
            const int c = j - j0;
            const T s = add(V[(size_t)j * ld + p], mul(sc, from_parts<T>(acc[2 * c], acc[2 * c + 1])));
            st.coef[j] = mul(ctau, s);
        }
    } else {
        const T tau = st.tau[p];
        for (int j = j0; j < jn; j++) {
            const int c = j - j0;
            const T s = add(V[(size_t)j * ld + p], from_parts<T>(acc[2 * c], acc[2 * c + 1]));
            st.coef[j] = mul(tau, s);
        }
    }
}

// Applies the step's reflector to the columns right of p (rows p..m-1) and
// finishes column p: factor step v_i = scale x_i stored below the diagonal; Q step
// V(i,p) = -tau v_i, V(p,p) = 1 - tau, rows above p zeroed.
template <class T>
__global__ __launch_bounds__(HH_BS) void k_hh_apply(int m, int n, T *V, int p, HhState<T> st, int form_q) {
    const size_t ld = (size_t)m;
    const T sc = *st.scale, tau = st.tau[p];
    for (long long i = (long long)blockIdx.x * HH_BS + threadIdx.x; i < m; i += (long long)gridDim.x * HH_BS) {
        if (i < p) {
            if (form_q) V[(size_t)p * ld + i] = from_parts<T>(0.0, 0.0);
            continue;
        }
        T v;
        if (i == p) {
            v = from_parts<T>(1.0, 0.0);
        } else {
            const T x = V[(size_t)p * ld + i];
            v = form_q ? x : mul(sc, x);
        }
        for (int j = p + 1; j < n; j++) {
            T *cij = V + (size_t)j * ld + i;
            *cij = sub(*cij, mul(st.coef[j], v));
        }
        if (form_q) {
            V[(size_t)p * ld + i] = (i == p) ? sub(from_parts<T>(1.0, 0.0), tau)
                                             : mul(from_parts<T>(-re_of(tau), -im_of(tau)), v);
        } else if (i > p) {
            V[(size_t)p * ld + i] = v;
        }
    }
}

template <class T>
int hh_qr(int m, int n, T *d_V, T *R) {
    if (m < 0) return -1;                     // dorgqr's argument checks (info = -1, -2)
    if (n < 0 || n > m) return -2;
    if (n == 0) return 0;
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    hipStream_t s = t->stream;
    const int G = std::max(1, std::min(1024, (int)(((long long)m + HH_BS * 4 - 1) / (HH_BS * 4))));
    const size_t bytes = carve_size(sizeof(double) * (size_t)G * 2 * HH_NC) + 2 * carve_size(sizeof(T) * n) +
                         carve_size(sizeof(T)) + carve_size(sizeof(T) * (size_t)n * n);
    void *ws = bkh_workspace(bytes);
    if (!ws) return BKH_DEVICE_ERROR;
    Carve cv(ws);
    double *part = cv.take<double>((size_t)G * 2 * HH_NC);
    HhState<T> st;
    st.tau = cv.take<T>(n);
    st.coef = cv.take<T>(n);
    st.scale = cv.take<T>(1);
    T *d_R = cv.take<T>((size_t)n * n);
    const int GA = vec_grid(m, HH_BS, 4);
    for (int p = 0; p < n; p++) {             // dgeqr2
        for (int j0 = p; j0 < n; j0 += HH_NC) {
            const int jn = std::min(n, j0 + HH_NC);
            hipLaunchKernelGGL(k_hh_dots<T>, dim3(G), dim3(HH_BS), 0, s, m, (const T *)d_V, p, j0, jn, part);
            hipLaunchKernelGGL(k_hh_fin<T>, dim3(1), dim3(HH_BS), 0, s, m, d_V, p, j0, jn, (const double *)part, G,
                               st, 0);
        }
        hipLaunchKernelGGL(k_hh_apply<T>, dim3(GA), dim3(HH_BS), 0, s, m, n, d_V, p, st, 0);
    }
    LAUNCH_OK();
    // R: the upper triangle of the factored panel (src/myqr.c:11-22 layout: r[i*n + j]
    // = R(j, i) for j <= i, 0 below the diagonal)
    HIP_OK(hipMemcpy2DAsync(d_R, sizeof(T) * n, d_V, sizeof(T) * (size_t)m, sizeof(T) * n, n,
                            hipMemcpyDeviceToDevice, s));
    for (int p = n - 1; p >= 0; p--) {        // dorg2r
        for (int j0 = p + 1; j0 < n; j0 += HH_NC) {
            const int jn = std::min(n, j0 + HH_NC);
            hipLaunchKernelGGL(k_hh_dots<T>, dim3(G), dim3(HH_BS), 0, s, m, (const T *)d_V, p, j0, jn, part);
            hipLaunchKernelGGL(k_hh_fin<T>, dim3(1), dim3(HH_BS), 0, s, m, d_V, p, j0, jn, (const double *)part, G,
                               st, 1);
        }
        hipLaunchKernelGGL(k_hh_apply<T>, dim3(GA), dim3(HH_BS), 0, s, m, n, d_V, p, st, 1);
    }
    LAUNCH_OK();
    HIP_OK(hipMemcpyAsync(R, d_R, sizeof(T) * (size_t)n * n, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++) R[(size_t)i * n + j] = T{};
    return 0;
}

}  // namespace

extern "C" int bkh_qr_householder(int m, int n, double *d_V, double *R) {
    if ((m > 0 && n > 0) && (!d_V || !R)) return BKH_BAD_ARGUMENT;
    return hh_qr<double>(m, n, d_V, R);
}

extern "C" int bkh_zqr_householder(int m, int n, double *d_V, double *R) {
    if ((m > 0 && n > 0) && (!d_V || !R)) return BKH_BAD_ARGUMENT;
    return hh_qr<Cx>(m, n, (Cx *)d_V, (Cx *)R);
}
