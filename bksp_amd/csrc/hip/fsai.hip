// fsai.hip -- learned factorised sparse approximate inverse (FSAI) preconditioner and
// preconditioned CG. No reference counterpart (the reference has no preconditioners).
//
// Model. For every edge (i, j), j <= i, of tril(A) an MLP 8 -> 64 -> 64 -> 1 (tanh)
// maps eight local features to h, and the factor G (pattern of tril(A)) is
//   g_ii = softplus(h + c0) / sqrt(a_ii),  g_ij = h / sqrt(a_jj),  c0 = log(e - 1),
// so M^-1 = G^T G is SPD for any weights and h = 0 gives Jacobi. Weights live in
// FP32 (Adam master copy) and enter the matrix cores as BF16 with FP32 accumulation.
//
// MLP on v_mfma_f32_16x16x16bf16_1k, one 16-edge tile per wave step, computed
// transposed (hidden x edges): the 16x16 result layout of one layer (lane L, register
// r: row 4*(L/16) + r, column L%16) is exactly the B-operand layout of the next, so
// activations never leave registers in the forward pass. The backward pass
// recomputes the forward, takes dL/dh per edge, runs dH1 = W2 dZ2 on the matrix cores
// the same way, and accumulates the weight gradients (sums over edges) with MFMAs
// whose operands are staged through a per-wave LDS transpose.
//
// Loss (Kaporin condition number of G A G^T, the classical FSAI objective, with the
// trace estimated on probe vectors z_p -- Rademacher vectors and residuals of the
// solver itself): L = log(T/n) - (2/n) sum_i log g_ii, T = (1/P) sum_p w_p.(A w_p),
// w_p = G^T z_p; dL/dg_ij = (2/(P T)) sum_p z_p[i] (A w_p)_j - [i == j] 2/(n g_ii).
#include <math.h>
#include <vector>
#include "bkh_cgfused.hpp"
#include "bkh_host.hpp"
#include "bkh_solver_common.hpp"
#include "bkh_spmv.hpp"

#define FS_H 64
#define FS_NF 8
#define FS_W1 0
#define FS_B1 (FS_W1 + FS_NF * FS_H)
#define FS_W2 (FS_B1 + FS_H)
#define FS_B2 (FS_W2 + FS_H * FS_H)
#define FS_W3 (FS_B2 + FS_H)
#define FS_B3 (FS_W3 + FS_H)
#define FS_NP (FS_B3 + 1)
#define FS_BLOCKS 256
#define FS_WAVES (FS_BLOCKS * 4)

typedef short s4v __attribute__((ext_vector_type(4)));
typedef short s8v __attribute__((ext_vector_type(8)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef float f4v __attribute__((ext_vector_type(4)));

static const double FS_C0 = 0.54132485461291810;   // log(e - 1)

struct bkh_fsai_model {
    float *params = nullptr, *m = nullptr, *v = nullptr;
    double *grad = nullptr;
    float *part = nullptr;        // per-wave gradient partials
    int step = 0;
};

struct bkh_fsai_field {
    int n = 0;
    long long nnzl = 0;
    bkh_matrix *A = nullptr;      // expanded symmetric matrix
    bkh_matrix *G = nullptr;      // factor (pattern tril(A)); G->T = G^T
    int *rpl = nullptr, *cil = nullptr, *erow = nullptr, *perm = nullptr;
    double *vl = nullptr;
    double4 *rows = nullptr;      // sqrt(a_ii), coupling sum, lower count, diagonal excess
    double2 *rowx = nullptr;      // 1 / sqrt(a_ii), log(a_ii): per-row parts of the edge features
    double *Z = nullptr, *W = nullptr, *U = nullptr, *gE = nullptr;
    int P = 0;
};

// FP32 -> BF16 with round-to-nearest-even (gfx950's v_cvt_pk_bf16_f32).
__device__ __forceinline__ short f2bf(float x) { return __builtin_bit_cast(short, (__bf16)x); }

__device__ __forceinline__ float bf_round(float x) { return __uint_as_float(((unsigned)(unsigned short)f2bf(x)) << 16); }

__device__ __forceinline__ f4v mfma_bf16(s4v a, s4v b, f4v c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// gfx950's v_mfma_f32_16x16x32_bf16: A lane L holds A[L%16][8(L/16) .. 8(L/16)+7],
// B lane L holds B[8(L/16) .. 8(L/16)+7][L%16], D as the 16x16x16 form.
__device__ __forceinline__ f4v mfma_bf16_k32(s8v a, s8v b, f4v c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8v, a), __builtin_bit_cast(bf8v, b), c, 0, 0, 0);
}

// tanh in FP32 from the hardware exp2 and reciprocal: tanh x = 1 - 2 / (exp(2x) + 1),
// and the series x - x^3/3 + 2x^5/15 below |x| = 1/16, where the subtraction would
// cancel (series error < 17/315 x^7). Absolute error ~2e-7: below half a BF16 step of
// every activation the matrix cores consume, and 5 to 8 instructions where the libm
// tanhf spends about 40 (the MLP evaluates 32 per lane per 16-edge tile).
// softplus and the logistic function in FP32 from the hardware exp2 / log2 (the factor
// is stored in FP32 for the fused PCG and the MLP's output h is FP32 already).
__device__ __forceinline__ float softplus_f(float x) {
    const float e = __builtin_amdgcn_exp2f(1.4426950408889634f * x);
    const float l = 0.69314718055994531f * __builtin_amdgcn_logf(1.0f + e);
    return x > 20.0f ? x : (x < -15.0f ? e : l);
}
__device__ __forceinline__ float sigmoid_f(float x) {
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.4426950408889634f * x));
}

__device__ __forceinline__ float tanh_fast(float x) {
    const float e = __builtin_amdgcn_exp2f(2.8853900817779268f * x);   // exp(2x) = 2^(2x / ln 2)
    const float t = 1.0f - 2.0f * __builtin_amdgcn_rcpf(e + 1.0f);
    const float x2 = x * x;
    const float s = x * (1.0f - x2 * (0.33333333f - 0.13333333f * x2));
    return fabsf(x) < 0.0625f ? s : t;
}

// Row data from the expanded matrix (the oracle's orc_fsai_rows).
__global__ __launch_bounds__(256) void k_fsai_rows(int n, const int *rp, const int *ci, const double *v,
                                                   const double *diag, double4 *rows, double2 *rowx) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        double acc = 0.0, asum = 0.0;
        int nl = 0;
        const double di = diag[i];
        for (int k = rp[i]; k < rp[i + 1]; k++) {
            const int j = ci[k];
            if (j == i) continue;
            acc += fabs(v[k]) / sqrt(di * diag[j]);
            asum += fabs(v[k]);
            if (j < i) nl++;
        }
        rows[i] = make_double4(sqrt(di), acc, (double)nl, (di - asum) / di);
        rowx[i] = make_double2(1.0 / sqrt(di), log(di));
    }
}

// Weights staged in LDS as BF16: W2[i][o] (the A operand of the backward's
// dH1 = W2 dZ2), and W2T with its input index permuted
// for the forward layer 2 on the K = 32 matrix-core form: lane group g of layer 1's
// output holds, for row blocks m = 2q and 2q + 1, hidden units 16m + 4g + r (r < 4),
// which are exactly the eight B-operand values of K-block q when K index 8g + j of
// that block stands for unit 32q + 16(j / 4) + 4g + j % 4. w2p[o][32q + 8g + j] is W2
// at that unit, so the activations stay in registers between the layers.
// The forward-only kernel keeps w2p alone (9 KB of LDS, so occupancy is set by
// registers); w2, which only the backward reads, is one element there. Rows are
// padded (FS_P2 / FS_P shorts) so that the 16 rows one read instruction touches fall
// in different LDS banks (an unpadded 128-byte row stride puts all 16 in one bank
// group).
#define FS_P2 (FS_H + 8)     // w2p row stride (16-byte reads)
#define FS_P (FS_H + 4)      // w2 row stride (8-byte reads)
#define FS_TS 20             // row stride of the backward's per-wave transpose tiles
template <bool BWD>
struct alignas(16) FsSmem {
    short w2p[FS_H * FS_P2];           // first: 16-byte aligned for its 16-byte reads
    short w2[BWD ? FS_H * FS_P : 1];
};

__device__ __forceinline__ int fs_k32_unit(int kk) {
    const int q = kk >> 5, g = (kk >> 3) & 3, j = kk & 7;
    return 32 * q + 16 * (j >> 2) + 4 * g + (j & 3);
}

// The four features of lane group gq (0: f0..f3, 1: f4..f7, 2-3: zero padding); the
// logarithms of the diagonal are per row (rowx), the index distance's log2 in FP32
// (the feature is rounded to FP32 and then to BF16 either way).
__device__ __forceinline__ void fs_features(int i, int j, double aij, const double4 *rows, const double2 *rowx,
                                            float inv_log2n1, int gq, float (&f)[4]) {
    if (gq == 0) {
        const double2 xi = rowx[i], xj = rowx[j];
        f[0] = (float)(i == j ? 1.0 : 0.0);
        f[1] = (float)(aij * xi.x * xj.x);
        f[2] = (float)(0.25 * (xi.y - xj.y));
        f[3] = (float)rows[i].y;
    } else if (gq == 1) {
        const double4 ri = rows[i];
        f[0] = (float)rows[j].y;
        f[1] = __builtin_amdgcn_logf(1.0f + (float)(i - j)) * inv_log2n1;
        f[2] = (float)(0.25 * ri.z);
        f[3] = (float)ri.w;
    } else {
        f[0] = f[1] = f[2] = f[3] = 0.0f;
    }
}

// 1 / log2(1 + n), the scale of the index-distance feature.
static inline float fs_inv_log2n1(int n) { return (float)(1.0 / log2(1.0 + (double)n)); }

// Forward (BWD = false: writes the factor values) or forward + backward (BWD = true:
// per-wave gradient partials). Persistent waves stride over 16-edge tiles.
template <bool BWD>
__global__ __launch_bounds__(256, BWD ? 1 : 4) void k_fsai_mlp(int n, long long nnzl, const int *erow, const int *cil,
                                                  const double *vl, const double4 *rows, const double2 *rowx,
                                                  float inv_log2n1, const float *params,
                                                  const int *perm, double *gval, float *gval32, double *gtval,
                                                  float *gtval32, const double *gE, float *part) {
    __shared__ FsSmem<BWD> sm;
    __shared__ __attribute__((aligned(16))) short lds_all[4][BWD ? 3 * FS_H * FS_TS + 16 * FS_TS : 1];
    const int L = threadIdx.x & 63, w = threadIdx.x >> 6, gq = L >> 4, el = L & 15;
    for (int e = threadIdx.x; e < FS_H * FS_H; e += 256) {
        const int i = e / FS_H, o = e % FS_H;
        if constexpr (BWD) sm.w2[i * FS_P + o] = f2bf(params[FS_W2 + e]);   // [i][o]
        const int o2 = e / FS_H, kk = e % FS_H;
        sm.w2p[o2 * FS_P2 + kk] = f2bf(params[FS_W2 + fs_k32_unit(kk) * FS_H + o2]);   // [o][kk]
    }
    __syncthreads();
    // layer-1 A operand: W1^T[16m + L%16][4gq + j] = W1[4gq + j][16m + L%16]
    s4v A1[4];
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int f = 4 * gq + j;
            A1[m][j] = f < FS_NF ? f2bf(params[FS_W1 + f * FS_H + 16 * m + el]) : (short)0;
        }
    float b1[4][4], b2[4][4], w3b[4][4];
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int hh = 16 * m + 4 * gq + r;
            b1[m][r] = params[FS_B1 + hh];
            b2[m][r] = params[FS_B2 + hh];
            w3b[m][r] = bf_round(params[FS_W3 + hh]);
        }
    const float b3 = params[FS_B3];
    short *lds = lds_all[w];
    short *Sdz2 = lds, *Sa1 = lds + FS_H * FS_TS, *Sdz1 = lds + 2 * FS_H * FS_TS, *Sx = lds + 3 * FS_H * FS_TS;
    f4v gW2[4][4], gW1[4];
    float gw3[4][4], gb2[4][4], gb1[4][4], gb3 = 0.0f;
    if (BWD) {
#pragma unroll
        for (int a = 0; a < 4; a++) {
            gW1[a] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int b = 0; b < 4; b++) { gW2[a][b] = f4v{0.f, 0.f, 0.f, 0.f}; gw3[a][b] = gb2[a][b] = gb1[a][b] = 0.f; }
        }
    }
    const long long ntile = (nnzl + 15) / 16;
    const int gw = blockIdx.x * 4 + w, nw = gridDim.x * 4;
    for (long long tile = gw; tile < ntile; tile += nw) {
        const long long e = tile * 16 + el;
        const bool valid = e < nnzl;
        const int i = valid ? erow[e] : 0, j = valid ? cil[e] : 0;
        const double aij = valid ? vl[e] : 0.0;
        float f[4];
        if (valid) fs_features(i, j, aij, rows, rowx, inv_log2n1, gq, f);
        else f[0] = f[1] = f[2] = f[3] = 0.0f;
        s4v xb;
#pragma unroll
        for (int q = 0; q < 4; q++) xb[q] = f2bf(f[q]);
        // layer 1: H1^T = W1^T X^T (4 row blocks of 16 hidden units)
        float a1[4][4];
        s4v B2[4];
#pragma unroll
        for (int m = 0; m < 4; m++) {
            const f4v d = mfma_bf16(A1[m], xb, f4v{0.f, 0.f, 0.f, 0.f});
#pragma unroll
            for (int r = 0; r < 4; r++) {
                a1[m][r] = tanh_fast(d[r] + b1[m][r]);
                B2[m][r] = f2bf(a1[m][r]);
            }
        }
        // layer 2: H2^T = W2^T H1^T, two K = 32 steps per 16 output units
        s8v B2k[2];
#pragma unroll
        for (int q = 0; q < 2; q++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                B2k[q][j] = B2[2 * q][j];
                B2k[q][4 + j] = B2[2 * q + 1][j];
            }
        float a2[4][4];
#pragma unroll
        for (int mo = 0; mo < 4; mo++) {
            f4v acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int q = 0; q < 2; q++) {
                const s8v A2 = *(const s8v *)&sm.w2p[(16 * mo + el) * FS_P2 + 32 * q + 8 * gq];
                acc = mfma_bf16_k32(A2, B2k[q], acc);
            }
#pragma unroll
            for (int r = 0; r < 4; r++) a2[mo][r] = tanh_fast(acc[r] + b2[mo][r]);
        }
        // layer 3: h = w3 . H2 + b3 (reduce the four lane groups of each edge)
        float ps = 0.0f;
#pragma unroll
        for (int mo = 0; mo < 4; mo++)
#pragma unroll
            for (int r = 0; r < 4; r++) ps += a2[mo][r] * w3b[mo][r];
        ps += __shfl_xor(ps, 16, 64);
        ps += __shfl_xor(ps, 32, 64);
        const float h = ps + b3;
        if constexpr (!BWD) {
            if (valid && gq == 0) {
                const double hd = (double)h;
                const double g = (i == j) ? (double)softplus_f(h + (float)FS_C0) * rowx[i].x : hd * rowx[j].x;
                gval[e] = g;
                gval32[e] = (float)g;
                const int pe = perm[e];
                gtval[pe] = g;
                gtval32[pe] = (float)g;
            }
            continue;
        }
        // ---- backward: dL/dh from dL/dg through the parametrisation
        float dh = 0.0f;
        if (valid) {
            const double dg = gE[e];
            dh = (float)((i == j) ? dg * (double)sigmoid_f(h + (float)FS_C0) * rowx[i].x : dg * rowx[j].x);
        }
        if (gq == 0) gb3 += dh;
        float dz2[4][4];
        s4v Bdz[4];
#pragma unroll
        for (int mo = 0; mo < 4; mo++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                gw3[mo][r] += a2[mo][r] * dh;
                dz2[mo][r] = w3b[mo][r] * dh * (1.0f - a2[mo][r] * a2[mo][r]);
                gb2[mo][r] += dz2[mo][r];
                Bdz[mo][r] = f2bf(dz2[mo][r]);
            }
        // dH1^T = W2 dZ2^T
        float dz1[4][4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            f4v acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int mo = 0; mo < 4; mo++) {
                const s4v A3 = *(const s4v *)&sm.w2[(16 * q + el) * FS_P + 16 * mo + 4 * gq];
                acc = mfma_bf16(A3, Bdz[mo], acc);
            }
#pragma unroll
            for (int r = 0; r < 4; r++) {
                dz1[q][r] = acc[r] * (1.0f - a1[q][r] * a1[q][r]);
                gb1[q][r] += dz1[q][r];
            }
        }
        // weight gradients (sums over the tile's edges) via an LDS transpose
#pragma unroll
        for (int m = 0; m < 4; m++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = 16 * m + 4 * gq + r;
                Sdz2[row * FS_TS + el] = Bdz[m][r];
                Sa1[row * FS_TS + el] = B2[m][r];
                Sdz1[row * FS_TS + el] = f2bf(dz1[m][r]);
            }
#pragma unroll
        for (int q = 0; q < 4; q++) Sx[(4 * gq + q) * FS_TS + el] = xb[q];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int mo = 0; mo < 4; mo++) {
            const s4v Ad = *(const s4v *)&Sdz2[(16 * mo + el) * FS_TS + 4 * gq];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const s4v Bh = *(const s4v *)&Sa1[(16 * q + el) * FS_TS + 4 * gq];
                gW2[mo][q] = mfma_bf16(Ad, Bh, gW2[mo][q]);        // dW2^T[o][i]
            }
        }
        const s4v Bx = *(const s4v *)&Sx[el * FS_TS + 4 * gq];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const s4v Ad1 = *(const s4v *)&Sdz1[(16 * q + el) * FS_TS + 4 * gq];
            gW1[q] = mfma_bf16(Ad1, Bx, gW1[q]);                    // dW1^T[i][f]
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    if (!BWD) return;
    // ---- per-wave partial gradient
    float *pw = part + (size_t)gw * FS_NP;
#pragma unroll
    for (int mo = 0; mo < 4; mo++)
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int o = 16 * mo + 4 * gq + r, ii = 16 * q + el;
                pw[FS_W2 + ii * FS_H + o] = gW2[mo][q][r];
            }
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int ii = 16 * q + 4 * gq + r;
            if (el < FS_NF) pw[FS_W1 + el * FS_H + ii] = gW1[q][r];
        }
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            float a = gw3[m][r], b = gb2[m][r], c = gb1[m][r];
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) {
                a += __shfl_xor(a, o, 64);
                b += __shfl_xor(b, o, 64);
                c += __shfl_xor(c, o, 64);
            }
            if (el == 0) {
                const int hh = 16 * m + 4 * gq + r;
                pw[FS_W3 + hh] = a;
                pw[FS_B2 + hh] = b;
                pw[FS_B1 + hh] = c;
            }
        }
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) gb3 += __shfl_xor(gb3, o, 64);
    if (L == 0) pw[FS_B3] = gb3;
}

// grad[k] = sum over waves of the partials, in wave order (double).
__global__ __launch_bounds__(256) void k_fsai_reduce(const float *part, int nw, double *grad) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= FS_NP) return;
    double acc = 0.0;
    for (int q = 0; q < nw; q++) acc += (double)part[(size_t)q * FS_NP + k];
    grad[k] = acc;
}

// dL/dg per edge from the probes: (2/(P T)) sum_p z_p[i] u_p[j] - [i == j] 2/(n g_ii).
__global__ __launch_bounds__(256) void k_fsai_edge_grad(int n, long long nnzl, const int *erow, const int *cil,
                                                        const double *g, int P, const double *Z, const double *U,
                                                        const Ctl *c, double *gE) {
    const double T = c->s[0] / P;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < nnzl; e += (long long)gridDim.x * 256) {
        const int i = erow[e], j = cil[e];
        double acc = 0.0;
        for (int p = 0; p < P; p++) acc += Z[(size_t)p * n + i] * U[(size_t)p * n + j];
        gE[e] = 2.0 * acc / (P * T) - ((i == j) ? 2.0 / (n * g[e]) : 0.0);
    }
}

// partial sums of log g_ii over the diagonal edges
__global__ __launch_bounds__(256) void k_fsai_logdiag(long long nnzl, const int *erow, const int *cil,
                                                      const double *g, double *part) {
    double v[1] = {0.0};
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < nnzl; e += (long long)gridDim.x * 256)
        if (erow[e] == cil[e]) v[0] += log(g[e]);
    block_partial<1, 256>(v, part);
}

struct FsAddFin {
    int slot;
    __device__ void operator()(Ctl *c, const double *tot) const { c->s[slot] += tot[0]; }
};

// Adam on the FP32 master weights (bias-corrected).
__global__ __launch_bounds__(256) void k_fsai_adam(float *p, float *m, float *v, const double *g, float lr,
                                                   float b1t, float b2t) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= FS_NP) return;
    const float gk = (float)g[k];
    const float mk = 0.9f * m[k] + 0.1f * gk, vk = 0.999f * v[k] + 0.001f * gk * gk;
    m[k] = mk;
    v[k] = vk;
    p[k] -= lr * (mk / (1.0f - b1t)) / (sqrtf(vk / (1.0f - b2t)) + 1e-8f);
}

// ---------------------------------------------------------------- host side

// Forward (factor build) grid: 1024 workgroups of 4 waves, the 4 waves per SIMD its
// 128 registers allow, each wave striding over 16-edge tiles (with 256 workgroups, one
// wave per SIMD, the MLP's dependent MFMA / transcendental / load chains had nothing
// to overlap with: SQ_WAVES 1024 and 2.7 % MFMA busy at 128^3). The backward keeps
// FS_BLOCKS workgroups: one gradient partial per wave.
static int fs_grid(long long nnzl) {
    long long tiles = (nnzl + 15) / 16;
    long long b = (tiles + 3) / 4;
    return (int)std::max(1LL, std::min(b, 1024LL));
}

extern "C" bkh_fsai_model *bkh_fsai_model_create(const float *init) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return nullptr;
    bkh_fsai_model *M = new bkh_fsai_model;
    bool ok = bkh_dmalloc(&M->params, sizeof(float) * FS_NP) == hipSuccess &&
              bkh_dmalloc(&M->m, sizeof(float) * FS_NP) == hipSuccess &&
              bkh_dmalloc(&M->v, sizeof(float) * FS_NP) == hipSuccess &&
              bkh_dmalloc(&M->grad, sizeof(double) * FS_NP) == hipSuccess &&
              bkh_dmalloc(&M->part, sizeof(float) * FS_NP * FS_WAVES) == hipSuccess;
    std::vector<float> h(FS_NP, 0.0f);
    if (init) for (int k = 0; k < FS_NP; k++) h[k] = init[k];
    ok = ok && hipMemcpy(M->params, h.data(), sizeof(float) * FS_NP, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemset(M->m, 0, sizeof(float) * FS_NP) == hipSuccess && hipMemset(M->v, 0, sizeof(float) * FS_NP) == hipSuccess;
    if (!ok) {
        bkh_set_error("FSAI model allocation failed");
        hipFree(M->params); hipFree(M->m); hipFree(M->v); hipFree(M->grad); hipFree(M->part);
        delete M;
        return nullptr;
    }
    return M;
}

extern "C" void bkh_fsai_model_destroy(bkh_fsai_model *M) {
    if (!M) return;
    hipFree(M->params); hipFree(M->m); hipFree(M->v); hipFree(M->grad); hipFree(M->part);
    delete M;
}

extern "C" int bkh_fsai_model_get(const bkh_fsai_model *M, float *params) {
    if (!M || !params) return BKH_BAD_ARGUMENT;
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    HIP_OK(hipMemcpyAsync(params, M->params, sizeof(float) * FS_NP, hipMemcpyDeviceToHost, t->stream));
    HIP_OK(hipStreamSynchronize(t->stream));
    return 0;
}

extern "C" int bkh_fsai_model_set(bkh_fsai_model *M, const float *params) {
    if (!M || !params) return BKH_BAD_ARGUMENT;
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    HIP_OK(hipMemcpyAsync(M->params, params, sizeof(float) * FS_NP, hipMemcpyHostToDevice, t->stream));
    HIP_OK(hipMemsetAsync(M->m, 0, sizeof(float) * FS_NP, t->stream));
    HIP_OK(hipMemsetAsync(M->v, 0, sizeof(float) * FS_NP, t->stream));
    HIP_OK(hipStreamSynchronize(t->stream));
    M->step = 0;
    return 0;
}

extern "C" int bkh_fsai_nparam(void) { return FS_NP; }

static void fs_field_free(bkh_fsai_field *F) {
    if (F->A) bkh_matrix_destroy(F->A);
    if (F->G) bkh_matrix_destroy(F->G);
    hipFree(F->rpl); hipFree(F->cil); hipFree(F->erow); hipFree(F->perm); hipFree(F->vl); hipFree(F->rows); hipFree(F->rowx);
    hipFree(F->Z); hipFree(F->W); hipFree(F->U); hipFree(F->gE);
    delete F;
}

extern "C" void bkh_fsai_field_destroy(bkh_fsai_field *F) {
    if (F) fs_field_free(F);
}

// A field: the lower triangle (with diagonal) of an SPD matrix and P probe vectors
// (row-major P x n, host; may be NULL with P = 0 for evaluation-only fields).
extern "C" bkh_fsai_field *bkh_fsai_field_create(int n, const int *rpl, const int *cil, const double *vl, int P,
                                                 const double *Z) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return nullptr;
    if (n < 1 || !rpl || !cil || !vl || P < 0 || (P > 0 && !Z)) { bkh_set_error("bad FSAI field"); return nullptr; }
    const long long nnzl = rpl[n];
    for (int i = 0; i < n; i++) {
        bool diag = false;
        for (int k = rpl[i]; k < rpl[i + 1]; k++) {
            if (cil[k] > i || cil[k] < 0) { bkh_set_error("FSAI field: row %d is not lower triangular", i); return nullptr; }
            if (cil[k] == i) diag = true;
        }
        if (!diag) { bkh_set_error("FSAI field: row %d has no diagonal", i); return nullptr; }
    }
    bkh_fsai_field *F = new bkh_fsai_field;
    F->n = n;
    F->nnzl = nnzl;
    F->P = P;
    F->A = bkh_matrix_create(n, rpl, cil, vl, 1, 0, 0);
    F->G = bkh_matrix_create(n, rpl, cil, vl, 0, 1, 1);
    if (!F->A || !F->G) { fs_field_free(F); return nullptr; }
    // the factor's values are rewritten in CSR order by the MLP: keep G and G^T on the
    // CSR-stream kernel (no SELL copy to keep in step)
    F->G->variant = SPMV_WAVE;
    F->G->T->variant = SPMV_WAVE;
    std::vector<int> erow((size_t)nnzl), perm((size_t)nnzl), cnt((size_t)n + 1, 0);
    for (int i = 0; i < n; i++)
        for (int k = rpl[i]; k < rpl[i + 1]; k++) { erow[k] = i; cnt[cil[k] + 1]++; }
    for (int j = 0; j < n; j++) cnt[j + 1] += cnt[j];
    for (int i = 0; i < n; i++)                            // G^T rows in the 't' order of the upload
        for (int k = rpl[i]; k < rpl[i + 1]; k++) perm[k] = cnt[cil[k]]++;
    const size_t np = (size_t)std::max(P, 1) * n;
    bool ok = bkh_dmalloc(&F->rpl, sizeof(int) * (n + 1)) == hipSuccess &&
              bkh_dmalloc(&F->cil, sizeof(int) * nnzl) == hipSuccess &&
              bkh_dmalloc(&F->erow, sizeof(int) * nnzl) == hipSuccess &&
              bkh_dmalloc(&F->perm, sizeof(int) * nnzl) == hipSuccess &&
              bkh_dmalloc(&F->vl, sizeof(double) * nnzl) == hipSuccess &&
              bkh_dmalloc(&F->rows, sizeof(double4) * n) == hipSuccess &&
              bkh_dmalloc(&F->rowx, sizeof(double2) * n) == hipSuccess &&
              bkh_dmalloc(&F->Z, sizeof(double) * np) == hipSuccess &&
              bkh_dmalloc(&F->W, sizeof(double) * np) == hipSuccess &&
              bkh_dmalloc(&F->U, sizeof(double) * np) == hipSuccess &&
              bkh_dmalloc(&F->gE, sizeof(double) * nnzl) == hipSuccess;
    hipStream_t s = t->stream;
    ok = ok && hipMemcpyAsync(F->rpl, rpl, sizeof(int) * (n + 1), hipMemcpyHostToDevice, s) == hipSuccess &&
         hipMemcpyAsync(F->cil, cil, sizeof(int) * nnzl, hipMemcpyHostToDevice, s) == hipSuccess &&
         hipMemcpyAsync(F->erow, erow.data(), sizeof(int) * nnzl, hipMemcpyHostToDevice, s) == hipSuccess &&
         hipMemcpyAsync(F->perm, perm.data(), sizeof(int) * nnzl, hipMemcpyHostToDevice, s) == hipSuccess &&
         hipMemcpyAsync(F->vl, vl, sizeof(double) * nnzl, hipMemcpyHostToDevice, s) == hipSuccess;
    if (ok && P > 0) ok = hipMemcpyAsync(F->Z, Z, sizeof(double) * np, hipMemcpyHostToDevice, s) == hipSuccess;
    if (ok) {
        const bkh_matrix *A = F->A;
        hipLaunchKernelGGL(k_fsai_rows, dim3(vec_grid(n, 256, 1)), dim3(256), 0, s, n, A->rp, A->ci, A->val, A->diag,
                           F->rows, F->rowx);
        ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(s) == hipSuccess;
    }
    if (!ok) { bkh_set_error("FSAI field upload failed"); fs_field_free(F); return nullptr; }
    return F;
}

// G's values from the model (G and G^T, FP64 and FP32 copies); g_out (host) optional.
// CSR-order values -> the SELL-64 copy (entry k of row r sits at so[r/64] + 64 k + r%64).
__global__ __launch_bounds__(256) void k_csr_to_sell(int n, const int *rp, const int *so, const float *v32,
                                                     float *sv32) {
    for (int r = blockIdx.x * 256 + threadIdx.x; r < n; r += gridDim.x * 256) {
        const int base = so[r >> 6] + (r & 63), k0 = rp[r], len = rp[r + 1] - k0;
        for (int k = 0; k < len; k++) sv32[base + 64 * k] = v32[k0 + k];
    }
}

// The fused PCG applies G and G^T through their slice-diagonal copies (FP32 values,
// refreshed after every factor) when both have one.
static bool pf_gdia(const bkh_matrix *G) {
    return G && G->dia_entries > 0 && G->dval32 && G->T && G->T->dia_entries > 0 && G->T->dval32;
}

// CSR-order values -> the slice-diagonal copy: entry k of row r (column c) sits at the
// position of offset c - r in its slice's ascending offset list (rows ascend, so one
// forward scan finds every entry's position); positions no entry holds stay 0.
__global__ __launch_bounds__(256) void k_csr_to_dia(int n, const int *rp, const int *ci, const int *dso,
                                                    const int *doff, const float *v32, float *dv32) {
    for (int r = blockIdx.x * 256 + threadIdx.x; r < n; r += gridDim.x * 256) {
        const int s0 = dso[r >> 6], w = (dso[(r >> 6) + 1] - s0) >> 6;
        const int *o = doff + (s0 >> 6);
        int p = 0;
        for (int k = rp[r]; k < rp[r + 1]; k++) {
            const int off = ci[k] - r;
            while (p < w - 1 && o[p] != off) p++;
            dv32[s0 + 64 * p + (r & 63)] = v32[k];
            p++;
        }
    }
}

static int fs_factor(ThreadCtx *t, const bkh_fsai_model *M, bkh_fsai_field *F) {
    hipLaunchKernelGGL((k_fsai_mlp<false>), dim3(fs_grid(F->nnzl)), dim3(256), 0, t->stream, F->n, F->nnzl, F->erow,
                       F->cil, F->vl, F->rows, F->rowx, fs_inv_log2n1(F->n), M->params, F->perm, F->G->val,
                       F->G->val32, F->G->T->val,
                       F->G->T->val32, (const double *)nullptr, (float *)nullptr);
    // the fused PCG applies G and G^T through their SELL copies (FP32 values)
    for (const bkh_matrix *Gm : {(const bkh_matrix *)F->G, (const bkh_matrix *)F->G->T})
        if (Gm->nsl > 0 && Gm->sval32)
            hipLaunchKernelGGL(k_csr_to_sell, dim3(vec_grid(F->n, 256, 1)), dim3(256), 0, t->stream, F->n, Gm->rp,
                               Gm->sso, (const float *)Gm->val32, Gm->sval32);
    // and through their slice-diagonal copies when both have one (FP32 values)
    if (pf_gdia(F->G))
        for (const bkh_matrix *Gm : {(const bkh_matrix *)F->G, (const bkh_matrix *)F->G->T})
            hipLaunchKernelGGL(k_csr_to_dia, dim3(vec_grid(F->n, 256, 1)), dim3(256), 0, t->stream, F->n, Gm->rp,
                               Gm->ci, Gm->dso, Gm->doff, (const float *)Gm->val32, Gm->dval32);
    LAUNCH_OK();
    return 0;
}

extern "C" int bkh_fsai_factor(const bkh_fsai_model *M, bkh_fsai_field *F, double *g_out) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (!M || !F) return BKH_BAD_ARGUMENT;
    int rc = fs_factor(t, M, F);
    if (rc) return rc;
    if (g_out) HIP_OK(hipMemcpyAsync(g_out, F->G->val, sizeof(double) * F->nnzl, hipMemcpyDeviceToHost, t->stream));
    HIP_OK(hipStreamSynchronize(t->stream));
    return 0;
}

// Loss and gradient (device, into M->grad); loss to *loss.
static int fs_loss_grad(ThreadCtx *t, bkh_fsai_model *M, bkh_fsai_field *F, double *loss) {
    if (F->P < 1) { bkh_set_error("FSAI field has no probe vectors"); return BKH_BAD_ARGUMENT; }
    hipStream_t s = t->stream;
    const int n = F->n;
    const int G = vec_grid(n, 256);
    const int Ge = vec_grid(F->nnzl, 256);
    const int nparts = std::max(std::max(spmv_grid(F->A), spmv_grid(F->G->T)), std::max(G, Ge)) * 2;
    void *ws = bkh_workspace(carve_size(sizeof(Ctl)) + carve_size(sizeof(double) * nparts));
    if (!ws) return BKH_DEVICE_ERROR;
    Carve cv(ws);
    Ctl *c = cv.take<Ctl>(1);
    double *part = cv.take<double>(nparts);
    int rc = fs_factor(t, M, F);
    if (rc) return rc;
    if (ctl_init(t, c, 0.0, 1)) return BKH_DEVICE_ERROR;
    Epi<0> e0;
    memset(&e0, 0, sizeof e0);
    for (int p = 0; p < F->P; p++) {
        double *z = F->Z + (size_t)p * n, *w = F->W + (size_t)p * n, *u = F->U + (size_t)p * n;
        const bkh_matrix *Gt = F->G->T;
        launch_spmv<double, double, 0>(s, Gt, Gt->val, z, w, 1.0, 0.0, e0, nullptr, nullptr, NoFin());
        Epi<1> e1;
        memset(&e1, 0, sizeof e1);
        e1.w[0] = w;
        e1.a[0] = 0;
        e1.b[0] = EPI_Y;
        launch_spmv<double, double, 1>(s, F->A, F->A->val, w, u, 1.0, 0.0, e1, c, part, FsAddFin{0});
    }
    hipLaunchKernelGGL(k_fsai_logdiag, dim3(Ge), dim3(256), 0, s, F->nnzl, F->erow, F->cil, (const double *)F->G->val,
                       part);
    launch_finalize<1>(s, part, Ge, c, FsAddFin{1});
    hipLaunchKernelGGL(k_fsai_edge_grad, dim3(Ge), dim3(256), 0, s, n, F->nnzl, F->erow, F->cil,
                       (const double *)F->G->val, F->P, (const double *)F->Z, (const double *)F->U, (const Ctl *)c,
                       F->gE);
    hipLaunchKernelGGL((k_fsai_mlp<true>), dim3(FS_BLOCKS), dim3(256), 0, s, n, F->nnzl, F->erow, F->cil, F->vl,
                       F->rows, F->rowx, fs_inv_log2n1(n), M->params, F->perm, (double *)nullptr, (float *)nullptr,
                       (double *)nullptr,
                       (float *)nullptr, (const double *)F->gE, M->part);
    hipLaunchKernelGGL(k_fsai_reduce, dim3((FS_NP + 255) / 256), dim3(256), 0, s, (const float *)M->part, FS_WAVES,
                       M->grad);
    LAUNCH_OK();
    if (loss) {
        if (ctl_fetch(t, c)) return BKH_DEVICE_ERROR;
        const double T = t->ctl_host->s[0] / F->P;
        *loss = log(T / n) - 2.0 * t->ctl_host->s[1] / n;
    }
    return 0;
}

extern "C" int bkh_fsai_loss_grad(bkh_fsai_model *M, bkh_fsai_field *F, double *loss, double *grad) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (!M || !F) return BKH_BAD_ARGUMENT;
    int rc = fs_loss_grad(t, M, F, loss);
    if (rc) return rc;
    if (grad) HIP_OK(hipMemcpyAsync(grad, M->grad, sizeof(double) * FS_NP, hipMemcpyDeviceToHost, t->stream));
    HIP_OK(hipStreamSynchronize(t->stream));
    return 0;
}

// One Adam step on the field's loss; the loss before the step to *loss (may be NULL,
// which saves the host round trip).
extern "C" int bkh_fsai_train_step(bkh_fsai_model *M, bkh_fsai_field *F, double lr, double *loss) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (!M || !F) return BKH_BAD_ARGUMENT;
    int rc = fs_loss_grad(t, M, F, loss);
    if (rc) return rc;
    M->step++;
    const float b1t = powf(0.9f, (float)M->step), b2t = powf(0.999f, (float)M->step);
    hipLaunchKernelGGL(k_fsai_adam, dim3((FS_NP + 255) / 256), dim3(256), 0, t->stream, M->params, M->m, M->v,
                       (const double *)M->grad, (float)lr, b1t, b2t);
    LAUNCH_OK();
    return 0;
}

// Average ms of one forward pass (factor build) over reps launches, HIP events on the
// kernel's stream.
extern "C" int bkh_fsai_bench_forward(const bkh_fsai_model *M, bkh_fsai_field *F, int reps, double *avg_ms) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (!M || !F || reps < 1) return BKH_BAD_ARGUMENT;
    fs_factor(t, M, F);
    HIP_OK(hipEventRecord(t->ev0, t->stream));
    for (int r = 0; r < reps; r++) fs_factor(t, M, F);
    HIP_OK(hipEventRecord(t->ev1, t->stream));
    HIP_OK(hipEventSynchronize(t->ev1));
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, t->ev0, t->ev1));
    if (avg_ms) *avg_ms = ms / reps;
    return 0;
}

// ------------------------------------------------------ preconditioned CG

enum { PC_RZ = 0, PC_ALPHA = 1, PC_BETA = 2 };

enum { PCK_NONE = 0, PCK_JACOBI = 1, PCK_FSAI = 2 };

// r = b - A x (A x in r on entry); z = D^-1 r (Jacobi) or z = r (none; FSAI forms z
// later). Partials {r.z, b.b}.
__global__ __launch_bounds__(256) void k_pcg_init(long long n, const double *b, double *r, double *z,
                                                  const double *diag, int kind, double *part) {
    double v[2] = {0.0, 0.0};
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const double ri = b[i] - r[i];
        r[i] = ri;
        const double zi = kind == PCK_JACOBI ? ri / diag[i] : ri;
        if (kind != PCK_FSAI) z[i] = zi;
        v[0] += ri * zi;
        v[1] += b[i] * b[i];
    }
    block_partial<2, 256>(v, part);
}

struct PcInitFin {
    int fsai;
    __device__ void operator()(Ctl *c, const double *tot) const {
        c->fnb = sqrt(tot[1]);
        if (!(c->fnb > 0.0)) { ctl_breakdown(c); return; }
        if (!fsai) c->s[PC_RZ] = tot[0];
    }
};

struct PcRzFin {            // after z = G^T G r: r.z (at the start), or the next beta
    int init;
    __device__ void operator()(Ctl *c, const double *tot) const {
        if (init) { c->s[PC_RZ] = tot[0]; return; }
        if (bad(c->s[PC_RZ])) { ctl_breakdown(c); return; }
        c->s[PC_BETA] = tot[0] / c->s[PC_RZ];
        c->s[PC_RZ] = tot[0];
    }
};

struct PcAlphaFin {
    __device__ void operator()(Ctl *c, const double *tot) const {
        if (bad(tot[0])) { ctl_breakdown(c); return; }
        c->s[PC_ALPHA] = c->s[PC_RZ] / tot[0];
    }
};

// x += alpha p, r -= alpha q; z = D^-1 r (Jacobi) or z = r (none). Partials {r.r, r.z}.
__global__ __launch_bounds__(256) void k_pcg_xr(long long n, double *x, double *r, double *z, const double *p,
                                                const double *q, const double *diag, int kind, const Ctl *c,
                                                double *part) {
    if (ctl_stopped(c)) return;
    const double al = c->s[PC_ALPHA];
    double v[2] = {0.0, 0.0};
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        x[i] = fma(al, p[i], x[i]);
        const double ri = fma(-al, q[i], r[i]);
        r[i] = ri;
        v[0] += ri * ri;
        if (kind != PCK_FSAI) {
            const double zi = kind == PCK_JACOBI ? ri / diag[i] : ri;
            z[i] = zi;
            v[1] += ri * zi;
        }
    }
    block_partial<2, 256>(v, part);
}

// ||r|| / ||b|| test (the iteration's index); Jacobi/none also form beta from r.z.
struct PcCheckFin {
    int fsai;
    __device__ void operator()(Ctl *c, const double *tot) const {
        ctl_check(c, tot[0]);
        if (c->done || fsai) return;
        if (bad(c->s[PC_RZ])) { ctl_breakdown(c); return; }
        c->s[PC_BETA] = tot[1] / c->s[PC_RZ];
        c->s[PC_RZ] = tot[1];
    }
};

// p = z + beta p (beta = 0 on the first call)
__global__ __launch_bounds__(256) void k_pcg_p(long long n, double *p, const double *z, const Ctl *c, int first) {
    if (ctl_stopped(c)) return;
    const double be = first ? 0.0 : c->s[PC_BETA];
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        p[i] = first ? z[i] : fma(be, p[i], z[i]);
}

// z = G^T (G r) with the epilogue r.z -> fin (FP32 factor values, FP64 vectors).
template <class Fin>
static void fs_apply(hipStream_t s, const bkh_matrix *G, const double *r, double *y, double *z, Ctl *c, double *part,
                     Fin fin) {
    Epi<0> e0;
    memset(&e0, 0, sizeof e0);
    launch_spmv<float, double, 0>(s, G, G->val32, r, y, 1.0, 0.0, e0, c, nullptr, NoFin());
    Epi<1> e1;
    memset(&e1, 0, sizeof e1);
    e1.w[0] = r;
    e1.a[0] = 0;
    e1.b[0] = EPI_Y;
    launch_spmv<float, double, 1>(s, G->T, G->T->val32, y, z, 1.0, 0.0, e1, c, part, fin);
}

// Preconditioned CG: q = A p, alpha = r.z / p.q, x += alpha p, r -= alpha q, stop on
// ||r|| / ||b|| < tol (the index convention of cg.c:56-59), z = M^-1 r, beta = r.z'/r.z,
// p = z + beta p.
static int pcg_fused(ThreadCtx *t, const bkh_matrix *A, const bkh_matrix *G, int kind, const double *b, double *x,
                     double tol, int max_iter, const bkh_opts &o, bkh_info *inf);

static int pcg_run(ThreadCtx *t, const bkh_matrix *A, const bkh_matrix *G, int kind, const double *b, double *x,
                   double tol, int max_iter, const bkh_opts &o, bkh_info *inf) {
    // The fused loop (A on its slice-diagonal copy, p formed in the gather) is the
    // default for every preconditioner: Jacobi PCG at 128^3 takes 65.5 ms against 78.5 ms
    // for the separate SpMV and vector kernels, FSAI 275 against 380 ms.
    const bool fused = o.path != 1;
    if (fused) {
        const int rc = pcg_fused(t, A, G, kind, b, x, tol, max_iter, o, inf);
        if (rc != 1) return rc;
    }
    hipStream_t s = t->stream;
    const long long n = A->n;
    const int Gv = vec_grid(n, 256);
    int nparts = std::max(spmv_grid(A), Gv);
    if (G) nparts = std::max(nparts, std::max(spmv_grid(G), spmv_grid(G->T)));
    nparts *= 2;
    void *ws = bkh_workspace(carve_size(sizeof(Ctl)) + carve_size(sizeof(double) * nparts) +
                             5 * carve_size(sizeof(double) * n));
    if (!ws) return BKH_DEVICE_ERROR;
    Carve cv(ws);
    Ctl *c = cv.take<Ctl>(1);
    double *part = cv.take<double>(nparts);
    double *r = cv.take<double>(n), *z = cv.take<double>(n), *p = cv.take<double>(n), *q = cv.take<double>(n),
           *y = cv.take<double>(n);
    const int K = o.graph_iters > 0 ? o.graph_iters : default_graph_iters(n);
    const bool fsai = kind == PCK_FSAI;
    HIP_OK(hipEventRecord(t->ev0, s));
    if (ctl_init(t, c, tol, max_iter)) return BKH_DEVICE_ERROR;
    Epi<0> e0;
    memset(&e0, 0, sizeof e0);
    launch_spmv<double, double, 0>(s, A, A->val, x, r, 1.0, 0.0, e0, nullptr, nullptr, NoFin());
    hipLaunchKernelGGL(k_pcg_init, dim3(Gv), dim3(256), 0, s, n, b, r, z, (const double *)A->diag, kind, part);
    launch_finalize<2>(s, part, Gv, c, PcInitFin{(int)fsai});
    if (fsai) fs_apply(s, G, r, y, z, c, part, PcRzFin{1});
    hipLaunchKernelGGL(k_pcg_p, dim3(Gv), dim3(256), 0, s, n, p, (const double *)z, (const Ctl *)c, 1);
    LAUNCH_OK();
    hipGraphExec_t ge = capture_cached(t, {0xF5A1, (uintptr_t)A, (uintptr_t)G, (uintptr_t)kind, (uintptr_t)b, (uintptr_t)x,
                                           (uintptr_t)r, (uintptr_t)z, (uintptr_t)p, (uintptr_t)q, (uintptr_t)y,
                                           (uintptr_t)c, (uintptr_t)part}, s, K, [&]() {
        Epi<1> e1;
        memset(&e1, 0, sizeof e1);
        e1.w[0] = p;
        e1.a[0] = 0;
        e1.b[0] = EPI_Y;
        launch_spmv<double, double, 1>(s, A, A->val, p, q, 1.0, 0.0, e1, c, part, PcAlphaFin());
        hipLaunchKernelGGL(k_pcg_xr, dim3(Gv), dim3(256), 0, s, n, x, r, z, (const double *)p, (const double *)q,
                           (const double *)A->diag, kind, (const Ctl *)c, part);
        launch_finalize<2>(s, part, Gv, c, PcCheckFin{(int)fsai});
        if (fsai) fs_apply(s, G, r, y, z, c, part, PcRzFin{0});
        hipLaunchKernelGGL(k_pcg_p, dim3(Gv), dim3(256), 0, s, n, p, (const double *)z, (const Ctl *)c, 0);
    });
    if (!ge) { bkh_set_error("hipGraph capture of the PCG iteration failed"); return BKH_DEVICE_ERROR; }
    int rc = run_graph_until_done(t, ge, c, max_iter / K + 2);
    if (rc) return rc;
    Ctl hc = *t->ctl_host;
    HIP_OK(hipEventRecord(t->ev1, s));
    HIP_OK(hipEventSynchronize(t->ev1));
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, t->ev0, t->ev1));
    memset(inf, 0, sizeof *inf);
    inf->time_ms = ms;
    bkh_note_solve_ms(ms);
    inf->ret = ctl_result(&hc);
    inf->status = hc.status;
    inf->iters = hc.result_it;
    inf->rec_res = hc.err;
    double tr = 0.0;
    if (true_residual(t, A, 0.0, b, x, &tr) == 0) inf->true_res = tr;
    return inf->ret;
}

// ------------------------------------------------ fused PCG (SELL-64 matrices)
//
// PCG in two (Jacobi, none) or three (FSAI) kernels per iteration and no finalize
// kernels, on the pattern of the fused CG: every workgroup of a kernel sums the
// previous kernel's partials itself in a fixed order.
//   PA(k): convergence test of iteration k-1 on r_k.r_k (cg.c:56-59 convention),
//          beta = (r_k.z_k)/(r_{k-1}.z_{k-1}); p_k = z_k + beta p_{k-1} formed in the
//          gather; q = A p_k; partials p_k.q.
//   PB(k): alpha = (r_k.z_k)/(p_k.q); x += alpha p_k; r_{k+1} = r_k - alpha q;
//          Jacobi/none: z = D^-1 r_{k+1} (or r_{k+1}), partials {r.z, r.r};
//          FSAI: y = G r_{k+1}, r_{k+1} formed in the gather from r_k and q.
//   PD(k): (FSAI) z = G^T y, partials {r_{k+1}.z, r_{k+1}.r_{k+1}}.
// r and p alternate between two buffers by the parity of k; r.z of iteration k lives
// in Ctl::s[PF_RZ + (k & 1)]. All three matrices are streamed through their SELL-64
// copies (lane per row, contiguous 64-entry loads); workgroup g owns `chunk` slices.
enum { PF_RZ = 20 };

// Workgroups of PF_BS threads (PF_BS / 64 waves, each taking every PF_W-th slice of
// the workgroup's range): few, large workgroups keep the partial arrays that every
// workgroup of the next kernel re-sums short (<= 512 entries).
#define PF_BS 1024
#define PF_W (PF_BS / 64)
#define PF_NPART 512
#define PF_SLICES(sl, row)                                                           \
    const int g_ = xcd_block(blockIdx.x, gridDim.x);                                 \
    const int w_ = threadIdx.x >> 6, l_ = threadIdx.x & 63;                          \
    const int send_ = min(nsl, (g_ + 1) * chunk);                                    \
    for (int sl = g_ * chunk + w_; sl < send_; sl += PF_W)                           \
        for (int row = sl * 64 + l_; row < n; row = n)

template <typename IT>
__device__ __forceinline__ int sell_col(const IT *col, int e, int row) {
    return sizeof(IT) == 2 ? row + (int)col[e] : (int)col[e];
}

// Scalars at the start of PA(k): false when the solve stops at k.
__device__ __forceinline__ bool pf_beta(Ctl *c, int k, const double *partPrev, int np, bool writer, double &beta) {
    if (k == c->first) {
        beta = 0.0;
        return true;
    }
    double d[2];
    {
        __shared__ double sh[2 * PF_W + 2];
        double a0 = 0.0, a1 = 0.0;
        for (int i = threadIdx.x; i < np; i += blockDim.x) { a0 += partPrev[2 * i]; a1 += partPrev[2 * i + 1]; }
        a0 = wave_sum(a0);
        a1 = wave_sum(a1);
        const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
        if ((threadIdx.x & 63) == 0) { sh[w] = a0; sh[PF_W + w] = a1; }
        __syncthreads();
        if (threadIdx.x == 0) {
            double t0 = sh[0], t1 = sh[PF_W];
            for (int q = 1; q < nw; q++) { t0 += sh[q]; t1 += sh[PF_W + q]; }
            sh[2 * PF_W] = t0;
            sh[2 * PF_W + 1] = t1;
        }
        __syncthreads();
        d[0] = sh[2 * PF_W];
        d[1] = sh[2 * PF_W + 1];
    }
    if (cgf_check(c, k, d[1], writer) != ST_RUNNING) return false;
    const double rz_old = c->s[PF_RZ + ((k + 1) & 1)];
    if (bad(rz_old)) {
        if (writer) { c->status = ST_BREAKDOWN; c->result_it = k - 1; c->done = 1; }
        return false;
    }
    if (writer) c->s[PF_RZ + (k & 1)] = d[0];
    beta = d[0] / rz_old;
    return true;
}

// PA(k): q = A p with p = z + beta p formed in the gather, p.q partials. DIA: on the
// slice-diagonal copy of A (so, val, col = its slice offsets, values and per-slice
// column offsets); its value-0 positions add fma(0, p, acc) = acc.
template <typename IT, typename ZT, bool DIA = false>
__global__ __launch_bounds__(PF_BS) void k_pf_a(int n, int nsl, int chunk, const int *__restrict__ rp,
                                              const int *__restrict__ so, const double *__restrict__ val,
                                              const IT *__restrict__ col, const ZT *z, double *P0, double *P1,
                                              double *q, Ctl *c, int j, const double *partPrev, int np,
                                              double *partA) {
    if (ctl_stopped(c)) return;
    const int k = c->base + j;
    const bool writer = blockIdx.x == 0 && threadIdx.x == 0;
    double be;
    if (!pf_beta(c, k, partPrev, np, writer, be)) return;
    const double *pold = (k & 1) ? P1 : P0;
    double *pnew = (k & 1) ? P0 : P1;
    double dv[1] = {0.0};
    PF_SLICES(sl, row) {
        int len, base;
        const int *dof = nullptr;
        if constexpr (DIA) {
            const int su = __builtin_amdgcn_readfirstlane(sl);
            const int s0 = so[su];
            len = (so[su + 1] - s0) >> 6;
            base = s0 + (row & 63);
            dof = (const int *)col + (s0 >> 6);
        } else {
            len = rp[row + 1] - rp[row];
            base = so[sl] + (row & 63);
        }
        double acc = 0.0;
#pragma unroll 4
        for (int kk = 0; kk < len; kk++) {
            const int e = base + 64 * kk;
            int cc;
            if constexpr (DIA) cc = min(max(row + dof[kk], 0), n - 1);
            else cc = sell_col(col, e, row);
            acc = fma(val[e], fma(be, pold[cc], (double)z[cc]), acc);
        }
        q[row] = acc;
        const double pn = fma(be, pold[row], (double)z[row]);
        pnew[row] = pn;
        dv[0] += pn * acc;
    }
    block_partial<1, PF_BS>(dv, partA);
}

__device__ __forceinline__ bool pf_alpha(Ctl *c, int k, const double *partA, int npA, double &al) {
    const double pq = block_sum_parts<PF_BS>(partA, npA);
    if (bad(pq)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) { c->status = ST_BREAKDOWN; c->result_it = k; c->done = 1; }
        return false;
    }
    al = c->s[PF_RZ + (k & 1)] / pq;
    return true;
}

// PB(k) for Jacobi (diag) or none (diag == nullptr).
__global__ __launch_bounds__(PF_BS) void k_pf_b_vec(int n, int nsl, int chunk, double *x, double *R0, double *R1,
                                                  const double *P0, const double *P1, const double *q, double *z,
                                                  const double *diag, Ctl *c, int j, const double *partA, int npA,
                                                  double *partB) {
    if (ctl_stopped(c)) return;
    const int k = c->base + j;
    double al;
    if (!pf_alpha(c, k, partA, npA, al)) return;
    const double *p = (k & 1) ? P0 : P1, *rk = (k & 1) ? R1 : R0;
    double *rn = (k & 1) ? R0 : R1;
    double dv[2] = {0.0, 0.0};
    PF_SLICES(sl, row) {
        x[row] = fma(al, p[row], x[row]);
        const double r = fma(-al, q[row], rk[row]);
        rn[row] = r;
        const double zz = diag ? r / diag[row] : r;
        z[row] = zz;
        dv[0] += r * zz;
        dv[1] += r * r;
    }
    block_partial<2, PF_BS>(dv, partB);
}

// PB(k) for FSAI: x, r_{k+1} own; y = G r_{k+1} with r_{k+1}[j] = r_k[j] - alpha q[j].
// y is stored in FP32: the factor G itself is held in FP32, so the preconditioner
// z = G^T G r is an FP32-accurate operator either way (r, x, p, q stay FP64).
template <typename IT, bool DIA = false>
__global__ __launch_bounds__(PF_BS) void k_pf_b_g(int n, int nsl, int chunk, const int *__restrict__ rp,
                                                const int *__restrict__ so, const float *__restrict__ val,
                                                const IT *__restrict__ col, double *x, double *R0, double *R1,
                                                const double *P0, const double *P1, const double *q, float *y,
                                                Ctl *c, int j, const double *partA, int npA) {
    if (ctl_stopped(c)) return;
    const int k = c->base + j;
    double al;
    if (!pf_alpha(c, k, partA, npA, al)) return;
    const double *p = (k & 1) ? P0 : P1, *rk = (k & 1) ? R1 : R0;
    double *rn = (k & 1) ? R0 : R1;
    PF_SLICES(sl, row) {
        double acc = 0.0;
        sell_row_entries<IT, DIA>(n, row, row & 63, __builtin_amdgcn_readfirstlane(sl), rp, so, col,
                                  [&](int e, int cc, bool) { acc = fma((double)val[e], fma(-al, q[cc], rk[cc]), acc); });
        y[row] = (float)acc;
        x[row] = fma(al, p[row], x[row]);
        rn[row] = fma(-al, q[row], rk[row]);
    }
}

// PD(k): z = G^T y (FP32 storage, as y); partials {r_{k+1}.z, r_{k+1}.r_{k+1}} with the
// stored z, so beta and the direction update see the same values.
template <typename IT, bool DIA = false>
__global__ __launch_bounds__(PF_BS) void k_pf_d(int n, int nsl, int chunk, const int *__restrict__ rp,
                                              const int *__restrict__ so, const float *__restrict__ val,
                                              const IT *__restrict__ col, const float *y, const double *R0,
                                              const double *R1, float *z, Ctl *c, int j, double *partD) {
    if (ctl_stopped(c)) return;
    const int k = c->base + j;
    const double *rn = (k & 1) ? R0 : R1;
    double dv[2] = {0.0, 0.0};
    PF_SLICES(sl, row) {
        double acc = 0.0;
        sell_row_entries<IT, DIA>(n, row, row & 63, __builtin_amdgcn_readfirstlane(sl), rp, so, col,
                                  [&](int e, int cc, bool) { acc = fma((double)val[e], (double)y[cc], acc); });
        const float zs = (float)acc;
        z[row] = zs;
        const double r = rn[row];
        dv[0] += r * (double)zs;
        dv[1] += r * r;
    }
    block_partial<2, PF_BS>(dv, partD);
}

// End of a replay of K iterations: the test of iteration base + K.
__global__ __launch_bounds__(PF_BS) void k_pf_tail(Ctl *c, int K, const double *partPrev, int np) {
    if (ctl_stopped(c)) return;
    const int k = c->base + K;
    double be;
    if (!pf_beta(c, k, partPrev, np, threadIdx.x == 0, be)) return;
    if (threadIdx.x == 0) c->base = k;
}

__global__ __launch_bounds__(256) void k_pf_d2f(long long n, const double *src, float *dst) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        dst[i] = (float)src[i];
}

// Start of the fused loop: r.z of iteration 0 from the unfused init (PC_RZ).
__global__ void k_pf_start(Ctl *c) {
    if (threadIdx.x == 0 && !c->done) {
        c->s[PF_RZ] = c->s[PC_RZ];
        c->base = 0;
        c->first = 0;
    }
}

static bool pf_ok(const bkh_matrix *M) {
    return M && M->nsl > 0 && (M->variant == SPMV_SELL || M->variant == SPMV_SELL32 || M->sval32);
}

// The system matrix's products run on its slice-diagonal copy when it has one and the
// default SELL-64 kernel is selected.
static bool pf_dia(const bkh_matrix *A) {
    return A->dia_entries > 0 && A->dval && A->variant == SPMV_SELL;
}

template <typename IT>
static const IT *sell_cols(const bkh_matrix *M) {
    if constexpr (sizeof(IT) == 2) return (const IT *)M->scd;
    else return (const IT *)M->scol;
}

template <typename ITA, typename ITG, typename ITT>
static void pf_capture_body(hipStream_t s, const bkh_matrix *A, const bkh_matrix *G, int kind, int K, int GA,
                            int chunk, double *x, double *R, double *P, double *q, double *z, double *y, Ctl *c,
                            double *partA, double *partB) {
    const int n = A->n, nsl = A->nsl;
    for (int j = 0; j < K; j++) {
        if (kind == PCK_FSAI) {
            // z and y in FP32 (their buffers hold n doubles each; the first 4n bytes are used)
            if (pf_dia(A))
                hipLaunchKernelGGL((k_pf_a<int, float, true>), dim3(GA), dim3(PF_BS), 0, s, n, nsl, chunk, A->rp,
                                   A->dso, (const double *)A->dval, (const int *)A->doff, (const float *)z, P, P + n,
                                   q, c, j, (const double *)partB, GA, partA);
            else
                hipLaunchKernelGGL((k_pf_a<ITA, float>), dim3(GA), dim3(PF_BS), 0, s, n, nsl, chunk, A->rp, A->sso,
                                   (const double *)A->sval, sell_cols<ITA>(A), (const float *)z, P, P + n, q, c, j,
                                   (const double *)partB, GA, partA);
            const bkh_matrix *Gt = G->T;
            if (pf_gdia(G)) {
                hipLaunchKernelGGL((k_pf_b_g<int, true>), dim3(GA), dim3(PF_BS), 0, s, n, nsl, chunk, G->rp, G->dso,
                                   (const float *)G->dval32, (const int *)G->doff, x, R, R + n, (const double *)P,
                                   (const double *)(P + n), (const double *)q, (float *)y, c, j,
                                   (const double *)partA, GA);
                hipLaunchKernelGGL((k_pf_d<int, true>), dim3(GA), dim3(PF_BS), 0, s, n, nsl, chunk, Gt->rp, Gt->dso,
                                   (const float *)Gt->dval32, (const int *)Gt->doff, (const float *)y,
                                   (const double *)R, (const double *)(R + n), (float *)z, c, j, partB);
            } else {
                hipLaunchKernelGGL((k_pf_b_g<ITG>), dim3(GA), dim3(PF_BS), 0, s, n, nsl, chunk, G->rp, G->sso,
                                   (const float *)G->sval32, sell_cols<ITG>(G), x, R, R + n, (const double *)P,
                                   (const double *)(P + n), (const double *)q, (float *)y, c, j,
                                   (const double *)partA, GA);
                hipLaunchKernelGGL((k_pf_d<ITT>), dim3(GA), dim3(PF_BS), 0, s, n, nsl, chunk, Gt->rp, Gt->sso,
                                   (const float *)Gt->sval32, sell_cols<ITT>(Gt), (const float *)y,
                                   (const double *)R, (const double *)(R + n), (float *)z, c, j, partB);
            }
        } else {
            if (pf_dia(A))
                hipLaunchKernelGGL((k_pf_a<int, double, true>), dim3(GA), dim3(PF_BS), 0, s, n, nsl, chunk, A->rp,
                                   A->dso, (const double *)A->dval, (const int *)A->doff, (const double *)z, P,
                                   P + n, q, c, j, (const double *)partB, GA, partA);
            else
                hipLaunchKernelGGL((k_pf_a<ITA, double>), dim3(GA), dim3(PF_BS), 0, s, n, nsl, chunk, A->rp, A->sso,
                                   (const double *)A->sval, sell_cols<ITA>(A), (const double *)z, P, P + n, q, c, j,
                                   (const double *)partB, GA, partA);
            hipLaunchKernelGGL(k_pf_b_vec, dim3(GA), dim3(PF_BS), 0, s, n, nsl, chunk, x, R, R + n, (const double *)P,
                               (const double *)(P + n), (const double *)q, z,
                               kind == PCK_JACOBI ? (const double *)A->diag : (const double *)nullptr, c, j,
                               (const double *)partA, GA, partB);
        }
    }
    hipLaunchKernelGGL(k_pf_tail, dim3(1), dim3(PF_BS), 0, s, c, K, (const double *)partB, GA);
}

// Fused PCG; returns 1 when the matrices are not all on the SELL layout (the caller
// then runs the unfused loop).
static int pcg_fused(ThreadCtx *t, const bkh_matrix *A, const bkh_matrix *G, int kind, const double *b, double *x,
                     double tol, int max_iter, const bkh_opts &o, bkh_info *inf) {
    if (!pf_ok(A) || !A->sval || (kind == PCK_FSAI && (!pf_ok(G) || !G->sval32 || !pf_ok(G->T) || !G->T->sval32)))
        return 1;
    if (kind == PCK_FSAI && (G->nsl != A->nsl || G->T->nsl != A->nsl)) return 1;
    hipStream_t s = t->stream;
    const long long n = A->n;
    const int Gv = vec_grid(n, 256);
    const int GA = std::max(1, std::min((A->nsl + PF_W - 1) / PF_W, PF_NPART));
    const int chunk = (A->nsl + GA - 1) / GA;
    int nparts = std::max(spmv_grid(A), Gv);
    if (G) nparts = std::max(nparts, std::max(spmv_grid(G), spmv_grid(G->T)));
    nparts *= 2;
    void *ws = bkh_workspace(carve_size(sizeof(Ctl)) + carve_size(sizeof(double) * nparts) +
                             2 * carve_size(sizeof(double) * 2 * PF_NPART) + 8 * carve_size(sizeof(double) * n));
    if (!ws) return BKH_DEVICE_ERROR;
    Carve cv(ws);
    Ctl *c = cv.take<Ctl>(1);
    double *part = cv.take<double>(nparts);
    double *partA = cv.take<double>(2 * PF_NPART), *partB = cv.take<double>(2 * PF_NPART);
    double *R = cv.take<double>(2 * n), *P = cv.take<double>(2 * n), *q = cv.take<double>(n), *z = cv.take<double>(n),
           *y = cv.take<double>(n);
    int K = o.graph_iters > 0 ? o.graph_iters : default_graph_iters(n);
    HIP_OK(hipEventRecord(t->ev0, s));
    if (ctl_init(t, c, tol, max_iter)) return BKH_DEVICE_ERROR;
    Epi<0> e0;
    memset(&e0, 0, sizeof e0);
    // r_0 = b - A x_0 (buffer 0), z_0, r_0.z_0 (the unfused init), then the loop
    launch_spmv<double, double, 0>(s, A, A->val, x, R, 1.0, 0.0, e0, nullptr, nullptr, NoFin());
    hipLaunchKernelGGL(k_pcg_init, dim3(Gv), dim3(256), 0, s, n, b, R, z, (const double *)A->diag, kind, part);
    launch_finalize<2>(s, part, Gv, c, PcInitFin{(int)(kind == PCK_FSAI)});
    if (kind == PCK_FSAI) fs_apply(s, G, R, y, z, c, part, PcRzFin{1});
    // P0 = z_0 (finite; the first PA forms p_0 = z_0 + 0 * P0)
    hipLaunchKernelGGL(k_pcg_p, dim3(Gv), dim3(256), 0, s, n, P, (const double *)z, (const Ctl *)c, 1);
    if (kind == PCK_FSAI) {            // the fused FSAI loop keeps z (and y) in FP32
        hipLaunchKernelGGL(k_pf_d2f, dim3(Gv), dim3(256), 0, s, (long long)n, (const double *)z, (float *)y);
        HIP_OK(hipMemcpyAsync(z, y, sizeof(float) * n, hipMemcpyDeviceToDevice, s));
    }
    hipLaunchKernelGGL(k_pf_start, dim3(1), dim3(64), 0, s, c);
    LAUNCH_OK();
    const bool a16 = A->scd != nullptr;
    const bool g16 = kind == PCK_FSAI && G->scd != nullptr, t16 = kind == PCK_FSAI && G->T->scd != nullptr;
    hipGraphExec_t ge = capture_cached(t, {0xF5A2, (uintptr_t)A, (uintptr_t)G, (uintptr_t)kind, (uintptr_t)x,
                                           (uintptr_t)R, (uintptr_t)P, (uintptr_t)q, (uintptr_t)z, (uintptr_t)y,
                                           (uintptr_t)c, (uintptr_t)partA, (uintptr_t)partB, (uintptr_t)K}, s, 1, [&]() {
        if (a16 && (kind != PCK_FSAI || (g16 && t16)))
            pf_capture_body<short, short, short>(s, A, G, kind, K, GA, chunk, x, R, P, q, z, y, c, partA, partB);
        else if (!a16 && (kind != PCK_FSAI || (!g16 && !t16)))
            pf_capture_body<int, int, int>(s, A, G, kind, K, GA, chunk, x, R, P, q, z, y, c, partA, partB);
        else if (a16)
            pf_capture_body<short, int, int>(s, A, G, kind, K, GA, chunk, x, R, P, q, z, y, c, partA, partB);
        else
            pf_capture_body<int, short, short>(s, A, G, kind, K, GA, chunk, x, R, P, q, z, y, c, partA, partB);
    });
    if (!ge) { bkh_set_error("hipGraph capture of the fused PCG iteration failed"); return BKH_DEVICE_ERROR; }
    int rc = run_graph_until_done(t, ge, c, max_iter / K + 2);
    if (rc) return rc;
    Ctl hc = *t->ctl_host;
    HIP_OK(hipEventRecord(t->ev1, s));
    HIP_OK(hipEventSynchronize(t->ev1));
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, t->ev0, t->ev1));
    memset(inf, 0, sizeof *inf);
    inf->time_ms = ms;
    bkh_note_solve_ms(ms);
    inf->ret = ctl_result(&hc);
    inf->status = hc.status;
    inf->iters = hc.result_it;
    inf->rec_res = hc.err;
    double tr = 0.0;
    if (true_residual(t, A, 0.0, b, x, &tr) == 0) inf->true_res = tr;
    return inf->ret;
}

// PCG with Jacobi (kind 1) or no (kind 0) preconditioning on any SPD matrix.
extern "C" int bkh_pcg(bkh_matrix *A, int kind, const double *b, double *x, double tol, int max_iter,
                       const bkh_opts *opts, bkh_info *info) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (!A || !b || !x || (kind != PCK_NONE && kind != PCK_JACOBI)) return BKH_BAD_ARGUMENT;
    bkh_opts o = opts ? *opts : bkh_opts{};
    bkh_info inf;
    if (max_iter <= 0) { if (info) { memset(info, 0, sizeof *info); info->status = ST_CONVERGED; } return 0; }
    int ret = pcg_run(t, A, nullptr, kind, b, x, tol, max_iter, o, &inf);
    if (ret < -4) return ret;
    if (info) *info = inf;
    return ret;
}

// PCG on the field's matrix with the learned factor (call bkh_fsai_factor first).
extern "C" int bkh_fsai_pcg(bkh_fsai_field *F, const double *b, double *x, double tol, int max_iter,
                            const bkh_opts *opts, bkh_info *info) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (!F || !b || !x) return BKH_BAD_ARGUMENT;
    bkh_opts o = opts ? *opts : bkh_opts{};
    bkh_info inf;
    if (max_iter <= 0) { if (info) { memset(info, 0, sizeof *info); info->status = ST_CONVERGED; } return 0; }
    int ret = pcg_run(t, F->A, F->G, PCK_FSAI, b, x, tol, max_iter, o, &inf);
    if (ret < -4) return ret;
    if (info) *info = inf;
    return ret;
}

extern "C" bkh_matrix *bkh_fsai_field_matrix(bkh_fsai_field *F) { return F ? F->A : nullptr; }
