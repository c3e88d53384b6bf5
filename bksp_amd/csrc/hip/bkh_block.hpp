// bkh_block.hpp -- kernels on m x s blocks of vectors (s <= 16), stored ROW-major
// inside the device layer (the s values of one row are contiguous, so an SpMM
// gather of X[c, :] is one contiguous read). The bksp.h dmat boundary is
// column-major; k_cm2rm / k_rm2cm convert.
//
//   k_spmm        Y = alpha A X + beta Y: CSR-stream of A through LDS, one thread
//                 per row, columns summed in stored order (src/matvec.c:130-172,
//                 :219-246 order, bit for bit with alpha = 1).
//   k_gram        G = X^T Y on FP64 MFMA (v_mfma_f64_16x16x4f64), one 16x16
//                 accumulator per wave, per-workgroup partials.
//   panel tiles   D(16 rows x 16) = X_tile(16 x s) * M(s x s) on MFMA: the skinny
//                 panel-times-small-matrix products of the block methods, fused
//                 with their axpy and, where the next step needs it, the Gram
//                 matrix of the result (whose D layout is already the MFMA A/B
//                 operand layout, so no data moves between the two products).
//
// MFMA f64 16x16x4 lane layout (checked on the device): A(16x4): lane L holds
// A[L%16][L/16]; B(4x16): lane L holds B[L/16][L%16]; D(16x16): lane L register r
// holds D[L/16 + 4r][L%16].
#pragma once
#include "bkh_core.hpp"
#include "bkh_host.hpp"
#include "bkh_small.hpp"

typedef double mf4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ mf4 mfma16x16x4(double a, double b, mf4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// ------------------------------------------------------------ layout changes

// rm[i*s + j] = cm[i + j*m] and back (one thread per element).
template <typename TI, typename TO>
__global__ __launch_bounds__(256) void k_cm2rm(long long m, int s, const TI *cm, TO *rm) {
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < m * s; e += (long long)gridDim.x * 256) {
        const long long i = e / s, j = e % s;
        rm[e] = (TO)cm[i + j * m];
    }
}
template <typename TI, typename TO>
__global__ __launch_bounds__(256) void k_rm2cm(long long m, int s, const TI *rm, TO *cm) {
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < m * s; e += (long long)gridDim.x * 256) {
        const long long i = e % m, j = e / m;
        cm[e] = (TO)rm[i * s + j];
    }
}

// ------------------------------------------------------------------- SpMM

template <int S, typename VT, typename T>
__global__ __launch_bounds__(SPMV_BS) void k_spmm(const int *__restrict__ rp, const int *__restrict__ ci,
                                                  const VT *__restrict__ val,
                                                  const int *__restrict__ blk, const T *X, T *Y,
                                                  double alpha, double beta, const Ctl *ctl) {
    if (ctl && ctl_stopped(ctl)) return;
    __shared__ T sv[SPMV_CAP];
    __shared__ int sc[SPMV_CAP];
    const int tid = threadIdx.x;
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int r0 = blk[b], r1 = blk[b + 1];
    const int k0 = rp[r0], k1 = rp[r1];
    const bool a1 = (alpha == 1.0);
    if (k1 - k0 <= SPMV_CAP) {
        for (int k = k0 + tid; k < k1; k += SPMV_BS) {
            sv[k - k0] = a1 ? (T)val[k] : (T)(alpha * (double)val[k]);
            sc[k - k0] = ci[k];
        }
        __syncthreads();
        const int row = r0 + tid;
        if (row < r1) {
            T acc[S];
#pragma unroll
            for (int j = 0; j < S; j++) acc[j] = (beta != 0.0) ? mul_rn((T)beta, Y[(size_t)row * S + j]) : (T)0;
            const int s0 = rp[row] - k0, s1 = rp[row + 1] - k0;
            for (int k = s0; k < s1; k++) {
                const T a = sv[k];
                const T *xr = X + (size_t)sc[k] * S;
#pragma unroll
                for (int j = 0; j < S; j++) acc[j] = add_rn(acc[j], mul_rn(a, xr[j]));
            }
#pragma unroll
            for (int j = 0; j < S; j++) Y[(size_t)row * S + j] = acc[j];
        }
    } else {
        // a single long row: column j accumulated by the workgroup, tree-reduced
        __shared__ double red[SPMV_BS / 64];
        for (int j = 0; j < S; j++) {
            double acc[1] = {0.0};
            for (int k = k0 + tid; k < k1; k += SPMV_BS)
                acc[0] += (a1 ? (double)val[k] : alpha * (double)val[k]) * (double)X[(size_t)ci[k] * S + j];
            block_sum<1, SPMV_BS>(acc, red);
            if (tid == 0)
                Y[(size_t)r0 * S + j] = (T)(acc[0] + ((beta != 0.0) ? beta * (double)Y[(size_t)r0 * S + j] : 0.0));
        }
    }
}

// ------------------------------------------------------------ MFMA helpers

// One 16-row tile product D = X_tile * M: X row-major (ld S), rows row0..row0+15
// (rows >= m read as 0); M s x s column-major in `Ms` (shared or global), padded
// with zeros beyond s. Lane L's A operand for K-step q is X[row0 + L%16][4q + L/16].
template <int S, typename T>
__device__ __forceinline__ mf4 tile_mul(const T *X, long long m, long long row0, const double *Ms) {
    const int L = threadIdx.x & 63;
    const long long row = row0 + (L & 15);
    mf4 d = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < (S + 3) / 4; q++) {
        const int k = 4 * q + (L >> 4);
        const double a = (row < m && k < S) ? (double)X[row * S + k] : 0.0;
        const int j = L & 15;
        const double bm = (k < S && j < S) ? Ms[k + j * S] : 0.0;
        d = mfma16x16x4(a, bm, d);
    }
    return d;
}

// Loads/stores the tile elements held in the D layout (row L/16 + 4r, col L%16).
template <int S, typename T>
__device__ __forceinline__ mf4 tile_load(const T *X, long long m, long long row0) {
    const int L = threadIdx.x & 63, j = L & 15;
    mf4 v;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const long long row = row0 + (L >> 4) + 4 * r;
        v[r] = (row < m && j < S) ? (double)X[row * S + j] : 0.0;
    }
    return v;
}

template <int S, typename T>
__device__ __forceinline__ void tile_store(T *X, long long m, long long row0, mf4 v) {
    const int L = threadIdx.x & 63, j = L & 15;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const long long row = row0 + (L >> 4) + 4 * r;
        if (row < m && j < S) X[row * S + j] = (T)v[r];
    }
}

// Gram accumulation of a tile held in the D layout: register r of every lane is
// row L/16 + 4r, column L%16, which is exactly the A operand (A[i][kk] with
// kk <-> row kk + 4r) and the B operand of an MFMA with K slot kk.
__device__ __forceinline__ mf4 tile_gram(mf4 v, mf4 w, mf4 acc) {
#pragma unroll
    for (int r = 0; r < 4; r++) acc = mfma16x16x4(v[r], w[r], acc);
    return acc;
}

// Sums the 16x16 wave accumulators of a workgroup and writes its s x s partial
// (column-major) to part[blockIdx.x * s * s ...]. `sh` holds 4 * 256 doubles.
template <int S>
__device__ __forceinline__ void gram_block_partial(mf4 acc, double *sh, double *part) {
    const int L = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int r = 0; r < 4; r++) sh[w * 256 + ((L >> 4) + 4 * r) * 16 + (L & 15)] = acc[r];
    __syncthreads();
    const int t = threadIdx.x;
    if (t < S * S) {
        const int i = t % S, j = t / S;
        double v = 0.0;
        for (int q = 0; q < 4; q++) v += sh[q * 256 + i * 16 + j];
        part[(size_t)blockIdx.x * S * S + t] = v;
    }
}

// G = X^T Y over m rows (row-major, ld S); grid-stride over 64-row chunks (16-row
// tiles per wave).
template <int S, typename T>
__global__ __launch_bounds__(256) void k_gram(long long m, const T *X, const T *Y,
                                              const Ctl *ctl, double *part) {
    if (ctl && ctl_stopped(ctl)) return;
    __shared__ double sh[4 * 256];
    const int w = threadIdx.x >> 6;
    mf4 acc = {0.0, 0.0, 0.0, 0.0};
    for (long long row0 = ((long long)blockIdx.x * 4 + w) * 16; row0 < m; row0 += (long long)gridDim.x * 64) {
        mf4 a = tile_load<S>(X, m, row0);
        mf4 b = (X == Y) ? a : tile_load<S>(Y, m, row0);
        acc = tile_gram(a, b, acc);
    }
    gram_block_partial<S>(acc, sh, part);
}

static inline int block_grid(long long m) {
    long long g = (m + 63) / 64;
    if (g > 2048) g = 2048;
    if (g < 1) g = 1;
    return (int)g;
}

// ------------------------------------------------ LDS-staged 16-row tiles
//
// A 16-row tile of a row-major m x S block is one contiguous run of 16 S elements.
// The staged helpers move it between global memory and a per-wave LDS tile (16 rows,
// stride TS_LD doubles) with 16-byte vectors on every lane, and the MFMA operands and
// results between LDS and registers -- instead of the 4- or 8-byte per-lane accesses
// in MFMA layout (with S < 16 lanes idle) of tile_load / tile_mul / tile_store.
// Values are converted to FP64 on the way in and rounded to T on the way out.
#define TS_LD 17

// Orders this wave's LDS writes before its other lanes' reads (and reads before a
// later overwrite) -- no workgroup barrier needed for a wave-private tile.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int S, typename T>
__device__ __forceinline__ void stage_fetch(const T *X, long long m, long long row0, double *tb) {
    constexpr int V = 16 / sizeof(T);
    constexpr int NV = (16 * S + V - 1) / V;          // vectors in the tile
    typedef T tvec __attribute__((ext_vector_type(V)));
    const int L = threadIdx.x & 63;
    const long long e0 = row0 * S, emax = m * S;
#pragma unroll
    for (int pass = 0; pass * 64 < NV; pass++) {
        const int lv = L + pass * 64;
        if (lv < NV) {
            const long long e = e0 + (long long)lv * V;
            tvec v;
            if (e + V <= emax) {
                v = *(const tvec *)(X + e);
            } else {
#pragma unroll
                for (int u = 0; u < V; u++) v[u] = (e + u < emax) ? X[e + u] : (T)0;
            }
#pragma unroll
            for (int u = 0; u < V; u++) {
                const int le = lv * V + u;
                tb[(le / S) * TS_LD + le % S] = (double)v[u];
            }
        }
    }
}

// The same tile in two steps: stage_load into registers (one 16-byte vector per lane
// and pass) and stage_put into LDS, so the global load of the next tile can be issued
// before the current one is processed.
template <int S, typename T>
struct StageRegs {
    static constexpr int V = 16 / sizeof(T);
    static constexpr int NV = (16 * S + V - 1) / V;
    static constexpr int NP = (NV + 63) / 64;
    typedef T tvec __attribute__((ext_vector_type(V)));
    tvec v[NP];
};

template <int S, typename T>
__device__ __forceinline__ void stage_load(const T *X, long long m, long long row0, StageRegs<S, T> &r) {
    typedef StageRegs<S, T> R;
    const int L = threadIdx.x & 63;
    const long long e0 = row0 * S, emax = m * S;
#pragma unroll
    for (int pass = 0; pass < R::NP; pass++) {
        const int lv = L + pass * 64;
        const long long e = e0 + (long long)lv * R::V;
        if (lv < R::NV && e + R::V <= emax) {
            r.v[pass] = *(const typename R::tvec *)(X + e);
        } else {
#pragma unroll
            for (int u = 0; u < R::V; u++) r.v[pass][u] = (lv < R::NV && e + u < emax) ? X[e + u] : (T)0;
        }
    }
}

template <int S, typename T>
__device__ __forceinline__ void stage_put(const StageRegs<S, T> &r, double *tb) {
    typedef StageRegs<S, T> R;
    const int L = threadIdx.x & 63;
#pragma unroll
    for (int pass = 0; pass < R::NP; pass++) {
        const int lv = L + pass * 64;
        if (lv < R::NV) {
#pragma unroll
            for (int u = 0; u < R::V; u++) {
                const int le = lv * R::V + u;
                tb[(le / S) * TS_LD + le % S] = (double)r.v[pass][u];
            }
        }
    }
}

template <int S, typename T>
__device__ __forceinline__ void stage_store(T *X, long long m, long long row0, const double *tb) {
    constexpr int V = 16 / sizeof(T);
    constexpr int NV = (16 * S + V - 1) / V;
    typedef T tvec __attribute__((ext_vector_type(V)));
    const int L = threadIdx.x & 63;
    const long long e0 = row0 * S, emax = m * S;
#pragma unroll
    for (int pass = 0; pass * 64 < NV; pass++) {
        const int lv = L + pass * 64;
        if (lv < NV) {
            const long long e = e0 + (long long)lv * V;
            tvec v;
#pragma unroll
            for (int u = 0; u < V; u++) {
                const int le = lv * V + u;
                v[u] = (T)tb[(le / S) * TS_LD + le % S];
            }
            if (e + V <= emax) {
                *(tvec *)(X + e) = v;
            } else {
#pragma unroll
                for (int u = 0; u < V; u++)
                    if (e + u < emax) X[e + u] = v[u];
            }
        }
    }
}

// D-layout registers (lane L register r: row L/16 + 4r, column L%16) from / to the tile.
template <int S>
__device__ __forceinline__ mf4 ts_load_d(const double *tb) {
    const int L = threadIdx.x & 63, j = L & 15;
    mf4 v;
#pragma unroll
    for (int r = 0; r < 4; r++) v[r] = (j < S) ? tb[((L >> 4) + 4 * r) * TS_LD + j] : 0.0;
    return v;
}

template <int S>
__device__ __forceinline__ void ts_put_d(double *tb, mf4 v) {
    const int L = threadIdx.x & 63, j = L & 15;
    if (j < S) {
#pragma unroll
        for (int r = 0; r < 4; r++) tb[((L >> 4) + 4 * r) * TS_LD + j] = v[r];
    }
}

// D = tile * M (M s x s column-major in Ms, zero beyond s), A operands from the tile.
template <int S>
__device__ __forceinline__ mf4 ts_mul(const double *tb, const double *Ms) {
    const int L = threadIdx.x & 63;
    mf4 d = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < (S + 3) / 4; q++) {
        const int k = 4 * q + (L >> 4), j = L & 15;
        const double a = (k < S) ? tb[(L & 15) * TS_LD + k] : 0.0;
        const double bm = (k < S && j < S) ? Ms[k + j * S] : 0.0;
        d = mfma16x16x4(a, bm, d);
    }
    return d;
}

// Rounds D-layout values as a store to T would (the Gram of what is stored).
template <typename T>
__device__ __forceinline__ mf4 as_stored(mf4 v) {
    if constexpr (sizeof(T) != sizeof(double)) {
#pragma unroll
        for (int r = 0; r < 4; r++) v[r] = (double)(T)v[r];
    }
    return v;
}

// G = X^T Y with staged tiles (same partial layout as k_gram); the next tile's rows
// are loaded while the current one is multiplied.
template <int S, typename T>
__global__ __launch_bounds__(256) void k_gram_st(long long m, const T *X, const T *Y, const Ctl *ctl, double *part) {
    if (ctl && ctl_stopped(ctl)) return;
    __shared__ double sh[4 * 256];
    __shared__ double tbs[4][2][16 * TS_LD];
    const int w = threadIdx.x >> 6;
    double *t0 = tbs[w][0], *t1 = tbs[w][1];
    const bool same = (X == Y);
    const long long step = (long long)gridDim.x * 64;
    long long row0 = ((long long)blockIdx.x * 4 + w) * 16;
    mf4 acc = {0.0, 0.0, 0.0, 0.0};
    StageRegs<S, T> rx, ry;
    if (row0 < m) {
        stage_load<S>(X, m, row0, rx);
        if (!same) stage_load<S>(Y, m, row0, ry);
    }
    for (; row0 < m; row0 += step) {
        wave_lds_sync();
        stage_put<S>(rx, t0);
        if (!same) stage_put<S>(ry, t1);
        if (row0 + step < m) {
            stage_load<S>(X, m, row0 + step, rx);
            if (!same) stage_load<S>(Y, m, row0 + step, ry);
        }
        wave_lds_sync();
        const mf4 a = ts_load_d<S>(t0);
        const mf4 b = same ? a : ts_load_d<S>(t1);
        acc = tile_gram(a, b, acc);
    }
    gram_block_partial<S>(acc, sh, part);
}

// ------------------------------------------------------------ SELL SpMM

// Y = alpha A X + beta Y on the SELL-64 copy of A: a lane owns a row and sums its
// entries in stored order (the CSR-stream kernel's arithmetic), reading the S values
// of each gathered row of X as contiguous vectors.
template <int S, typename VT, typename T, typename IT>
__global__ __launch_bounds__(256) void k_spmm_sell(int n, const int *__restrict__ rp, const int *__restrict__ so,
                                                   const VT *__restrict__ val, const IT *__restrict__ col,
                                                   const T *X, T *Y, double alpha, double beta, const Ctl *ctl) {
    if (ctl && ctl_stopped(ctl)) return;
    const int row = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (row >= n) return;
    const int len = rp[row + 1] - rp[row];
    const int base = so[row >> 6] + (row & 63);
    const bool a1 = (alpha == 1.0);
    T acc[S];
#pragma unroll
    for (int j = 0; j < S; j++) acc[j] = (beta != 0.0) ? mul_rn((T)beta, Y[(size_t)row * S + j]) : (T)0;
    for (int kk = 0; kk < len; kk++) {
        const int e = base + 64 * kk;
        const int c = sizeof(IT) == 2 ? row + (int)col[e] : (int)col[e];
        const T a = a1 ? (T)val[e] : (T)(alpha * (double)val[e]);
        const T *xr = X + (size_t)c * S;
#pragma unroll
        for (int j = 0; j < S; j++) acc[j] = add_rn(acc[j], mul_rn(a, xr[j]));
    }
#pragma unroll
    for (int j = 0; j < S; j++) Y[(size_t)row * S + j] = acc[j];
}

#ifndef BKH_SPMM_KB
#define BKH_SPMM_KB 8      // entries per row loaded before use in k_spmm_sell_v
#endif

// Lanes per row of k_spmm_sell_v: each lane owns 16 bytes of a row of X and Y
// (S * sizeof(T) = 16 * LPR), 0 when S does not split that way.
template <int S, typename T>
constexpr int spmm_lpr() {
    constexpr int b = S * (int)sizeof(T);
    return (b == 16 || b == 32 || b == 64 || b == 128) ? b / 16 : 0;
}

// Lane k of the calling lane's group of G consecutive lanes (G a power of two, k < G,
// k a constant after unrolling): DPP quad permutations for G = 2 and 4, ds_bpermute
// above. Every lane of a group must be active.
template <int CTRL>
__device__ __forceinline__ int dpp_mov(int v) {
    return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
    return __int_as_float(dpp_mov<CTRL>(__float_as_int(v)));
}
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
    return __hiloint2double(dpp_mov<CTRL>(__double2hiint(v)), dpp_mov<CTRL>(__double2loint(v)));
}
template <int G, typename V>
__device__ __forceinline__ V grp_bcast(V v, int k) {
    if constexpr (G == 1) {
        return v;
    } else if constexpr (G == 2) {       // quad_perm (k, k, 2 + k, 2 + k)
        return (k & 1) ? dpp_mov<0xF5>(v) : dpp_mov<0xA0>(v);
    } else if constexpr (G == 4) {       // quad_perm (k, k, k, k)
        switch (k & 3) {
            case 0: return dpp_mov<0x00>(v);
            case 1: return dpp_mov<0x55>(v);
            case 2: return dpp_mov<0xAA>(v);
            default: return dpp_mov<0xFF>(v);
        }
    } else {
        return __shfl(v, (int)(threadIdx.x & 63 & ~(G - 1)) + k, 64);
    }
}

// KB entries k0 .. k0 + KB - 1 of one SELL / slice-diagonal row (entry kk at
// val[base + 64 kk]; last = the row's final entry), accumulated into the calling
// lane's CPL = S / LPR columns in entry order. Lane `sub` of the row's LPR lanes loads
// the values (and SELL columns) of entries k0 + sub + LPR i, and the row's lanes take
// them from it by lane broadcast, so a row costs ceil(KB / LPR) value loads per block
// instead of KB; then all KB gathers of X are issued before the first is used.
// Entries past `last` re-read the final one (in bounds) and are not accumulated.
template <int S, int LPR, int KB, typename VT, typename T, typename IT, bool DIA>
__device__ __forceinline__ void spmm_row_block(int n, int row, int sub, int k0, int last, int base,
                                               const int *dof, const VT *__restrict__ val,
                                               const IT *__restrict__ col, const T *X, bool a1,
                                               double alpha, T (&acc)[S / LPR]) {
    constexpr int CPL = S / LPR, NV = (KB + LPR - 1) / LPR;
    typedef T tvec __attribute__((ext_vector_type(CPL)));
    VT vl[NV];
    int cl[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) {
        const int e = base + 64 * min(k0 + sub + LPR * i, last);
        vl[i] = __builtin_nontemporal_load(val + e);          // the matrix streams once:
        if constexpr (!DIA) cl[i] = (int)__builtin_nontemporal_load(col + e);   // keep X in L2
    }
    VT av[KB];
    tvec xv[KB];
#pragma unroll
    for (int q = 0; q < KB; q++) {
        av[q] = grp_bcast<LPR>(vl[q / LPR], q % LPR);
        int c;
        if constexpr (DIA) {
            c = min(max(row + dof[min(k0 + q, last)], 0), n - 1);
        } else {
            const int ce = grp_bcast<LPR>(cl[q / LPR], q % LPR);
            c = sizeof(IT) == 2 ? row + ce : ce;
        }
        xv[q] = *(const tvec *)(X + (size_t)c * S + sub * CPL);
    }
#pragma unroll
    for (int q = 0; q < KB; q++)
        if (k0 + q <= last) {
            const T a = a1 ? (T)av[q] : (T)(alpha * (double)av[q]);
#pragma unroll
            for (int j = 0; j < CPL; j++) acc[j] = add_rn(acc[j], mul_rn(a, xv[q][j]));
        }
}

// Y = alpha A X + beta Y on the SELL-64 copy of A with LPR lanes per row: lane `sub`
// of a row owns its columns [sub*CPL, (sub+1)*CPL), so every gather of a neighbour's
// row of X is one 16-byte load per lane and a wave's loads cover 64/LPR whole rows
// (contiguous for neighbouring rows) instead of a quarter of each of 64 rows' cache
// lines. Per column the sum runs over the row's entries in stored order with the same
// roundings as k_spmm_sell, so both give identical results. X, Y 16-byte aligned.
// DIA: the slice-diagonal copy (so, val, col = its slice offsets, values and per-slice
// column offsets; a wave's rows lie in one slice, so the offsets are uniform loads);
// its value-0 positions add 0 products at clamped columns, which leaves every sum as
// on the SELL copy for finite X.
template <int S, typename VT, typename T, typename IT, bool DIA = false, int KB = BKH_SPMM_KB>
__global__ __launch_bounds__(256) void k_spmm_sell_v(int n, const int *__restrict__ rp, const int *__restrict__ so,
                                                     const VT *__restrict__ val, const IT *__restrict__ col,
                                                     const T *X, T *Y, double alpha, double beta, const Ctl *ctl) {
    constexpr int LPR = spmm_lpr<S, T>();
    static_assert(LPR > 0, "k_spmm_sell_v needs S * sizeof(T) in {16, 32, 64, 128}");
    constexpr int CPL = S / LPR, RPB = 256 / LPR;
    typedef T tvec __attribute__((ext_vector_type(CPL)));
    if (ctl && ctl_stopped(ctl)) return;
    const int row = xcd_block(blockIdx.x, gridDim.x) * RPB + (int)threadIdx.x / LPR;
    const int sub = (int)threadIdx.x % LPR;
    if (row >= n) return;
    int len, base;
    const int *dof = nullptr;
    if constexpr (DIA) {
        const int sl = __builtin_amdgcn_readfirstlane(row >> 6);
        const int s0 = so[sl];
        len = (so[sl + 1] - s0) >> 6;
        base = s0 + (row & 63);
        dof = (const int *)col + (s0 >> 6);
    } else {
        len = rp[row + 1] - rp[row];
        base = so[row >> 6] + (row & 63);
    }
    const bool a1 = (alpha == 1.0);
    T acc[CPL];
    if (beta != 0.0) {
        const tvec y = *(const tvec *)(Y + (size_t)row * S + sub * CPL);
#pragma unroll
        for (int j = 0; j < CPL; j++) acc[j] = mul_rn((T)beta, y[j]);
    } else {
#pragma unroll
        for (int j = 0; j < CPL; j++) acc[j] = (T)0;
    }
    for (int k0 = 0; k0 < len; k0 += KB)
        spmm_row_block<S, LPR, KB, VT, T, IT, DIA>(n, row, sub, k0, len - 1, base, dof, val, col, X, a1, alpha, acc);
    tvec out;
#pragma unroll
    for (int j = 0; j < CPL; j++) out[j] = acc[j];
    __builtin_nontemporal_store(out, (tvec *)(Y + (size_t)row * S + sub * CPL));
}

// AP = A P with the Gram matrix P^T AP in the same pass (block CG, src/blcg.c:74-79):
// the SpMM of k_spmm_sell_v (LPR lanes per row, alpha = 1, beta = 0, each row summed
// in stored order, so AP is bitwise the plain SpMM's) over a contiguous range of
// slices per workgroup, and while a wave's rows of P and AP are still in registers
// they go through a wave-private LDS tile into v_mfma_f64_16x16x4f64: A operand
// P^T (lane L: P[row 4q + L/16][L%16]), B operand AP (lane L: AP[row 4q + L/16][L%16]),
// columns >= S zero. The 16 x 16 accumulator holds G[a][b] = sum P_a AP_b in lane
// L register r at a = L/16 + 4r, b = L%16; one G (column-major S x S) per workgroup
// goes to part. This replaces the separate Gram pass, which re-read P and AP (16 m S
// bytes) right after the SpMM wrote AP. T = float: products in FP32 as the plain
// SpMM, the Gram of the stored FP32 values accumulated in FP64.
template <int S, typename VT, typename T, typename IT, bool DIA>
__device__ __forceinline__ void spmm_gram_row(int n, int sl, int row, int sub, bool live, const int *__restrict__ rp,
                                              const int *__restrict__ so, const VT *__restrict__ val,
                                              const IT *__restrict__ col, int &len, int &base, const int *&dof) {
    len = 0;
    base = 0;
    dof = nullptr;
    if (!live) return;
    if constexpr (DIA) {
        const int s0 = so[sl];
        len = (so[sl + 1] - s0) >> 6;
        base = s0 + (row & 63);
        dof = (const int *)col + (s0 >> 6);
    } else {
        len = rp[row + 1] - rp[row];
        base = so[row >> 6] + (row & 63);
    }
}

// UNR rows per thread (from slices SPB apart) are multiplied together; KB entries of
// each row are in flight at once (spmm_row_block). MINW: waves per SIMD the register
// budget is set for (8: two 1024-thread workgroups per CU, 4: one).
template <int S, typename VT, typename T, typename IT, bool DIA, int BS, int UNR, int KB, int MINW>
__global__ __launch_bounds__(BS, MINW) void k_spmm_gram(int n, int nsl, int chunk, const int *__restrict__ rp,
                                                  const int *__restrict__ so, const VT *__restrict__ val,
                                                  const IT *__restrict__ col, const T *X, T *Y, const Ctl *ctl,
                                                  double *__restrict__ part) {
    constexpr int LPR = spmm_lpr<S, T>();
    static_assert(LPR > 0 && S <= 16, "k_spmm_gram needs S * sizeof(T) in {16, 32, 64, 128} and S <= 16");
    constexpr int CPL = S / LPR, RW = 64 / LPR, NWV = BS / 64;   // RW: rows per wave and row slot
    constexpr int SPB = BS / (64 * LPR);                           // slices per row slot of the workgroup
    static_assert(SPB >= 1 && BS % (64 * LPR) == 0, "a pass must cover whole slices");
    typedef T tvec __attribute__((ext_vector_type(CPL)));
    if (ctl && ctl_stopped(ctl)) return;
    __shared__ T tp[NWV][UNR * RW * S], ta[NWV][UNR * RW * S];
    __shared__ double red[NWV][S * S];
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int sub = tid % LPR, rp_ = tid / LPR;                 // row within the slot
    const int rw = (tid % 64) / LPR;                             // row within the wave's tile
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int q0 = lb * chunk, q1 = min(nsl, q0 + chunk);
    mf4 g = {0.0, 0.0, 0.0, 0.0};
    for (int qb = q0; qb < q1; qb += UNR * SPB) {
        int row[UNR], len[UNR], base[UNR];
        const int *dof[UNR];
        bool live[UNR];
        T acc[UNR][CPL];
        int L = 0;
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            const int sl = __builtin_amdgcn_readfirstlane(qb + u * SPB + rp_ / 64);   // uniform per wave
            row[u] = sl * 64 + (rp_ & 63);
            live[u] = sl < q1 && row[u] < n;
            spmm_gram_row<S, VT, T, IT, DIA>(n, sl, row[u], sub, live[u], rp, so, val, col, len[u], base[u], dof[u]);
            L = max(L, len[u]);
#pragma unroll
            for (int j = 0; j < CPL; j++) acc[u][j] = (T)0;
        }
        for (int k0 = 0; k0 < L; k0 += KB) {
#pragma unroll
            for (int u = 0; u < UNR; u++)
                if (k0 < len[u])                                  // rows that are not live have len 0
                    spmm_row_block<S, LPR, KB, VT, T, IT, DIA>(n, row[u], sub, k0, len[u] - 1, base[u], dof[u], val,
                                                               col, X, true, 1.0, acc[u]);
        }
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            tvec pv;
#pragma unroll
            for (int j = 0; j < CPL; j++) pv[j] = (T)0;
            if (live[u]) {
                tvec out;
#pragma unroll
                for (int j = 0; j < CPL; j++) out[j] = acc[u][j];
                __builtin_nontemporal_store(out, (tvec *)(Y + (size_t)row[u] * S + sub * CPL));
                pv = *(const tvec *)(X + (size_t)row[u] * S + sub * CPL);   // P's own row (just gathered: L2)
            }
#pragma unroll
            for (int j = 0; j < CPL; j++) {
                tp[w][(u * RW + rw) * S + sub * CPL + j] = pv[j];
                ta[w][(u * RW + rw) * S + sub * CPL + j] = acc[u][j];
            }
        }
        wave_lds_sync();
        const int a = lane & 15;
#pragma unroll
        for (int q = 0; q < UNR * RW / 4; q++) {
            const int r = 4 * q + (lane >> 4);
            const double av = a < S ? (double)tp[w][r * S + a] : 0.0;
            const double bv = a < S ? (double)ta[w][r * S + a] : 0.0;
            g = mfma16x16x4(av, bv, g);
        }
        wave_lds_sync();
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int a = (lane >> 4) + 4 * r, b = lane & 15;
        if (a < S && b < S) red[w][a + b * S] = g[r];
    }
    __syncthreads();
    for (int t = tid; t < S * S; t += BS) {
        double v = red[0][t];
#pragma unroll
        for (int k = 1; k < NWV; k++) v += red[k][t];
        part[(size_t)lb * S * S + t] = v;
    }
}

// The same product as k_spmm_gram on the slice-diagonal copy, software-pipelined: a
// slice's record (DiaSlice) arrives with one scalar load, and the values and the KB
// gathers of a wave's next pass are issued before the current pass is summed, so a
// wave keeps two passes of loads in flight. Interior slices (every row + offset in
// [0, n)) address each gather as the uniform X + o S plus the lane's byte offset, with
// no clamp. P's own row for the Gram is the gathered row of the slice's offset-0
// position. Needs every slice with 1..KB positions and an offset 0 (A->dmeta). Each
// row sums its positions in order with the roundings of k_spmm_sell_v, so AP is
// bitwise the plain SpMM's, and the MFMA order per workgroup is k_spmm_gram's at the
// same chunking. 128^3 Poisson, s = 8: 117.8 -> 102.5 us FP64, 66.4 -> 53.3 us FP32.
template <int S, typename T, int KB, int LPR>
struct DiaStage {
    static constexpr int CPL = S / LPR, NV = (KB + LPR - 1) / LPR;
    typedef T tvec __attribute__((ext_vector_type(CPL)));
    T vl[NV];
    tvec xv[KB];
};

template <int S, typename T, int BS, int KB, int MINW>
__global__ __launch_bounds__(BS, MINW) void k_spmm_gram_dia(int n, int nsl, int chunk,
                                                            const DiaSlice *__restrict__ meta,
                                                            const T *__restrict__ val, const T *X, T *Y,
                                                            const Ctl *ctl, double *__restrict__ part) {
    constexpr int LPR = spmm_lpr<S, T>();
    static_assert(LPR > 0 && S <= 16 && KB <= DIA_META_POS, "k_spmm_gram_dia: 16-byte rows, S <= 16");
    constexpr int CPL = S / LPR, RW = 64 / LPR, NWV = BS / 64, NV = (KB + LPR - 1) / LPR;
    constexpr int SPB = NWV / LPR;                                  // slices per pass
    static_assert(SPB >= 1 && NWV % LPR == 0, "a pass must cover whole slices");
    typedef DiaStage<S, T, KB, LPR> St;
    typedef typename St::tvec tvec;
    if (ctl && ctl_stopped(ctl)) return;
    __shared__ T tp[NWV][RW * S], ta[NWV][RW * S];
    __shared__ double red[NWV][S * S];
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int sub = lane % LPR, rw = lane / LPR;
    const int row_in = (w % LPR) * RW + rw;                         // row within the slice
    const int wsl = __builtin_amdgcn_readfirstlane(w / LPR);         // slice within the pass
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int q0 = lb * chunk, q1 = min(nsl, q0 + chunk);
    mf4 g = {0.0, 0.0, 0.0, 0.0};
    // loads of one pass (slice index clamped: a pass past the range loads a real slice
    // and is never summed)
    auto load = [&](St &st, const DiaSlice &m, int sl) {
        const int row = sl * 64 + row_in;
        const int base = m.s0 + row_in;
#pragma unroll
        for (int i = 0; i < NV; i++)
            st.vl[i] = __builtin_nontemporal_load(val + base + 64 * min(sub + LPR * i, m.len - 1));
        if (m.inter) {
            const unsigned lo = ((unsigned)row * S + sub * CPL) * (unsigned)sizeof(T);
#pragma unroll
            for (int q = 0; q < KB; q++)
                st.xv[q] = *(const tvec *)((const char *)(X + (long long)m.o[q] * S) + lo);
        } else {
#pragma unroll
            for (int q = 0; q < KB; q++) {
                const int c = min(max(row + m.o[q], 0), n - 1);
                st.xv[q] = *(const tvec *)(X + (size_t)c * S + sub * CPL);
            }
        }
    };
    auto compute = [&](St &st, const DiaSlice &m, int qb, int sl) {
        const int row = sl * 64 + row_in;
        const bool live = qb + wsl < q1 && row < n;
        T acc[CPL];
#pragma unroll
        for (int j = 0; j < CPL; j++) acc[j] = (T)0;
#pragma unroll
        for (int q = 0; q < KB; q++) {
            const T a = grp_bcast<LPR>(st.vl[q / LPR], q % LPR);
            if (q < m.len) {
#pragma unroll
                for (int j = 0; j < CPL; j++) acc[j] = add_rn(acc[j], mul_rn(a, st.xv[q][j]));
            }
        }
        tvec out;
#pragma unroll
        for (int j = 0; j < CPL; j++) out[j] = acc[j];
        if (live) __builtin_nontemporal_store(out, (tvec *)(Y + (size_t)row * S + sub * CPL));
        tvec pv = st.xv[0];
#pragma unroll
        for (int q = 1; q < KB; q++)
            if (q == m.cq) pv = st.xv[q];
#pragma unroll
        for (int j = 0; j < CPL; j++) {
            tp[w][rw * S + sub * CPL + j] = live ? pv[j] : (T)0;
            ta[w][rw * S + sub * CPL + j] = live ? out[j] : (T)0;
        }
        wave_lds_sync();
        const int a = lane & 15;
#pragma unroll
        for (int q = 0; q < RW / 4; q++) {
            const int r = 4 * q + (lane >> 4);
            const double av = a < S ? (double)tp[w][r * S + a] : 0.0;
            const double bv = a < S ? (double)ta[w][r * S + a] : 0.0;
            g = mfma16x16x4(av, bv, g);
        }
        wave_lds_sync();
    };
    auto slice_of = [&](int idx) { return __builtin_amdgcn_readfirstlane(min(idx, nsl - 1)); };
    St A, B;
    int qb = q0;
    int slA = slice_of(qb + wsl);
    DiaSlice mA = meta[slA];
    if (qb < q1) load(A, mA, slA);
    for (; qb < q1; qb += 2 * SPB) {
        const int slB = slice_of(qb + SPB + wsl);
        const DiaSlice mB = meta[slB];
        load(B, mB, slB);
        compute(A, mA, qb, slA);
        if (qb + SPB >= q1) break;
        slA = slice_of(qb + 2 * SPB + wsl);
        mA = meta[slA];
        load(A, mA, slA);
        compute(B, mB, qb + SPB, slB);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int a = (lane >> 4) + 4 * r, b = lane & 15;
        if (a < S && b < S) red[w][a + b * S] = g[r];
    }
    __syncthreads();
    for (int t = tid; t < S * S; t += BS) {
        double v = red[0][t];
#pragma unroll
        for (int k = 1; k < NWV; k++) v += red[k][t];
        part[(size_t)lb * S * S + t] = v;
    }
}

// Y = alpha A X + beta Y for S row-major columns: on the slice-diagonal copy when asked
// (dia) and present, else the SELL-64 copy (16-byte row segments when X and Y allow),
// else CSR-stream. Every form sums a row's entries in stored order.
template <int S, typename VT, typename T>
static void launch_spmm(hipStream_t s, const bkh_matrix *A, const VT *val, const T *X, T *Y,
                        double alpha, double beta, const Ctl *c, bool dia = false) {
    constexpr int LPRD = spmm_lpr<S, T>();
    if constexpr (LPRD > 0) {
        const VT *dv = nullptr;
        if (dia && A->dia_entries > 0 && A->nsl > 0 && A->variant == SPMV_SELL) {
            if constexpr (sizeof(VT) == sizeof(float)) dv = (val == (const VT *)A->val32) ? (const VT *)A->dval32 : nullptr;
            else dv = (val == (const VT *)A->val) ? (const VT *)A->dval : nullptr;
        }
        if (dv && (((uintptr_t)X | (uintptr_t)Y) & 15) == 0) {
            const int G = (int)((A->n + 256 / LPRD - 1) / (256 / LPRD));
            hipLaunchKernelGGL((k_spmm_sell_v<S, VT, T, int, true>), dim3(G), dim3(256), 0, s, A->n, A->rp, A->dso,
                               dv, (const int *)A->doff, X, Y, alpha, beta, c);
            return;
        }
    }
    const VT *sv = nullptr;
    if (A->nsl > 0) {
        if constexpr (sizeof(VT) == sizeof(float)) sv = (val == (const VT *)A->val32) ? (const VT *)A->sval32 : nullptr;
        else sv = (val == (const VT *)A->val) ? (const VT *)A->sval : nullptr;
    }
    if (sv && (A->scd || A->scol)) {
        constexpr int LPR = spmm_lpr<S, T>();
        if constexpr (LPR > 0) {
            if ((((uintptr_t)X | (uintptr_t)Y) & 15) == 0) {
                const int G = (int)((A->n + 256 / LPR - 1) / (256 / LPR));
                if (A->scd)
                    hipLaunchKernelGGL((k_spmm_sell_v<S, VT, T, short>), dim3(G), dim3(256), 0, s, A->n, A->rp,
                                       A->sso, sv, (const short *)A->scd, X, Y, alpha, beta, c);
                else
                    hipLaunchKernelGGL((k_spmm_sell_v<S, VT, T, int>), dim3(G), dim3(256), 0, s, A->n, A->rp,
                                       A->sso, sv, (const int *)A->scol, X, Y, alpha, beta, c);
                return;
            }
        }
        const int G = (A->n + 255) / 256;
        if (A->scd)
            hipLaunchKernelGGL((k_spmm_sell<S, VT, T, short>), dim3(G), dim3(256), 0, s, A->n, A->rp, A->sso, sv,
                               (const short *)A->scd, X, Y, alpha, beta, c);
        else
            hipLaunchKernelGGL((k_spmm_sell<S, VT, T, int>), dim3(G), dim3(256), 0, s, A->n, A->rp, A->sso, sv,
                               (const int *)A->scol, X, Y, alpha, beta, c);
        return;
    }
    hipLaunchKernelGGL((k_spmm<S, VT, T>), dim3(A->nblk), dim3(SPMV_BS), 0, s, A->rp, A->ci, val, A->blk,
                       X, Y, alpha, beta, c);
}

